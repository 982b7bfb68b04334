// Harness that exposes the compiled reference (its own sources, built in
// place by oracle/Makefile into oracle/_ref/libmcrt_ref.so) through a small
// C API, so the tests can pin the oracle and the product against the real
// reference code.  TEST INFRASTRUCTURE ONLY.
//
// Everything here calls the reference's public classes; nothing restates
// its algorithms.  The SplitBVH root is built with the same steps as
// SplitBVH::SplitBVH (splitbvh.cpp:1410-1433) through the public static
// SplitBVH::build, because the root pointer itself is private.

#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

#include <glm/glm.hpp>
#include <glm/gtc/matrix_transform.hpp>

#include "aabb.h"
#include "boundedtriangle.h"
#include "camera.h"
#include "constants.h"
#include "debug_mesh.h"
#include "envmap.h"
#include "film.h"
#include "material.h"
#include "mesh.h"
#include "ray.h"
#include "sample.h"
#include "scenebvh.h"
#include "splitbvh.h"
#include "triangle.h"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace
{
struct RefBVH
{
    std::vector<std::shared_ptr<Triangle>> tris;
    std::shared_ptr<SplitBVHNode> root;
};

struct Fnv
{
    uint64_t h = 1469598103934665603ull;
    void bytes(const void* p, size_t n)
    {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    }
    void f(float x) { if (x == 0.0f) x = 0.0f; bytes(&x, 4); }
    void i(int32_t x) { bytes(&x, 4); }
    void box(const AABB& b)
    {
        f(b.min().x); f(b.min().y); f(b.min().z); f(b.max().x); f(b.max().y); f(b.max().z);
    }
};

struct Walk
{
    long long nodes = 0, leaves = 0, refs = 0, empty = 0;
    int max_depth = 0;
};

void digest(const std::shared_ptr<SplitBVHNode>& n, Fnv& h, Walk& w, int depth)
{
    w.nodes++;
    if (depth > w.max_depth) w.max_depth = depth;
    auto leaf = std::dynamic_pointer_cast<SplitBVHLeafNode>(n);
    if (leaf)
    {
        w.leaves++;
        int cnt = 0;
        for (int i = 0; i < k_sbvh_leaf_node_size; i++) if (leaf->primitive[i]) cnt++;
        if (cnt == 0) w.empty++;
        w.refs += cnt;
        h.i(1);
        h.box(*leaf->bounding_box);
        h.i(cnt);
        for (int i = 0; i < k_sbvh_leaf_node_size; i++)
        {
            if (!leaf->primitive[i]) continue;
            h.i(leaf->primitive[i]->originalTriangle()->primitiveID());
            h.box(leaf->primitive[i]->clippedBoundingBox());
        }
        return;
    }
    auto br = std::dynamic_pointer_cast<SplitBVHBranchNode>(n);
    h.i(0);
    h.box(*br->bounding_box);
    h.i(br->split_axis);
    digest(br->left_child, h, w, depth + 1);
    digest(br->right_child, h, w, depth + 1);
}

vec3 V(const float* p) { return vec3(p[0], p[1], p[2]); }
}  // namespace

REF_API void* ref_bvh_build(const float* tris9, int n)
{
    RefBVH* b = new RefBVH();
    std::vector<std::shared_ptr<BoundedTriangle>> refs;
    for (int i = 0; i < n; i++)
    {
        auto t = std::make_shared<Triangle>(V(tris9 + 9 * i), V(tris9 + 9 * i + 3), V(tris9 + 9 * i + 6));
        t->setPrimitiveID(i);
        t->setMaterialID(0);
        b->tris.push_back(t);
    }
    for (const auto& t : b->tris) refs.push_back(std::make_shared<BoundedTriangle>(t, t->boundingBox()));
    b->root = SplitBVH::build(refs, 0u, "root");
    return b;
}

REF_API void ref_bvh_free(void* h) { delete (RefBVH*)h; }

// stats: [nodes, leaves, refs, max_depth, empty_leaves]
REF_API uint64_t ref_bvh_digest(void* h, long long* stats)
{
    RefBVH* b = (RefBVH*)h;
    Fnv f;
    Walk w;
    digest(b->root, f, w, 0);
    if (stats)
    {
        stats[0] = w.nodes; stats[1] = w.leaves; stats[2] = w.refs; stats[3] = w.max_depth;
        stats[4] = w.empty;
    }
    return f.h;
}

// rays n x 8 (ox oy oz tmin dx dy dz tmax) -> SplitBVH::intersect semantics.
REF_API void ref_bvh_intersect(void* h, const float* rays, long long n, int* prim, float* tuv)
{
    RefBVH* b = (RefBVH*)h;
#pragma omp parallel for schedule(dynamic, 1024)
    for (long long i = 0; i < n; i++)
    {
        const float* r = rays + 8 * i;
        Ray ray(V(r), V(r + 4), r[3], r[7]);
        HitInfo hit;
        hit.reset();
        bool ok = b->root->intersect(ray, hit);
        prim[i] = ok ? hit.m_objID : -1;
        tuv[3 * i] = ok ? hit.m_t : INFINITY;
        tuv[3 * i + 1] = ok ? hit.m_u : 0.0f;
        tuv[3 * i + 2] = ok ? hit.m_v : 0.0f;
    }
}

// The plug-in boundary itself: SceneBVH over TriangleObjects (scenebvh.cpp:340-445).
// Triangles tris9 with vertex normals vn9 (n x 9) or, for NULL, the face
// normal (TriangleObject(a, b, c)); rays n x 6 (o, d); out: prim, t/u/v,
// shading normal n x 3; occluded with tfar.
REF_API void ref_scenebvh_query(const float* tris9, const float* vn9, int ntri, const float* rays6, long long n,
                                float tfar, int* prim, float* tuv, float* nrm, uint8_t* occ)
{
    std::vector<TriangleObject> tris;
    for (int i = 0; i < ntri; i++)
    {
        const float* p = tris9 + 9 * i;
        TriangleObject t = vn9 ? TriangleObject(V(p), V(p + 3), V(p + 6), V(vn9 + 9 * i), V(vn9 + 9 * i + 3),
                                                V(vn9 + 9 * i + 6))
                               : TriangleObject(V(p), V(p + 3), V(p + 6));
        t.setMaterialID(0);
        tris.push_back(t);
    }
    SceneBVH bvh(tris);
    for (long long i = 0; i < n; i++)
    {
        Ray r(V(rays6 + 6 * i), V(rays6 + 6 * i + 3));
        HitInfo hh = bvh.intersect(r);
        prim[i] = hh.primitiveID();
        tuv[3 * i] = hh.m_t; tuv[3 * i + 1] = hh.m_u; tuv[3 * i + 2] = hh.m_v;
        nrm[3 * i] = hh.m_nl.x; nrm[3 * i + 1] = hh.m_nl.y; nrm[3 * i + 2] = hh.m_nl.z;
        occ[i] = bvh.occluded(r, tfar) ? 1 : 0;
    }
}

REF_API int ref_debug_mesh(float* tris9)
{
    DebugMesh m;
    const auto& t = m.triangles();
    if (tris9)
        for (size_t i = 0; i < t.size(); i++)
        {
            const vec3* p[3] = {&t[i].a, &t[i].b, &t[i].c};
            for (int k = 0; k < 3; k++)
            {
                tris9[9 * i + 3 * k] = p[k]->x;
                tris9[9 * i + 3 * k + 1] = p[k]->y;
                tris9[9 * i + 3 * k + 2] = p[k]->z;
            }
        }
    return (int)t.size();
}

REF_API int ref_ray_box(const float* ray8, const float* box6)
{
    Ray r(V(ray8), V(ray8 + 4), ray8[3], ray8[7]);
    AABB b(V(box6), V(box6 + 3));
    return b.overlap(r) ? 1 : 0;
}

REF_API int ref_ray_tri(const float* ray8, const float* tri9, float* tuv)
{
    Ray r(V(ray8), V(ray8 + 4), ray8[3], ray8[7]);
    TriangleObject t(V(tri9), V(tri9 + 3), V(tri9 + 6));
    t.setMaterialID(0);
    t.setPrimitiveID(0);
    HitInfo h;
    bool ok = t.intersect(r, h);
    if (ok) { tuv[0] = h.m_t; tuv[1] = h.m_u; tuv[2] = h.m_v; }
    return ok ? 1 : 0;
}

REF_API void ref_clip_bound_tight(const float* tri9, const float* box6, float lo, float hi,
                                  int axis, float* out6)
{
    auto t = std::make_shared<Triangle>(V(tri9), V(tri9 + 3), V(tri9 + 6));
    BoundedTriangle bt(t, AABB(V(box6), V(box6 + 3)));
    AABB r = bt.clipBoundTight(lo, hi, axis);
    out6[0] = r.min().x; out6[1] = r.min().y; out6[2] = r.min().z;
    out6[3] = r.max().x; out6[4] = r.max().y; out6[5] = r.max().z;
}

// Mesh + genNormals (mesh.cpp:24-118) and the loadModel vertex transform
// (scene.cpp:374-381) with the given column-major matrix (NULL = identity).
// Faces with an index outside the vertex array (the phantom face Mesh
// appends at end of file) are not returned.
static bool face_ok(const Mesh& m, size_t i)
{
    const long long nv = (long long)m.m_verts.size();
    const int idx[3] = {m.m_faces[i].x, m.m_faces[i].y, m.m_faces[i].z};
    for (int k = 0; k < 3; k++)
        if (idx[k] < 0 || idx[k] >= nv) return false;
    return true;
}

REF_API long long ref_load_obj(const char* path, const float* m16, float* tris9, float* vn9)
{
    Mesh obj(path);
    long long nf = 0;
    for (size_t i = 0; i < obj.m_faces.size(); i++) nf += face_ok(obj, i) ? 1 : 0;
    if (!tris9) return nf;
    if (m16)
    {
        glm::mat4 m;
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 4; r++) m[c][r] = m16[4 * c + r];
        for (size_t k = 0; k < obj.m_verts.size(); k++)
        {
            glm::vec4 v = glm::vec4(obj.m_verts[k].x, obj.m_verts[k].y, obj.m_verts[k].z, 1.0f);
            v = m * v;
            obj.m_verts[k] = vec3(v.x, v.y, v.z);
        }
        obj.genNormals();
    }
    long long i = -1;
    for (size_t f = 0; f < obj.m_faces.size(); f++)
    {
        if (!face_ok(obj, f)) continue;
        i++;
        int idx[3] = {obj.m_faces[f].x, obj.m_faces[f].y, obj.m_faces[f].z};
        for (int k = 0; k < 3; k++)
        {
            tris9[9 * i + 3 * k] = obj.m_verts[idx[k]].x;
            tris9[9 * i + 3 * k + 1] = obj.m_verts[idx[k]].y;
            tris9[9 * i + 3 * k + 2] = obj.m_verts[idx[k]].z;
            vn9[9 * i + 3 * k] = obj.m_vertex_normals[idx[k]].x;
            vn9[9 * i + 3 * k + 1] = obj.m_vertex_normals[idx[k]].y;
            vn9[9 * i + 3 * k + 2] = obj.m_vertex_normals[idx[k]].z;
        }
    }
    return nf;
}

// The two loadModel transforms of the demo (demo_pt.cpp:42-61), made by
// GLM: out32[0..15] = translate(mat4(1), vec3(0.5, 0, 0)) and
// out32[16..31] = scale(mat4(1), vec3(0.015)) * rotate(mat4(1),
// radians(300), vec3(0, 1, 0)), each column-major (out[4c + r] = m[c][r]).
REF_API void ref_demo_transforms(float* out32)
{
    const glm::mat4 shift = glm::translate(glm::mat4(1.0f), glm::vec3(0.5f, 0.0f, 0.0f));
    const glm::mat4 rotation = glm::rotate(glm::mat4(1.0f), glm::radians(300.0f), glm::vec3(0.0f, 1.0f, 0.0f));
    const glm::mat4 scaling = glm::scale(glm::mat4(1.0f), glm::vec3(0.015f));
    const glm::mat4 sr = scaling * rotation;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
        {
            out32[4 * c + r] = shift[c][r];
            out32[16 + 4 * c + r] = sr[c][r];
        }
}

// TriangleObject(a, b, c, na, nb, nc).normal(u, v) (triangle.cpp:90-93) for
// n triangles (tris9, vn9) at barycentrics uv2: out n x 3.
REF_API void ref_shading_normal(const float* tris9, const float* vn9, const float* uv2, long long n, float* out3)
{
    for (long long i = 0; i < n; i++)
    {
        const float* t = tris9 + 9 * i;
        const float* q = vn9 + 9 * i;
        TriangleObject tri(V(t), V(t + 3), V(t + 6), V(q), V(q + 3), V(q + 6));
        const vec3 r = tri.normal(uv2[2 * i], uv2[2 * i + 1]);
        out3[3 * i] = r.x; out3[3 * i + 1] = r.y; out3[3 * i + 2] = r.z;
    }
}

// TriangleObject::setIntersectionEpsilon (process-global in the reference).
REF_API void ref_set_intersection_epsilon(float eps) { TriangleObject::setIntersectionEpsilon(eps); }

// Camera::makeRay at (mx, my) = pxy / (w, h); rays6 n x 6.
REF_API void ref_camera_rays(const float* o, const float* at, int w, int h, float fov,
                             const float* pxy, long long n, float* rays6)
{
    Camera cam(V(o), V(at), vec3(0, 1, 0), w, h, fov);
    for (long long i = 0; i < n; i++)
    {
        Ray r = cam.makeRay(pxy[2 * i] / w, pxy[2 * i + 1] / h);
        rays6[6 * i] = r.orig.x; rays6[6 * i + 1] = r.orig.y; rays6[6 * i + 2] = r.orig.z;
        rays6[6 * i + 3] = r.dir.x; rays6[6 * i + 4] = r.dir.y; rays6[6 * i + 5] = r.dir.z;
    }
}

// Material::sample / pdfW / evaluate for type (1 DIFF, 4 GLSY) with colour
// rgb: in n x (wi3, nl3, r1, r2); out n x (wo3, weight3, pdf_of_wo, eval3).
REF_API void ref_bsdf(int type, const float* rgb, const float* in8, long long n, float* out10)
{
    Material m(V(rgb), 0.0f, (MaterialType)type);
    for (long long i = 0; i < n; i++)
    {
        const float* q = in8 + 8 * i;
        vec3 wi = V(q), nl = V(q + 3), w;
        vec3 wo = m.sample(wi, nl, q[6], q[7], w);
        float pdf = m.pdfW(wi, nl, wo);
        vec3 ev = m.evaluate(wi, nl, wo);
        float* o = out10 + 10 * i;
        o[0] = wo.x; o[1] = wo.y; o[2] = wo.z; o[3] = w.x; o[4] = w.y; o[5] = w.z;
        o[6] = pdf; o[7] = ev.x; o[8] = ev.y; o[9] = ev.z;
    }
}

// DiscreteSampler over `data` (sample.cpp:103-118) evaluated at r.
REF_API void ref_discrete(const float* data, int n, const float* r, long long nr, int* idx, float* pdf)
{
    std::vector<float> d(data, data + n);
    DiscreteSampler s(d);
    for (int i = 0; i < n; i++) pdf[i] = s.pdf(i);
    for (long long k = 0; k < nr; k++) idx[k] = s.sample(r[k]);
}

// FrameBuffer holding a W x H x 3 film (row y = 0 at the bottom), filled
// through accumulatePixel.
static void fill_film(FrameBuffer& fb, const float* film, int w, int h)
{
    fb.resize(w, h);
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++)
        {
            const float* p = film + 3 * ((size_t)j * w + i);
            fb.accumulatePixel(i, j, vec3(p[0], p[1], p[2]));
        }
}

// FrameBuffer::dumpHDR / dumpPPM of `film` to `path`.
REF_API void ref_film_dump(const char* path, const float* film, int w, int h, int ppm)
{
    FrameBuffer fb;
    fill_film(fb, film, w, h);
    if (ppm)
        fb.dumpPPM(path);
    else
        fb.dumpHDR(path);
}

// FrameBuffer::tonemapGamma in place on `film`.
REF_API void ref_film_tonemap_gamma(float* film, int w, int h, float gamma)
{
    FrameBuffer fb;
    fill_film(fb, film, w, h);
    fb.tonemapGamma(gamma);
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++)
        {
            vec3 c = fb.pixel(i, j);
            float* p = film + 3 * ((size_t)j * w + i);
            p[0] = c.x; p[1] = c.y; p[2] = c.z;
        }
}

// EnvmapLoader(path) of a `{int w, h}` + float RGB file.
REF_API void* ref_envmap_create(const char* path) { return new EnvmapLoader(path); }
REF_API void ref_envmap_free(void* h) { delete (EnvmapLoader*)h; }

// getPdf and sampleRadiance of n directions: out n x 4 = {pdf, r, g, b}.
REF_API void ref_envmap_query(void* h, const float* dirs, long long n, float* out4)
{
    const EnvmapLoader* e = (const EnvmapLoader*)h;
    for (long long i = 0; i < n; i++)
    {
        float d[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        out4[4 * i] = e->getPdf(d);
        e->sampleRadiance(d, out4 + 4 * i + 1);
    }
}

// sample(r1, r2) for n pairs (the sub-pixel jitter comes from the
// reference's own randf()): out n x 7 = {dir, rad, pdf}.
REF_API void ref_envmap_sample(void* h, const float* r2, long long n, float* out7)
{
    const EnvmapLoader* e = (const EnvmapLoader*)h;
    for (long long i = 0; i < n; i++)
    {
        DirLight dl = e->sample(r2[2 * i], r2[2 * i + 1]);
        float* o = out7 + 7 * i;
        o[0] = dl.dir[0]; o[1] = dl.dir[1]; o[2] = dl.dir[2];
        o[3] = dl.rad[0]; o[4] = dl.rad[1]; o[5] = dl.rad[2];
        o[6] = dl.pdf;
    }
}
