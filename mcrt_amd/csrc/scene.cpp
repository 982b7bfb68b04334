// Scene ingestion (OBJ), finalize (SBVH build + collapse + upload), light
// tables, camera set-up and the query dispatcher.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "capi_internal.h"
#include "sbvh.h"
#include "wide.h"

namespace mcrt
{
// Mesh::Mesh (mesh.cpp:24-60): whitespace tokens; `v x y z`, `f a b c`
// (1-based), everything else skipped token by token.  A record is kept only
// when all of its numbers were extracted: at end of file the reference's
// failed `>> token` repeats the previous token and, after a face line, adds
// a phantom face with indices -1 that it then reads and writes out of
// bounds; every real face is loaded identically.  The optional
// column-major transform is applied per vertex in GLM's order
// (m0*x + m1*y) + (m2*z + m3*w) (type_mat4x4.inl:532-543), then
// Mesh::genNormals (mesh.cpp:66-92): vertex normal = normalize of the sum
// of the unnormalised face normals (vb-va) x (vc-va) in face order.
// A `v` or `f` token followed by something that is not a number before the
// end of the file (`f 1/1/1 ...`, a `v` word inside a comment) leaves the
// reference's stream failed but not at its end, so its loop never ends; here
// it is an MCRT_ERR_IO naming the line.  A record cut off by the end of the
// file is dropped, as the reference's loop ends there too.
int load_obj(const char* path, const float* m16, ObjMesh& m)
{
    std::ifstream ifs(path, std::ios::binary);
    if (!ifs.good()) return fail(MCRT_ERR_IO, "cannot open %s", path);
    std::stringstream reader;
    reader << ifs.rdbuf();
    std::string token;
    auto bad_record = [&](const char* what) {
        reader.clear();
        const std::string& text = reader.str();
        const std::streamoff pos = reader.tellg();
        const size_t end = pos < 0 ? text.size() : std::min(text.size(), (size_t)pos);
        const long line = 1 + (long)std::count(text.begin(), text.begin() + (std::ptrdiff_t)end, '\n');
        return fail(MCRT_ERR_IO, "%s:%ld: unsupported %s record (only 'v x y z' and 'f a b c' are read)", path,
                    line, what);
    };
    while (reader >> token)
    {
        if (token == "v")
        {
            float a, b, c;
            if (reader >> a >> b >> c) m.v.push_back(f3(a, b, c));
            else if (!reader.eof()) return bad_record("vertex");
        }
        else if (token == "f")
        {
            unsigned int a, b, c;
            if (reader >> a >> b >> c)
            {
                m.f.push_back((int)(a - 1));
                m.f.push_back((int)(b - 1));
                m.f.push_back((int)(c - 1));
            }
            else if (!reader.eof()) return bad_record("face");
        }
    }
    for (int idx : m.f)
        if (idx < 0 || (size_t)idx >= m.v.size())
            return fail(MCRT_ERR_IO, "%s: face index %d out of range (%zu vertices)", path, idx + 1, m.v.size());
    if (m16)
        for (F3& p : m.v)
        {
            const float x = p[0], y = p[1], z = p[2], w = 1.0f;
            F3 q;
            for (int r = 0; r < 3; r++)
                q[r] = (m16[0 * 4 + r] * x + m16[1 * 4 + r] * y) + (m16[2 * 4 + r] * z + m16[3 * 4 + r] * w);
            p = q;
        }
    m.vn.assign(m.v.size(), f3(0, 0, 0));
    for (size_t i = 0; i < m.f.size() / 3; i++)
    {
        int a = m.f[3 * i], b = m.f[3 * i + 1], c = m.f[3 * i + 2];
        F3 fn = cross3(sub(m.v[b], m.v[a]), sub(m.v[c], m.v[a]));
        m.vn[a] = add(m.vn[a], fn);
        m.vn[b] = add(m.vn[b], fn);
        m.vn[c] = add(m.vn[c], fn);
    }
    for (F3& n : m.vn) n = normalize3(n);
    return MCRT_OK;
}

// Scene::loadModel's per-face TriangleObject(a, b, c, na, nb, nc)
// (scene.cpp:384-406): corners and vertex normals by the face's indices.
void obj_soup(const ObjMesh& m, float* v9, float* vn9)
{
    const size_t nf = m.f.size() / 3;
    for (size_t i = 0; i < nf; i++)
        for (int c = 0; c < 3; c++)
        {
            const int vi = m.f[3 * i + c];
            for (int k = 0; k < 3; k++)
            {
                v9[9 * i + 3 * c + k] = m.v[vi][k];
                vn9[9 * i + 3 * c + k] = m.vn[vi][k];
            }
        }
}

// Camera::Camera (camera.cpp:9-35), up = (0,1,0) (scene.cpp:231).
DevCamera make_camera(const mcrt_scene* s)
{
    DevCamera c;
    F3 o = f3(s->cam_o[0], s->cam_o[1], s->cam_o[2]);
    F3 at = f3(s->cam_t[0], s->cam_t[1], s->cam_t[2]);
    c.ww = float(s->cam_w) / float(s->cam_h);
    float rad = (s->cam_fov * 0.5f) / 180.0f * 3.14159265358979323846f;  // toRadians (numeric.cpp:47-50)
    c.dist = 1.0f / (2.0f * std::tan(rad));
    F3 fwd = normalize3(sub(at, o));
    F3 fx = normalize3(cross3(fwd, f3(0, 1, 0)));
    F3 fy = cross3(fx, fwd);
    for (int k = 0; k < 3; k++) { c.o[k] = o[k]; c.fwd[k] = fwd[k]; c.fx[k] = fx[k]; c.fy[k] = fy[k]; }
    c.w = s->cam_w;
    c.h = s->cam_h;
    return c;
}

int upload_materials(mcrt_scene* s)
{
    std::vector<DevMaterial> dm(s->mats.size());
    for (size_t i = 0; i < s->mats.size(); i++)
    {
        for (int k = 0; k < 3; k++) dm[i].rgb[k] = s->mats[i].rgb[k];
        dm[i].type = s->mats[i].type;
        dm[i].gloss = s->mats[i].gloss;
        dm[i].pad[0] = dm[i].pad[1] = dm[i].pad[2] = 0.0f;
    }
    int rc = s->d_mats.upload(dm.data(), dm.size() * sizeof(DevMaterial));
    if (rc) return rc;
    s->dev.mats = s->d_mats.as<DevMaterial>();
    return MCRT_OK;
}

// ------------------------------------------------------------ BVH cache --
// File: CacheHeader, then the QNode array, then the WRef array.  A file is
// used only when magic, version, record sizes, triangle count, the hash of
// the triangles' bits, the file size and the hash of the payload all match.
namespace
{
struct CacheHeader
{
    char magic[8];
    uint32_t version, qnode_size, wref_size, pad;
    uint64_t ntri, tri_hash, nq, nr;
    int32_t stack_bound, depth;
    int64_t binary_nodes, leaves, refs, max_depth;
    uint64_t tree_digest, payload_hash;
};
const char kCacheMagic[8] = {'M', 'C', 'R', 'T', 'B', 'V', 'H', '\0'};
const uint32_t kCacheVersion = 2;  // 2: rank tables, single-reference children

// FNV-1a over 64-bit words (bytes beyond a multiple of 8 folded in last).
uint64_t hash_bytes(const void* p, size_t n, uint64_t h = 0xcbf29ce484222325ull)
{
    const unsigned char* b = static_cast<const unsigned char*>(p);
    size_t i = 0;
    for (; i + 8 <= n; i += 8)
    {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        h = (h ^ w) * 0x100000001b3ull;
    }
    for (; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

uint64_t triangles_hash(const std::vector<Tri>& tris)
{
    return hash_bytes(tris.data(), tris.size() * sizeof(Tri), 0xcbf29ce484222325ull ^ (uint64_t)tris.size());
}

// Structure of a loaded payload: every primitive id below ntri, every leaf
// link inside the reference array, every interior link after its parent
// (the order collapse_wide writes) and inside the node array, and the
// recorded stack bound at least the one the links imply (collapse_wide's
// recursion: a node pushes its passing children and visits one).  The
// hashes catch accidental damage; this keeps a file that passes them but
// was written by something else from sending the traversal outside its
// arrays or beyond its stack.
bool cache_structure_valid(const WideBVH& wb, uint64_t ntri, int32_t stack_bound)
{
    const size_t nq = wb.qnodes.size(), nr = wb.refs.size();
    for (const WRef& r : wb.refs)
        if (r.prim < 0 || (uint64_t)r.prim >= ntri) return false;
    std::vector<int32_t> need(nq, 0);
    for (size_t i = nq; i-- > 0;)
    {
        const QNode& q = wb.qnodes[i];
        int32_t nchild = 0, child_need = 0;
        for (int c = 0; c < 4; c++)
        {
            const uint32_t l = q.child[c];
            if (l == MCRT_CHILD_EMPTY) continue;
            nchild++;
            if (l & MCRT_CHILD_LEAF)
            {
                if ((uint64_t)(l & MCRT_LEAF_REF_MASK) + ((l >> 29) & 3u) + 1u > nr) return false;
            }
            else
            {
                if (l <= i || l >= nq) return false;
                child_need = std::max(child_need, need[l]);
            }
        }
        need[i] = std::max(nchild, nchild - 1 + child_need);
    }
    return nq == 0 || need[0] <= stack_bound;
}

bool load_bvh_cache(const char* path, uint64_t ntri, uint64_t tri_hash, WideBVH& wb, mcrt_build_stats& st)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    CacheHeader h;
    bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kCacheMagic, 8) == 0 &&
              h.version == kCacheVersion && h.qnode_size == sizeof(QNode) && h.wref_size == sizeof(WRef) &&
              h.ntri == ntri && h.tri_hash == tri_hash;
    if (ok)
    {
        std::fseek(f, 0, SEEK_END);
        const long size = std::ftell(f);
        ok = size == (long)(sizeof(h) + h.nq * sizeof(QNode) + h.nr * sizeof(WRef));
        std::fseek(f, (long)sizeof(h), SEEK_SET);
    }
    if (ok)
    {
        wb.qnodes.resize(h.nq);
        wb.refs.resize(h.nr);
        ok = std::fread(wb.qnodes.data(), sizeof(QNode), h.nq, f) == h.nq &&
             std::fread(wb.refs.data(), sizeof(WRef), h.nr, f) == h.nr;
        ok = ok && hash_bytes(wb.refs.data(), h.nr * sizeof(WRef),
                              hash_bytes(wb.qnodes.data(), h.nq * sizeof(QNode))) == h.payload_hash;
        ok = ok && cache_structure_valid(wb, ntri, h.stack_bound);
    }
    std::fclose(f);
    if (!ok)
    {
        wb.qnodes.clear();
        wb.refs.clear();
        return false;
    }
    wb.stack_bound = h.stack_bound;
    wb.depth = h.depth;
    st.binary_nodes = h.binary_nodes;
    st.leaves = h.leaves;
    st.refs = h.refs;
    st.max_depth = h.max_depth;
    st.tree_digest = h.tree_digest;
    return true;
}

// Writes to path + ".tmp" and renames, so readers never see a partial file.
int store_bvh_cache(const char* path, uint64_t ntri, uint64_t tri_hash, const WideBVH& wb,
                    const mcrt_build_stats& st)
{
    CacheHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kCacheMagic, 8);
    h.version = kCacheVersion;
    h.qnode_size = sizeof(QNode);
    h.wref_size = sizeof(WRef);
    h.ntri = ntri;
    h.tri_hash = tri_hash;
    h.nq = wb.qnodes.size();
    h.nr = wb.refs.size();
    h.stack_bound = wb.stack_bound;
    h.depth = wb.depth;
    h.binary_nodes = st.binary_nodes;
    h.leaves = st.leaves;
    h.refs = st.refs;
    h.max_depth = st.max_depth;
    h.tree_digest = st.tree_digest;
    h.payload_hash = hash_bytes(wb.refs.data(), h.nr * sizeof(WRef), hash_bytes(wb.qnodes.data(), h.nq * sizeof(QNode)));
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return fail(MCRT_ERR_IO, "cannot write BVH cache %s", tmp.c_str());
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
              std::fwrite(wb.qnodes.data(), sizeof(QNode), h.nq, f) == h.nq &&
              std::fwrite(wb.refs.data(), sizeof(WRef), h.nr, f) == h.nr;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0)
    {
        std::remove(tmp.c_str());
        return fail(MCRT_ERR_IO, "cannot write BVH cache %s", path);
    }
    return MCRT_OK;
}
}  // namespace

int finalize_scene(mcrt_scene* s, int threads, const char* cache)
{
    using clk = std::chrono::steady_clock;
    const size_t n = s->tris.size();
    mcrt_build_stats st{};
    st.triangles = (int64_t)n;
    auto t0 = clk::now();
    WideBVH wb;
    const uint64_t th = cache ? triangles_hash(s->tris) : 0;
    if (cache && load_bvh_cache(cache, n, th, wb, st))
        st.from_cache = 1;
    else
    {
        BTree bt;
        build_sbvh(s->tris, threads, bt);
        st.binary_nodes = (int64_t)bt.nodes.size();
        for (const BNode& b : bt.nodes) st.leaves += b.leaf ? 1 : 0;
        st.refs = (int64_t)bt.refs.size();
        st.max_depth = bt.max_depth;
        st.tree_digest = tree_digest(bt);
        collapse_wide(bt, s->tris, wb);
        { BTree().nodes.swap(bt.nodes); std::vector<SRef>().swap(bt.refs); }
        std::vector<WNode>().swap(wb.nodes);
        if (cache)
        {
            const int rc = store_bvh_cache(cache, n, th, wb, st);
            if (rc) return rc;
        }
    }
    auto t1 = clk::now();
    st.wide_nodes = (int64_t)wb.qnodes.size();
    st.stack_bound = wb.stack_bound;

    // Per-triangle shading records and the light list (Scene::finalize,
    // scene.cpp:151-178): lights are LGHT triangles in index order, picked
    // proportionally to area with DiscreteSampler's fp32 normalisation.
    std::vector<TriShade> shade(n);
    std::vector<float> verts(9 * n);
    std::vector<DevLight> lights;
    for (size_t i = 0; i < n; i++)
    {
        TriShade& t = shade[i];
        for (int k = 0; k < 3; k++)
        {
            t.na[k] = s->vn[9 * i + k];
            t.nb[k] = s->vn[9 * i + 3 + k];
            t.nc[k] = s->vn[9 * i + 6 + k];
            verts[9 * i + k] = s->tris[i].a[k];
            verts[9 * i + 3 + k] = s->tris[i].b[k];
            verts[9 * i + 6 + k] = s->tris[i].c[k];
        }
        t.mat = s->tri_mat[i];
        t.light = -1;
        t.hot = -1;
        if (s->mats[t.mat].type == MCRT_LGHT)
        {
            t.light = (int32_t)lights.size();
            DevLight L;
            L.tri = (int32_t)i;
            const Tri& tr = s->tris[i];
            F3 cr = cross3(sub(tr.b, tr.a), sub(tr.c, tr.a));  // TriangleObject::area (triangle.cpp:129-132)
            L.area = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) * 0.5f;
            lights.push_back(L);
        }
    }
    // Gradient hot slots: every light, then the largest triangles by area.
    std::vector<int32_t> hot_tri;
    {
        int slot = 0;
        for (const DevLight& L : lights)
            if (slot < MCRT_HOT_SLOTS) shade[L.tri].hot = slot++;
        std::vector<std::pair<float, int32_t>> area;
        area.reserve(n);
        for (size_t i = 0; i < n; i++)
            if (shade[i].light < 0)
            {
                const Tri& tr = s->tris[i];
                F3 cr = cross3(sub(tr.b, tr.a), sub(tr.c, tr.a));
                area.emplace_back(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2], (int32_t)i);
            }
        const size_t k = std::min(area.size(), (size_t)std::max(0, MCRT_HOT_SLOTS - slot));
        std::partial_sort(area.begin(), area.begin() + k, area.end(),
                          [](const std::pair<float, int32_t>& a, const std::pair<float, int32_t>& b) {
                              return a.first > b.first || (a.first == b.first && a.second < b.second);
                          });
        for (size_t j = 0; j < k; j++) shade[area[j].second].hot = slot++;
        hot_tri.assign(slot, -1);
        for (size_t i = 0; i < n; i++)
            if (shade[i].hot >= 0) hot_tri[shade[i].hot] = (int32_t)i;
    }
    if (!lights.empty())  // DiscreteSampler::update (sample.cpp:103-118)
    {
        float norm = 0.0f;
        for (const DevLight& L : lights) norm = norm + L.area;
        const float k = 1.0f / norm;
        float sum = 0.0f;
        for (DevLight& L : lights)
        {
            L.pdf = L.area * k;
            sum += L.pdf;
            L.cdf = sum;
        }
    }

    // Box around every emitter reference (the clipped boxes the traversal
    // tests before Moller-Trumbore), for the path tracer's next-event cull.
    bool emit_any = false;
    float emit_lo[3] = {0.0f, 0.0f, 0.0f}, emit_hi[3] = {0.0f, 0.0f, 0.0f};
    for (const WRef& r : wb.refs)
    {
        if (s->mats[s->tri_mat[r.prim]].type != MCRT_LGHT) continue;
        for (int k = 0; k < 3; k++)
        {
            emit_lo[k] = emit_any ? std::min(emit_lo[k], r.lo[k]) : r.lo[k];
            emit_hi[k] = emit_any ? std::max(emit_hi[k], r.hi[k]) : r.hi[k];
        }
        emit_any = true;
    }

    int rc;
    {
        // one array of 64 B records, nodes then references (layout.h)
        std::vector<QNode> dn;
        if (!device_records(wb, dn))
            return fail(MCRT_ERR_OUT_OF_MEMORY, "BVH of %zu nodes and %zu references exceeds 2^29 records",
                        wb.qnodes.size(), wb.refs.size());
        const size_t nb = dn.size() * sizeof(QNode), rb = wb.refs.size() * sizeof(WRef);
        if ((rc = s->d_bvh.ensure(nb + rb))) return rc;
        hipError_t e = hipSuccess;
        if (nb) e = hipMemcpy(s->d_bvh.p, dn.data(), nb, hipMemcpyHostToDevice);
        if (e == hipSuccess && rb)
            e = hipMemcpy(static_cast<char*>(s->d_bvh.p) + nb, wb.refs.data(), rb, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(MCRT_ERR_DEVICE, "hipMemcpy H2D: %s", hipGetErrorString(e));
    }
    if ((rc = s->d_shade.upload(shade.data(), shade.size() * sizeof(TriShade)))) return rc;
    if ((rc = s->d_verts.upload(verts.data(), verts.size() * sizeof(float)))) return rc;
    if ((rc = s->d_lights.upload(lights.data(), lights.size() * sizeof(DevLight)))) return rc;
    if ((rc = s->d_hot.upload(hot_tri.data(), hot_tri.size() * sizeof(int32_t)))) return rc;
    if ((rc = upload_materials(s))) return rc;

    DevScene& d = s->dev;
    d.nodes = wb.qnodes.empty() ? nullptr : s->d_bvh.as<QNode>();
    d.shade = s->d_shade.as<TriShade>();
    d.verts = s->d_verts.as<float>();
    d.lights = s->d_lights.as<DevLight>();
    d.nlights = (int32_t)lights.size();
    d.hot_tri = s->d_hot.as<int32_t>();
    d.nhot = (int32_t)hot_tri.size();
    d.stack_bound = wb.stack_bound;
    d.node_span = qnode_span(wb.qnodes);
    d.ntri = (int64_t)n;
    d.emit_any = emit_any ? 1 : 0;
    for (int k = 0; k < 3; k++)
    {
        d.emit_lo[k] = emit_lo[k];
        d.emit_hi[k] = emit_hi[k];
    }
    // Origin cells of the query reorder: the triangles' box in kCellsPerAxis
    // cells per axis.
    {
        float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++)
            {
                const float a = s->tris[i].a[k], b = s->tris[i].b[k], c = s->tris[i].c[k];
                const float mn = std::min(a, std::min(b, c)), mx = std::max(a, std::max(b, c));
                lo[k] = i ? std::min(lo[k], mn) : mn;
                hi[k] = i ? std::max(hi[k], mx) : mx;
            }
        for (int k = 0; k < 3; k++)
        {
            const float ext = hi[k] - lo[k];
            d.cell_lo[k] = lo[k];
            d.cell_sc[k] = ext > 0.0f && std::isfinite(ext) ? (float)kCellsPerAxis / ext : 0.0f;
        }
    }
    // Unless set: reorder large queries over a BVH larger than the last-level
    // cache (256 MB on MI355X), where the caller's ray order sends every
    // wavefront across the whole tree (the 558 MB soup's 16M incoherent rays:
    // 8.5 instead of 10.7 ms, sort included).  Over a cache-resident BVH the
    // sort costs more than it saves (the 101 MB one of the 1.3M-triangle
    // mesh: 2.67 instead of 2.08 ms).
    // Unless set: 32 traversal waves per CU (8 per SIMD, all resident) over
    // a cache-resident BVH, 28 (7 per SIMD) over a larger one, where more
    // rays in flight evict more of each other's nodes (the 1.3M-triangle
    // mesh's queries 1.5% and its frame 2% faster at 32; the soup's
    // incoherent closest-hit queries 3% slower).
    const bool big_bvh = s->d_bvh.bytes > (size_t(192) << 20);
    if (s->trace_reorder_rays == 0) s->trace_reorder_rays = big_bvh ? int64_t(1) << 20 : -1;
    if (s->trace_waves_per_cu == 0) s->trace_waves_per_cu = big_bvh ? 28 : 4 * kTraceWavesPerSimd;
    if (s->pt_trace_waves_per_cu == 0) s->pt_trace_waves_per_cu = big_bvh ? 28 : 4 * kTraceWavesPerSimd;
    apply_trace_tuning(s);
    auto t2 = clk::now();
    st.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    st.flatten_seconds = std::chrono::duration<double>(t2 - t1).count();
    st.device_bytes = (int64_t)(s->d_bvh.bytes + s->d_shade.bytes + s->d_verts.bytes +
                                s->d_lights.bytes + s->d_mats.bytes + s->d_hot.bytes);
    s->stats = st;
    return MCRT_OK;
}

// EnvmapLoader::buildCdf (envmap.cpp:208-331) for a latitude-longitude
// map of w x h RGB pixels, rows top-down: luminance 0.299 r + 0.587 g +
// 0.114 b (numeric.cpp:62-66) per pixel, formed in double; a copy pre-warped
// by sin(phi) of the row centre, phi = pi (m + 0.5) / h, multiplied in
// double (the reference's unqualified sin of a float resolves to the double
// function); norm = 1 / the in-order float sum of the warped values; the
// DiscreteSampler CDFs (sample.cpp:103-118) of every column and of the
// column sums.  Device layout: {r, g, b, luminance} per pixel (one 16 B
// fetch gives radiance and pdf), the column CDFs column-major (each binary
// search reads one contiguous run).
int set_envmap(mcrt_scene* s, const float* rgb, int w, int h)
{
    DevScene& d = s->dev;
    HIPCHK(hipSetDevice(s->ctx->device));
    // renders in flight may read the current tables
    HIPCHK(hipDeviceSynchronize());
    if (!rgb)
    {
        s->env_gen++;
        d.env_px = d.env_cdf_u = d.env_cdf_v = nullptr;
        d.env_w = d.env_h = 0;
        d.env_norm = 0.0f;
        s->d_env_px.release();
        s->d_env_cdf_u.release();
        s->d_env_cdf_v.release();
        return MCRT_OK;
    }
    const size_t n = (size_t)w * (size_t)h;
    std::vector<float> px(4 * n), warped(n);
    for (size_t i = 0; i < n; i++)
    {
        const float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        px[4 * i] = r;
        px[4 * i + 1] = g;
        px[4 * i + 2] = b;
        px[4 * i + 3] = (float)(0.299 * (double)r + 0.587 * (double)g + 0.114 * (double)b);
    }
    const float pi = 3.14159265358979323846f;
    for (int m = 0; m < h; m++)
    {
        const float phi = pi * ((float)m + 0.5f) / (float)h;
        const double sp = std::sin((double)phi);
        for (int c = 0; c < w; c++)
        {
            const size_t i = (size_t)c + (size_t)m * w;
            warped[i] = (float)((double)px[4 * i + 3] * sp);
        }
    }
    float total = 0.0f;
    for (float x : warped) total += x;
    const float norm = 1.0f / total;
    if (!(total > 0.0f) || !std::isfinite(total) || !std::isfinite(norm))
        return fail(MCRT_ERR_INVALID_ARG, "environment map: the luminance must sum to a positive finite value");
    auto cdf = [](const float* v, int k, float* out) {
        float z = 0.0f;
        for (int i = 0; i < k; i++) z += v[i];
        const float inv = 1.0f / z;
        float run = 0.0f;
        for (int i = 0; i < k; i++)
        {
            run += v[i] * inv;
            out[i] = run;
        }
    };
    std::vector<float> col(h), colsum(w), cdf_v(n), cdf_u(w);
    for (int c = 0; c < w; c++)
    {
        float acc = 0.0f;
        for (int m = 0; m < h; m++)
        {
            col[m] = warped[(size_t)c + (size_t)m * w];
            acc += col[m];
        }
        colsum[c] = acc;
        cdf(col.data(), h, &cdf_v[(size_t)c * h]);
    }
    cdf(colsum.data(), w, cdf_u.data());
    s->env_gen++;
    int rc;
    if ((rc = s->d_env_px.upload(px.data(), px.size() * sizeof(float)))) return rc;
    if ((rc = s->d_env_cdf_u.upload(cdf_u.data(), cdf_u.size() * sizeof(float)))) return rc;
    if ((rc = s->d_env_cdf_v.upload(cdf_v.data(), cdf_v.size() * sizeof(float)))) return rc;
    d.env_px = s->d_env_px.as<float>();
    d.env_cdf_u = s->d_env_cdf_u.as<float>();
    d.env_cdf_v = s->d_env_cdf_v.as<float>();
    d.env_w = w;
    d.env_h = h;
    d.env_norm = norm;
    return MCRT_OK;
}

// The radiance of an environment-map file (EnvmapLoader's constructors,
// envmap.cpp:98-140): {int32 width, int32 height} then the floats, or, with
// width and height given, the floats alone.
int read_envmap(const char* path, int& w, int& h, std::vector<float>& rgb)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(MCRT_ERR_IO, "cannot open environment map %s", path);
    int32_t hd[2] = {w, h};
    if (w == 0 && h == 0 && std::fread(hd, sizeof(int32_t), 2, f) != 2)
    {
        std::fclose(f);
        return fail(MCRT_ERR_IO, "%s: no environment-map header", path);
    }
    if (hd[0] <= 0 || hd[1] <= 0 || (int64_t)hd[0] * hd[1] > (int64_t(1) << 28))
    {
        std::fclose(f);
        return fail(MCRT_ERR_INVALID_ARG, "%s: environment map size %d x %d out of range", path, hd[0], hd[1]);
    }
    w = hd[0];
    h = hd[1];
    const size_t n = (size_t)w * h * 3;
    rgb.resize(n);
    const size_t got = std::fread(rgb.data(), sizeof(float), n, f);
    std::fclose(f);
    if (got != n) return fail(MCRT_ERR_IO, "%s: %zu of %zu floats", path, got, n);
    return MCRT_OK;
}

void trace_geometry(DevScene& d, int num_cus, int waves_per_cu, int lds_depth)
{
    d.lds_depth = lds_depth;
    d.spill_lanes = (int64_t)num_cus * waves_per_cu * 64;
    const int any_waves = 4 * kTraceAnyWavesPerSimd;
    d.lanes_any = (int64_t)num_cus * (waves_per_cu < any_waves ? waves_per_cu : any_waves) * 64;
    d.spill_depth = d.stack_bound > d.lds_depth ? d.stack_bound - d.lds_depth : 0;
    d.spill = nullptr;  // per stream, see launch_view
}

void apply_trace_tuning(mcrt_scene* s)
{
    trace_geometry(s->dev, s->ctx->num_cus, s->trace_waves_per_cu, s->trace_lds_depth);
    s->dev.reorder_rays = s->trace_reorder_rays > 0 ? s->trace_reorder_rays : 0;
    s->dev.refill_lanes = s->trace_refill_lanes;
    s->dev.refill_lanes_any = s->trace_any_refill_lanes;
    s->dev.det_eps = s->det_eps;
}

StreamRes* stream_res(mcrt_scene* s, hipStream_t stream)
{
    std::lock_guard<std::mutex> g(s->ctx->res_mu);
    std::unique_ptr<StreamRes>& r = s->ctx->res[stream];
    if (!r) r.reset(new StreamRes());
    return r.get();
}

int render_view(mcrt_scene* s, StreamRes* r, DevScene& view)
{
    int rc = launch_view(s, r, view);
    if (rc) return rc;
    trace_geometry(view, s->ctx->num_cus, s->pt_trace_waves_per_cu, s->pt_trace_lds_depth);
    view.refill_lanes = s->pt_trace_refill_lanes;
    if (view.spill_depth > 0)
    {
        if ((rc = r->spill.ensure((size_t)view.spill_lanes * view.spill_depth * sizeof(uint64_t)))) return rc;
        view.spill = r->spill.as<uint64_t>();
    }
    return MCRT_OK;
}

int launch_view(mcrt_scene* s, StreamRes* r, DevScene& view, int64_t n)
{
    view = s->dev;
    if (view.spill_depth > 0)
    {
        int rc = r->spill.ensure((size_t)view.spill_lanes * view.spill_depth * sizeof(uint64_t));
        if (rc) return rc;
        view.spill = r->spill.as<uint64_t>();
    }
    int rc = r->wq.ensure(kWqWords * sizeof(uint32_t));
    if (rc) return rc;
    view.wq = r->wq.as<uint32_t>();
    view.order = nullptr;
    view.cell_work = nullptr;
    if (view.reorder_rays > 0 && n >= view.reorder_rays)
    {
        // one launch covers at most 2^30 rays (launch_trace)
        const int64_t m = n < (int64_t(1) << 30) ? n : (int64_t(1) << 30);
        if ((rc = r->order.ensure((size_t)m * sizeof(uint32_t)))) return rc;
        if ((rc = r->cells.ensure(kCellWorkWords * sizeof(uint32_t)))) return rc;
        view.order = r->order.as<uint32_t>();
        view.cell_work = r->cells.as<uint32_t>();
    }
    return MCRT_OK;
}

int run_query(mcrt_scene* s, const float* rays, size_t n, mcrt_hit* hits, uint8_t* occ, int mode, int flags,
              hipStream_t stream, uint64_t* counters)
{
    if (!s->finalized) return fail(MCRT_ERR_STATE, "ray query before mcrt_scene_finalize");
    if (n == 0) return MCRT_OK;
    HIPCHK(hipSetDevice(s->ctx->device));
    StreamRes* r = stream_res(s, stream);
    std::lock_guard<std::mutex> gr(r->mu);
    DevScene view;
    int rc = launch_view(s, r, view, (int64_t)n);
    if (rc) return rc;
    if (flags & MCRT_DEVICE_PTRS)
    {
        HIPCHK(launch_trace(view, rays, (int64_t)n, hits, occ, mode, counters, stream));
        return MCRT_OK;
    }
    std::lock_guard<std::mutex> g(s->q_mu);
    const size_t in_bytes = n * 8 * sizeof(float);
    const size_t out_bytes = mode == TRACE_ANY ? n : n * sizeof(mcrt_hit);
    if ((rc = s->q_in.ensure(in_bytes))) return rc;
    if ((rc = s->q_out.ensure(out_bytes))) return rc;
    HIPCHK(hipMemcpyAsync(s->q_in.p, rays, in_bytes, hipMemcpyHostToDevice, stream));
    HIPCHK(launch_trace(view, s->q_in.as<float>(), (int64_t)n, mode == TRACE_ANY ? nullptr : s->q_out.p,
                        mode == TRACE_ANY ? s->q_out.as<uint8_t>() : nullptr, mode, counters, stream));
    HIPCHK(hipMemcpyAsync(mode == TRACE_ANY ? (void*)occ : (void*)hits, s->q_out.p, out_bytes,
                          hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return MCRT_OK;
}
}  // namespace mcrt
