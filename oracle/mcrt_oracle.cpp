// mcrt oracle -- CPU restatement of the reference's ray-query core and path
// tracer.  TEST INFRASTRUCTURE ONLY: this file is built into
// oracle/liboracle.so, which only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load.  The product (mcrt_amd/) never links it.
//
// Every function restates the reference arithmetic in the reference order
// (left-to-right vec3 ops, no FMA: build with -ffp-contract=off and no -march)
// and cites the reference file:line it follows.  Reference paths are relative
// to the reference repository root (src/core/...).
//
// Parity pins (see tests/): the builder's tree hash and every hit against the
// compiled reference (oracle/_ref), brute force, textbook special cases of
// the slab and Moller-Trumbore tests, closed forms for camera/BSDF/light
// sampling, an analytic furnace for the integrator, and an FP64 render.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace orc
{
static const float kInf = std::numeric_limits<float>::infinity();
static const float kPi = 3.14159265358979323846f;  // constants.cpp:11,19
// TriangleObject::s_intersection_epsilon (triangle.cpp:9): NUM_EPS_DET
// (constants.cpp:15) unless set, process-global as in the reference.
static float kEpsDet = 1e-6f;
static const float kEpsRay = 1e-4f;                 // constants.cpp:17
static const float kEpsCos = 1e-6f;                 // constants.cpp:19
static const int kLeafSize = 4;                     // splitbvh.h:27
static const int kBins = 16;                        // splitbvh.h:28
static const float kEdgeSplitThreshold = 1e-3f;     // splitbvh.h:30

// ---------------------------------------------------------------- vec3 ----
// vectors.cpp:114-231 -- component-wise ops, dot is (x*x'+y*y')+z*z'.
struct V3
{
    float x, y, z;
    float operator[](int i) const { return (&x)[i]; }
    float& operator[](int i) { return (&x)[i]; }
};
static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
static inline V3 operator*(V3 a, float f) { return V3{a.x * f, a.y * f, a.z * f}; }
static inline V3 operator*(float f, V3 a) { return a * f; }  // vectors.cpp:248
static inline V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b)  // vectors.cpp:180-185
{
    return V3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z),
              (a.x * b.y) - (a.y * b.x)};
}
static inline float length(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
static inline V3 normalize(V3 v)  // vectors.cpp:227-231
{
    float inv = 1.0f / length(v);
    return V3{v.x * inv, v.y * inv, v.z * inv};
}
static inline float fmin_(float a, float b) { return a < b ? a : b; }  // numeric.cpp:24
static inline float fmax_(float a, float b) { return a > b ? a : b; }  // numeric.cpp:29
static inline float clampf_(float x, float a, float b) { return x < a ? a : x > b ? b : x; }
static inline V3 vmin(V3 a, V3 b) { return V3{fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
static inline V3 vmax(V3 a, V3 b) { return V3{fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }

// ---------------------------------------------------------------- AABB ----
// aabb.cpp:260-346
struct Box
{
    V3 mn{kInf, kInf, kInf};
    V3 mx{-kInf, -kInf, -kInf};
    void expand(const Box& b)  // aabb.cpp:285-303
    {
        mn.x = fmin_(mn.x, b.mn.x); mn.y = fmin_(mn.y, b.mn.y); mn.z = fmin_(mn.z, b.mn.z);
        mx.x = fmax_(mx.x, b.mx.x); mx.y = fmax_(mx.y, b.mx.y); mx.z = fmax_(mx.z, b.mx.z);
    }
    void expand(V3 p)  // aabb.cpp:305-322
    {
        mn.x = fmin_(mn.x, p.x); mn.y = fmin_(mn.y, p.y); mn.z = fmin_(mn.z, p.z);
        mx.x = fmax_(mx.x, p.x); mx.y = fmax_(mx.y, p.y); mx.z = fmax_(mx.z, p.z);
    }
    bool valid() const { return mx.x >= mn.x && mx.y >= mn.y && mx.z >= mn.z; }
    float sa_cost() const  // aabb.cpp:335-341 (half surface area)
    {
        V3 e = mx - mn;
        return e.x * e.y + e.y * e.z + e.x * e.z;
    }
    V3 center() const { return (mn + mx) * 0.5f; }  // aabb.cpp:480-483
    int largest_dim() const  // aabb.cpp:348-371
    {
        float e0 = mx.x - mn.x, e1 = mx.y - mn.y, e2 = mx.z - mn.z;
        int d = 0;
        float m = e0;
        if (e1 > m) { m = e1; d = 1; }
        if (e2 > m) { m = e2; d = 2; }
        return d;
    }
    void reset() { *this = Box(); }
    void intersect_by(const Box& b)  // aabb.cpp:451-478
    {
        mn.x = fmax_(mn.x, b.mn.x); mx.x = fmin_(mx.x, b.mx.x);
        if (mn.x > mx.x) { reset(); return; }
        mn.y = fmax_(mn.y, b.mn.y); mx.y = fmin_(mx.y, b.mx.y);
        if (mn.y > mx.y) { reset(); return; }
        mn.z = fmax_(mn.z, b.mn.z); mx.z = fmin_(mx.z, b.mx.z);
        if (mn.z > mx.z) { reset(); return; }
    }
    Box intersect(const Box& b) const  // aabb.cpp:395-426
    {
        Box r = *this;
        r.intersect_by(b);
        return r;
    }
    Box clip_axis(float lo, float hi, int axis) const  // aabb.cpp:495-531
    {
        Box r = *this;
        r.mn[axis] = clampf_(r.mn[axis], lo, hi);
        r.mx[axis] = clampf_(r.mx[axis], lo, hi);
        return r;
    }
};

// ----------------------------------------------------------------- Ray ----
// ray.cpp:66-85: inv_dir = (d==0 ? 0 : 1/d), dir_sign = d<0
struct Ray
{
    V3 o, d;
    float tmin, tmax;
    V3 inv;
    int sign[3];
    Ray() {}
    Ray(V3 o_, V3 d_, float t0, float t1) : o(o_), d(d_), tmin(t0), tmax(t1)
    {
        inv.x = d.x == 0.0f ? 0.0f : 1.0f / d.x;
        inv.y = d.y == 0.0f ? 0.0f : 1.0f / d.y;
        inv.z = d.z == 0.0f ? 0.0f : 1.0f / d.z;
        sign[0] = d.x < 0.0f; sign[1] = d.y < 0.0f; sign[2] = d.z < 0.0f;
    }
};

// aabb.cpp:128-191 (live branch of rayAABBOverlap): far-cull first.
static inline bool ray_box_overlap(const Ray& r, const V3& bmin, const V3& bmax)
{
    const V3 b[2] = {bmin, bmax};
    float tmin = (b[r.sign[0]].x - r.o.x) * r.inv.x;
    float tymin = (b[r.sign[1]].y - r.o.y) * r.inv.y;
    float tmax = (b[1 - r.sign[0]].x - r.o.x) * r.inv.x;
    float tymax = (b[1 - r.sign[1]].y - r.o.y) * r.inv.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tmin > r.tmax) return false;
    if (tymax < tmax) tmax = tymax;
    if (tmax < r.tmin) return false;
    float tzmin = (b[r.sign[2]].z - r.o.z) * r.inv.z;
    float tzmax = (b[1 - r.sign[2]].z - r.o.z) * r.inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tmin > r.tmax) return false;
    if (tzmax < tmax) tmax = tzmax;
    if (tmax < r.tmin) return false;
    return true;
}

// triangle.cpp:13-51 (Moller-Trumbore, no culling, absolute det epsilon).
static inline bool ray_tri(V3 o, V3 d, V3 v0, V3 v1, V3 v2, float& t, float& u, float& v)
{
    V3 e1 = v1 - v0;
    V3 e2 = v2 - v0;
    V3 p = cross(d, e2);
    float det = dot(e1, p);
    if (std::fabs(det) < kEpsDet) return false;
    float inv = 1 / det;
    V3 tv = o - v0;
    u = dot(tv, p) * inv;
    if (u < 0 || u > 1) return false;
    V3 q = cross(tv, e1);
    v = dot(d, q) * inv;
    if (v < 0 || u + v > 1) return false;
    t = dot(e2, q) * inv;
    return true;
}

struct Tri
{
    V3 a, b, c;
};

static inline Box tri_box(const Tri& t)  // triangle.cpp:256-261
{
    Box r;
    r.mn = vmin(t.a, t.b);
    r.mx = vmax(t.a, t.b);
    r.expand(t.c);
    return r;
}

// boundedtriangle.cpp:3-124 (live branch): tight bound of the part of the
// triangle inside the [lo, hi] slab of `axis`, intersected with the clipped
// bound of the reference.
static Box clip_bound_tight(const Tri& tri, const Box& clipped, float lo, float hi, int axis)
{
    if (clipped.mn[axis] > hi || clipped.mx[axis] < lo) return Box();
    const V3& a = tri.a;
    const V3& b = tri.b;
    const V3& c = tri.c;
    Box tight;
    {
        V3 ba = b - a;
        float t0 = (lo - a[axis]) / ba[axis];
        float t1 = (hi - a[axis]) / ba[axis];
        if (t0 >= 0.0f && t0 <= 1.0f) tight.expand(a + ba * t0);
        if (t1 >= 0.0f && t1 <= 1.0f) tight.expand(a + ba * t1);
    }
    {
        V3 cb = c - b;
        float t0 = (lo - b[axis]) / cb[axis];
        float t1 = (hi - b[axis]) / cb[axis];
        if (t0 >= 0.0f && t0 <= 1.0f) tight.expand(b + cb * t0);
        if (t1 >= 0.0f && t1 <= 1.0f) tight.expand(b + cb * t1);
    }
    {
        V3 ac = a - c;
        float t0 = (lo - c[axis]) / ac[axis];
        float t1 = (hi - c[axis]) / ac[axis];
        if (t0 >= 0.0f && t0 <= 1.0f) tight.expand(c + ac * t0);
        if (t1 >= 0.0f && t1 <= 1.0f) tight.expand(c + ac * t1);
    }
    if (a[axis] >= lo && a[axis] <= hi) tight.expand(a);
    if (b[axis] >= lo && b[axis] <= hi) tight.expand(b);
    if (c[axis] >= lo && c[axis] <= hi) tight.expand(c);
    Box more = clipped.clip_axis(lo, hi, axis);
    tight.intersect_by(more);
    if (tight.valid()) return more.intersect(tight);
    if (more.valid()) return more;
    return Box();
}

// ------------------------------------------------------------ builder ----
// splitbvh.cpp:178-1625.  A reference is (triangle index, clipped box).
struct Ref
{
    Box box;
    int tri;
};

struct Node
{
    Box box;
    int leaf = 0;
    int axis = -1;
    int left = -1, right = -1;
    int nref = 0;
    Ref ref[kLeafSize];
};

struct BVH
{
    std::vector<Tri> tris;
    std::vector<Node> nodes;  // nodes[0] = root
    int max_depth = 0;
    long long n_refs = 0;
};

static inline int bin_of(float v, float nmin, float delta)  // splitbvh.cpp:286-287
{
    int i = int((v - nmin) / delta);
    return i < kBins - 1 ? i : kBins - 1;
}

// MedianSplitBVHSplitter (splitbvh.cpp:178-236): sort a copy by clipped-box
// centre on the node's largest dimension, split at n/2.
struct MedianSplit
{
    std::vector<Ref> refs;
    int axis = -1, mid = -1;
    float cost = kInf;
    MedianSplit(const std::vector<Ref>& in, const Box& nb) : refs(in)
    {
        axis = nb.largest_dim();
        const int ax = axis;
        std::sort(refs.begin(), refs.end(), [ax](const Ref& a, const Ref& b) {
            return a.box.center()[ax] < b.box.center()[ax];
        });
        mid = (int)refs.size() / 2;
        int total = (int)refs.size();
        Box lb, rb;
        for (int i = 0; i < mid; i++) lb.expand(refs[i].box);
        for (int i = mid; i < total; i++) rb.expand(refs[i].box);
        cost = lb.sa_cost() * mid + rb.sa_cost() * (total - mid);
    }
};

struct Bin
{
    Box box;
    int entries = 0, exits = 0, valid = 0;
};

// Sweep of the 15 candidate planes shared by the object and spatial
// splitters (splitbvh.cpp:307-367, :791-854): early skip on the left cost
// alone (>=), accept on <= so later planes and axes win ties.
static void sweep_bins(const Bin* bins, bool spatial, int axis, float& min_cost, int& min_axis,
                       int& min_bin)
{
    for (int i = 1; i < kBins; i++)
    {
        Box lb;
        int cl = 0;
        for (int j = 0; j < i; j++)
            if (bins[j].valid) { cl += bins[j].entries; lb.expand(bins[j].box); }
        float cost = 0.0f;
        if (cl > 0) cost += lb.sa_cost() * cl;
        else cost = kInf;
        if (cost >= min_cost) continue;
        Box rb;
        int cr = 0;
        for (int j = i; j < kBins; j++)
            if (bins[j].valid) { cr += spatial ? bins[j].exits : bins[j].entries; rb.expand(bins[j].box); }
        if (cr > 0) cost += rb.sa_cost() * cr;
        else cost = kInf;
        if (cost <= min_cost) { min_cost = cost; min_axis = axis; min_bin = i; }
    }
}

// SAHBVHSplitter (splitbvh.cpp:238-662): 16 bins by clipped-box MIN corner.
struct SAHSplit
{
    const std::vector<Ref>& refs;
    const Box& nb;
    float edge[3];
    float cost = kInf;
    int axis = -1, bin = -1;
    SAHSplit(const std::vector<Ref>& in, const Box& b) : refs(in), nb(b)
    {
        V3 e = nb.mx - nb.mn;
        edge[0] = e.x; edge[1] = e.y; edge[2] = e.z;
        for (int ax = 0; ax < 3; ax++)
        {
            if (!(edge[ax] > kEdgeSplitThreshold)) continue;
            Bin bins[kBins];
            float delta = edge[ax] / kBins;
            for (const Ref& r : refs)
            {
                int k = bin_of(r.box.mn[ax], nb.mn[ax], delta);
                bins[k].entries++;
                bins[k].box.expand(r.box);
                bins[k].valid = 1;
            }
            sweep_bins(bins, false, ax, cost, axis, bin);
        }
    }
    void split(std::vector<Ref>& l, std::vector<Ref>& r) const  // :624-662
    {
        float delta = edge[axis] / kBins;
        for (const Ref& x : refs)
        {
            int k = bin_of(x.box.mn[axis], nb.mn[axis], delta);
            if (k < bin) l.push_back(x);
            else r.push_back(x);
        }
    }
};

// SplitBVHSplitter (splitbvh.cpp:664-1408): spatial bins; each spanned bin is
// grown by the tight clip of the triangle to the bin slab.
struct SpatialSplit
{
    const std::vector<Ref>& refs;
    const std::vector<Tri>& tris;
    const Box& nb;
    float edge[3];
    float cost = kInf;
    int axis = -1, bin = -1;
    SpatialSplit(const std::vector<Ref>& in, const std::vector<Tri>& t, const Box& b)
        : refs(in), tris(t), nb(b)
    {
        V3 e = nb.mx - nb.mn;
        edge[0] = e.x; edge[1] = e.y; edge[2] = e.z;
        for (int ax = 0; ax < 3; ax++)
        {
            if (!(edge[ax] > kEdgeSplitThreshold)) continue;
            Bin bins[kBins];
            float delta = edge[ax] / kBins;
            float nmin = nb.mn[ax];
            for (const Ref& r : refs)
            {
                const Tri& tri = tris[r.tri];
                int i0 = bin_of(r.box.mn[ax], nmin, delta);
                int i1 = bin_of(r.box.mx[ax], nmin, delta);
                float bmin = nmin + i0 * delta;
                Box cb = clip_bound_tight(tri, r.box, bmin, bmin + delta, ax);
                if (cb.valid()) { bins[i0].entries++; bins[i0].box.expand(cb); bins[i0].valid = 1; }
                // The z sweep starts its middle loop at i0 (splitbvh.cpp:1047),
                // x and y at i0+1 (:757, :902); re-clipping bin i0 is idempotent.
                for (int j = (ax == 2 ? i0 : i0 + 1); j < i1; j++)
                {
                    bmin = nmin + j * delta;
                    cb = clip_bound_tight(tri, r.box, bmin, bmin + delta, ax);
                    if (cb.valid()) { bins[j].box.expand(cb); bins[j].valid = 1; }
                }
                bmin = nmin + i1 * delta;
                cb = clip_bound_tight(tri, r.box, bmin, bmin + delta, ax);
                if (cb.valid()) { bins[i1].exits++; bins[i1].box.expand(cb); bins[i1].valid = 1; }
            }
            sweep_bins(bins, true, ax, cost, axis, bin);
        }
    }
    void split(std::vector<Ref>& l, std::vector<Ref>& r) const  // :1216-1408
    {
        float delta = edge[axis] / kBins;
        float nmin = nb.mn[axis], nmax = nb.mx[axis];
        float clip = nmin + bin * delta;
        for (const Ref& x : refs)
        {
            int i0 = bin_of(x.box.mn[axis], nmin, delta);
            int i1 = bin_of(x.box.mx[axis], nmin, delta);
            const Tri& tri = tris[x.tri];
            if (i0 < bin)
            {
                Box cb = clip_bound_tight(tri, x.box, nmin, clip, axis);
                if (cb.valid()) l.push_back(Ref{cb, x.tri});
            }
            if (i1 >= bin)
            {
                Box cb = clip_bound_tight(tri, x.box, clip, nmax, axis);
                if (cb.valid()) r.push_back(Ref{cb, x.tri});
            }
        }
    }
};

// SplitBVH::build (splitbvh.cpp:1435-1625).
static int build_rec(BVH& bvh, const std::vector<Ref>& refs, int depth)
{
    if (depth > bvh.max_depth) bvh.max_depth = depth;
    int id = (int)bvh.nodes.size();
    bvh.nodes.emplace_back();
    Box box;
    for (const Ref& r : refs) box.expand(r.box);
    if ((int)refs.size() <= kLeafSize)
    {
        Node& n = bvh.nodes[id];
        n.leaf = 1;
        n.box = box;
        n.nref = (int)refs.size();
        for (int i = 0; i < n.nref; i++) n.ref[i] = refs[i];
        bvh.n_refs += n.nref;
        return id;
    }
    std::vector<Ref> l, r;
    int axis;
    {
        MedianSplit med(refs, box);
        SAHSplit sah(refs, box);
        SpatialSplit spa(refs, bvh.tris, box);
        // sort low to high cost with SAH first (splitbvh.cpp:1520-1537)
        int types[3] = {1, 0, 2};
        float costs[3] = {sah.cost, med.cost, spa.cost};
        if (costs[0] > costs[1]) { std::swap(costs[0], costs[1]); std::swap(types[0], types[1]); }
        if (costs[1] > costs[2]) { std::swap(costs[1], costs[2]); std::swap(types[1], types[2]); }
        if (costs[0] > costs[1]) { std::swap(costs[0], costs[1]); std::swap(types[0], types[1]); }
        if (types[0] == 0)
        {
            axis = med.axis;
            l.assign(med.refs.begin(), med.refs.begin() + med.mid);
            r.assign(med.refs.begin() + med.mid, med.refs.end());
        }
        else if (types[0] == 1) { axis = sah.axis; sah.split(l, r); }
        else { axis = spa.axis; spa.split(l, r); }
    }
    int li = build_rec(bvh, l, depth + 1);
    { std::vector<Ref>().swap(l); }
    int ri = build_rec(bvh, r, depth + 1);
    Node& n = bvh.nodes[id];
    n.box = box;
    n.axis = axis;
    n.left = li;
    n.right = ri;
    return id;
}

// SplitBVH::SplitBVH (splitbvh.cpp:1410-1433): primID = scene index; root
// references carry the full triangle box.
static BVH* build(const float* v, int n)
{
    BVH* b = new BVH();
    b->tris.resize(n);
    for (int i = 0; i < n; i++)
        b->tris[i] = Tri{v3(v[9 * i + 0], v[9 * i + 1], v[9 * i + 2]),
                         v3(v[9 * i + 3], v[9 * i + 4], v[9 * i + 5]),
                         v3(v[9 * i + 6], v[9 * i + 7], v[9 * i + 8])};
    std::vector<Ref> refs(n);
    for (int i = 0; i < n; i++) refs[i] = Ref{tri_box(b->tris[i]), i};
    b->nodes.reserve(2 * (size_t)n + 1);
    if (n > 0) build_rec(*b, refs, 0);
    return b;
}

// ----------------------------------------------------------- traversal ----
struct Hit
{
    float t = kInf, u = 0, v = 0;
    int prim = -1;
};
struct Counters
{
    long long boxes = 0, leaf_refs = 0, mt = 0;
};

// SplitBVHLeafNode::intersect / SplitBVHBranchNode::intersect
// (splitbvh.cpp:5-145): depth-first, near child first iff dir[axis] > 0,
// ray.tmax shrinks to each accepted hit, strict t < tmax acceptance.
static bool trav_rec(const BVH& b, int ni, Ray& ray, Hit& hit, Counters* c)
{
    const Node& n = b.nodes[ni];
    if (c) c->boxes++;
    if (!ray_box_overlap(ray, n.box.mn, n.box.mx)) return false;
    if (n.leaf)
    {
        bool any = false;
        for (int i = 0; i < n.nref; i++)
        {
            if (c) c->leaf_refs++;
            if (!ray_box_overlap(ray, n.ref[i].box.mn, n.ref[i].box.mx)) continue;
            // BoundedTriangle::intersect repeats the clipped-box test
            // (boundedtriangle.cpp:131-139); it cannot fail here.
            const Tri& t = b.tris[n.ref[i].tri];
            float tt = kInf, uu = 0, vv = 0;
            if (c) c->mt++;
            bool h = ray_tri(ray.o, ray.d, t.a, t.b, t.c, tt, uu, vv);
            if (h && tt > ray.tmin && tt < ray.tmax)  // triangle.cpp:207
            {
                if (tt < hit.t)
                {
                    hit.t = tt; hit.u = uu; hit.v = vv; hit.prim = n.ref[i].tri;
                    any = true;
                    ray.tmax = tt;
                }
            }
        }
        return any;
    }
    bool any = false;
    Hit child;
    int first = ray.d[n.axis] > 0.0f ? n.left : n.right;
    int second = ray.d[n.axis] > 0.0f ? n.right : n.left;
    if (first >= 0 && trav_rec(b, first, ray, child, c) && child.t < hit.t)
    {
        hit = child; any = true; ray.tmax = hit.t;
    }
    if (second >= 0 && trav_rec(b, second, ray, child, c) && child.t < hit.t)
    {
        hit = child; any = true; ray.tmax = hit.t;
    }
    return any;
}

static Hit intersect(const BVH& b, V3 o, V3 d, float tmin, float tmax, Counters* c)
{
    Hit h;
    if (b.nodes.empty()) return h;
    Ray r(o, d, tmin, tmax);
    if (!trav_rec(b, 0, r, h, c)) h = Hit();
    return h;
}

// DebugMesh::intersect (debug_mesh.cpp:780-810): all triangles, tmax never
// shrinks, first index wins exact ties.
static Hit brute(const std::vector<Tri>& tris, V3 o, V3 d, float tmin, float tmax)
{
    Hit h;
    for (size_t i = 0; i < tris.size(); i++)
    {
        float t = kInf, u = 0, v = 0;
        bool hit = ray_tri(o, d, tris[i].a, tris[i].b, tris[i].c, t, u, v);
        if (hit && t > tmin && t < tmax && t < h.t)
        {
            h.t = t; h.u = u; h.v = v; h.prim = (int)i;
        }
    }
    return h;
}

// --------------------------------------------------------- tree digest ----
// Canonical pre-order byte stream of the binary tree: kind, box, axis or
// (count, [prim, clipped box]...).  -0.0f is folded to +0.0f because the
// builder's min/max reductions may legitimately differ in zero sign.
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
    void box(const Box& b) { f(b.mn.x); f(b.mn.y); f(b.mn.z); f(b.mx.x); f(b.mx.y); f(b.mx.z); }
};

static void digest_rec(const BVH& b, int ni, Fnv& h)
{
    const Node& n = b.nodes[ni];
    h.i(n.leaf ? 1 : 0);
    h.box(n.box);
    if (n.leaf)
    {
        h.i(n.nref);
        for (int i = 0; i < n.nref; i++) { h.i(n.ref[i].tri); h.box(n.ref[i].box); }
        return;
    }
    h.i(n.axis);
    digest_rec(b, n.left, h);
    digest_rec(b, n.right, h);
}

// ---------------------------------------------------------------- OBJ ----
// Mesh::Mesh (mesh.cpp:24-60): whitespace tokens, `v x y z` and `f a b c`.
// At end of file the reference's last `reader >> token` fails and leaves the
// previous token in place; if that was "f" it appends a phantom face with
// indices -1 (extraction failures store 0) and later reads and writes
// m_verts[-1] / m_vertex_normals[-1].  Only completely extracted records are
// kept here, which is the reference's result for every real face.
struct Mesh
{
    std::vector<V3> v;
    std::vector<int> f;  // 3 per face, 0-based
    std::vector<V3> vn;
};

// A `v`/`f` token followed by a non-number before the end of the file (the
// reference's loop never ends on it) is reported as failure (-2).
static int load_obj(const char* path, Mesh& m)
{
    std::ifstream ifs(path, std::ios::binary);
    if (!ifs.good()) return -1;
    std::stringstream reader;
    reader << ifs.rdbuf();
    std::string token;
    while (reader >> token)
    {
        if (token == "v")
        {
            float a, b, c;
            if (reader >> a >> b >> c) m.v.push_back(v3(a, b, c));
            else if (!reader.eof()) return -2;
        }
        else if (token == "f")
        {
            unsigned int a, b, c;
            if (reader >> a >> b >> c)
            {
                m.f.push_back((int)(a - 1)); m.f.push_back((int)(b - 1)); m.f.push_back((int)(c - 1));
            }
            else if (!reader.eof()) return -2;
        }
    }
    for (int idx : m.f)
        if (idx < 0 || (size_t)idx >= m.v.size()) return -2;
    return 0;
}

// Scene::loadModel (scene.cpp:368-381): v' = transform * vec4(v, 1) with
// GLM's tmat4x4 * tvec4 (type_mat4x4.inl:532-543): per output row r,
// (m[0][r]*x + m[1][r]*y) + (m[2][r]*z + m[3][r]*w); m16 is column-major,
// m16[4*c + r] = m[c][r].
static void transform_vertices(Mesh& m, const float* m16)
{
    for (auto& p : m.v)
    {
        const float x = p.x, y = p.y, z = p.z, w = 1.0f;
        float q[3];
        for (int r = 0; r < 3; r++)
        {
            const float mul0 = m16[0 + r] * x, mul1 = m16[4 + r] * y;
            const float mul2 = m16[8 + r] * z, mul3 = m16[12 + r] * w;
            const float add0 = mul0 + mul1, add1 = mul2 + mul3;
            q[r] = add0 + add1;
        }
        p = v3(q[0], q[1], q[2]);
    }
}

// Mesh::genNormals (mesh.cpp:66-92): vertex normal = normalize(sum of
// unnormalised (vb-va) x (vc-va)) in face order.
static void gen_normals(Mesh& m)
{
    m.vn.assign(m.v.size(), v3(0, 0, 0));
    for (size_t i = 0; i < m.f.size() / 3; i++)
    {
        int a = m.f[3 * i], b = m.f[3 * i + 1], c = m.f[3 * i + 2];
        V3 fnrm = cross(m.v[b] - m.v[a], m.v[c] - m.v[a]);
        m.vn[a] = m.vn[a] + fnrm;
        m.vn[b] = m.vn[b] + fnrm;
        m.vn[c] = m.vn[c] + fnrm;
    }
    for (auto& n : m.vn) n = normalize(n);
}

}  // namespace orc

using namespace orc;

// =============================================================== C API ====

ORC_API int orc_ray_box(const float* ray8, const float* box6)
{
    Ray r(v3(ray8[0], ray8[1], ray8[2]), v3(ray8[4], ray8[5], ray8[6]), ray8[3], ray8[7]);
    return ray_box_overlap(r, v3(box6[0], box6[1], box6[2]), v3(box6[3], box6[4], box6[5])) ? 1 : 0;
}

// Returns 1 and writes (t,u,v) if MT accepts with tmin < t < tmax.
ORC_API int orc_ray_tri(const float* ray8, const float* tri9, float* tuv)
{
    V3 o = v3(ray8[0], ray8[1], ray8[2]), d = v3(ray8[4], ray8[5], ray8[6]);
    float t = kInf, u = 0, v = 0;
    bool h = ray_tri(o, d, v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]),
                     v3(tri9[6], tri9[7], tri9[8]), t, u, v);
    if (h && t > ray8[3] && t < ray8[7]) { tuv[0] = t; tuv[1] = u; tuv[2] = v; return 1; }
    return 0;
}

ORC_API void orc_clip_bound_tight(const float* tri9, const float* box6, float lo, float hi,
                                  int axis, float* out6)
{
    Tri t{v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]), v3(tri9[6], tri9[7], tri9[8])};
    Box b;
    b.mn = v3(box6[0], box6[1], box6[2]);
    b.mx = v3(box6[3], box6[4], box6[5]);
    Box r = clip_bound_tight(t, b, lo, hi, axis);
    out6[0] = r.mn.x; out6[1] = r.mn.y; out6[2] = r.mn.z;
    out6[3] = r.mx.x; out6[4] = r.mx.y; out6[5] = r.mx.z;
}

ORC_API void* orc_bvh_build(const float* tris9, int n) { return build(tris9, n); }
ORC_API void orc_bvh_free(void* h) { delete (BVH*)h; }

// stats: [nodes, leaves, refs, max_depth, empty_leaves]
ORC_API uint64_t orc_bvh_digest(void* h, long long* stats)
{
    BVH* b = (BVH*)h;
    Fnv f;
    if (!b->nodes.empty()) digest_rec(*b, 0, f);
    if (stats)
    {
        long long leaves = 0, empty = 0;
        for (const Node& n : b->nodes) if (n.leaf) { leaves++; if (n.nref == 0) empty++; }
        stats[0] = (long long)b->nodes.size();
        stats[1] = leaves;
        stats[2] = b->n_refs;
        stats[3] = b->max_depth;
        stats[4] = empty;
    }
    return f.h;
}

// rays: n x 8 floats (ox oy oz tmin dx dy dz tmax); prim: n int32; tuv: n x 3.
// counters (optional): 3 int64 totals (box tests, leaf refs, MT tests).
ORC_API void orc_bvh_intersect(void* h, const float* rays, long long n, int* prim, float* tuv,
                               long long* counters, int nthreads)
{
    const BVH* b = (const BVH*)h;
    long long cb = 0, cl = 0, cm = 0;
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads > 0 ? nthreads : 1) \
    reduction(+ : cb, cl, cm)
    for (long long i = 0; i < n; i++)
    {
        const float* r = rays + 8 * i;
        Counters c;
        Hit hh = intersect(*b, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7],
                           counters ? &c : nullptr);
        prim[i] = hh.prim;
        tuv[3 * i] = hh.t; tuv[3 * i + 1] = hh.u; tuv[3 * i + 2] = hh.v;
        cb += c.boxes; cl += c.leaf_refs; cm += c.mt;
    }
    if (counters) { counters[0] = cb; counters[1] = cl; counters[2] = cm; }
}

ORC_API void orc_brute_intersect(const float* tris9, int ntri, const float* rays, long long n,
                                 int* prim, float* tuv, int nthreads)
{
    std::vector<Tri> tris(ntri);
    for (int i = 0; i < ntri; i++)
        tris[i] = Tri{v3(tris9[9 * i], tris9[9 * i + 1], tris9[9 * i + 2]),
                      v3(tris9[9 * i + 3], tris9[9 * i + 4], tris9[9 * i + 5]),
                      v3(tris9[9 * i + 6], tris9[9 * i + 7], tris9[9 * i + 8])};
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long i = 0; i < n; i++)
    {
        const float* r = rays + 8 * i;
        Hit hh = brute(tris, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7]);
        prim[i] = hh.prim;
        tuv[3 * i] = hh.t; tuv[3 * i + 1] = hh.u; tuv[3 * i + 2] = hh.v;
    }
}

// OBJ ingest (mesh.cpp:24-118 + scene.cpp:368-408): returns the face count,
// -1 if the file cannot be opened, -2 for a record that is not `v x y z` /
// `f a b c` or a face index out of range.  m16 = column-major transform or
// NULL (identity: the reference's transform leaves the bits unchanged; its
// second genNormals gives the same normals as the first).  With out
// buffers NULL only counts.  tris: nf x 9, vn: nf x 9 (vertex normals per
// corner).
ORC_API long long orc_load_obj(const char* path, const float* m16, float* tris, float* vn)
{
    Mesh m;
    const int rc = load_obj(path, m);
    if (rc) return rc;
    long long nf = (long long)m.f.size() / 3;
    if (!tris) return nf;
    if (m16) transform_vertices(m, m16);
    gen_normals(m);
    for (long long i = 0; i < nf; i++)
        for (int k = 0; k < 3; k++)
        {
            int vi = m.f[3 * i + k];
            tris[9 * i + 3 * k + 0] = m.v[vi].x;
            tris[9 * i + 3 * k + 1] = m.v[vi].y;
            tris[9 * i + 3 * k + 2] = m.v[vi].z;
            if (vn)
            {
                vn[9 * i + 3 * k + 0] = m.vn[vi].x;
                vn[9 * i + 3 * k + 1] = m.vn[vi].y;
                vn[9 * i + 3 * k + 2] = m.vn[vi].z;
            }
        }
    return nf;
}

// TriangleObject::setIntersectionEpsilon (triangle.cpp:53-56).
ORC_API void orc_set_intersection_epsilon(float eps) { kEpsDet = eps; }

ORC_API int orc_num_procs()
{
#ifdef _OPENMP
    return omp_get_num_procs();
#else
    return 1;
#endif
}

#include "oracle_pt.inc"
