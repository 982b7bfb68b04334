// Device traversal core: reference-order closest-hit / any-hit over the
// 4-wide SBVH (layout.h).  Included by every kernel that casts rays.
//
// Semantics are those of SplitBVH::intersect (splitbvh.cpp:5-145,
// :1627-1642): depth-first, near child first iff dir[split_axis] > 0, a
// node's box tested at entry against the *current* tmax, leaf references
// tested in slot order against their clipped boxes and then by
// Moller-Trumbore, a hit accepted iff tmin < t < tmax, tmax shrinking to
// every accepted t (so among equal t the first visited wins).  The 4-wide
// node tests its children when it is visited and a child's entry distance
// `lo` travels on the stack: the reference's box test depends on tmax only
// through `lo <= tmax`, so re-checking `lo` when the entry is popped is the
// reference's entry-time test (conservative, as every node test is: see
// below).  Any-hit stops at the first accepted
// hit, which is the first hit of the same traversal, so its boolean equals
// SceneBVH::occluded's (scenebvh.cpp:379-396).  Node tests only have to be
// conservative (node_margin below): the `lo` carried on the stack is at most
// the reference's, so its re-test passes wherever the reference's passes.
//
// Leaf arithmetic is the reference's: slab t = (bound - o) * inv_dir with
// inv_dir = (d == 0 ? 0 : 1/d) (ray.cpp:77-84), MT with the absolute
// determinant cut (1e-6 unless set; triangle.cpp:13-56); all fp32 with no
// contraction and correctly rounded division.
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"

// Host+device so the identical traversal can also run in a host test build.
#define MCRT_HD __host__ __device__ __forceinline__

namespace mcrt
{
MCRT_HD uint32_t bits_of(float f) { return __builtin_bit_cast(uint32_t, f); }
MCRT_HD float float_of(uint32_t u) { return __builtin_bit_cast(float, u); }
MCRT_HD float4 f4_of(uint4 u) { return make_float4(float_of(u.x), float_of(u.y), float_of(u.z), float_of(u.w)); }

struct RayPrep
{
    float ox, oy, oz, dx, dy, dz;
    float ix, iy, iz;
    float R0;
    float w;     // node-test margin scale (node_margin)
    // per axis, -RN(o * inv) -/+ RN(w * |inv|): the node-test plane map's
    // offsets before the node origin's term (node_margin)
    float cnx, cny, cnz, cfx, cfy, cfz;
    bool sx, sy, sz;
    bool exact;  // an inv_dir component is infinite (or the node-test
                 // margins would leave their bounds): use the reference's
                 // sequential test with its NaN behaviour and descend every
                 // child of a node
    bool oct_hi;         // dz > 0: the octant's rank byte is in the node's upper rank word
    uint32_t oct_shift;  // 8 * (octant & 3): its position in that word
};

MCRT_HD float inv_or_zero(float d) { return d == 0.0f ? 0.0f : 1.0f / d; }

MCRT_HD RayPrep prep_ray(float ox, float oy, float oz, float dx, float dy, float dz,
                                            float tmin)
{
    RayPrep r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.dx = dx; r.dy = dy; r.dz = dz;
    r.ix = inv_or_zero(dx); r.iy = inv_or_zero(dy); r.iz = inv_or_zero(dz);
    r.sx = dx < 0.0f; r.sy = dy < 0.0f; r.sz = dz < 0.0f;
    r.R0 = tmin;
    r.w = 0.0f;
    r.cnx = r.cny = r.cnz = r.cfx = r.cfy = r.cfz = 0.0f;
    r.exact = __builtin_isinf(r.ix) || __builtin_isinf(r.iy) || __builtin_isinf(r.iz);
    // octant (dx > 0) | (dy > 0) << 1 | (dz > 0) << 2 of the nodes' rank tables
    r.oct_hi = dz > 0.0f;
    r.oct_shift = (dx > 0.0f ? 8u : 0u) | (dy > 0.0f ? 16u : 0u);
    return r;
}

// rayAABBOverlap (aabb.cpp:128-191).  Returns pass and the entry distance
// `lo` (the final tmin of the reference's sequence).  With finite inv_dir no
// NaN can arise and the sequence reduces to lo <= hi && lo <= R1 && hi >= R0.
MCRT_HD bool slab(const RayPrep& r, float lx, float hx, float ly, float hy, float lz,
                                     float hz, float R1, float& lo_out)
{
    float ax = (lx - r.ox) * r.ix, bx = (hx - r.ox) * r.ix;
    float ay = (ly - r.oy) * r.iy, by = (hy - r.oy) * r.iy;
    float az = (lz - r.oz) * r.iz, bz = (hz - r.oz) * r.iz;
    float t0x = r.sx ? bx : ax, t1x = r.sx ? ax : bx;
    float t0y = r.sy ? by : ay, t1y = r.sy ? ay : by;
    float t0z = r.sz ? bz : az, t1z = r.sz ? az : bz;
    if (!r.exact)
    {
        float lo = fmaxf(fmaxf(t0x, t0y), t0z);
        float hi = fminf(fminf(t1x, t1y), t1z);
        lo_out = lo;
        return lo <= hi && lo <= R1 && hi >= r.R0;
    }
    if ((t0x > t1y) || (t0y > t1x)) return false;
    float tmin = t0y > t0x ? t0y : t0x;
    if (tmin > R1) return false;
    float tmax = t1y < t1x ? t1y : t1x;
    if (tmax < r.R0) return false;
    if ((tmin > t1z) || (t0z > tmax)) return false;
    if (t0z > tmin) tmin = t0z;
    if (tmin > R1) return false;
    if (t1z < tmax) tmax = t1z;
    if (tmax < r.R0) return false;
    lo_out = tmin;
    return true;
}

// Node tests are conservative.  A node test only decides which subtrees are
// opened, and opening one more changes no result: its references' exact
// clipped boxes then fail (the argument that makes the quantised boxes
// exact, layout.h: QNode).  So a node test may use cheaper arithmetic as long
// as it passes wherever the reference's slab arithmetic on the decoded box,
// t = RN(RN(plane - o) * inv), passes.  With plane = RN(origin + q * s) it
// maps the plane code q straight to
//     t_near = fma(q, s * inv, fma(origin, inv, c_near))
//     t_far  = fma(q, s * inv, fma(origin, inv, c_far)),
// c_near / c_far = RN(-RN(o * inv) -/+ m), m = RN(w * |inv|), formed once per
// ray: per node one fused multiply-add per axis and side, and per plane one
// conversion and one fused multiply-add (s * inv is exact: s = 2^k), where
// the reference arithmetic takes four.  Against it the rounding error of
// t_near is below u * |inv| * (6 max |o| + 5 span) + 3 u m to first order
// (u = 2^-24; with span >= |origin| + 255 s, hence >= |plane|), plus
// subnormal terms below 2^-146: the four roundings here are bounded by
// u (|o inv| + |o inv| + m + |origin - o| |inv| + m + |plane - o| |inv| + m),
// the reference's three by u |inv| (|plane| + 2 |plane - o|).  The margin m
// covers that with room when w = 2^-21 (max |o| + 2 span) = 8 u (...),
// span = qnode_span over the nodes, and w * |inv| >= 2^-100 for every
// non-zero inv.  So t_near is at most, and t_far at least, the reference
// value, and the node test passes wherever the reference-arithmetic test
// passes.  With inv = 0 both are exactly the reference's 0.  Every quantity
// stays finite while w * |inv| < 2^100.  A ray outside these bounds is
// marked exact: it descends every child and is decided by the exact
// reference tests.
MCRT_HD void node_margin(RayPrep& r, float span)
{
    const float om = fmaxf(fmaxf(fabsf(r.ox), fabsf(r.oy)), fabsf(r.oz));
    const float w = fmaxf(0x1p-21f * __builtin_fmaf(2.0f, span, om), 0x1p-60f);
    const float px = w * fabsf(r.ix), py = w * fabsf(r.iy), pz = w * fabsf(r.iz);
    const bool okx = r.ix == 0.0f || (px >= 0x1p-100f && px < 0x1p100f);
    const bool oky = r.iy == 0.0f || (py >= 0x1p-100f && py < 0x1p100f);
    const bool okz = r.iz == 0.0f || (pz >= 0x1p-100f && pz < 0x1p100f);
    r.w = w;
    r.exact = r.exact || !(okx && oky && okz);
    const float qx = r.ox * r.ix, qy = r.oy * r.iy, qz = r.oz * r.iz;
    r.cnx = -qx - px;
    r.cfx = px - qx;
    r.cny = -qy - py;
    r.cfy = py - qy;
    r.cnz = -qz - pz;
    r.cfz = pz - qz;
}

// Conservative test of child `shift / 8` of a QNode (node_margin), given the
// plane words already ordered near / far per axis by the ray's signs, and
// per axis the plane map's slope s * inv (a*) and its lower / upper offsets
// (n*, f*).
MCRT_HD bool qslab_c(const RayPrep& r, uint32_t child, uint32_t nx, uint32_t fx, uint32_t ny, uint32_t fy,
                     uint32_t nz, uint32_t fz, float ax, float ay, float az, float nbx, float fbx, float nby,
                     float fby, float nbz, float fbz, int shift, float R1, float& lo)
{
    // computed for every slot (an empty slot's planes decode to finite
    // values), so the four tests need no branches
    const float t0x = __builtin_fmaf((float)((nx >> shift) & 0xffu), ax, nbx);
    const float t1x = __builtin_fmaf((float)((fx >> shift) & 0xffu), ax, fbx);
    const float t0y = __builtin_fmaf((float)((ny >> shift) & 0xffu), ay, nby);
    const float t1y = __builtin_fmaf((float)((fy >> shift) & 0xffu), ay, fby);
    const float t0z = __builtin_fmaf((float)((nz >> shift) & 0xffu), az, nbz);
    const float t1z = __builtin_fmaf((float)((fz >> shift) & 0xffu), az, fbz);
    lo = fmaxf(fmaxf(t0x, t0y), t0z);
    const float hi = fminf(fminf(t1x, t1y), t1z);
    return (lo <= hi) & (lo <= R1) & (hi >= r.R0) & (child != MCRT_CHILD_EMPTY);
}

// rayTriangleIntersect (triangle.cpp:13-51), e1/e2 precomputed; `eps` is
// the absolute determinant cut (TriangleObject::s_intersection_epsilon,
// 1e-6 unless set, triangle.cpp:9, :53-56).
MCRT_HD bool mt(const RayPrep& r, float v0x, float v0y, float v0z, float e1x, float e1y,
                                   float e1z, float e2x, float e2y, float e2z, float eps, float& t, float& u, float& v)
{
    float px = (r.dy * e2z) - (r.dz * e2y);
    float py = (r.dz * e2x) - (r.dx * e2z);
    float pz = (r.dx * e2y) - (r.dy * e2x);
    float det = e1x * px + e1y * py + e1z * pz;
    if (fabsf(det) < eps) return false;
    float inv = 1.0f / det;
    float tx = r.ox - v0x, ty = r.oy - v0y, tz = r.oz - v0z;
    u = (tx * px + ty * py + tz * pz) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    float qx = (ty * e1z) - (tz * e1y);
    float qy = (tz * e1x) - (tx * e1z);
    float qz = (tx * e1y) - (ty * e1x);
    v = (r.dx * qx + r.dy * qy + r.dz * qz) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = (e2x * qx + e2y * qy + e2z * qz) * inv;
    return true;
}

struct TraceStack
{
    uint32_t* node;  // LDS, [k * stride + lane], depth + 1 rows: row `depth` discards writes
    float* lo;
    int stride, lane, depth;
    uint64_t* spill;  // global, [(k - depth) * spill_stride + spill_lane]: the lanes of a
                      // wavefront spilling the same entry write one contiguous run
    uint32_t spill_lane;
    uint32_t spill_stride;  // lanes the spill area covers

    // Byte offset of this lane's entry in row k (of both arrays).
    MCRT_HD uint32_t off(int k) const { return (uint32_t)k * (uint32_t)(stride * 4) + (uint32_t)(lane * 4); }
    // Unconditional LDS write at off(k), k <= depth (row depth: discarded).
    MCRT_HD void put(uint32_t o, uint32_t n, float l)
    {
        *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(node) + o) = n;
        *reinterpret_cast<float*>(reinterpret_cast<char*>(lo) + o) = l;
    }
    // The same without the entry distance (any-hit never re-tests it).
    MCRT_HD void put_node(uint32_t o, uint32_t n) { *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(node) + o) = n; }
    MCRT_HD void push(int sp, uint32_t n, float l)
    {
        if (sp < depth)
        {
            node[sp * stride + lane] = n;
            lo[sp * stride + lane] = l;
        }
        else
        {
            spill[(size_t)(sp - depth) * spill_stride + spill_lane] =
                ((uint64_t)bits_of(l) << 32) | (uint64_t)n;
        }
    }
    MCRT_HD void pop(int sp, uint32_t& n, float& l) const
    {
        if (sp < depth)
        {
            n = node[sp * stride + lane];
            l = lo[sp * stride + lane];
        }
        else
        {
            uint64_t x = spill[(size_t)(sp - depth) * spill_stride + spill_lane];
            n = (uint32_t)x;
            l = float_of((uint32_t)(x >> 32));
        }
    }
};

struct TraceCount
{
    uint32_t nodes = 0, refs = 0, tris = 0;
};

// The traversal of one ray as a sequence of steps (one wide node or one
// leaf per step), so a kernel can hand a lane its next ray as soon as the
// current one ends instead of waiting for the wavefront's slowest lane.
template <bool ANY, bool COUNT>
struct Tracer
{
    RayPrep r;
    float R1, bu, bv;
    int prim;
    uint32_t cur;
    int sp;

    MCRT_HD void start(const RayPrep& rp, float tmax)
    {
        r = rp;
        R1 = tmax;
        bu = bv = 0.0f;
        prim = -1;
        cur = 0;
        sp = 0;
    }

    // The 64 B record `cur` names: a QNode, or a leaf reference, both in the
    // scene's one record array (layout.h).  Sc is any type with the scene's
    // `nodes` and `det_eps` (DevScene, SceneRoot).
    template <class Sc>
    MCRT_HD const uint4* record(const Sc& S) const
    {
        return reinterpret_cast<const uint4*>(S.nodes) + 4 * (size_t)(cur & MCRT_LEAF_REF_MASK);
    }

    // Visits `cur`; returns true once the traversal has ended.  Node and
    // leaf records are both 64 B, so one fetch serves every lane of a
    // wavefront whichever kind each lane visits: a step costs one memory
    // round trip.
    template <class Sc>
    MCRT_HD bool step(const Sc& S, TraceStack& st, TraceCount& cnt)
    {
        const uint4* p = record(S);
        return visit(p[0], p[1], p[2], p[3], st, cnt, S.det_eps);
    }

    // Visits `cur` given its record w0..w3 (eps: the scene's determinant cut).
    MCRT_HD bool visit(const uint4 w0, const uint4 w1, const uint4 w2, const uint4 w3, TraceStack& st,
                       TraceCount& cnt, float eps)
    {
        if (!(cur & MCRT_CHILD_LEAF))
        {
            // QNode (layout.h): four 16 B rows, child planes origin + q * 2^k
            const uint4 hd = w0, ch = w1, qa = w2, qb = w3;
            const float ox = float_of(hd.x), oy = float_of(hd.y), oz = float_of(hd.z);
            const float sx = float_of((hd.w & 0xffu) << 23);
            const float sy = float_of(((hd.w >> 8) & 0xffu) << 23);
            const float sz = float_of(((hd.w >> 16) & 0xffu) << 23);
            if (COUNT) cnt.nodes++;
            float l0, l1, l2, l3;
            bool p0, p1, p2, p3;
            if (!r.exact)
            {
                // near and far planes by the ray's signs (the swap of
                // aabb.cpp's slab test), chosen once for all four children
                const uint32_t nx = r.sx ? qa.y : qa.x, fx = r.sx ? qa.x : qa.y;
                const uint32_t ny = r.sy ? qa.w : qa.z, fy = r.sy ? qa.z : qa.w;
                const uint32_t nz = r.sz ? qb.y : qb.x, fz = r.sz ? qb.x : qb.y;
                // the per-axis plane maps (node_margin), shared by the four children
                const float ax = sx * r.ix, ay = sy * r.iy, az = sz * r.iz;
                const float nbx = __builtin_fmaf(ox, r.ix, r.cnx), fbx = __builtin_fmaf(ox, r.ix, r.cfx);
                const float nby = __builtin_fmaf(oy, r.iy, r.cny), fby = __builtin_fmaf(oy, r.iy, r.cfy);
                const float nbz = __builtin_fmaf(oz, r.iz, r.cnz), fbz = __builtin_fmaf(oz, r.iz, r.cfz);
                p0 = qslab_c(r, ch.x, nx, fx, ny, fy, nz, fz, ax, ay, az, nbx, fbx, nby, fby, nbz, fbz, 0, R1, l0);
                p1 = qslab_c(r, ch.y, nx, fx, ny, fy, nz, fz, ax, ay, az, nbx, fbx, nby, fby, nbz, fbz, 8, R1, l1);
                p2 = qslab_c(r, ch.z, nx, fx, ny, fy, nz, fz, ax, ay, az, nbx, fbx, nby, fby, nbz, fbz, 16, R1, l2);
                p3 = qslab_c(r, ch.w, nx, fx, ny, fy, nz, fz, ax, ay, az, nbx, fbx, nby, fby, nbz, fbz, 24, R1, l3);
            }
            else
            {
                // an infinite inv_dir component: descend every child and let
                // the exact reference tests decide (results are unchanged)
                p0 = ch.x != MCRT_CHILD_EMPTY;
                p1 = ch.y != MCRT_CHILD_EMPTY;
                p2 = ch.z != MCRT_CHILD_EMPTY;
                p3 = ch.w != MCRT_CHILD_EMPTY;
                l0 = l1 = l2 = l3 = -__builtin_inff();
            }
            // Each slot's position in this octant's visit order; the passing
            // children as a mask over positions.  The nearest is visited
            // next; the others go on the stack farthest first, i.e. a child
            // lands above as many entries as passing children follow it.
            // (Any-hit's boolean does not depend on the order, but this
            // near-first order finds occluders sooner than slot order.)
            const uint32_t rk = (r.oct_hi ? qb.w : qb.z) >> r.oct_shift;
            const uint32_t r0 = rk & 3u, r1 = (rk >> 2) & 3u, r2 = (rk >> 4) & 3u, r3 = (rk >> 6) & 3u;
            const uint32_t m = (p0 ? 1u << r0 : 0u) | (p1 ? 1u << r1 : 0u) | (p2 ? 1u << r2 : 0u) |
                               (p3 ? 1u << r3 : 0u);
            if (m != 0)
            {
                const uint32_t first = (uint32_t)__builtin_ctz(m);
                const int c = __builtin_popcount(m);
                const bool q0 = p0 && r0 != first, q1 = p1 && r1 != first;
                const bool q2 = p2 && r2 != first, q3 = p3 && r3 != first;
                // the nearest passing child is the passing one not pushed
                cur = (p0 && !q0) ? ch.x : (p1 && !q1) ? ch.y : (p2 && !q2) ? ch.z : ch.w;
                const uint32_t m1 = m >> 1;
                const int i0 = sp + __builtin_popcount(m1 >> r0), i1 = sp + __builtin_popcount(m1 >> r1);
                const int i2 = sp + __builtin_popcount(m1 >> r2), i3 = sp + __builtin_popcount(m1 >> r3);
                if (sp + c - 1 <= st.depth)
                {
                    // every push lands in LDS: four unconditional writes, those
                    // of slots not pushed into the discard row
                    const uint32_t dis = st.off(st.depth);
                    if (ANY)
                    {
                        st.put_node(q0 ? st.off(i0) : dis, ch.x);
                        st.put_node(q1 ? st.off(i1) : dis, ch.y);
                        st.put_node(q2 ? st.off(i2) : dis, ch.z);
                        st.put_node(q3 ? st.off(i3) : dis, ch.w);
                    }
                    else
                    {
                        st.put(q0 ? st.off(i0) : dis, ch.x, l0);
                        st.put(q1 ? st.off(i1) : dis, ch.y, l1);
                        st.put(q2 ? st.off(i2) : dis, ch.z, l2);
                        st.put(q3 ? st.off(i3) : dis, ch.w, l3);
                    }
                }
                else
                {
                    if (q0) st.push(i0, ch.x, l0);
                    if (q1) st.push(i1, ch.y, l1);
                    if (q2) st.push(i2, ch.z, l2);
                    if (q3) st.push(i3, ch.w, l3);
                }
                sp += c - 1;
                return false;
            }
        }
        else
        {
            // One leaf reference per step, fetched whole (clipped box and
            // triangle, 64 B, one memory round trip): a wavefront whose lanes
            // sit in different leaves never serialises several dependent
            // fetches in one step.  The leaf's remaining references stay in
            // `cur` and are tested in slot order by the following steps.
            const float4 a = f4_of(w0), b = f4_of(w1), c = f4_of(w2), d = f4_of(w3);
            if (COUNT) cnt.refs++;
            float lo;
            if (slab(r, a.x, a.w, a.y, b.x, a.z, b.y, R1, lo))
            {
                float t, u, v;
                if (COUNT) cnt.tris++;
                if (mt(r, b.z, b.w, c.x, c.y, c.z, c.w, d.x, d.y, d.z, eps, t, u, v) && t > r.R0 && t < R1)
                {
                    R1 = t;
                    bu = u;
                    bv = v;
                    prim = (int)bits_of(d.w);
                    if (ANY) return true;
                }
            }
            if (cur >= (MCRT_CHILD_LEAF | (1u << 29)))
            {
                // more references follow: one fewer, the next one
                cur = cur - (1u << 29) + 1u;
                return false;
            }
        }
        if (ANY)
        {
            // tmax has not changed since the entry passed its test at push
            // (it only shrinks at the hit that ends an any-hit traversal),
            // so the entry is taken without the re-test (its LDS slot holds
            // no distance)
            if (sp == 0) return true;
            --sp;
            float lo;
            if (sp < st.depth)
                cur = st.node[sp * st.stride + st.lane];
            else
                st.pop(sp, cur, lo);
            return false;
        }
        while (sp > 0)
        {
            --sp;
            uint32_t n;
            float lo;
            st.pop(sp, n, lo);
            if (!(lo > R1))
            {
                cur = n;
                return false;
            }
        }
        return true;
    }

    // Primitive id (-1 on a miss) and t, u, v (inf, 0, 0 on a miss).
    MCRT_HD int result(float& t_out, float& u_out, float& v_out) const
    {
        if (prim < 0)
        {
            t_out = __builtin_inff();
            u_out = v_out = 0.0f;
        }
        else
        {
            t_out = R1;
            u_out = bu;
            v_out = bv;
        }
        return prim;
    }
};

template <bool ANY, bool COUNT>
MCRT_HD int trace(const DevScene& S, TraceStack& st, const RayPrep& r, float tmax,
                  float& t_out, float& u_out, float& v_out, TraceCount& cnt)
{
    Tracer<ANY, COUNT> tr;
    RayPrep rm = r;
    node_margin(rm, S.node_span);
    tr.start(rm, tmax);
    if (S.nodes != nullptr)
        while (!tr.step(S, st, cnt)) {}
    return tr.result(t_out, u_out, v_out);
}
}  // namespace mcrt
