// Host-side fp32 geometry for the SBVH builder.  Arithmetic follows the
// reference operation for operation (left-to-right vec3 ops, no FMA: the
// host sources are compiled with -ffp-contract=off and no -march), because
// the builder must reproduce SplitBVH::build's tree exactly.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace mcrt
{
static const float kInf = std::numeric_limits<float>::infinity();

struct F3
{
    float v[3];
    float operator[](int i) const { return v[i]; }
    float& operator[](int i) { return v[i]; }
};
static inline F3 f3(float x, float y, float z) { return F3{{x, y, z}}; }
static inline F3 sub(const F3& a, const F3& b) { return f3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
static inline F3 add(const F3& a, const F3& b) { return f3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
static inline F3 mul(const F3& a, float s) { return f3(a[0] * s, a[1] * s, a[2] * s); }
static inline float dot3(const F3& a, const F3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline F3 cross3(const F3& a, const F3& b)  // vectors.cpp:180-185
{
    return f3((a[1] * b[2]) - (a[2] * b[1]), (a[2] * b[0]) - (a[0] * b[2]), (a[0] * b[1]) - (a[1] * b[0]));
}
static inline F3 normalize3(const F3& a)  // vectors.cpp:227-231
{
    float inv = 1.0f / std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    return f3(a[0] * inv, a[1] * inv, a[2] * inv);
}
static inline float mn_(float a, float b) { return a < b ? a : b; }  // f_min, numeric.cpp:22-25
static inline float mx_(float a, float b) { return a > b ? a : b; }  // f_max, numeric.cpp:27-30

// AABB (aabb.cpp:260-531): empty box is (+inf, -inf).
struct Box
{
    F3 lo{{kInf, kInf, kInf}};
    F3 hi{{-kInf, -kInf, -kInf}};
    void grow(const Box& b)
    {
        for (int k = 0; k < 3; k++) { lo[k] = mn_(lo[k], b.lo[k]); hi[k] = mx_(hi[k], b.hi[k]); }
    }
    void grow(const F3& p)
    {
        for (int k = 0; k < 3; k++) { lo[k] = mn_(lo[k], p[k]); hi[k] = mx_(hi[k], p[k]); }
    }
    bool valid() const { return hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2]; }
    // surfaceAreaCost (aabb.cpp:335-341): half the surface area.
    float half_area() const
    {
        F3 e = sub(hi, lo);
        return e[0] * e[1] + e[1] * e[2] + e[0] * e[2];
    }
    int largest_axis() const  // aabb.cpp:348-371
    {
        float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
        int d = 0;
        float m = e0;
        if (e1 > m) { m = e1; d = 1; }
        if (e2 > m) { d = 2; }
        return d;
    }
    float center(int k) const { return (lo[k] + hi[k]) * 0.5f; }
};

// AABB::intersectBy (aabb.cpp:451-478): empty on the first disjoint axis.
static inline void box_intersect_by(Box& a, const Box& b)
{
    for (int k = 0; k < 3; k++)
    {
        a.lo[k] = mx_(a.lo[k], b.lo[k]);
        a.hi[k] = mn_(a.hi[k], b.hi[k]);
        if (a.lo[k] > a.hi[k]) { a = Box(); return; }
    }
}

static inline float clamp_(float x, float a, float b) { return x < a ? a : x > b ? b : x; }

struct Tri
{
    F3 a, b, c;
};

// Triangle bounding box (triangle.cpp:256-261).
static inline Box tri_box(const Tri& t)
{
    Box r;
    for (int k = 0; k < 3; k++) { r.lo[k] = mn_(t.a[k], t.b[k]); r.hi[k] = mx_(t.a[k], t.b[k]); }
    r.grow(t.c);
    return r;
}

// BoundedTriangle::clipBoundTight (boundedtriangle.cpp:3-124): bound of the
// triangle part inside the [lo, hi] slab of `axis` (edge/plane crossings
// plus vertices inside, inclusive), intersected with the clipped bound.
static inline void edge_cross(Box& tight, const F3& p, const F3& q, float lo, float hi, int axis)
{
    F3 d = sub(q, p);
    float t0 = (lo - p[axis]) / d[axis];
    float t1 = (hi - p[axis]) / d[axis];
    if (t0 >= 0.0f && t0 <= 1.0f) tight.grow(add(p, mul(d, t0)));
    if (t1 >= 0.0f && t1 <= 1.0f) tight.grow(add(p, mul(d, t1)));
}

static inline Box clip_tight(const Tri& tri, const Box& clipped, float lo, float hi, int axis)
{
    if (clipped.lo[axis] > hi || clipped.hi[axis] < lo) return Box();
    Box tight;
    edge_cross(tight, tri.a, tri.b, lo, hi, axis);
    edge_cross(tight, tri.b, tri.c, lo, hi, axis);
    edge_cross(tight, tri.c, tri.a, lo, hi, axis);
    if (tri.a[axis] >= lo && tri.a[axis] <= hi) tight.grow(tri.a);
    if (tri.b[axis] >= lo && tri.b[axis] <= hi) tight.grow(tri.b);
    if (tri.c[axis] >= lo && tri.c[axis] <= hi) tight.grow(tri.c);
    Box more = clipped;  // AABB::clipAxis (aabb.cpp:523-529)
    more.lo[axis] = clamp_(more.lo[axis], lo, hi);
    more.hi[axis] = clamp_(more.hi[axis], lo, hi);
    box_intersect_by(tight, more);
    if (tight.valid())
    {
        Box r = more;  // more.intersect(tight)
        box_intersect_by(r, tight);
        return r;
    }
    if (more.valid()) return more;
    return Box();
}
}  // namespace mcrt
