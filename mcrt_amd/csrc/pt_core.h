// Path-tracer device functions: counter-based RNG, the project's own fp32
// sin/cos/log/exp, vec3 helpers, the orthonormal frame, BSDF sampling and
// area-light sampling.  Every expression keeps the reference's operation
// order (vectors.cpp, geometry.cpp, bsdf.cpp, scene.cpp, triangle.cpp) and
// is compiled without contraction, so results are reproducible bit for bit.
#pragma once
#include "trace_core.h"

namespace mcrt
{
// ------------------------------------------------------------ RNG -------
// splitmix64 finaliser (Steele, Lea, Flood 2014) as a counter hash; a key
// per (seed, pixel, sample, bounce) and one hash per dimension.  The camera
// uses bounce 255.
MCRT_HD uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
MCRT_HD uint64_t rng_key(uint64_t seed, uint64_t pixel, uint64_t sample, uint32_t bounce)
{
    return mix64(mix64(mix64(seed) ^ pixel) ^ ((sample << 8) | (uint64_t)bounce));
}
MCRT_HD float rng_u01(uint64_t key, uint32_t dim)
{
    uint64_t h = mix64(key + (uint64_t)(dim + 1) * 0x9e3779b97f4a7c15ull);
    return (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------- own fp32 math -----
// sin/cos of x in [0, 2pi]: Cody-Waite reduction by pi/4 and the Cephes
// single-precision minimax polynomials (S. L. Moshier, Cephes sinf/cosf).
MCRT_HD void own_sincos(float x, float* s, float* c)
{
    int j = (int)(x * 1.27323954473516f);
    if (j & 1) j += 1;
    float y = (float)j;
    float r = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    float z = r * r;
    float ps = ((((-1.9515295891e-4f * z) + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z) * r + r;
    float pc = ((((((2.443315711809948e-5f * z) - 1.388731625493765e-3f) * z) + 4.166664568298827e-2f) * z) * z -
                0.5f * z) + 1.0f;
    switch ((j >> 1) & 3)
    {
        case 0: *s = ps; *c = pc; break;
        case 1: *s = pc; *c = -ps; break;
        case 2: *s = -ps; *c = -pc; break;
        default: *s = -pc; *c = ps; break;
    }
}
// ln(x), x > 0 (Cephes logf) and e^x (Cephes expf).
MCRT_HD float own_log(float x)
{
    int e;
    x = frexpf(x, &e);
    if (x < 0.707106781186547524f) { e -= 1; x = (x + x) - 1.0f; }
    else x = x - 1.0f;
    float z = x * x;
    float y = ((((((((7.0376836292e-2f * x - 1.1514610310e-1f) * x + 1.1676998740e-1f) * x -
                    1.2420140846e-1f) * x + 1.4249322787e-1f) * x - 1.6668057665e-1f) * x +
                 2.0000714765e-1f) * x - 2.4999993993e-1f) * x + 3.3333331174e-1f) * x * z;
    float fe = (float)e;
    y = y + (-2.12194440e-4f * fe);
    y = y + (-0.5f * z);
    z = x + y;
    z = z + 0.693359375f * fe;
    return z;
}
MCRT_HD float own_exp(float x)
{
    if (x > 88.0f) return __builtin_inff();
    if (x < -87.0f) return 0.0f;
    float z = floorf(1.44269504088896341f * x + 0.5f);
    x = x - z * 0.693359375f;
    x = x - z * -2.12194440e-4f;
    int n = (int)z;
    float zz = x * x;
    float p = (((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x + 4.1665795894e-2f) * x +
                1.6666665459e-1f) * x + 5.0000001201e-1f) * zz + x + 1.0f;
    return ldexpf(p, n);
}
MCRT_HD float own_pow(float x, float y)
{
    if (!(x > 0.0f)) return 0.0f;
    return own_exp(y * own_log(x));
}

// ------------------------------------------------------------ vec3 ------
struct V3f
{
    float x, y, z;
};
MCRT_HD V3f v3f(float x, float y, float z) { return V3f{x, y, z}; }
MCRT_HD V3f operator+(V3f a, V3f b) { return V3f{a.x + b.x, a.y + b.y, a.z + b.z}; }
MCRT_HD V3f operator-(V3f a, V3f b) { return V3f{a.x - b.x, a.y - b.y, a.z - b.z}; }
MCRT_HD V3f operator-(V3f a) { return V3f{-a.x, -a.y, -a.z}; }
MCRT_HD V3f operator*(V3f a, float f) { return V3f{a.x * f, a.y * f, a.z * f}; }
MCRT_HD V3f operator*(V3f a, V3f b) { return V3f{a.x * b.x, a.y * b.y, a.z * b.z}; }
MCRT_HD float dot(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MCRT_HD float len2(V3f a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
MCRT_HD V3f nrm(V3f a)  // vectors.cpp:227-231
{
    float inv = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return V3f{a.x * inv, a.y * inv, a.z * inv};
}
MCRT_HD float fmax_ref(float a, float b) { return a > b ? a : b; }   // f_max (numeric.cpp:27-30)
MCRT_HD float stdmax(float a, float b) { return (a < b) ? b : a; }   // std::max
MCRT_HD V3f ld3(const float* p) { return V3f{p[0], p[1], p[2]}; }

constexpr float kPi = 3.14159265358979323846f;  // NUM_PI (constants.cpp:11)
constexpr float kTwoPi = 2.0f * kPi;            // `2 * NUM_PI` in sample.cpp:134
constexpr float kEpsCos = 1e-6f;                // NUM_EPS_COSINE
constexpr float kEpsRay = 1e-4f;                // NUM_EPS_RAY

// OrthonormalFrame::setFromNormal + toWorld (geometry.cpp:74-105;
// Duff et al., "Building an Orthonormal Basis, Revisited", JCGT 2017).
MCRT_HD V3f to_world(V3f n, V3f c)
{
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    V3f mx = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    V3f my = {b, sign + n.y * n.y * a, -n.y};
    return (mx * c.x + my * c.y) + n * c.z;
}

// TriangleObject::normal(u, v) (triangle.cpp:90-93).
MCRT_HD V3f shading_normal(const TriShade& t, float u, float v)
{
    float w = (1.0f - u) - v;
    return nrm((ld3(t.na) * w + ld3(t.nb) * u) + ld3(t.nc) * v);
}

// LambertianBSDF::sample direction (bsdf.cpp:51-81; sampleHemisphereCosine,
// sample.cpp:132-137).
MCRT_HD V3f lambert_dir(V3f n, float r1, float r2)
{
    float x = r1 * kTwoPi;
    float r2s = sqrtf(r2);
    float sx, cx;
    own_sincos(x, &sx, &cx);
    V3f c = {cx * r2s, sx * r2s, sqrtf(1.0f - r2)};
    return to_world(n, c);
}

// GlossyBSDF::sample (bsdf.cpp:209-263; sampleHemisphereCosinePower,
// sample.cpp:139-145; reflectWorld, geometry.cpp:45-55).
MCRT_HD V3f glossy_dir(V3f wi, V3f n, float g, float r1, float r2, float& weight)
{
    float x = r1 * kTwoPi;
    float r2p = own_pow(r2, 1.0f / (g + 1.0f));
    float sx, cx;
    own_sincos(x, &sx, &cx);
    float t = sqrtf(1.0f - r2p * r2p);
    V3f c = {sx * t, cx * t, r2p};
    V3f refl = n * (2.0f * dot(n, wi)) - wi;
    V3f wo = to_world(refl, c);
    if (dot(wo, n) < 0) wo = v3f(0.0f, 0.0f, 0.0f);
    float cs = fmax_ref(dot(wo, n), kEpsCos);
    float wo_n = dot(wo, n), wi_n = dot(wi, n);
    float wgt = (g + 2.0f) / (g + 1.0f) * cs / fmax_ref(fabsf(wi_n), fabsf(wo_n));
    if (wgt != wgt) wgt = 0;
    weight = wgt;
    return wo;
}

// DiscreteSampler::sample(float) (sample.cpp:36-52): lower bound, c >= r.
MCRT_HD int sample_light(const DevLight* L, int n, float r)
{
    int b = 0, e = n - 1;
    while (e > b)
    {
        int mid = b + (e - b) / 2;
        if (L[mid].cdf >= r) e = mid;
        else b = mid + 1;
    }
    return b;
}
}  // namespace mcrt
