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
// atan (Cephes atanf: reduction at tan(3pi/8) and tan(pi/8), odd minimax
// polynomial), asin (Cephes asinf: half-angle reduction above 1/2) and acos
// from asin.
MCRT_HD float own_atan(float x)
{
    const float sign = x < 0.0f ? -1.0f : 1.0f;
    if (x < 0.0f) x = -x;
    float y = 0.0f;
    if (x > 2.414213562373095f)
    {
        y = 1.5707963267948966f;
        x = -(1.0f / x);
    }
    else if (x > 0.4142135623730950f)
    {
        y = 0.7853981633974483f;
        x = (x - 1.0f) / (x + 1.0f);
    }
    const float z = x * x;
    y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x +
             x);
    return sign * y;
}
MCRT_HD float own_asin(float xx)
{
    const float a = fabsf(xx);
    if (a > 1.0f) return __builtin_nanf("");
    if (a < 1.0e-4f) return xx;
    float x, z;
    const bool half = a > 0.5f;
    if (half)
    {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
    }
    else
    {
        x = a;
        z = x * x;
    }
    z = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z +
         1.6666752422e-1f) * z * x + x;
    if (half)
    {
        z = z + z;
        z = 1.5707963267948966f - z;
    }
    return xx < 0.0f ? -z : z;
}
MCRT_HD float own_acos(float x)
{
    if (x < -0.5f) return 3.14159265358979f - 2.0f * own_asin(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * own_asin(sqrtf(0.5f * (1.0f - x)));
    return 1.5707963267948966f - own_asin(x);
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
MCRT_HD V3f cross(V3f a, V3f b)  // vectors.cpp:180-185
{
    return V3f{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
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
MCRT_HD int sample_cdf(const float* cdf, int n, float r)
{
    int b = 0, e = n - 1;
    while (e > b)
    {
        int mid = b + (e - b) / 2;
        if (cdf[mid] >= r) e = mid;
        else b = mid + 1;
    }
    return b;
}

// ------------------------------------------------ environment map -------
// Latitude-longitude map (EnvmapLoader, envmap.cpp): row v <-> polar angle
// phi = pi v from +y, column u <-> azimuth theta = 2 pi u, direction
// (sin phi sin theta, cos phi, -sin phi cos theta).

// static_cast<int>(x) clamped to [0, hi] (clampi, envmap.cpp:27-30); NaN and
// negatives give 0.
MCRT_HD int env_trunc(double x, int hi)
{
    if (!(x >= 0.0)) return 0;
    if (x >= (double)hi) return hi;
    return (int)x;
}

// Pixel of a direction (getPdf / sampleRadiance, envmap.cpp:440-484, with
// dir_to_theta :44-53): theta = atan(z / x) + pi/2 (+ pi when x < 0), the
// sum formed in double as the reference's double atan does.
MCRT_HD int env_pixel(const DevScene& D, V3f d)
{
    const float phi = own_acos(d.y);
    float theta = (float)((double)own_atan(d.z / d.x) + 1.5707963267948966);
    if (d.x < 0) theta += kPi;
    const int u = env_trunc((double)(theta / (kPi * 2.0f) * (float)D.env_w), D.env_w - 1);
    const int v = env_trunc((double)(phi / kPi * (float)D.env_h), D.env_h - 1);
    return u + v * D.env_w;
}

// Solid-angle pdf of a pixel: norm * lum * W * H / (2 pi^2); the sin(phi)
// of the pre-warp cancels the pixel's solid angle (envmap.cpp:434-441).
MCRT_HD float env_pdf(const DevScene& D, int pix)
{
    return (((D.env_norm * D.env_px[4 * (int64_t)pix + 3]) * (float)D.env_w) * (float)D.env_h) /
           ((2.0f * kPi) * kPi);
}
MCRT_HD V3f env_rad(const DevScene& D, int pix) { return ld3(D.env_px + 4 * (int64_t)pix); }

// EnvmapLoader::sample (envmap.cpp:332-456): column by the column-sum CDF,
// row by that column's CDF, (j1, j2) the jitter inside the pixel.
MCRT_HD void env_sample(const DevScene& D, float r1, float r2, float j1, float j2, V3f& dir, V3f& rad, float& pdf)
{
    const int W = D.env_w, H = D.env_h;
    const int ui = sample_cdf(D.env_cdf_u, W, r1);
    const int vi = sample_cdf(D.env_cdf_v + (int64_t)ui * H, H, r2);
    const float u = ((float)ui + j1) / (float)W;
    const float v = ((float)vi + j2) / (float)H;
    const int pix = env_trunc((double)(u * (float)W), W - 1) + env_trunc((double)(v * (float)H), H - 1) * W;
    const float phi = v * kPi, theta = (u * 2.0f) * kPi;
    float sp, cp, st, ct;
    own_sincos(phi, &sp, &cp);
    own_sincos(theta, &st, &ct);
    dir = v3f(sp * st, cp, sp * -ct);
    pdf = env_pdf(D, pix);
    rad = env_rad(D, pix);
}
}  // namespace mcrt
