// Film output formats of the reference's FrameBuffer (film.cpp), for films
// in host memory (W*H*3 floats, row y = 0 at the image bottom):
//   * Radiance .hdr (dumpHDR, film.cpp:65-108): RGBE pixels (toRGBE,
//     film.cpp:52-63), rows from y = H-1 down to 0, each row as the new-style
//     scanline header {2, 2, W >> 8, W & 255} followed by the four byte
//     planes r, g, b, e, each cut into literal runs of at most 127 bytes
//     (count byte then the bytes; no repeat runs);
//   * binary PPM (dumpPPM, film.cpp:9-33): "P6\nW H\n255\n", rows from the
//     top, byte = (uint8)(min(1, v) * 255);
//   * tonemapGamma (film.cpp:195-202): v = powf(clamp(v, 0, 1), 1/gamma).
// Byte-for-byte identical to the reference on finite non-negative films; a
// NaN or negative channel (which the reference converts with undefined
// behaviour) is written as 0.
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "capi_internal.h"

namespace
{
// f_max / f_min / clampf of numeric.cpp:22-45
inline float ref_max(float a, float b) { return a > b ? a : b; }
inline float ref_min(float a, float b) { return a < b ? a : b; }
inline float ref_clamp(float x, float a, float b) { return x < a ? a : x > b ? b : x; }

// float -> uint8 as the reference's static_cast for values in [0, 256);
// NaN and negatives -> 0.
inline uint8_t to_u8(float x)
{
    if (!(x >= 0.0f)) return 0;
    if (x >= 255.0f) return 255;
    return static_cast<uint8_t>(x);
}

void rgbe(float r, float g, float b, uint8_t out[4])
{
    float d = ref_max(ref_max(r, g), b);
    if ((double)d <= 1e-32 || !(d == d))
    {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e = 0;
    const float m = std::frexp(d, &e);  // d = m * 2^e, m in [0.5, 1)
    const float s = m * 256.0f / d;
    out[0] = to_u8(r * s);
    out[1] = to_u8(g * s);
    out[2] = to_u8(b * s);
    out[3] = static_cast<uint8_t>(e + 128);
}

struct File
{
    FILE* f;
    explicit File(const char* path) : f(std::fopen(path, "wb")) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int check_film(const char* who, const char* path, const float* film, int32_t w, int32_t h)
{
    if (!path || !film) return mcrt::fail(MCRT_ERR_INVALID_ARG, "%s: NULL argument", who);
    if (w <= 0 || h <= 0) return mcrt::fail(MCRT_ERR_INVALID_ARG, "%s: width and height must be > 0", who);
    if (w > 65535) return mcrt::fail(MCRT_ERR_INVALID_ARG, "%s: width must be < 65536", who);
    return MCRT_OK;
}
}  // namespace

extern "C" {

MCRT_API int mcrt_film_write_hdr(const char* path, const float* film, int32_t w, int32_t h)
{
    int rc = check_film("mcrt_film_write_hdr", path, film, w, h);
    if (rc) return rc;
    File out(path);
    if (!out.f) return mcrt::fail(MCRT_ERR_IO, "mcrt_film_write_hdr: cannot write to %s", path);
    std::string head = "#?RADIANCE\n# Made with custom writer\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y " +
                       std::to_string(h) + " +X " + std::to_string(w) + "\n";
    std::vector<uint8_t> buf(head.begin(), head.end());
    std::vector<uint8_t> px(4 * (size_t)w);
    for (int32_t j = h - 1; j >= 0; --j)
    {
        for (int32_t i = 0; i < w; i++)
        {
            const float* p = film + 3 * ((size_t)j * w + i);
            rgbe(p[0], p[1], p[2], &px[4 * (size_t)i]);
        }
        buf.push_back(2);
        buf.push_back(2);
        buf.push_back((uint8_t)((w >> 8) & 0xff));
        buf.push_back((uint8_t)(w & 0xff));
        for (int k = 0; k < 4; k++)
            for (int32_t c = 0; c < w;)
            {
                const int32_t run = std::min<int32_t>(127, w - c);
                buf.push_back((uint8_t)run);
                for (int32_t i = c; i < c + run; i++) buf.push_back(px[4 * (size_t)i + k]);
                c += run;
            }
    }
    if (std::fwrite(buf.data(), 1, buf.size(), out.f) != buf.size())
        return mcrt::fail(MCRT_ERR_IO, "mcrt_film_write_hdr: short write to %s", path);
    return MCRT_OK;
}

MCRT_API int mcrt_film_write_ppm(const char* path, const float* film, int32_t w, int32_t h)
{
    int rc = check_film("mcrt_film_write_ppm", path, film, w, h);
    if (rc) return rc;
    File out(path);
    if (!out.f) return mcrt::fail(MCRT_ERR_IO, "mcrt_film_write_ppm: cannot write to %s", path);
    std::string head = "P6\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
    std::vector<uint8_t> buf(head.begin(), head.end());
    buf.reserve(buf.size() + 3 * (size_t)w * h);
    for (int32_t j = h - 1; j >= 0; --j)
        for (int32_t i = 0; i < w; i++)
        {
            const float* p = film + 3 * ((size_t)j * w + i);
            for (int c = 0; c < 3; c++) buf.push_back(to_u8(ref_min(1.0f, p[c]) * 255));
        }
    if (std::fwrite(buf.data(), 1, buf.size(), out.f) != buf.size())
        return mcrt::fail(MCRT_ERR_IO, "mcrt_film_write_ppm: short write to %s", path);
    return MCRT_OK;
}

MCRT_API int mcrt_film_tonemap_gamma(float* film, size_t n, float gamma)
{
    if (!film && n) return mcrt::fail(MCRT_ERR_INVALID_ARG, "mcrt_film_tonemap_gamma: NULL film");
    if (!(gamma > 0.0f)) return mcrt::fail(MCRT_ERR_INVALID_ARG, "mcrt_film_tonemap_gamma: gamma must be > 0");
    const float inv = 1.0f / gamma;
    for (size_t i = 0; i < n; i++) film[i] = powf(ref_clamp(film[i], 0.0f, 1.0f), inv);
    return MCRT_OK;
}

}  // extern "C"
