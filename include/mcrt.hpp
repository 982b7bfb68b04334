/*
 * mcrt.hpp -- C++ host facade over the C ABI (mcrt.h), mirroring the
 * reference's scene / film interface so that code written against it
 * changes little when it moves to the GPU path:
 *   mcrt::Scene        Scene (src/core/scene.h:28-103): add, loadModel,
 *                      setEnvmap, setCamera, finalize, intersect, occluded,
 *                      numTrangles,
 *                      plus render() for the demo_pt render loop
 *                      (src/examples/pt/demo_pt.cpp:657-720)
 *   mcrt::FrameBuffer  FrameBuffer (src/core/film.h:6-31): resize, width,
 *                      height, pixel, scale, tonemapGamma, dumpHDR, dumpPPM
 *   mcrt::Material     Material(color, glossiness, type) (material.h); ids
 *                      come from Scene::addMaterial (the reference keeps one
 *                      process-global library, todo_addMaterial)
 * Header-only; link with libmcrt_hip.so.  Failures throw mcrt::Error with
 * the library's status and message (the reference prints and continues or
 * exits; the facade never exits).
 */
#ifndef MCRT_HPP
#define MCRT_HPP

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mcrt.h"

namespace mcrt
{
class Error : public std::runtime_error
{
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc)
{
    if (rc != MCRT_OK) throw Error(rc, mcrt_last_error());
}

struct Vec3
{
    float x = 0.0f, y = 0.0f, z = 0.0f;
    Vec3() = default;
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    Vec3 operator*(float s) const { return Vec3(x * s, y * s, z * s); }
    Vec3 operator+(const Vec3& o) const { return Vec3(x + o.x, y + o.y, z + o.z); }
};

/* Ray (src/core/ray.h:29-37): origin, normalised direction, [tmin, tmax]. */
struct Ray
{
    Vec3 orig, dir;
    float tmin = 1e-4f;  // NUM_EPS_RAY
    float tmax = std::numeric_limits<float>::infinity();
    Ray() = default;
    Ray(const Vec3& o, const Vec3& d) : orig(o), dir(d) {}
};

/* HitInfo (src/core/intersection.h): primitive -1 and t = inf on a miss. */
struct HitInfo
{
    int32_t prim = -1;
    float t = std::numeric_limits<float>::infinity(), u = 0.0f, v = 0.0f;
    explicit operator bool() const { return prim >= 0; }
    int32_t primitiveID() const { return prim; }
};

enum MaterialType
{
    LGHT = MCRT_LGHT,
    DIFF = MCRT_DIFF,
    GLSY = MCRT_GLSY
};

struct Material
{
    Vec3 color;
    float glossiness = 25.0f;
    MaterialType type = DIFF;
    Material() = default;
    Material(const Vec3& c, float g, MaterialType t) : color(c), glossiness(g), type(t) {}
};

struct RenderParams
{
    int spp = 16;
    int max_depth = 8;
    int rr_start = 3;
    float rr_prob = 0.8f;
    uint64_t seed = 1;
    int64_t sample_offset = 0;
    bool nee = true;
};

class FrameBuffer
{
public:
    FrameBuffer() = default;
    FrameBuffer(int w, int h) { resize(w, h); }
    void resize(int w, int h)
    {
        w_ = w;
        h_ = h;
        buf_.assign((size_t)w * h * 3, 0.0f);
    }
    int width() const { return w_; }
    int height() const { return h_; }
    Vec3 pixel(int i, int j) const
    {
        const float* p = &buf_[3 * ((size_t)j * w_ + i)];
        return Vec3(p[0], p[1], p[2]);
    }
    float* data() { return buf_.data(); }
    const float* data() const { return buf_.data(); }
    void scale(float s)
    {
        for (float& v : buf_) v *= s;
    }
    void tonemapGamma(float gamma) { check(mcrt_film_tonemap_gamma(buf_.data(), buf_.size(), gamma)); }
    void dumpHDR(const std::string& path) const { check(mcrt_film_write_hdr(path.c_str(), buf_.data(), w_, h_)); }
    void dumpPPM(const std::string& path) const { check(mcrt_film_write_ppm(path.c_str(), buf_.data(), w_, h_)); }

private:
    int w_ = 0, h_ = 0;
    std::vector<float> buf_;
};

class Context
{
public:
    explicit Context(int device = 0) { check(mcrt_create(device, &h_)); }
    ~Context()
    {
        if (h_) mcrt_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    mcrt_ctx* handle() const { return h_; }
    std::string deviceName() const
    {
        char buf[256];
        check(mcrt_device_name(h_, buf, sizeof(buf)));
        return buf;
    }

private:
    mcrt_ctx* h_ = nullptr;
};

class Scene
{
public:
    explicit Scene(Context& ctx) { check(mcrt_scene_create(ctx.handle(), &h_)); }
    ~Scene()
    {
        if (h_) mcrt_scene_destroy(h_);
    }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    mcrt_scene* handle() const { return h_; }

    int addMaterial(const Material& m)
    {
        const float rgb[3] = {m.color.x, m.color.y, m.color.z};
        int32_t id = -1;
        check(mcrt_material_add(h_, m.type, rgb, m.glossiness, &id));
        return id;
    }
    /* Scene::add(TriangleObject(a, b, c)) with the face normal at every corner. */
    void add(const Vec3& a, const Vec3& b, const Vec3& c, int material)
    {
        const float v[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
        const int32_t m = material;
        check(mcrt_scene_add_triangles(h_, v, nullptr, &m, 1));
    }
    /* n triangles (9 floats each), optional corner normals, material ids. */
    void add(const float* v9, const float* vn9, const int32_t* mat, size_t n)
    {
        check(mcrt_scene_add_triangles(h_, v9, vn9, mat, n));
    }
    /* Scene::loadModel(filename, transform, material); m16 column-major or
     * NULL for the identity.  Returns the number of faces added. */
    size_t loadModel(const std::string& path, const float* m16, int material)
    {
        size_t n = 0;
        check(mcrt_scene_load_obj(h_, path.c_str(), m16, material, &n));
        return n;
    }
    /* Scene::setEnvmap(filename) (file with its size header) and
     * setEnvmap(filename, width, height) (floats only), scene.h:44-45. */
    void setEnvmap(const std::string& path) { check(mcrt_scene_set_envmap(h_, path.c_str(), 0, 0)); }
    void setEnvmap(const std::string& path, int width, int height)
    {
        check(mcrt_scene_set_envmap(h_, path.c_str(), width, height));
    }
    /* From memory: width * height * 3 floats, rows top-down; NULL removes it. */
    void setEnvmap(const float* rgb, int width, int height)
    {
        check(mcrt_scene_set_envmap_rgb(h_, rgb, width, height));
    }
    void setCamera(const Vec3& o, const Vec3& target, int width, int height, float fov_y)
    {
        const float a[3] = {o.x, o.y, o.z}, b[3] = {target.x, target.y, target.z};
        check(mcrt_scene_set_camera(h_, a, b, width, height, fov_y));
        cam_w_ = width;
        cam_h_ = height;
    }
    mcrt_build_stats finalize(int threads = 0)
    {
        mcrt_build_opts o{threads};
        mcrt_build_stats st{};
        check(mcrt_scene_finalize(h_, &o, &st));
        return st;
    }
    size_t numTrangles() const
    {
        size_t n = 0;
        check(mcrt_scene_num_triangles(h_, &n));
        return n;
    }
    /* Scene::intersect (one ray; a compatibility path -- batch with the
     * overload below). */
    HitInfo intersect(const Ray& r) const
    {
        const float q[8] = {r.orig.x, r.orig.y, r.orig.z, r.tmin, r.dir.x, r.dir.y, r.dir.z, r.tmax};
        mcrt_hit h;
        check(mcrt_intersect(h_, q, 1, &h, 0, nullptr));
        HitInfo out;
        out.prim = h.prim;
        out.t = h.t;
        out.u = h.u;
        out.v = h.v;
        return out;
    }
    /* Scene::occluded(ray, tfar): any hit in (tmin, tfar - 1e-4). */
    bool occluded(const Ray& r, float tfar = std::numeric_limits<float>::infinity()) const
    {
        const float q[8] = {r.orig.x, r.orig.y, r.orig.z, r.tmin, r.dir.x, r.dir.y, r.dir.z, tfar - 1e-4f};
        uint8_t o = 0;
        check(mcrt_occluded(h_, q, 1, &o, 0, nullptr));
        return o != 0;
    }
    /* Batched queries over n rays of 8 floats (host or, with
     * MCRT_DEVICE_PTRS, device buffers). */
    void intersect(const float* rays, size_t n, mcrt_hit* hits, int flags = 0, void* stream = nullptr) const
    {
        check(mcrt_intersect(h_, rays, n, hits, flags, stream));
    }
    void occluded(const float* rays, size_t n, uint8_t* out, int flags = 0, void* stream = nullptr) const
    {
        check(mcrt_occluded(h_, rays, n, out, flags, stream));
    }
    /* The demo_pt render loop: film receives the per-pixel mean of p.spp
     * samples (resized to the camera's W x H). */
    void render(const RenderParams& p, FrameBuffer& film) const
    {
        if (film.width() != cam_w_ || film.height() != cam_h_) film.resize(cam_w_, cam_h_);
        mcrt_render_params q{};
        q.spp = p.spp;
        q.max_depth = p.max_depth;
        q.rr_start = p.rr_start;
        q.rr_prob = p.rr_prob;
        q.seed = p.seed;
        q.sample_offset = p.sample_offset;
        q.nee = p.nee ? 1 : 0;
        check(mcrt_render(h_, &q, film.data(), 0, nullptr));
    }

private:
    mcrt_scene* h_ = nullptr;
    int cam_w_ = 0, cam_h_ = 0;
};
}  // namespace mcrt

#endif /* MCRT_HPP */
