// C-ABI implementation (include/mcrt.h): scene ingestion, host SBVH build,
// 4-wide collapse, HBM upload and the query / render entry points.
#include "mcrt.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "capi_internal.h"

using namespace mcrt;

namespace
{
thread_local std::string g_err;
}

int mcrt::fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int mcrt::ok()
{
    return MCRT_OK;
}

mcrt::DevBuf::~DevBuf() { release(); }

void mcrt::DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int mcrt::DevBuf::ensure(size_t want)
{
    if (want <= bytes && p) return MCRT_OK;
    release();
    if (want == 0) return MCRT_OK;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess)
    {
        p = nullptr;
        (void)hipGetLastError();
        return fail(MCRT_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return MCRT_OK;
}

int mcrt::DevBuf::upload(const void* src, size_t n)
{
    int rc = ensure(n);
    if (rc) return rc;
    if (n == 0) return MCRT_OK;
    hipError_t e = hipMemcpy(p, src, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(MCRT_ERR_DEVICE, "hipMemcpy H2D: %s", hipGetErrorString(e));
    return MCRT_OK;
}

mcrt_batch::~mcrt_batch()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap)
    {
        (void)hipStreamSynchronize(cap);
        (void)hipStreamDestroy(cap);
    }
    // the workspace may still be in use by a replay enqueued on a user stream
    (void)hipDeviceSynchronize();
}

extern "C" {

MCRT_API const char* mcrt_last_error(void) { return g_err.c_str(); }
MCRT_API const char* mcrt_version(void) { return "mcrt-mi355x 0.1 (gfx950)"; }

MCRT_API int mcrt_create(int device, mcrt_ctx** out)
{
    if (!out) return fail(MCRT_ERR_INVALID_ARG, "mcrt_create: out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
    {
        (void)hipGetLastError();
        return fail(MCRT_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    }
    if (device < 0 || device >= n) return fail(MCRT_ERR_INVALID_ARG, "device %d out of range [0,%d)", device, n);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    auto* c = new (std::nothrow) mcrt_ctx();
    if (!c) return fail(MCRT_ERR_OUT_OF_MEMORY, "mcrt_create: out of host memory");
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    // the allocation-free query's scratch array on this device (the device
    // is current); a failure here only disables that entry point
    if (scratch_address(device, &c->scratch) != MCRT_OK)
    {
        c->scratch = nullptr;
        c->scratch_err = mcrt_last_error();
    }
    *out = c;
    return MCRT_OK;
}

MCRT_API int mcrt_destroy(mcrt_ctx* ctx)
{
    if (!ctx) return fail(MCRT_ERR_INVALID_ARG, "mcrt_destroy: NULL context");
    (void)hipSetDevice(ctx->device);  // the per-stream workspaces are freed here
    delete ctx;
    return MCRT_OK;
}

MCRT_API int mcrt_device_name(mcrt_ctx* ctx, char* buf, size_t cap)
{
    if (!ctx || !buf || cap == 0) return fail(MCRT_ERR_INVALID_ARG, "mcrt_device_name: bad arguments");
    snprintf(buf, cap, "%s (%d CUs)", ctx->name.c_str(), ctx->num_cus);
    return MCRT_OK;
}

MCRT_API int mcrt_scene_create(mcrt_ctx* ctx, mcrt_scene** out)
{
    if (!ctx || !out) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_create: NULL argument");
    auto* s = new (std::nothrow) mcrt_scene();
    if (!s) return fail(MCRT_ERR_OUT_OF_MEMORY, "mcrt_scene_create: out of host memory");
    s->ctx = ctx;
    *out = s;
    return MCRT_OK;
}

MCRT_API int mcrt_scene_destroy(mcrt_scene* s)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_destroy: NULL scene");
    (void)hipSetDevice(s->ctx->device);
    delete s;
    return MCRT_OK;
}

MCRT_API int mcrt_material_add(mcrt_scene* s, int type, const float rgb[3], float glossiness, int32_t* out_id)
{
    if (!s || !rgb) return fail(MCRT_ERR_INVALID_ARG, "mcrt_material_add: NULL argument");
    if (s->finalized) return fail(MCRT_ERR_STATE, "mcrt_material_add: scene already finalized");
    if (type != MCRT_LGHT && type != MCRT_DIFF && type != MCRT_GLSY)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_material_add: unsupported material type %d", type);
    if (type == MCRT_GLSY && !(glossiness >= 0.0f))
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_material_add: glossiness must be >= 0");
    HostMaterial m;
    m.type = type;
    m.rgb[0] = rgb[0]; m.rgb[1] = rgb[1]; m.rgb[2] = rgb[2];
    m.gloss = glossiness;
    s->mats.push_back(m);
    if (out_id) *out_id = (int32_t)s->mats.size() - 1;
    return MCRT_OK;
}

MCRT_API int mcrt_materials_add(mcrt_scene* s, size_t n, const int32_t* types, const float* rgb,
                                const float* gloss, int32_t* first_id)
{
    if (!s || (n && (!types || !rgb || !gloss))) return fail(MCRT_ERR_INVALID_ARG, "mcrt_materials_add: NULL argument");
    if (s->finalized) return fail(MCRT_ERR_STATE, "mcrt_materials_add: scene already finalized");
    for (size_t i = 0; i < n; i++)
    {
        int t = types[i];
        if (t != MCRT_LGHT && t != MCRT_DIFF && t != MCRT_GLSY)
            return fail(MCRT_ERR_INVALID_ARG, "mcrt_materials_add: unsupported material type %d at %zu", t, i);
        if (t == MCRT_GLSY && !(gloss[i] >= 0.0f))
            return fail(MCRT_ERR_INVALID_ARG, "mcrt_materials_add: glossiness must be >= 0 (material %zu)", i);
    }
    if (first_id) *first_id = (int32_t)s->mats.size();
    for (size_t i = 0; i < n; i++)
    {
        HostMaterial m;
        m.type = types[i];
        m.rgb[0] = rgb[3 * i]; m.rgb[1] = rgb[3 * i + 1]; m.rgb[2] = rgb[3 * i + 2];
        m.gloss = gloss[i];
        s->mats.push_back(m);
    }
    return MCRT_OK;
}

MCRT_API int mcrt_material_set_rgb(mcrt_scene* s, const float* rgb, size_t nmat)
{
    if (!s || (!rgb && nmat)) return fail(MCRT_ERR_INVALID_ARG, "mcrt_material_set_rgb: NULL argument");
    if (nmat != s->mats.size())
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_material_set_rgb: %zu colours for %zu materials", nmat,
                    s->mats.size());
    for (size_t i = 0; i < nmat; i++)
        for (int k = 0; k < 3; k++) s->mats[i].rgb[k] = rgb[3 * i + k];
    if (s->finalized)
    {
        HIPCHK(hipSetDevice(s->ctx->device));
        // renders, adjoints and batch replays enqueued on any stream read the
        // live table: let them finish before it is overwritten
        HIPCHK(hipDeviceSynchronize());
        int rc = upload_materials(s);
        if (rc) return rc;
    }
    return MCRT_OK;
}

MCRT_API int mcrt_scene_set_envmap_rgb(mcrt_scene* s, const float* rgb, int32_t width, int32_t height)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_envmap_rgb: NULL scene");
    if (rgb && (width <= 0 || height <= 0 || (int64_t)width * height > (int64_t(1) << 28)))
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_envmap_rgb: size %d x %d out of range", width, height);
    return set_envmap(s, rgb, width, height);
}

MCRT_API int mcrt_scene_set_envmap(mcrt_scene* s, const char* path, int32_t width, int32_t height)
{
    if (!s || !path) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_envmap: NULL argument");
    if ((width == 0) != (height == 0) || width < 0 || height < 0)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_envmap: width and height both 0 (file header) or both > 0");
    std::vector<float> rgb;
    int w = width, h = height;
    const int rc = read_envmap(path, w, h, rgb);
    if (rc) return rc;
    return set_envmap(s, rgb.data(), w, h);
}

MCRT_API int mcrt_scene_add_triangles(mcrt_scene* s, const float* v9, const float* vn9, const int32_t* mat,
                                      size_t n)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_add_triangles: NULL scene");
    if (n && (!v9 || !mat)) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_add_triangles: NULL array");
    if (s->finalized) return fail(MCRT_ERR_STATE, "mcrt_scene_add_triangles: scene already finalized");
    for (size_t i = 0; i < n; i++)
        if (mat[i] < 0 || (size_t)mat[i] >= s->mats.size())
            return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_add_triangles: material id %d of triangle %zu is unknown",
                        mat[i], i);
    for (size_t i = 0; i < n; i++)
    {
        Tri t;
        for (int k = 0; k < 3; k++) { t.a[k] = v9[9 * i + k]; t.b[k] = v9[9 * i + 3 + k]; t.c[k] = v9[9 * i + 6 + k]; }
        s->tris.push_back(t);
        if (vn9)
            for (int k = 0; k < 9; k++) s->vn.push_back(vn9[9 * i + k]);
        else
        {
            F3 fn = normalize3(cross3(sub(t.a, t.c), sub(t.b, t.c)));  // triangle.cpp:67-74
            for (int c = 0; c < 3; c++)
                for (int k = 0; k < 3; k++) s->vn.push_back(fn[k]);
        }
        s->tri_mat.push_back(mat[i]);
    }
    return MCRT_OK;
}

MCRT_API int mcrt_scene_load_obj(mcrt_scene* s, const char* path, const float* m16, int32_t mat, size_t* out_n)
{
    if (!s || !path) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_load_obj: NULL argument");
    if (s->finalized) return fail(MCRT_ERR_STATE, "mcrt_scene_load_obj: scene already finalized");
    if (mat < 0 || (size_t)mat >= s->mats.size())
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_load_obj: material id %d is unknown", mat);
    ObjMesh m;
    int rc = load_obj(path, m16, m);
    if (rc) return rc;
    const size_t nf = m.f.size() / 3;
    std::vector<float> v9(9 * nf), vn9(9 * nf);
    obj_soup(m, v9.data(), vn9.data());
    std::vector<int32_t> mats(nf, mat);
    rc = mcrt_scene_add_triangles(s, v9.data(), vn9.data(), mats.data(), nf);
    if (rc) return rc;
    if (out_n) *out_n = nf;
    return MCRT_OK;
}

MCRT_API int mcrt_obj_read(const char* path, const float* m16, float* v9, float* vn9, size_t cap, size_t* out_n)
{
    if (!path || !out_n || (v9 && !vn9)) return fail(MCRT_ERR_INVALID_ARG, "mcrt_obj_read: NULL argument");
    ObjMesh m;
    const int rc = load_obj(path, m16, m);
    if (rc) return rc;
    const size_t nf = m.f.size() / 3;
    *out_n = nf;
    if (!v9) return MCRT_OK;
    if (nf > cap) return fail(MCRT_ERR_INVALID_ARG, "mcrt_obj_read: %s has %zu faces, buffers hold %zu", path, nf, cap);
    obj_soup(m, v9, vn9);
    return MCRT_OK;
}

MCRT_API int mcrt_scene_set_camera(mcrt_scene* s, const float o[3], const float target[3], int32_t width,
                                   int32_t height, float fov_y)
{
    if (!s || !o || !target) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_camera: NULL argument");
    if (width <= 0 || height <= 0) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_camera: bad film size");
    if (!(fov_y > 0.0f && fov_y < 180.0f)) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_camera: bad fov");
    s->has_cam = true;
    for (int k = 0; k < 3; k++) { s->cam_o[k] = o[k]; s->cam_t[k] = target[k]; }
    s->cam_w = width;
    s->cam_h = height;
    s->cam_fov = fov_y;
    s->dev_cam = make_camera(s);
    return MCRT_OK;
}

MCRT_API int mcrt_scene_set_intersection_epsilon(mcrt_scene* s, float eps)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_intersection_epsilon: NULL scene");
    if (!(eps >= 0.0f && eps < __builtin_inff()))
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_intersection_epsilon: epsilon must be finite and >= 0");
    s->det_eps = eps;
    if (s->finalized) apply_trace_tuning(s);
    return MCRT_OK;
}

MCRT_API int mcrt_scene_num_triangles(const mcrt_scene* s, size_t* out)
{
    if (!s || !out) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_num_triangles: NULL argument");
    *out = s->tris.size();
    return MCRT_OK;
}

MCRT_API int mcrt_scene_finalize(mcrt_scene* s, const mcrt_build_opts* opts, mcrt_build_stats* stats)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_finalize: NULL scene");
    if (s->finalized) return fail(MCRT_ERR_STATE, "mcrt_scene_finalize: scene already finalized");
    HIPCHK(hipSetDevice(s->ctx->device));
    int rc = finalize_scene(s, opts ? opts->threads : 0, opts ? opts->bvh_cache : nullptr);
    if (rc) return rc;
    s->finalized = true;
    if (stats) *stats = s->stats;
    return MCRT_OK;
}

MCRT_API int mcrt_intersect(mcrt_scene* s, const float* rays, size_t n, mcrt_hit* hits, int flags, void* stream)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_intersect: NULL scene");
    if (n && (!rays || !hits)) return fail(MCRT_ERR_INVALID_ARG, "mcrt_intersect: NULL buffer");
    return run_query(s, rays, n, hits, nullptr, TRACE_CLOSEST, flags, (hipStream_t)stream, nullptr);
}

MCRT_API int mcrt_intersect_noalloc(mcrt_scene* s, const float* rays, size_t n, mcrt_hit* hits, void* stream)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_intersect_noalloc: NULL scene");
    if (n && (!rays || !hits)) return fail(MCRT_ERR_INVALID_ARG, "mcrt_intersect_noalloc: NULL buffer");
    if (!s->finalized) return fail(MCRT_ERR_STATE, "mcrt_intersect_noalloc: scene not finalized");
    if (!s->ctx->scratch)
        return fail(MCRT_ERR_DEVICE, "mcrt_intersect_noalloc: no scratch memory: %s", s->ctx->scratch_err.c_str());
    if (n == 0) return MCRT_OK;
    HIPCHK(hipSetDevice(s->ctx->device));
    // the scene's launch geometry, with the work counters and the stack spill
    // area in the scratch array and no ray reorder (results do not depend on
    // the order; the reorder's buffers would be allocations)
    DevScene view = s->dev;
    char* base = static_cast<char*>(s->ctx->scratch);
    view.wq = reinterpret_cast<uint32_t*>(base);
    view.order = nullptr;
    view.cell_work = nullptr;
    view.reorder_rays = 0;
    view.spill = nullptr;
    if (view.spill_depth > 0)
    {
        // fewer resident lanes when the spill area would not fit: waves claim
        // rays from the work counters, so any grid covers every ray
        const size_t per_lane = (size_t)view.spill_depth * sizeof(uint64_t);
        const int64_t lanes = (int64_t)((kScratchBytes - kScratchWqBytes) / per_lane) / 64 * 64;
        if (lanes < 64)
            return fail(MCRT_ERR_INVALID_ARG,
                        "mcrt_intersect_noalloc: the scene's traversal stack needs %zu bytes of spill per "
                        "wavefront, more than the %zu bytes of scratch memory",
                        per_lane * 64, kScratchBytes - kScratchWqBytes);
        if (lanes < view.spill_lanes) view.spill_lanes = lanes;
        view.spill = reinterpret_cast<uint64_t*>(base + kScratchWqBytes);
    }
    HIPCHK(launch_trace(view, rays, (int64_t)n, hits, nullptr, TRACE_CLOSEST, nullptr, (hipStream_t)stream));
    return MCRT_OK;
}

MCRT_API int mcrt_occluded(mcrt_scene* s, const float* rays, size_t n, uint8_t* occluded, int flags, void* stream)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_occluded: NULL scene");
    if (n && (!rays || !occluded)) return fail(MCRT_ERR_INVALID_ARG, "mcrt_occluded: NULL buffer");
    return run_query(s, rays, n, nullptr, occluded, TRACE_ANY, flags, (hipStream_t)stream, nullptr);
}

static int check_render(const mcrt_scene* s, const mcrt_render_params* p, const char* who)
{
    if (!s->finalized) return fail(MCRT_ERR_STATE, "%s: scene not finalized", who);
    if (!s->has_cam) return fail(MCRT_ERR_STATE, "%s: no camera (mcrt_scene_set_camera)", who);
    // one pixel's samples stay in one wavefront chunk (at most 2^28 paths)
    if (p->spp < 1 || p->spp > (1 << 28)) return fail(MCRT_ERR_INVALID_ARG, "%s: spp must be in [1, 2^28]", who);
    if (p->max_depth < 0 || p->max_depth > 250) return fail(MCRT_ERR_INVALID_ARG, "%s: max_depth in [0, 250]", who);
    if (!(p->rr_prob > 0.0f && p->rr_prob <= 1.0f)) return fail(MCRT_ERR_INVALID_ARG, "%s: rr_prob in (0, 1]", who);
    if (p->sample_offset < 0) return fail(MCRT_ERR_INVALID_ARG, "%s: sample_offset must be >= 0", who);
    return MCRT_OK;
}

MCRT_API int mcrt_render(mcrt_scene* s, const mcrt_render_params* p, float* film, int flags, void* stream)
{
    if (!s || !p || !film) return fail(MCRT_ERR_INVALID_ARG, "mcrt_render: NULL argument");
    int rc = check_render(s, p, "mcrt_render");
    if (rc) return rc;
    HIPCHK(hipSetDevice(s->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    StreamRes* r = stream_res(s, st);
    std::lock_guard<std::mutex> g(r->mu);
    if (flags & MCRT_DEVICE_PTRS) return render_scene(s, r, p, film, nullptr, st);
    const size_t bytes = (size_t)s->dev_cam.w * s->dev_cam.h * 3 * sizeof(float);
    if ((rc = r->ptw.film.ensure(bytes))) return rc;
    if ((rc = render_scene(s, r, p, r->ptw.film.as<float>(), nullptr, st))) return rc;
    HIPCHK(hipMemcpyAsync(film, r->ptw.film.p, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MCRT_OK;
}

MCRT_API int mcrt_render_adjoint(mcrt_scene* s, const mcrt_render_params* p, const float* dL_dfilm,
                                 float* d_albedo, float* d_emission, float* film, int flags, void* stream)
{
    if (!s || !p || !dL_dfilm || !d_albedo || !d_emission)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_render_adjoint: NULL argument");
    int rc = check_render(s, p, "mcrt_render_adjoint");
    if (rc) return rc;
    HIPCHK(hipSetDevice(s->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    StreamRes* r = stream_res(s, st);
    std::lock_guard<std::mutex> g(r->mu);
    if (flags & MCRT_DEVICE_PTRS)
    {
        AdjointIO io{dL_dfilm, d_albedo, d_emission};
        return render_scene(s, r, p, film, &io, st);
    }
    PTWork& w = r->ptw;
    const size_t fb = (size_t)s->dev_cam.w * s->dev_cam.h * 3 * sizeof(float);
    const size_t gb = (size_t)s->dev.ntri * 3 * sizeof(float);
    if ((rc = w.dfilm.ensure(fb)) || (rc = w.galb.ensure(gb)) || (rc = w.gem.ensure(gb))) return rc;
    if (film && (rc = w.film.ensure(fb))) return rc;
    HIPCHK(hipMemcpyAsync(w.dfilm.p, dL_dfilm, fb, hipMemcpyHostToDevice, st));
    AdjointIO io{w.dfilm.as<float>(), w.galb.as<float>(), w.gem.as<float>()};
    if ((rc = render_scene(s, r, p, film ? w.film.as<float>() : nullptr, &io, st))) return rc;
    HIPCHK(hipMemcpyAsync(d_albedo, w.galb.p, gb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(d_emission, w.gem.p, gb, hipMemcpyDeviceToHost, st));
    if (film) HIPCHK(hipMemcpyAsync(film, w.film.p, fb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MCRT_OK;
}

MCRT_API int mcrt_query_counters(mcrt_scene* s, const float* rays, size_t n, int any_hit, uint64_t counters[4],
                                 int flags, void* stream)
{
    if (!s || !counters || (n && !rays)) return fail(MCRT_ERR_INVALID_ARG, "mcrt_query_counters: NULL argument");
    if (!s->finalized) return fail(MCRT_ERR_STATE, "mcrt_query_counters: scene not finalized");
    HIPCHK(hipSetDevice(s->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    DevBuf d_cnt, d_rays, d_out;
    int rc;
    if ((rc = d_cnt.ensure(4 * sizeof(uint64_t)))) return rc;
    HIPCHK(hipMemsetAsync(d_cnt.p, 0, 4 * sizeof(uint64_t), st));
    const float* r = rays;
    if (!(flags & MCRT_DEVICE_PTRS))
    {
        if ((rc = d_rays.ensure(n * 8 * sizeof(float)))) return rc;
        HIPCHK(hipMemcpyAsync(d_rays.p, rays, n * 8 * sizeof(float), hipMemcpyHostToDevice, st));
        r = d_rays.as<float>();
    }
    if ((rc = d_out.ensure(any_hit ? n : n * sizeof(mcrt_hit)))) return rc;
    StreamRes* res = stream_res(s, st);
    std::lock_guard<std::mutex> g(res->mu);
    DevScene view;
    if ((rc = launch_view(s, res, view, (int64_t)n))) return rc;
    HIPCHK(launch_trace(view, r, (int64_t)n, any_hit ? nullptr : d_out.p, any_hit ? d_out.as<uint8_t>() : nullptr,
                        any_hit ? TRACE_ANY : TRACE_CLOSEST, d_cnt.as<uint64_t>(), st));
    HIPCHK(hipMemcpyAsync(counters, d_cnt.p, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MCRT_OK;
}

static int check_tuning(const mcrt_tuning* t)
{
    if (t->trace_waves_per_cu < 0 || t->trace_waves_per_cu > 1024)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: trace_waves_per_cu in [1, 1024] (0 = keep)");
    if (t->trace_lds_depth != 0 && !valid_lds_depth(t->trace_lds_depth))
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: trace_lds_depth must be 8, 12, 16 or 24 (0 = keep)");
    if (t->pt_chunk_paths < 0 || t->pt_chunk_paths > (int64_t(1) << 28))
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: pt_chunk_paths in [1, 2^28] (0 = keep)");
    if (t->pt_trace_waves_per_cu < 0 || t->pt_trace_waves_per_cu > 1024)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: pt_trace_waves_per_cu in [1, 1024] (0 = keep)");
    if (t->pt_trace_lds_depth != 0 && !valid_lds_depth(t->pt_trace_lds_depth))
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: pt_trace_lds_depth must be 8, 12, 16 or 24 (0 = keep)");
    if (t->trace_reorder_rays < -1)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: trace_reorder_rays is -1 (never) or a ray count (0 = keep)");
    if (t->trace_refill_lanes < 0 || t->trace_refill_lanes > 64)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: trace_refill_lanes in [1, 64] (0 = keep)");
    if (t->pt_trace_refill_lanes < 0 || t->pt_trace_refill_lanes > 64)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: pt_trace_refill_lanes in [1, 64] (0 = keep)");
    if (t->trace_any_refill_lanes < 0 || t->trace_any_refill_lanes > 64)
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_tuning: trace_any_refill_lanes in [1, 64] (0 = keep)");
    return MCRT_OK;
}

MCRT_API int mcrt_scene_set_tuning(mcrt_scene* s, const mcrt_tuning* t)
{
    if (!s || !t) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_set_tuning: NULL argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    if (t->trace_waves_per_cu) s->trace_waves_per_cu = t->trace_waves_per_cu;
    if (t->trace_lds_depth) s->trace_lds_depth = t->trace_lds_depth;
    if (t->pt_chunk_paths) s->pt_chunk_paths = t->pt_chunk_paths;
    if (t->pt_trace_waves_per_cu) s->pt_trace_waves_per_cu = t->pt_trace_waves_per_cu;
    if (t->pt_trace_lds_depth) s->pt_trace_lds_depth = t->pt_trace_lds_depth;
    if (t->trace_reorder_rays) s->trace_reorder_rays = t->trace_reorder_rays;
    if (t->trace_refill_lanes) s->trace_refill_lanes = t->trace_refill_lanes;
    if (t->pt_trace_refill_lanes) s->pt_trace_refill_lanes = t->pt_trace_refill_lanes;
    if (t->trace_any_refill_lanes) s->trace_any_refill_lanes = t->trace_any_refill_lanes;
    if (s->finalized) apply_trace_tuning(s);
    return MCRT_OK;
}

MCRT_API int mcrt_scene_get_tuning(const mcrt_scene* s, mcrt_tuning* t)
{
    if (!s || !t) return fail(MCRT_ERR_INVALID_ARG, "mcrt_scene_get_tuning: NULL argument");
    t->trace_waves_per_cu = s->trace_waves_per_cu;
    t->trace_lds_depth = s->trace_lds_depth;
    t->pt_chunk_paths = s->pt_chunk_paths;
    t->pt_trace_waves_per_cu = s->pt_trace_waves_per_cu;
    t->pt_trace_lds_depth = s->pt_trace_lds_depth;
    t->trace_reorder_rays = s->trace_reorder_rays;
    t->trace_refill_lanes = s->trace_refill_lanes;
    t->pt_trace_refill_lanes = s->pt_trace_refill_lanes;
    t->trace_any_refill_lanes = s->trace_any_refill_lanes;
    return MCRT_OK;
}

static int autotune(mcrt_scene* s, const float* rays, size_t n, const mcrt_render_params* render, int flags,
                    void* stream, mcrt_tuning* best, double* best_mrays_per_s);

MCRT_API int mcrt_autotune(mcrt_scene* s, const float* rays, size_t n, const mcrt_render_params* render, int flags,
                           void* stream, mcrt_tuning* best, double* best_mrays_per_s)
{
    if (!s) return fail(MCRT_ERR_INVALID_ARG, "mcrt_autotune: NULL argument");
    mcrt_tuning before;
    mcrt_scene_get_tuning(s, &before);
    const int rc = autotune(s, rays, n, render, flags, stream, best, best_mrays_per_s);
    if (rc)
    {
        // a failed sweep leaves the scene as it found it
        const std::string err = mcrt_last_error();
        s->trace_waves_per_cu = before.trace_waves_per_cu;
        s->trace_lds_depth = before.trace_lds_depth;
        s->pt_chunk_paths = before.pt_chunk_paths;
        s->pt_trace_waves_per_cu = before.pt_trace_waves_per_cu;
        s->pt_trace_lds_depth = before.pt_trace_lds_depth;
        s->trace_reorder_rays = before.trace_reorder_rays;
        s->trace_refill_lanes = before.trace_refill_lanes;
        s->pt_trace_refill_lanes = before.pt_trace_refill_lanes;
        s->trace_any_refill_lanes = before.trace_any_refill_lanes;
        if (s->finalized) apply_trace_tuning(s);
        fail(rc, "%s", err.c_str());
    }
    return rc;
}

static int autotune(mcrt_scene* s, const float* rays, size_t n, const mcrt_render_params* render, int flags,
                    void* stream, mcrt_tuning* best, double* best_mrays_per_s)
{
    if (n && !rays) return fail(MCRT_ERR_INVALID_ARG, "mcrt_autotune: NULL argument");
    if (!s->finalized) return fail(MCRT_ERR_STATE, "mcrt_autotune: scene not finalized");
    if (n == 0 && !render) return fail(MCRT_ERR_INVALID_ARG, "mcrt_autotune: nothing to time (no rays, no render)");
    int rc;
    if (render && (rc = check_render(s, render, "mcrt_autotune"))) return rc;
    HIPCHK(hipSetDevice(s->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    StreamRes* res = stream_res(s, st);
    std::lock_guard<std::mutex> g(res->mu);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    struct Ev
    {
        hipEvent_t a, b;
        ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};
    double best_rate = 0.0;
    // Traversal: closest-hit over the rays for each (waves per CU, LDS depth).
    if (n)
    {
        DevBuf d_rays, d_hits;
        const float* r = rays;
        if (!(flags & MCRT_DEVICE_PTRS))
        {
            if ((rc = d_rays.ensure(n * 8 * sizeof(float)))) return rc;
            HIPCHK(hipMemcpyAsync(d_rays.p, rays, n * 8 * sizeof(float), hipMemcpyHostToDevice, st));
            r = d_rays.as<float>();
        }
        if ((rc = d_hits.ensure(n * sizeof(mcrt_hit)))) return rc;
        static const int kWaves[] = {8, 16, 24, 28, 32, 64, 128};
        static const int kDepth[] = {8, 12, 16};
        int bw = s->trace_waves_per_cu, bd = s->trace_lds_depth;
        for (int w : kWaves)
            for (int d : kDepth)
            {
                s->trace_waves_per_cu = w;
                s->trace_lds_depth = d;
                apply_trace_tuning(s);
                DevScene view;
                if ((rc = launch_view(s, res, view, (int64_t)n))) return rc;
                HIPCHK(launch_trace(view, r, (int64_t)n, d_hits.p, nullptr, TRACE_CLOSEST, nullptr, st));
                HIPCHK(hipEventRecord(e0, st));
                for (int k = 0; k < 3; k++)
                    HIPCHK(launch_trace(view, r, (int64_t)n, d_hits.p, nullptr, TRACE_CLOSEST, nullptr, st));
                HIPCHK(hipEventRecord(e1, st));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                const double rate = 3.0 * (double)n / (ms * 1e-3) / 1e6;
                if (rate > best_rate)
                {
                    best_rate = rate;
                    bw = w;
                    bd = d;
                }
            }
        s->trace_waves_per_cu = bw;
        s->trace_lds_depth = bd;
        // then the compaction threshold with the chosen pair
        static const int kRefill[] = {1, 8, 16, 32};
        int br = s->trace_refill_lanes;
        double refill_rate = 0.0;
        for (int k : kRefill)
        {
            s->trace_refill_lanes = k;
            apply_trace_tuning(s);
            DevScene view;
            if ((rc = launch_view(s, res, view, (int64_t)n))) return rc;
            HIPCHK(launch_trace(view, r, (int64_t)n, d_hits.p, nullptr, TRACE_CLOSEST, nullptr, st));
            HIPCHK(hipEventRecord(e0, st));
            for (int j = 0; j < 3; j++)
                HIPCHK(launch_trace(view, r, (int64_t)n, d_hits.p, nullptr, TRACE_CLOSEST, nullptr, st));
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            const double rate = 3.0 * (double)n / (ms * 1e-3) / 1e6;
            if (rate > refill_rate)
            {
                refill_rate = rate;
                br = k;
            }
        }
        s->trace_refill_lanes = br;
        // then the any-hit queries' threshold, on the same rays
        {
            DevBuf d_occ;
            if ((rc = d_occ.ensure(n))) return rc;
            static const int kAnyRefill[] = {16, 32, 48};
            int ba = s->trace_any_refill_lanes;
            double any_rate = 0.0;
            for (int k : kAnyRefill)
            {
                s->trace_any_refill_lanes = k;
                apply_trace_tuning(s);
                DevScene view;
                if ((rc = launch_view(s, res, view, (int64_t)n))) return rc;
                uint8_t* o = d_occ.as<uint8_t>();
                HIPCHK(launch_trace(view, r, (int64_t)n, nullptr, o, TRACE_ANY, nullptr, st));
                HIPCHK(hipEventRecord(e0, st));
                for (int j = 0; j < 3; j++) HIPCHK(launch_trace(view, r, (int64_t)n, nullptr, o, TRACE_ANY, nullptr, st));
                HIPCHK(hipEventRecord(e1, st));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                const double rate = 3.0 * (double)n / (ms * 1e-3) / 1e6;
                if (rate > any_rate)
                {
                    any_rate = rate;
                    ba = k;
                }
            }
            s->trace_any_refill_lanes = ba;
        }
        // then the ray order: the caller's, or by origin cell from min(n, 2^20) rays
        const int64_t order[2] = {-1, (int64_t)n < (int64_t(1) << 20) ? (int64_t)n : int64_t(1) << 20};
        double order_rate[2] = {0.0, 0.0};
        for (int c = 0; c < 2; c++)
        {
            s->trace_reorder_rays = order[c];
            apply_trace_tuning(s);
            DevScene view;
            if ((rc = launch_view(s, res, view, (int64_t)n))) return rc;
            HIPCHK(launch_trace(view, r, (int64_t)n, d_hits.p, nullptr, TRACE_CLOSEST, nullptr, st));
            HIPCHK(hipEventRecord(e0, st));
            for (int k = 0; k < 3; k++)
                HIPCHK(launch_trace(view, r, (int64_t)n, d_hits.p, nullptr, TRACE_CLOSEST, nullptr, st));
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            order_rate[c] = 3.0 * (double)n / (ms * 1e-3) / 1e6;
        }
        s->trace_reorder_rays = order_rate[1] > order_rate[0] ? order[1] : order[0];
        best_rate = std::max(order_rate[0], order_rate[1]);
        apply_trace_tuning(s);
    }
    // Path tracer: its traversal launches (waves per CU x LDS depth), then the
    // paths per wavefront chunk; one timed render of the camera each.
    if (render)
    {
        const int64_t total = (int64_t)s->dev_cam.w * s->dev_cam.h * render->spp;
        DevBuf film;
        if ((rc = film.ensure((size_t)s->dev_cam.w * s->dev_cam.h * 3 * sizeof(float)))) return rc;
        auto time_render = [&](float& ms) -> int {
            int e = render_scene(s, res, render, film.as<float>(), nullptr, st);
            if (e) return e;
            HIPCHK(hipEventRecord(e0, st));
            if ((e = render_scene(s, res, render, film.as<float>(), nullptr, st))) return e;
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            return MCRT_OK;
        };
        static const int kPtWaves[] = {16, 28, 32, 64};
        static const int kPtDepth[] = {8, 16};
        int pw = s->pt_trace_waves_per_cu, pd = s->pt_trace_lds_depth;
        float best_ms = 0.0f;
        for (int w : kPtWaves)
            for (int d : kPtDepth)
            {
                s->pt_trace_waves_per_cu = w;
                s->pt_trace_lds_depth = d;
                float ms = 0.0f;
                if ((rc = time_render(ms))) return rc;
                if (best_ms == 0.0f || ms < best_ms)
                {
                    best_ms = ms;
                    pw = w;
                    pd = d;
                }
            }
        s->pt_trace_waves_per_cu = pw;
        s->pt_trace_lds_depth = pd;
        static const int kPtRefill[] = {1, 8, 16, 32};
        int pr = s->pt_trace_refill_lanes;
        best_ms = 0.0f;
        for (int k : kPtRefill)
        {
            s->pt_trace_refill_lanes = k;
            float ms = 0.0f;
            if ((rc = time_render(ms))) return rc;
            if (best_ms == 0.0f || ms < best_ms)
            {
                best_ms = ms;
                pr = k;
            }
        }
        s->pt_trace_refill_lanes = pr;
        int64_t bc = s->pt_chunk_paths;
        best_ms = 0.0f;
        for (int lg = 22; lg <= 26; lg++)
        {
            const int64_t c = int64_t(1) << lg;
            if (lg > 22 && (c >> 1) >= total) break;  // larger chunks change nothing
            s->pt_chunk_paths = c;
            float ms = 0.0f;
            if ((rc = time_render(ms))) return rc;
            if (best_ms == 0.0f || ms < best_ms)
            {
                best_ms = ms;
                bc = c;
            }
        }
        s->pt_chunk_paths = bc;
    }
    HIPCHK(hipStreamSynchronize(st));
    if (best) mcrt_scene_get_tuning(s, best);
    if (best_mrays_per_s) *best_mrays_per_s = best_rate;
    return MCRT_OK;
}

static int check_batch(size_t n, mcrt_scene* const* scenes, const mcrt_render_params* params,
                       float* const* films, const char* who)
{
    if (n == 0 || !scenes || !params || !films) return fail(MCRT_ERR_INVALID_ARG, "%s: empty batch or NULL array", who);
    if (n > 65535) return fail(MCRT_ERR_INVALID_ARG, "%s: at most 65535 scenes per batch", who);
    for (size_t i = 0; i < n; i++)
    {
        if (!scenes[i] || !films[i]) return fail(MCRT_ERR_INVALID_ARG, "%s: NULL scene or film %zu", who, i);
        if (scenes[i]->ctx != scenes[0]->ctx)
            return fail(MCRT_ERR_INVALID_ARG, "%s: scene %zu belongs to another context", who, i);
        int rc = check_render(scenes[i], &params[i], who);
        if (rc) return rc;
    }
    return MCRT_OK;
}

MCRT_API int mcrt_batch_create(size_t n, mcrt_scene* const* scenes, const mcrt_render_params* params,
                               float* const* films, int flags, mcrt_batch** out)
{
    if (!out) return fail(MCRT_ERR_INVALID_ARG, "mcrt_batch_create: NULL out");
    *out = nullptr;
    int rc = check_batch(n, scenes, params, films, "mcrt_batch_create");
    if (rc) return rc;
    if (!(flags & MCRT_DEVICE_PTRS))
        return fail(MCRT_ERR_INVALID_ARG, "mcrt_batch_create: films must be device pointers (MCRT_DEVICE_PTRS)");
    HIPCHK(hipSetDevice(scenes[0]->ctx->device));
    std::unique_ptr<mcrt_batch> b(new mcrt_batch());
    b->ctx = scenes[0]->ctx;
    b->scenes.assign(scenes, scenes + n);
    b->params.assign(params, params + n);
    b->films.assign(films, films + n);
    b->flags = flags;
    if ((rc = batch_init(b.get()))) return rc;
    *out = b.release();
    return MCRT_OK;
}

MCRT_API int mcrt_batch_render(mcrt_batch* b, void* stream)
{
    if (!b) return fail(MCRT_ERR_INVALID_ARG, "mcrt_batch_render: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    for (size_t i = 0; i < b->scenes.size(); i++)
    {
        if (b->scenes[i]->dev.det_eps != b->det_eps[i])
            return fail(MCRT_ERR_STATE,
                        "mcrt_batch_render: the intersection epsilon of scene %zu changed after mcrt_batch_create", i);
        if (b->scenes[i]->env_gen != b->env_gen[i])
            return fail(MCRT_ERR_STATE,
                        "mcrt_batch_render: the environment map of scene %zu changed after mcrt_batch_create", i);
    }
    std::lock_guard<std::mutex> g(b->res.mu);
    if (!(b->flags & MCRT_BATCH_GRAPH)) return batch_enqueue(b, st);
    if (!b->exec)
    {
        if (!b->cap) HIPCHK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        const int rc = batch_enqueue(b, b->cap);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cap, &graph);
        if (rc) return rc;
        HIPCHK(e);
        b->graph = graph;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->exec, st));
    return MCRT_OK;
}

MCRT_API int mcrt_batch_info(const mcrt_batch* b, int64_t* total_paths, int64_t* chunks)
{
    if (!b) return fail(MCRT_ERR_INVALID_ARG, "mcrt_batch_info: NULL batch");
    if (total_paths) *total_paths = b->total_paths;
    if (chunks) *chunks = (int64_t)b->chunks.size();
    return MCRT_OK;
}

// The render half of mcrt_autotune on a batch: each setting applied to the
// batch's first scene (whose tuning a batch uses), the batch re-initialised
// (chunks, workspace) and rendered once, timed with events on `st`.
static int batch_autotune(mcrt_batch* b, hipStream_t st, double* best_ms_out)
{
    mcrt_scene* s0 = b->scenes[0];
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    struct Ev
    {
        hipEvent_t a, b;
        ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};
    int rc;
    auto time_batch = [&](float& ms) -> int {
        int e = batch_init(b);
        if (e) return e;
        HIPCHK(hipEventRecord(e0, st));
        if ((e = batch_enqueue(b, st))) return e;
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        return MCRT_OK;
    };
    float ms = 0.0f, best_ms = 0.0f;
    if ((rc = time_batch(ms))) return rc;  // warm-up: code objects, workspace
    static const int kPtWaves[] = {16, 28, 32, 64};
    static const int kPtDepth[] = {8, 16};
    int pw = s0->pt_trace_waves_per_cu, pd = s0->pt_trace_lds_depth;
    for (int w : kPtWaves)
        for (int d : kPtDepth)
        {
            s0->pt_trace_waves_per_cu = w;
            s0->pt_trace_lds_depth = d;
            if ((rc = time_batch(ms))) return rc;
            if (best_ms == 0.0f || ms < best_ms)
            {
                best_ms = ms;
                pw = w;
                pd = d;
            }
        }
    s0->pt_trace_waves_per_cu = pw;
    s0->pt_trace_lds_depth = pd;
    static const int kPtRefill[] = {1, 8, 16, 32};
    int pr = s0->pt_trace_refill_lanes;
    best_ms = 0.0f;
    for (int k : kPtRefill)
    {
        s0->pt_trace_refill_lanes = k;
        if ((rc = time_batch(ms))) return rc;
        if (best_ms == 0.0f || ms < best_ms)
        {
            best_ms = ms;
            pr = k;
        }
    }
    s0->pt_trace_refill_lanes = pr;
    int64_t bc = s0->pt_chunk_paths;
    best_ms = 0.0f;
    for (int lg = 22; lg <= 26; lg++)
    {
        const int64_t c = int64_t(1) << lg;
        if (lg > 22 && (c >> 1) >= b->total_paths) break;  // larger chunks change nothing
        s0->pt_chunk_paths = c;
        if ((rc = time_batch(ms))) return rc;
        if (best_ms == 0.0f || ms < best_ms)
        {
            best_ms = ms;
            bc = c;
        }
    }
    s0->pt_chunk_paths = bc;
    if ((rc = batch_init(b))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    if (best_ms_out) *best_ms_out = best_ms;
    return MCRT_OK;
}

MCRT_API int mcrt_batch_autotune(mcrt_batch* b, void* stream, mcrt_tuning* best, double* best_ms)
{
    if (!b) return fail(MCRT_ERR_INVALID_ARG, "mcrt_batch_autotune: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> g(b->res.mu);
    // a captured graph holds the old settings: drop it (after any replay in
    // flight), the next mcrt_batch_render captures the tuned sequence
    HIPCHK(hipDeviceSynchronize());
    if (b->exec) HIPCHK(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
    if (b->graph) HIPCHK(hipGraphDestroy(b->graph));
    b->graph = nullptr;
    mcrt_scene* s0 = b->scenes[0];
    mcrt_tuning before;
    mcrt_scene_get_tuning(s0, &before);
    const int rc = batch_autotune(b, st, best_ms);
    if (rc)
    {
        const std::string err = mcrt_last_error();
        s0->pt_trace_waves_per_cu = before.pt_trace_waves_per_cu;
        s0->pt_trace_lds_depth = before.pt_trace_lds_depth;
        s0->pt_trace_refill_lanes = before.pt_trace_refill_lanes;
        s0->pt_chunk_paths = before.pt_chunk_paths;
        (void)batch_init(b);
        return fail(rc, "%s", err.c_str());
    }
    if (best) mcrt_scene_get_tuning(s0, best);
    return MCRT_OK;
}

MCRT_API int mcrt_batch_destroy(mcrt_batch* b)
{
    if (!b) return MCRT_OK;
    HIPCHK(hipSetDevice(b->ctx->device));
    delete b;
    return MCRT_OK;
}

MCRT_API int mcrt_render_batch(size_t n, mcrt_scene* const* scenes, const mcrt_render_params* params,
                               float* const* films, int flags, void* stream)
{
    int rc = check_batch(n, scenes, params, films, "mcrt_render_batch");
    if (rc) return rc;
    HIPCHK(hipSetDevice(scenes[0]->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    std::vector<float*> dev(films, films + n);
    DevBuf staging;
    std::vector<size_t> off(n + 1, 0);
    for (size_t i = 0; i < n; i++)
        off[i + 1] = off[i] + (size_t)scenes[i]->dev_cam.w * scenes[i]->dev_cam.h * 3;
    if (!(flags & MCRT_DEVICE_PTRS))
    {
        if ((rc = staging.ensure(off[n] * sizeof(float)))) return rc;
        for (size_t i = 0; i < n; i++) dev[i] = staging.as<float>() + off[i];
    }
    mcrt_batch* b = nullptr;
    if ((rc = mcrt_batch_create(n, scenes, params, dev.data(), MCRT_DEVICE_PTRS, &b))) return rc;
    std::unique_ptr<mcrt_batch> hold(b);
    if ((rc = batch_enqueue(b, st))) return rc;
    if (!(flags & MCRT_DEVICE_PTRS))
        for (size_t i = 0; i < n; i++)
            HIPCHK(hipMemcpyAsync(films[i], dev[i], (off[i + 1] - off[i]) * sizeof(float), hipMemcpyDeviceToHost, st));
    // the batch's workspace is freed on return: wait for its kernels
    HIPCHK(hipStreamSynchronize(st));
    return MCRT_OK;
}

}  // extern "C"
