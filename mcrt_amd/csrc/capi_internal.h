// Internal types behind the opaque handles of include/mcrt.h.
#pragma once
#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "device.h"
#include "geom.h"
#include "mcrt.h"

namespace mcrt
{
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int ok();

#define HIPCHK(x)                                                                              \
    do                                                                                         \
    {                                                                                          \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
        {                                                                                      \
            (void)hipGetLastError();                                                           \
            return ::mcrt::fail(MCRT_ERR_DEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
        }                                                                                      \
    } while (0)

// Owned device allocation.
struct DevBuf
{
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
    void release();
    int ensure(size_t want);
    int upload(const void* src, size_t n);
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct HostMaterial
{
    int type;
    float rgb[3];
    float gloss;
};

struct ObjMesh
{
    std::vector<F3> v, vn;
    std::vector<int> f;  // 3 per face, 0-based
};
int load_obj(const char* path, const float* m16, ObjMesh& out);
// The faces of `m` as a triangle soup: v9 / vn9 = nf x (a, b, c) positions / vertex normals.
void obj_soup(const ObjMesh& m, float* v9, float* vn9);
}  // namespace mcrt

namespace mcrt
{
// Path-tracer workspace (SoA path state), grown on demand.
struct PTWork
{
    DevBuf T, L, rays[2], act[2], rec[2], bsl[2], cnt, hits, film;
    DevBuf Ltot, delta;                    // adjoint
    DevBuf dfilm, galb, gem;               // adjoint staging for host pointers
    DevBuf sid, pkey, skey, rsid[2];       // batched rendering
};

// Device buffers of one adjoint render.
struct AdjointIO
{
    const float* dfilm;  // W*H*3 dLoss/dfilm
    float* d_albedo;     // ntri*3
    float* d_emission;   // ntri*3
};

// Device resources of one stream in a context.  Work enqueued on distinct
// streams may overlap on the GPU, so each stream has its own traversal
// spill stacks and path-tracer workspace (grown to the largest scene that
// used it); `mu` serialises host threads enqueueing on the same stream (a
// render's kernels must not interleave with another's on one stream).
struct StreamRes
{
    std::mutex mu;
    DevBuf spill;
    DevBuf wq;  // ray-query work counter
    DevBuf order, cells;  // traversal order of large queries (DevScene::order, cell_work)
    PTWork ptw;
};
}  // namespace mcrt

struct mcrt_ctx
{
    int device = 0;
    int num_cus = 0;
    std::string name;
    // device address of the allocation-free query's scratch array on this
    // context's device (nullptr: unavailable, scratch_err says why)
    void* scratch = nullptr;
    std::string scratch_err;
    // per-stream spill stacks and path-tracer workspaces, shared by the
    // scenes of the context (work on one stream runs in order)
    std::mutex res_mu;
    std::map<hipStream_t, std::unique_ptr<mcrt::StreamRes>> res;
};

struct mcrt_scene
{
    mcrt_ctx* ctx = nullptr;
    // host description
    std::vector<mcrt::Tri> tris;
    std::vector<float> vn;  // 9 per triangle
    std::vector<int32_t> tri_mat;
    std::vector<mcrt::HostMaterial> mats;
    bool has_cam = false;
    float cam_o[3] = {0, 0, 0}, cam_t[3] = {0, 0, 1};
    int32_t cam_w = 0, cam_h = 0;
    float cam_fov = 45.0f;
    mcrt::DevCamera dev_cam;
    bool finalized = false;
    mcrt_build_stats stats{};
    // device residency
    mcrt::DevScene dev;
    mcrt::DevBuf d_bvh, d_shade, d_verts, d_mats, d_lights, d_hot;
    // Moller-Trumbore determinant cut (TriangleObject::setIntersectionEpsilon)
    float det_eps = 1e-6f;
    // environment map (dev.env_*); env_gen counts replacements
    mcrt::DevBuf d_env_px, d_env_cdf_u, d_env_cdf_v;
    uint64_t env_gen = 0;
    // staging for host-pointer queries
    std::mutex q_mu;
    mcrt::DevBuf q_in, q_out;
    // launch tuning (mcrt_scene_set_tuning / mcrt_autotune)
    // defaults measured on MI355X (1.3M-triangle displaced icosphere, 16M
    // incoherent rays; 1024^2 renders of it and of an outdoor scene): waves
    // claim rays from work counters, so the grid only has to fill the
    // machine once, with 8-entry LDS stacks.  Waves per CU: 0 = not set,
    // decided at finalize by the BVH's size (32, all resident, over a
    // cache-resident BVH; 28 over a larger one)
    int32_t trace_waves_per_cu = 0;
    int32_t trace_lds_depth = 8;
    int32_t pt_trace_waves_per_cu = 0;
    int32_t pt_trace_lds_depth = 8;
    // 2^26 paths per chunk (a whole 1024^2 x 64-spp frame; about 28.4 GB
    // of path state, 42.3 GB with an environment map): fewer, larger launches
    // whose ragged ends weigh less (the frame 109 ms, against 121 ms in 2^24
    // chunks and 140 ms in 2^23)
    int64_t pt_chunk_paths = int64_t(1) << 26;
    // queries of at least this many rays are traversed in origin-cell order;
    // -1: never; 0: not set, decided at finalize by the BVH's size
    int64_t trace_reorder_rays = 0;
    // compaction threshold of the traversal launches (queries / renders): a
    // wave refills its idle lanes once at least this many of its 64 are idle.
    // 16 measured fastest for the 16M incoherent queries of the 1.3M-triangle
    // mesh (1.74 ms against 2.04 ms refilling every lane at once), its
    // 1024^2 x 64-spp frame (100 against 109 ms) and the 4M-triangle soup's
    // queries; results do not depend on it
    int32_t trace_refill_lanes = 16;
    int32_t pt_trace_refill_lanes = 16;
    // any-hit queries' threshold: their rays end at the first hit, so lanes
    // free up faster; 32 measured fastest (the 1.3M-triangle mesh's 16M rays
    // 0.724 against 0.749 ms at 16, the soup's 6.43 against 7.09 ms)
    int32_t trace_any_refill_lanes = 32;
};

namespace mcrt
{
DevCamera make_camera(const mcrt_scene* s);
// Builds (or, with `cache`, loads when it matches) the BVH, collapses and
// uploads the scene.
int finalize_scene(mcrt_scene* s, int threads, const char* cache);
int upload_materials(mcrt_scene* s);
// Environment map from w x h RGB floats (rows top-down); rgb = NULL removes
// it.  Waits for the device before replacing the tables.
int set_envmap(mcrt_scene* s, const float* rgb, int w, int h);
// Reads an environment-map file; w = h = 0: the file has the size header.
int read_envmap(const char* path, int& w, int& h, std::vector<float>& rgb);
int run_query(mcrt_scene* s, const float* rays, size_t n, mcrt_hit* hits, uint8_t* occ, int mode, int flags,
              hipStream_t stream, uint64_t* counters);
// The resources of `stream` in s's context (created on first use; live as
// long as the context).
StreamRes* stream_res(mcrt_scene* s, hipStream_t stream);
// Sets the traversal tuning of s->dev (LDS depth, spill geometry).
void apply_trace_tuning(mcrt_scene* s);
// The scene's device view for a launch with r's spill stacks (caller holds
// r->mu); for a query of n rays also r's reorder buffers when n reaches the
// scene's reorder threshold.
int launch_view(mcrt_scene* s, StreamRes* r, DevScene& view, int64_t n = 0);
// The same for the traversal launches of renders (pt_trace_* tuning).
int render_view(mcrt_scene* s, StreamRes* r, DevScene& view);
// Sets the traversal geometry of `view` (LDS depth, spill lanes and depth).
void trace_geometry(DevScene& view, int num_cus, int waves_per_cu, int lds_depth);
int pt_workspace(StreamRes* r, int64_t P, bool adjoint, bool batch, int rs);
// Renders into a device film (W*H*3 floats, may be NULL with adj) on
// `stream` (asynchronous); with adj also the path-replay gradients.  The
// caller holds r->mu.
int render_scene(mcrt_scene* s, StreamRes* r, const mcrt_render_params* p, float* film, const AdjointIO* adj,
                 hipStream_t stream);

// One wavefront chunk of a batched render: P paths over nseg segments
// starting at segment seg0, npix pixels.
struct BatchChunk
{
    int64_t P, nseg, seg0, npix;
};
}  // namespace mcrt

// A batched render: scenes, parameters and device films fixed at creation;
// device tables, chunking, workspace and (optionally) the captured graph.
struct mcrt_batch
{
    mcrt_ctx* ctx = nullptr;
    std::vector<mcrt_scene*> scenes;
    std::vector<mcrt_render_params> params;
    std::vector<float*> films;
    int flags = 0;
    int max_depth = 0;
    int rs = 2;                      // next-event rays per record (4: a scene has an environment map)
    std::vector<uint64_t> env_gen;   // each scene's environment generation at creation
    std::vector<float> det_eps;      // each scene's intersection epsilon at creation
    int64_t total_paths = 0;
    std::vector<mcrt::BatchChunk> chunks;
    mcrt::DevBuf d_tab, d_roots, d_seg;
    mcrt::StreamRes res;
    mcrt::DevScene trace_view;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~mcrt_batch();
};

namespace mcrt
{
// Builds the device tables, chunks and workspace of a validated batch.
int batch_init(mcrt_batch* b);
// Enqueues every kernel of the batch on `stream` (launches only).
int batch_enqueue(mcrt_batch* b, hipStream_t stream);
}  // namespace mcrt
