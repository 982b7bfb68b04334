/*
 * mcrt.h -- C ABI of the MI355X (gfx950) ray-query core and path tracer.
 *
 * The boundary replaces the reference's ray-query plug-in point and its
 * scene/integrator interface:
 *   - SceneBVH(triangles) / intersect / occluded      src/core/scenebvh.h:11-25
 *   - Scene::add / loadModel / setCamera / finalize   src/core/scene.h:28-103
 *   - the demo_pt render loop + rayTrace             src/examples/pt/demo_pt.cpp:512-720
 *   - todo_addMaterial (global material library)      src/core/material.cpp:130-139
 *
 * Conventions: opaque handles, every call returns an mcrt_status (0 = OK)
 * and never exits or throws; mcrt_last_error() gives a thread-local message
 * for the last failing call on this thread.  Caller-owned buffers; device
 * memory for scenes is owned by the library per handle.  Buffers passed to
 * queries and renders are host pointers unless MCRT_DEVICE_PTRS is set in
 * `flags`, in which case they are device pointers on the context's device
 * and the call is asynchronous on `stream` (a hipStream_t, NULL = default).
 * Host-pointer calls are synchronous.
 *
 * Threading: a finalized scene is immutable and may be queried and rendered
 * from several host threads on distinct streams.  Scene mutation and
 * mcrt_scene_finalize are single-threaded per handle.
 */
#ifndef MCRT_H
#define MCRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCRT_API __attribute__((visibility("default")))

typedef enum mcrt_status {
    MCRT_OK = 0,
    MCRT_ERR_INVALID_ARG = 1,  /* NULL handle/pointer, bad size or enum     */
    MCRT_ERR_STATE = 2,        /* e.g. query before finalize, add after     */
    MCRT_ERR_IO = 3,           /* file cannot be opened or written          */
    MCRT_ERR_DEVICE = 4,       /* HIP runtime error (message has details)   */
    MCRT_ERR_OUT_OF_MEMORY = 5 /* host or device allocation failed         */
} mcrt_status;

/* MaterialType values of src/core/material.h:7 that the path supports. */
typedef enum mcrt_material_type {
    MCRT_LGHT = 0, /* area light: emitted radiance = rgb, one-sided        */
    MCRT_DIFF = 1, /* Lambertian, albedo = rgb             bsdf.cpp:51-116 */
    MCRT_GLSY = 4  /* Phong lobe, exponent = glossiness    bsdf.cpp:204-301 */
} mcrt_material_type;

enum {
    MCRT_DEVICE_PTRS = 1, /* buffers are device pointers, call is async   */
    MCRT_BATCH_GRAPH = 2  /* mcrt_batch_create: capture the batch's launch
                             sequence in a hipGraph on the first render and
                             replay the graph afterwards                   */
};

typedef struct mcrt_ctx mcrt_ctx;
typedef struct mcrt_scene mcrt_scene;
typedef struct mcrt_batch mcrt_batch;

/* One ray = 8 floats {ox, oy, oz, tmin, dx, dy, dz, tmax} (32 bytes). */
/* One hit = {prim, t, u, v}; prim = -1 and t = +inf on a miss, as HitInfo
 * after reset() (src/core/intersection.cpp:23-31).  u weights vertex b and
 * v vertex c (src/core/triangle.cpp:11-51). */
typedef struct mcrt_hit {
    int32_t prim;
    float t, u, v;
} mcrt_hit;

typedef struct mcrt_build_opts {
    int32_t threads;       /* host threads for the SBVH build; 0 = all cores  */
    int32_t reserved;
    const char* bvh_cache; /* optional file: when it holds the flattened BVH of
                              these exact triangles (same bits, same order) it
                              is loaded instead of building; otherwise the BVH
                              is built and the file (re)written.  NULL = build */
} mcrt_build_opts;

typedef struct mcrt_build_stats {
    int64_t triangles;
    int64_t binary_nodes;   /* SplitBVH nodes (branches + leaves)            */
    int64_t leaves;
    int64_t refs;           /* leaf references (spatial splits duplicate)    */
    int64_t wide_nodes;     /* 4-wide nodes after collapsing                 */
    int64_t max_depth;      /* binary depth                                  */
    int64_t stack_bound;    /* traversal stack entries that always suffice   */
    uint64_t tree_digest;   /* FNV-1a of the canonical binary tree           */
    double build_seconds;   /* host SBVH build                               */
    double flatten_seconds; /* collapse + upload                             */
    int64_t device_bytes;   /* scene bytes resident in HBM                   */
    int64_t from_cache;     /* 1: the BVH was loaded from opts->bvh_cache    */
} mcrt_build_stats;

typedef struct mcrt_render_params {
    int32_t spp;           /* samples per pixel in this call                */
    int32_t max_depth;     /* path vertices n = 0..max_depth (demo_pt.cpp:520-523) */
    int32_t rr_start;      /* Russian roulette from vertex rr_start on      */
    float rr_prob;         /* survival probability (1 = no roulette)        */
    uint64_t seed;         /* counter-based RNG key                         */
    int64_t sample_offset; /* index of the first sample (progressive runs)  */
    int32_t nee;           /* 1 = light sampling + MIS at diffuse vertices
                              (demo_pt default), 0 = passive (BSDF only)    */
    int32_t reserved;
} mcrt_render_params;

MCRT_API const char* mcrt_last_error(void);
MCRT_API const char* mcrt_version(void);

/* Lifecycle.  `device` is a HIP device ordinal.  The context owns one
 * path-tracer workspace and traversal spill area per stream it has run on
 * (shared by its scenes); destroy its scenes and batches before it. */
MCRT_API int mcrt_create(int device, mcrt_ctx** out);
MCRT_API int mcrt_destroy(mcrt_ctx* ctx);
MCRT_API int mcrt_device_name(mcrt_ctx* ctx, char* buf, size_t cap);

/* Scene building (Scene / todo_addMaterial). */
MCRT_API int mcrt_scene_create(mcrt_ctx* ctx, mcrt_scene** out);
MCRT_API int mcrt_scene_destroy(mcrt_scene* scene);
/* Material(c, e, t) + todo_addMaterial: returns the material id. */
MCRT_API int mcrt_material_add(mcrt_scene* scene, int type, const float rgb[3], float glossiness,
                               int32_t* out_id);
/* n materials at once (types, 3 floats rgb, glossiness per material); ids
 * are first_id .. first_id + n - 1. */
MCRT_API int mcrt_materials_add(mcrt_scene* scene, size_t n, const int32_t* types, const float* rgb,
                                const float* glossiness, int32_t* first_id);
/* Bulk replacement of material colours (albedo / emission) after finalize,
 * for optimisation loops; rgb holds 3 floats per material id.  After
 * finalize it waits for the device (work already enqueued on any stream,
 * batch replays included, reads the old colours) and then uploads the new
 * table synchronously. */
MCRT_API int mcrt_material_set_rgb(mcrt_scene* scene, const float* rgb, size_t nmat);
/* Scene::setEnvmap(filename) and setEnvmap(filename, width, height)
 * (scene.h:44-45; EnvmapLoader, envmap.cpp:98-140): a latitude-longitude map
 * of float RGB rows, top-down (row v: polar angle pi v from +y; column u:
 * azimuth 2 pi u).  width = height = 0 reads the file's {int32 width, int32
 * height} header; otherwise the file holds the floats alone.  The map lights
 * the scene from infinity in every render: escaping paths see it, and with
 * nee on each diffuse vertex samples it (map pdf and cosine lobe, MIS;
 * demo_pt.cpp:421-499, 564-589).  Allowed before or after finalize; waits
 * for the device.  A batch keeps the map its scenes had at creation:
 * mcrt_batch_render fails with MCRT_ERR_STATE once a scene's map changed. */
MCRT_API int mcrt_scene_set_envmap(mcrt_scene* scene, const char* path, int32_t width, int32_t height);
/* The same from host memory: rgb holds width * height * 3 floats, rows
 * top-down (width * height <= 2^28); rgb = NULL removes the map.  The
 * luminance must sum to a positive finite value. */
MCRT_API int mcrt_scene_set_envmap_rgb(mcrt_scene* scene, const float* rgb, int32_t width, int32_t height);
/* Scene::add per triangle: v9 = n x (a, b, c), vn9 = n x (na, nb, nc) or
 * NULL for the face normal normalize(cross(a-c, b-c)) at every corner
 * (triangle.cpp:67-74), mat = material id per triangle. */
MCRT_API int mcrt_scene_add_triangles(mcrt_scene* scene, const float* v9, const float* vn9,
                                      const int32_t* mat, size_t n);
/* Scene::loadModel(path, transform, mat): `v x y z` / `f a b c` OBJ,
 * column-major 4x4 transform (NULL = identity), area-weighted vertex
 * normals; out_n (optional) receives the number of faces added. */
MCRT_API int mcrt_scene_load_obj(mcrt_scene* scene, const char* path, const float* m16,
                                 int32_t mat, size_t* out_n);
/* Mesh(path) with the transform and genNormals of Scene::loadModel
 * (src/core/mesh.cpp:24-118, src/core/scene.cpp:368-406), without a scene:
 * the faces as a triangle soup, for callers that give each triangle its own
 * material through mcrt_scene_add_triangles (the reference's OBJ path has
 * one material per file).  v9 and vn9 receive n x 9 floats (corners a, b,
 * c; vertex normals na, nb, nc); `cap` is the number of triangles they
 * hold.  With v9 = NULL only *out_n (the face count) is written.  A `v` or
 * `f` token followed by something other than numbers is MCRT_ERR_IO (the
 * reference's loop never ends on it).  No device or context is needed. */
MCRT_API int mcrt_obj_read(const char* path, const float* m16, float* v9, float* vn9, size_t cap,
                           size_t* out_n);
/* Scene::setCamera(origin, target, W, H, fov_y), up = (0, 1, 0). */
MCRT_API int mcrt_scene_set_camera(mcrt_scene* scene, const float o[3], const float target[3],
                                   int32_t width, int32_t height, float fov_y);
/* Scene::finalize: SBVH build on the host, 4-wide collapse, upload. */
MCRT_API int mcrt_scene_finalize(mcrt_scene* scene, const mcrt_build_opts* opts,
                                 mcrt_build_stats* stats);
MCRT_API int mcrt_scene_num_triangles(const mcrt_scene* scene, size_t* out);
/* TriangleObject::setIntersectionEpsilon (src/core/triangle.cpp:53-56): the
 * absolute determinant cut of the Moller-Trumbore test (a triangle is
 * missed when |det| < eps; default 1e-6, NUM_EPS_DET).  The reference's
 * value is process-global; here it belongs to the scene.  Allowed before or
 * after finalize, but not while work on the scene is in flight; a batch
 * keeps the value its scenes had at creation (mcrt_batch_render fails with
 * MCRT_ERR_STATE once it changed). */
MCRT_API int mcrt_scene_set_intersection_epsilon(mcrt_scene* scene, float eps);

/* Ray queries.  mcrt_intersect = SplitBVH::intersect over each ray's own
 * [tmin, tmax] (SceneBVH::intersect uses tmin = 1e-4, tmax = +inf).
 * mcrt_occluded writes 1 if any hit lies in (tmin, tmax)
 * (SceneBVH::occluded = the same with tmax = tfar - 1e-4). */
MCRT_API int mcrt_intersect(mcrt_scene* scene, const float* rays, size_t n, mcrt_hit* hits,
                            int flags, void* stream);
MCRT_API int mcrt_occluded(mcrt_scene* scene, const float* rays, size_t n, uint8_t* occluded,
                           int flags, void* stream);
/* mcrt_intersect without allocating (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  rays and
 * hits are device pointers (as with MCRT_DEVICE_PTRS) and the call is
 * asynchronous on `stream`; results equal mcrt_intersect's bit for bit.
 * The launch's work counters and traversal stack spill area live in a
 * 64 MiB array of device memory that each GPU the library runs on holds
 * from mcrt_create until the process ends; the ray reorder of large
 * queries is not used.  A scene whose deepest traversal stack needs more
 * spill per wavefront than the array holds is rejected (MCRT_ERR_INVALID_ARG);
 * deep stacks otherwise run on fewer resident waves.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (several streams or host
 * threads) race and must be serialised by the caller. */
MCRT_API int mcrt_intersect_noalloc(mcrt_scene* scene, const float* rays, size_t n, mcrt_hit* hits,
                                    void* stream);
/* Instrumented query (no reference counterpart; measurement only): the
 * same traversal as mcrt_intersect (any_hit = 0) or mcrt_occluded
 * (any_hit = 1), writing no results, accumulating into counters[4] the
 * totals of 4-wide node fetches (128 B each), leaf-reference box fetches
 * (32 B), triangle fetches + Moller-Trumbore tests (32 B) and rays.
 * Synchronous; rays follow `flags` like the queries. */
MCRT_API int mcrt_query_counters(mcrt_scene* scene, const float* rays, size_t n, int any_hit,
                                 uint64_t counters[4], int flags, void* stream);

/* Wavefront path tracer.  film = W*H*3 floats, row y = 0 is the image
 * bottom (film.cpp:204-215); each pixel receives the mean of its samples'
 * clamped radiance (demo_pt.cpp:700-710). */
MCRT_API int mcrt_render(mcrt_scene* scene, const mcrt_render_params* params, float* film,
                         int flags, void* stream);

/* Path-replay adjoint of mcrt_render (no reference counterpart; the
 * differentiable-rendering extension of the integrator).  Given dLoss/dfilm
 * (W*H*3) for the film that mcrt_render(params) produces, writes the
 * gradients with respect to the colour of each triangle's material:
 * d_albedo[3*t+c] for diffuse / glossy triangles and d_emission[3*t+c] for
 * light triangles (ntri*3 floats each, overwritten; the other array holds 0
 * for that triangle).  A material shared by several triangles has the sum
 * of their entries as its gradient.  The same paths are replayed (the RNG is
 * counter-based), so memory does not grow with path length.  `film`
 * (optional, W*H*3) receives the primal film.  At each path vertex the
 * albedo gradient of channel c is formed as delta_c * (radiance gathered
 * after the vertex) / albedo_c in fp32, so channels whose albedo is exactly
 * 0 get no gradient (0, though the true derivative is not), and for small
 * albedos the difference of the path's radiance sums loses about
 * 6e-8 * L_path / (albedo * L_after) relative accuracy (1e-3 at albedo 1e-4
 * when the radiance after the vertex is comparable to the path's). */
MCRT_API int mcrt_render_adjoint(mcrt_scene* scene, const mcrt_render_params* params,
                                 const float* dL_dfilm, float* d_albedo, float* d_emission,
                                 float* film, int flags, void* stream);

/* Batched rendering (no reference counterpart: the reference renders one
 * scene per process): n finalized scenes of one context, each with its own
 * camera (film size), parameters and film, rendered together -- the paths
 * of all scenes share the wavefront chunks and every stage is one launch
 * for all of them.  Every film equals what mcrt_render gives for that
 * scene alone, bit for bit.  Traversal tuning and chunk size are those of
 * scenes[0].  The scenes must stay alive and unmodified (material colours
 * aside) while the batch exists.
 *   mcrt_batch_create: films[i] are device pointers (W_i*H_i*3 floats) and
 *     `flags` must include MCRT_DEVICE_PTRS; MCRT_BATCH_GRAPH captures the
 *     launch sequence in a hipGraph on the first mcrt_batch_render.
 *   mcrt_batch_render: renders every film, asynchronously on `stream`.
 *   mcrt_batch_info: total paths and wavefront chunks of the batch.
 *   mcrt_render_batch: one-shot create + render + destroy; films are host
 *     pointers (synchronous) or, with MCRT_DEVICE_PTRS, device pointers. */
MCRT_API int mcrt_batch_create(size_t n, mcrt_scene* const* scenes, const mcrt_render_params* params,
                               float* const* films, int flags, mcrt_batch** out);
MCRT_API int mcrt_batch_render(mcrt_batch* batch, void* stream);
MCRT_API int mcrt_batch_info(const mcrt_batch* batch, int64_t* total_paths, int64_t* chunks);
MCRT_API int mcrt_batch_destroy(mcrt_batch* batch);
MCRT_API int mcrt_render_batch(size_t n, mcrt_scene* const* scenes, const mcrt_render_params* params,
                               float* const* films, int flags, void* stream);

/* Film output (host memory, W*H*3 floats, row y = 0 at the image bottom);
 * byte-identical to the reference's FrameBuffer writers on finite,
 * non-negative films (NaN or negative channels are written as 0, where the
 * reference's conversion is undefined).  No device or context is needed.
 *   mcrt_film_write_hdr     FrameBuffer::dumpHDR       src/core/film.cpp:65-108
 *   mcrt_film_write_ppm     FrameBuffer::dumpPPM       src/core/film.cpp:9-33
 *   mcrt_film_tonemap_gamma FrameBuffer::tonemapGamma  src/core/film.cpp:195-202
 *                           (in place, n floats)                               */
MCRT_API int mcrt_film_write_hdr(const char* path, const float* film, int32_t width, int32_t height);
MCRT_API int mcrt_film_write_ppm(const char* path, const float* film, int32_t width, int32_t height);
MCRT_API int mcrt_film_tonemap_gamma(float* film, size_t n, float gamma);

/* Launch tuning of a scene (no reference counterpart: the reference's CPU
 * traversal has no launch parameters).  Results never depend on it.  In
 * mcrt_scene_set_tuning a zero field keeps the current value; it may be
 * called before or after finalize, but not while work on the scene is in
 * flight. */
typedef struct mcrt_tuning {
    int32_t trace_waves_per_cu; /* traversal workgroups (one wavefront each) per CU,
                                   1..1024; unset, finalize picks 32 (all resident)
                                   when the BVH fits in 192 MB, else 28.  The waves
                                   claim rays from work counters, so more than fit
                                   at once only adds launch overhead; any-hit
                                   queries use at most 28, what their kernel keeps
                                   resident                                          */
    int32_t trace_lds_depth;    /* traversal stack entries per lane kept in LDS: 8, 12,
                                   16 or 24 (8); deeper entries go to a per-stream
                                   global spill area                                    */
    int64_t pt_chunk_paths;     /* paths per wavefront chunk of the renders, 1..2^28
                                   (2^26); path state is ~0.5 KB per path              */
    int32_t pt_trace_waves_per_cu; /* the same two settings for the traversal
                                      launches of renders, whose ray sets (many
                                      short shadow rays) favour more waves: (as
                                      trace_waves_per_cu)                        */
    int32_t pt_trace_lds_depth;    /* (8)                                         */
    int64_t trace_reorder_rays;    /* a query of at least this many rays is
                                      traversed in the order of its rays' origin
                                      cells (16^3 over the scene's box, a counting
                                      sort on the device; results unchanged);
                                      -1: never.  Unset, finalize picks 2^20 when
                                      the BVH exceeds 192 MB (larger than the GPU's
                                      last-level cache), else -1                 */
    int32_t trace_refill_lanes;    /* compaction threshold of the queries' traversal,
                                      1..64 (16): a wave hands new rays to its idle
                                      lanes once at least this many of its 64 lanes
                                      are idle, so refills come in fuller rounds    */
    int32_t pt_trace_refill_lanes; /* the same for the traversal launches of renders (16) */
    int32_t trace_any_refill_lanes; /* the same for any-hit queries (mcrt_occluded), whose
                                       rays end at their first hit (32)            */
} mcrt_tuning;
MCRT_API int mcrt_scene_set_tuning(mcrt_scene* scene, const mcrt_tuning* tuning);
MCRT_API int mcrt_scene_get_tuning(const mcrt_scene* scene, mcrt_tuning* tuning);

/* Autotuner (no reference counterpart): times closest-hit over the n rays
 * for trace_waves_per_cu in {8,16,24,28,32,64,128} x trace_lds_depth in
 * {8,12,16} and keeps the fastest, then trace_refill_lanes in {1,8,16,32}
 * with the chosen pair, then trace_any_refill_lanes in {16,32,48} (any-hit
 * over the same rays), then trace_reorder_rays -1 against min(n, 2^20)
 * (the caller's ray order against origin-cell order), keeping the fastest
 * each time; with `render` != NULL also times one render per
 * pt_trace_waves_per_cu in {16,28,32,64} x pt_trace_lds_depth in {8,16}, then
 * per pt_trace_refill_lanes in {1,8,16,32}, then per pt_chunk_paths in
 * 2^22..2^26, keeping the fastest.  The chosen tuning is applied to the
 * scene and returned in `best`, the best closest-hit rate in
 * `best_mrays_per_s` (either may be NULL); a failure restores the scene's
 * tuning.  Rays follow `flags` like the queries; synchronous. */
MCRT_API int mcrt_autotune(mcrt_scene* scene, const float* rays, size_t n,
                           const mcrt_render_params* render, int flags, void* stream,
                           mcrt_tuning* best, double* best_mrays_per_s);
/* Autotuner of a batch (no reference counterpart): the render half of
 * mcrt_autotune timed on the whole batch -- pt_trace_waves_per_cu in
 * {16,28,32,64} x pt_trace_lds_depth in {8,16}, then pt_trace_refill_lanes in
 * {1,8,16,32}, then pt_chunk_paths in 2^22..2^26, one batch render per
 * setting into the batch's films, keeping the fastest each time.  The
 * chosen tuning is applied to scenes[0] (whose tuning the batch uses) and
 * the batch re-initialised with it; a captured graph is dropped and
 * captured again by the next mcrt_batch_render.  Returns the tuning in
 * `best` and the fastest render's milliseconds in `best_ms` (either may be
 * NULL).  Waits for the device first; synchronous.  A failure restores
 * scenes[0]'s tuning. */
MCRT_API int mcrt_batch_autotune(mcrt_batch* batch, void* stream, mcrt_tuning* best, double* best_ms);

#ifdef __cplusplus
}
#endif
#endif /* MCRT_H */
