// Device-side scene handle and launch helpers shared by the HIP sources
// and the host C-ABI layer.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "layout.h"

namespace mcrt
{
struct DevMaterial
{
    float rgb[3];
    int32_t type;  // mcrt_material_type
    float gloss;
    float pad[3];
};

struct DevLight
{
    int32_t tri;
    float area;  // TriangleObject::area
    float pdf;   // DiscreteSampler pdf
    float cdf;   // DiscreteSampler cdf
};

struct DevCamera
{
    float o[3], fwd[3], fx[3], fy[3];
    float dist, ww;
    int32_t w, h;
};

struct DevScene
{
    const QNode* nodes = nullptr;      // the BVH's 64 B records: QNodes, then WRefs (layout.h)
    const TriShade* shade = nullptr;   // per triangle
    const float* verts = nullptr;      // per triangle a, b, c (9 floats)
    const DevMaterial* mats = nullptr;
    const DevLight* lights = nullptr;
    int32_t nlights = 0;
    int32_t stack_bound = 0;
    const int32_t* hot_tri = nullptr;  // triangle of each gradient hot slot
    int32_t nhot = 0;
    int32_t refill_lanes = 1;          // a traversal wave claims new rays once this many of its lanes are idle
    int32_t refill_lanes_any = 1;      // the same for any-hit queries
    float det_eps = 1e-6f;             // Moller-Trumbore |det| cut (TriangleObject::s_intersection_epsilon)
    float node_span = 0.0f;            // qnode_span of the nodes (wide.h): scale of the node-test margin
    int64_t ntri = 0;
    uint64_t* spill = nullptr;         // traversal stack overflow storage
    int64_t spill_lanes = 0;           // lanes the spill buffer covers = resident lanes of a launch
    int64_t lanes_any = 0;             // lanes of an any-hit launch: at most spill_lanes, and at most
                                       // what the any-hit kernel keeps resident (kTraceAnyWavesPerSimd)
    int32_t spill_depth = 0;           // entries per lane beyond the LDS stack
    int32_t lds_depth = 16;            // LDS stack entries per lane (8, 12, 16 or 24)
    uint32_t* wq = nullptr;            // ray-query work counters, kWqWords words (per stream)
    // environment map (env_w == 0: none)
    const float* env_px = nullptr;     // per pixel {r, g, b, luminance}, rows top-down
    const float* env_cdf_u = nullptr;  // CDF over the columns (env_w)
    const float* env_cdf_v = nullptr;  // CDF over the rows of each column (env_w x env_h, column-major)
    int32_t env_w = 0, env_h = 0;
    float env_norm = 0.0f;             // 1 / sum of the sin-warped luminance
    // Ray order of large queries (traverse.hip): a query of at least
    // reorder_rays rays (0: never) is traversed in the order of its rays'
    // origin cells, a kCellsPerAxis^3 grid over the scene's box (cell_lo,
    // and cells per unit length in cell_sc).  order (a permutation of the
    // rays) and cell_work (kCellWorkWords) are per-stream buffers.
    int64_t reorder_rays = 0;
    float cell_lo[3] = {0.0f, 0.0f, 0.0f};
    float cell_sc[3] = {0.0f, 0.0f, 0.0f};
    uint32_t* order = nullptr;
    uint32_t* cell_work = nullptr;
    // Union of the boxes of every reference to an emitter (LGHT) triangle
    // (emit_any == 0: there is none).  A ray that fails the traversal's slab
    // test against it fails it against every emitter reference, so its
    // closest hit cannot be an emitter.
    int32_t emit_any = 0;
    float emit_lo[3] = {0.0f, 0.0f, 0.0f};
    float emit_hi[3] = {0.0f, 0.0f, 0.0f};
};

// The part of a scene a traversal reads (batched launches look it up per ray).
struct SceneRoot
{
    const QNode* nodes = nullptr;  // the scene's BVH records (DevScene::nodes)
    float det_eps = 1e-6f;  // the scene's Moller-Trumbore determinant cut
    float node_span = 0.0f; // qnode_span of the scene's nodes
};

// Waves per SIMD the traversal kernels with an 8-entry LDS stack are
// compiled to keep resident (four SIMDs per CU): closest hit 8 (64 VGPRs),
// any hit 7 (held to eight waves' registers it spills scalar registers).
constexpr int kTraceWavesPerSimd = 8;
constexpr int kTraceAnyWavesPerSimd = 7;

// Work counters of a traversal launch: one per slice of the rays, 64 B
// apart (trace_kernel.h).
constexpr uint32_t kSlices = 8;
constexpr uint32_t kSliceStride = 16;  // uint32 words between slice counters
constexpr uint32_t kWqWords = kSlices * kSliceStride;

// Origin cells of the query reorder: 16 per axis (Morton-ordered), counted
// and scattered by kCellBlocks workgroups.  cell_work holds the workgroups'
// offsets inside each cell ([kCellBlocks][kCells]), then the cell totals
// and the cell bases (kCells each), then a flag set when one cell holds
// every ray (the order would not change: the reorder is skipped).
constexpr uint32_t kCellsPerAxis = 16;
constexpr uint32_t kCells = kCellsPerAxis * kCellsPerAxis * kCellsPerAxis;
constexpr uint32_t kCellBlocks = 256;
constexpr size_t kCellTotals = (size_t)kCellBlocks * kCells;
constexpr size_t kCellBases = kCellTotals + kCells;
constexpr size_t kCellSkip = kCellBases + kCells;
constexpr size_t kCellWorkWords = kCellSkip + 64;

// Traversal modes.
enum TraceMode
{
    TRACE_CLOSEST = 0,
    TRACE_ANY = 1
};

// Launches a query over n rays (8 floats each, device memory).  Closest:
// hits = int4 {prim, t, u, v} per ray.  Any: occ = 1 byte per ray.
// counters (optional, device, 4 x uint64): node fetches, ref fetches,
// MT tests, rays.
hipError_t launch_trace(const DevScene& s, const float* rays, int64_t n, void* hits, uint8_t* occ,
                        int mode, uint64_t* counters, hipStream_t stream);

// Device memory of the allocation-free closest-hit query
// (mcrt_intersect_noalloc): a file-scope array in traverse.hip, one copy per
// GPU, holding the launch's work counters (the first kScratchWqBytes) and
// then its traversal stack spill area.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchWqBytes = 512;
static_assert(kWqWords * sizeof(uint32_t) <= kScratchWqBytes, "work counters fit their scratch slot");

// The device address of that array on `device` (which must be current),
// looked up once per device; 0 = MCRT_OK, else a status with the message
// set.
int scratch_address(int device, void** out);

// LDS stack depths the traversal kernels are instantiated for.
inline bool valid_lds_depth(int d) { return d == 8 || d == 12 || d == 16 || d == 24; }
}  // namespace mcrt
