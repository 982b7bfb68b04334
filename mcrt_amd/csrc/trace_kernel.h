// The traversal kernel template, shared by the query entry points
// (traverse.hip) and the path tracer's extend step (pt.hip).
//
// One lane per ray at a time, grid-strided over the batch, one wavefront
// per workgroup (its LDS stack slab is private to the wave).  Two ray sources:
//   * SEG = false: rays[0, n) with n known to the host;
//   * SEG = true : two segments whose sizes live in device memory (the
//     path tracer's compacted continuation rays [0, cnt[0]) and its
//     next-event rays [seg1, seg1 + cnt[1])), so the bounce loop needs no
//     host round trip and can be captured in a graph.
// BATCH = true: the rays belong to several scenes (batched rendering); ray
// i is traversed in the BVH of scene rsid[i].
#pragma once
#include "trace_core.h"

namespace mcrt
{
constexpr int kWave = 64;

// Rays of several scenes in one launch: ray i belongs to scene rsid[i].
struct TraceBatch
{
    const SceneRoot* roots = nullptr;
    const uint16_t* rsid = nullptr;
};

template <int D, bool ANY, bool COUNT, bool SEG, bool BATCH>
__global__ __launch_bounds__(kWave) void k_trace(DevScene S, const float4* __restrict__ rays, int64_t n,
                                                 const uint32_t* __restrict__ seg_cnt, int64_t seg1,
                                                 int4* __restrict__ hits, uint8_t* __restrict__ occ,
                                                 unsigned long long* __restrict__ counters, TraceBatch B)
{
    __shared__ uint32_t s_node[D * kWave];
    __shared__ float s_lo[D * kWave];
    TraceStack st;
    st.node = s_node;
    st.lo = s_lo;
    st.stride = kWave;
    st.lane = threadIdx.x;
    st.depth = D;
    st.spill = S.spill;
    st.spill_lane = blockIdx.x * kWave + threadIdx.x;
    st.spill_depth = S.spill_depth;
    // 32-bit ray indices: a launch covers fewer than 2^31 rays (launch_trace
    // splits larger batches; path-tracer chunks are at most 3 * 2^28 rays).
    uint32_t n0 = (uint32_t)n, total = (uint32_t)n;
    if (SEG)
    {
        n0 = seg_cnt[0];
        total = n0 + seg_cnt[1];
    }
    TraceCount cnt;
    uint32_t nrays = 0;
    const uint32_t stride = gridDim.x * kWave;
    uint32_t k = blockIdx.x * kWave + threadIdx.x;
    // Lane-persistent loop: each lane walks its rays k, k + stride, ... and
    // starts the next one in the same iteration in which the previous ends,
    // so divergent ray lengths do not idle the wavefront.
    Tracer<ANY, COUNT> tr;
    SceneRoot root;
    root.nodes = S.nodes;
    root.refs = S.refs;
    uint32_t i = 0;
    bool live = k < total;
    if (live)
    {
        i = (SEG && k >= n0) ? (uint32_t)seg1 + (k - n0) : k;
        if (BATCH) root = B.roots[B.rsid[i]];
        const float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
        tr.start(prep_ray(a.x, a.y, a.z, b.x, b.y, b.z, a.w), b.w);
    }
    while (live)
    {
        if (root.nodes != nullptr && !tr.step(root, st, cnt)) continue;
        float t, u, v;
        const int prim = tr.result(t, u, v);
        if (ANY)
            occ[i] = prim >= 0 ? 1 : 0;
        else
            hits[i] = make_int4(prim, (int)bits_of(t), (int)bits_of(u), (int)bits_of(v));
        nrays++;
        k += stride;
        live = k < total;
        if (live)
        {
            i = (SEG && k >= n0) ? (uint32_t)seg1 + (k - n0) : k;
            if (BATCH) root = B.roots[B.rsid[i]];
            const float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
            tr.start(prep_ray(a.x, a.y, a.z, b.x, b.y, b.z, a.w), b.w);
        }
    }
    if (COUNT)
    {
        atomicAdd(&counters[0], (unsigned long long)cnt.nodes);
        atomicAdd(&counters[1], (unsigned long long)cnt.refs);
        atomicAdd(&counters[2], (unsigned long long)cnt.tris);
        atomicAdd(&counters[3], (unsigned long long)nrays);
    }
}

// Workgroups a trace launch uses: enough to cover n rays, at most the
// number of lanes the spill buffer covers.
inline int64_t trace_blocks(const DevScene& s, int64_t n)
{
    int64_t max_blocks = s.spill_lanes / kWave;
    int64_t blocks = (n + kWave - 1) / kWave;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    return blocks;
}

// Launches k_trace with the scene's LDS stack depth on `blocks` workgroups.
template <bool ANY, bool COUNT, bool SEG, bool BATCH = false>
inline void launch_k_trace(const DevScene& S, int64_t blocks, hipStream_t st, const float4* rays, int64_t n,
                           const uint32_t* seg_cnt, int64_t seg1, int4* hits, uint8_t* occ,
                           unsigned long long* counters, TraceBatch B = TraceBatch())
{
    const dim3 g((unsigned)blocks), b(kWave);
    switch (S.lds_depth)
    {
    case 8:
        hipLaunchKernelGGL((k_trace<8, ANY, COUNT, SEG, BATCH>), g, b, 0, st, S, rays, n, seg_cnt, seg1, hits, occ,
                           counters, B);
        break;
    case 12:
        hipLaunchKernelGGL((k_trace<12, ANY, COUNT, SEG, BATCH>), g, b, 0, st, S, rays, n, seg_cnt, seg1, hits, occ,
                           counters, B);
        break;
    case 24:
        hipLaunchKernelGGL((k_trace<24, ANY, COUNT, SEG, BATCH>), g, b, 0, st, S, rays, n, seg_cnt, seg1, hits, occ,
                           counters, B);
        break;
    default:
        hipLaunchKernelGGL((k_trace<16, ANY, COUNT, SEG, BATCH>), g, b, 0, st, S, rays, n, seg_cnt, seg1, hits, occ,
                           counters, B);
        break;
    }
}
}  // namespace mcrt
