// The traversal kernel template, shared by the query entry points
// (traverse.hip) and the path tracer's extend step (pt.hip).
//
// One lane per ray, grid-strided over the batch, one wavefront per
// workgroup (its LDS stack slab is private to the wave).  Two ray sources:
//   * SEG = false: rays[0, n) with n known to the host;
//   * SEG = true : two segments whose sizes live in device memory (the
//     path tracer's compacted continuation rays [0, cnt[0]) and its
//     next-event rays [seg1, seg1 + cnt[1])), so the bounce loop needs no
//     host round trip and can be captured in a graph.
#pragma once
#include "trace_core.h"

namespace mcrt
{
constexpr int kWave = 64;

template <int D, bool ANY, bool COUNT, bool SEG>
__global__ __launch_bounds__(kWave) void k_trace(DevScene S, const float4* __restrict__ rays, int64_t n,
                                                 const uint32_t* __restrict__ seg_cnt, int64_t seg1,
                                                 int4* __restrict__ hits, uint8_t* __restrict__ occ,
                                                 unsigned long long* __restrict__ counters)
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
    st.spill_lane = (int64_t)blockIdx.x * kWave + threadIdx.x;
    st.spill_depth = S.spill_depth;
    int64_t n0 = n, total = n;
    if (SEG)
    {
        n0 = seg_cnt[0];
        total = n0 + seg_cnt[1];
    }
    TraceCount cnt;
    uint32_t nrays = 0;
    for (int64_t k = (int64_t)blockIdx.x * kWave + threadIdx.x; k < total; k += (int64_t)gridDim.x * kWave)
    {
        const int64_t i = (SEG && k >= n0) ? seg1 + (k - n0) : k;
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        RayPrep r = prep_ray(a.x, a.y, a.z, b.x, b.y, b.z, a.w);
        float t, u, v;
        int prim = trace<ANY, COUNT>(S, st, r, b.w, t, u, v, cnt);
        if (ANY)
            occ[i] = prim >= 0 ? 1 : 0;
        else
            hits[i] = make_int4(prim, (int)bits_of(t), (int)bits_of(u), (int)bits_of(v));
        nrays++;
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
}  // namespace mcrt
