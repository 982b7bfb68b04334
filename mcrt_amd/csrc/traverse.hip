// Ray-query kernels (mcrt_intersect / mcrt_occluded): one lane per ray,
// grid-strided over the batch, one wavefront per workgroup so each wave
// owns its LDS stack slab.
#include "trace_core.h"

namespace mcrt
{
namespace
{
constexpr int kWave = 64;
constexpr int kLdsDepth = 16;
constexpr int kBlocksPerCU = 16;

template <int D, bool ANY, bool COUNT>
__global__ __launch_bounds__(kWave) void k_trace(DevScene S, const float4* __restrict__ rays, int64_t n,
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
    TraceCount cnt;
    uint32_t nrays = 0;
    for (int64_t i = (int64_t)blockIdx.x * kWave + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWave)
    {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        RayPrep r = prep_ray(a.x, a.y, a.z, b.x, b.y, b.z, a.w);
        float t, u, v;
        int prim = trace<ANY, COUNT>(S, st, r, b.w, t, u, v, cnt);
        if (ANY)
            occ[i] = prim >= 0 ? 1 : 0;
        else
            hits[i] = make_int4(prim, __float_as_int(t), __float_as_int(u), __float_as_int(v));
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

template <bool ANY, bool COUNT>
hipError_t launch(const DevScene& s, const float* rays, int64_t n, void* hits, uint8_t* occ,
                  uint64_t* counters, hipStream_t stream)
{
    int64_t max_blocks = s.spill_lanes / kWave;
    int64_t blocks = (n + kWave - 1) / kWave;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_trace<kLdsDepth, ANY, COUNT>), dim3((unsigned)blocks), dim3(kWave), 0, stream, s,
                       reinterpret_cast<const float4*>(rays), n, reinterpret_cast<int4*>(hits), occ,
                       reinterpret_cast<unsigned long long*>(counters));
    return hipGetLastError();
}
}  // namespace

int64_t trace_lanes(int num_cus) { return (int64_t)num_cus * kBlocksPerCU * kWave; }

hipError_t launch_trace(const DevScene& s, const float* rays, int64_t n, void* hits, uint8_t* occ, int mode,
                        uint64_t* counters, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const bool any = mode == TRACE_ANY;
    if (counters)
        return any ? launch<true, true>(s, rays, n, hits, occ, counters, stream)
                   : launch<false, true>(s, rays, n, hits, occ, counters, stream);
    return any ? launch<true, false>(s, rays, n, hits, occ, counters, stream)
               : launch<false, false>(s, rays, n, hits, occ, counters, stream);
}
}  // namespace mcrt
