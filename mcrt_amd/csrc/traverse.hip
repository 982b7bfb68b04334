// Ray-query launches (mcrt_intersect / mcrt_occluded / mcrt_query_counters).
#include "trace_kernel.h"

namespace mcrt
{
namespace
{
constexpr int kLdsDepth = 16;
constexpr int kBlocksPerCU = 16;

template <bool ANY, bool COUNT>
hipError_t launch(const DevScene& s, const float* rays, int64_t n, void* hits, uint8_t* occ, uint64_t* counters,
                  hipStream_t stream)
{
    hipLaunchKernelGGL((k_trace<kLdsDepth, ANY, COUNT, false>), dim3((unsigned)trace_blocks(s, n)), dim3(kWave), 0,
                       stream, s, reinterpret_cast<const float4*>(rays), n, nullptr, (int64_t)0,
                       reinterpret_cast<int4*>(hits), occ, reinterpret_cast<unsigned long long*>(counters));
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
