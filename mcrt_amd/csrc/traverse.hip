// Ray-query launches (mcrt_intersect / mcrt_occluded / mcrt_query_counters).
#include "trace_kernel.h"

namespace mcrt
{
namespace
{
__global__ void k_wq_reset(uint32_t* wq)
{
    if (threadIdx.x < kSlices) wq[threadIdx.x * kSliceStride] = 0;
}

template <bool ANY, bool COUNT>
hipError_t launch(const DevScene& s, const float* rays, int64_t n, void* hits, uint8_t* occ, uint64_t* counters,
                  hipStream_t stream)
{
    // a kernel, not hipMemsetAsync, so a caller may capture the query in a graph
    hipLaunchKernelGGL(k_wq_reset, dim3(1), dim3(64), 0, stream, s.wq);
    launch_k_trace<ANY, COUNT, false>(s, trace_blocks(s, n), stream, reinterpret_cast<const float4*>(rays), n,
                                      nullptr, (int64_t)0, reinterpret_cast<int4*>(hits), occ,
                                      reinterpret_cast<unsigned long long*>(counters), s.wq);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_trace(const DevScene& s, const float* rays, int64_t n, void* hits, uint8_t* occ, int mode,
                        uint64_t* counters, hipStream_t stream)
{
    const bool any = mode == TRACE_ANY;
    const int64_t kMaxLaunch = int64_t(1) << 30;  // the kernel indexes rays with 32 bits
    for (int64_t off = 0; off < n; off += kMaxLaunch)
    {
        const int64_t m = n - off < kMaxLaunch ? n - off : kMaxLaunch;
        const float* r = rays + 8 * off;
        void* h = any ? nullptr : static_cast<void*>(static_cast<int4*>(hits) + off);
        uint8_t* o = any ? occ + off : nullptr;
        hipError_t e = counters ? (any ? launch<true, true>(s, r, m, h, o, counters, stream)
                                       : launch<false, true>(s, r, m, h, o, counters, stream))
                                : (any ? launch<true, false>(s, r, m, h, o, counters, stream)
                                       : launch<false, false>(s, r, m, h, o, counters, stream));
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
}  // namespace mcrt
