// Ray-query launches (mcrt_intersect / mcrt_occluded / mcrt_query_counters).
#include <map>
#include <mutex>

#include "capi_internal.h"
#include "trace_kernel.h"

namespace mcrt
{
namespace
{
// The allocation-free query's device memory (kScratchBytes, device.h).  In
// host code `scratch` is only a stand-in for the symbol: its device address
// comes from hipGetSymbolAddress (scratch_address), per GPU.
static __device__ __align__(256) uint64_t scratch[kScratchBytes / sizeof(uint64_t)];

__global__ void k_wq_reset(uint32_t* wq)
{
    if (threadIdx.x < kSlices) wq[threadIdx.x * kSliceStride] = 0;
}

// ---- traversal order of large queries -------------------------------------
// Over a BVH larger than the GPU's caches (the 4M-triangle soup's is 558 MB,
// the Infinity Cache 256 MB), rays in the caller's order make every
// wavefront fetch nodes from all over the tree.  Traversed in the order of
// their origins' cells instead (a counting sort into 16^3 Morton-ordered
// cells over the scene's box), neighbouring lanes and neighbouring waves
// share most of their upper nodes.  A ray's result does not depend on when
// it is traversed, so results are unchanged; hits are written at the ray's
// own index.
constexpr uint32_t kCellTPB = 512;
static_assert(kCellsPerAxis == 16, "spread_cell interleaves 4 bits per axis");

// The 4 bits of v moved to bits 0, 3, 6, 9 (Morton interleave).
__device__ __forceinline__ uint32_t spread_cell(uint32_t v)
{
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6);
}

// Cell of a ray origin; origins outside the box go to the border cells, and
// fmaxf returns its other operand for a NaN, so every origin has a cell.
__device__ __forceinline__ uint32_t origin_cell(const DevScene& S, const float4 o)
{
    const float top = (float)(kCellsPerAxis - 1);
    const uint32_t x = (uint32_t)fminf(fmaxf((o.x - S.cell_lo[0]) * S.cell_sc[0], 0.0f), top);
    const uint32_t y = (uint32_t)fminf(fmaxf((o.y - S.cell_lo[1]) * S.cell_sc[1], 0.0f), top);
    const uint32_t z = (uint32_t)fminf(fmaxf((o.z - S.cell_lo[2]) * S.cell_sc[2], 0.0f), top);
    return spread_cell(x) | (spread_cell(y) << 1) | (spread_cell(z) << 2);
}

// Rays [b0, b1) of workgroup blockIdx.x in the count and scatter passes
// (the same split in both, so their per-workgroup counts agree).
__device__ __forceinline__ void cell_chunk(uint32_t n, uint32_t& b0, uint32_t& b1)
{
    const uint32_t chunk = (n + kCellBlocks - 1) / kCellBlocks;
    b0 = min(n, blockIdx.x * chunk);
    b1 = min(n, b0 + chunk);
}

// Zeroes the cell totals and the skip flag (a kernel, not hipMemsetAsync:
// graph-capturable).
__global__ void k_cell_reset(uint32_t* work)
{
    for (uint32_t c = threadIdx.x; c < kCells; c += blockDim.x) work[kCellTotals + c] = 0;
    if (threadIdx.x == 0) work[kCellSkip] = 0;
}

// Sets the skip flag when 1024 rays spread evenly over the batch share one
// cell (a camera's rays, which are in order already): then no pass below
// does any work.  Any order gives the same results, so a sample decides.
__global__ __launch_bounds__(1024) void k_cell_sample(DevScene S, const float4* __restrict__ rays, uint32_t n)
{
    __shared__ uint32_t c0, differ;
    const uint32_t i = (uint32_t)(((uint64_t)threadIdx.x * n) / 1024u);  // < n
    const uint32_t c = origin_cell(S, rays[2 * (size_t)i]);
    if (threadIdx.x == 0)
    {
        c0 = c;
        differ = 0;
    }
    __syncthreads();
    if (c != c0) differ = 1;
    __syncthreads();
    if (threadIdx.x == 0 && differ == 0) S.cell_work[kCellSkip] = 1;
}

// Per workgroup: its rays per cell (LDS histogram), then one global atomic
// per non-empty cell, whose old value is the workgroup's offset inside the
// cell (the ranges of the workgroups are disjoint, in any order).
__global__ __launch_bounds__(kCellTPB) void k_cell_count(DevScene S, const float4* __restrict__ rays, uint32_t n)
{
    __shared__ uint32_t h[kCells];
    if (S.cell_work[kCellSkip] != 0) return;  // uniform: every workgroup returns
    uint32_t* offs = S.cell_work;
    uint32_t* totals = S.cell_work + kCellTotals;
    for (uint32_t c = threadIdx.x; c < kCells; c += kCellTPB) h[c] = 0;
    __syncthreads();
    uint32_t b0, b1;
    cell_chunk(n, b0, b1);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += kCellTPB) atomicAdd(&h[origin_cell(S, rays[2 * (size_t)i])], 1u);
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < kCells; c += kCellTPB)
    {
        const uint32_t k = h[c];
        offs[(size_t)blockIdx.x * kCells + c] = k != 0 ? atomicAdd(&totals[c], k) : 0u;
    }
}

// Exclusive prefix sum of the cell totals: the first slot of every cell.
// When one cell holds all n rays, sets the skip flag: the order would not
// change.
__global__ __launch_bounds__(1024) void k_cell_bases(DevScene S, uint32_t n)
{
    constexpr uint32_t per = kCells / 1024;
    __shared__ uint32_t part[1024];
    // every thread reads the flag before the barrier, and threads set it
    // only after it, so all of them see the same value and return together
    const bool skip = S.cell_work[kCellSkip] != 0;
    __syncthreads();
    if (skip) return;
    const uint32_t* totals = S.cell_work + kCellTotals;
    uint32_t* bases = S.cell_work + kCellBases;
    const uint32_t t = threadIdx.x;
    uint32_t v[per], sum = 0;
    for (uint32_t k = 0; k < per; k++)
    {
        v[k] = totals[t * per + k];
        sum += v[k];
        if (v[k] == n) S.cell_work[kCellSkip] = 1;
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1)  // inclusive scan of the parts
    {
        const uint32_t add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint32_t run = t > 0 ? part[t - 1] : 0u;
    for (uint32_t k = 0; k < per; k++)
    {
        bases[t * per + k] = run;
        run += v[k];
    }
}

// Each ray's slot: its cell's base + its workgroup's offset in the cell + a
// running count in LDS.  order becomes a permutation of [0, n).
__global__ __launch_bounds__(kCellTPB) void k_cell_scatter(DevScene S, const float4* __restrict__ rays, uint32_t n)
{
    __shared__ uint32_t cur[kCells];
    if (S.cell_work[kCellSkip] != 0) return;  // uniform: every workgroup returns
    const uint32_t* offs = S.cell_work + (size_t)blockIdx.x * kCells;
    const uint32_t* bases = S.cell_work + kCellBases;
    for (uint32_t c = threadIdx.x; c < kCells; c += kCellTPB) cur[c] = bases[c] + offs[c];
    __syncthreads();
    uint32_t b0, b1;
    cell_chunk(n, b0, b1);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += kCellTPB)
        S.order[atomicAdd(&cur[origin_cell(S, rays[2 * (size_t)i])], 1u)] = i;
}

template <bool ANY, bool COUNT>
hipError_t launch(const DevScene& s, const float* rays, int64_t n, void* hits, uint8_t* occ, uint64_t* counters,
                  hipStream_t stream)
{
    // kernels, not hipMemsetAsync, so a caller may capture the query in a graph
    hipLaunchKernelGGL(k_wq_reset, dim3(1), dim3(64), 0, stream, s.wq);
    const float4* r4 = reinterpret_cast<const float4*>(rays);
    int4* h4 = reinterpret_cast<int4*>(hits);
    unsigned long long* c = reinterpret_cast<unsigned long long*>(counters);
    if (s.order != nullptr && s.reorder_rays > 0 && n >= s.reorder_rays)
    {
        hipLaunchKernelGGL(k_cell_reset, dim3(1), dim3(1024), 0, stream, s.cell_work);
        hipLaunchKernelGGL(k_cell_sample, dim3(1), dim3(1024), 0, stream, s, r4, (uint32_t)n);
        hipLaunchKernelGGL(k_cell_count, dim3(kCellBlocks), dim3(kCellTPB), 0, stream, s, r4, (uint32_t)n);
        hipLaunchKernelGGL(k_cell_bases, dim3(1), dim3(1024), 0, stream, s, (uint32_t)n);
        hipLaunchKernelGGL(k_cell_scatter, dim3(kCellBlocks), dim3(kCellTPB), 0, stream, s, r4, (uint32_t)n);
        TraceBatch tb;
        tb.order = s.order;
        tb.skip = s.cell_work + kCellSkip;
        launch_k_trace<ANY, COUNT, false, false, true>(s, trace_blocks(s, n), stream, r4, n, nullptr, (int64_t)0, (int64_t)0, h4,
                                                       occ, c, s.wq, tb);
    }
    else
        launch_k_trace<ANY, COUNT, false>(s, trace_blocks(s, n), stream, r4, n, nullptr, (int64_t)0, (int64_t)0, h4, occ, c,
                                          s.wq);
    return hipGetLastError();
}
}  // namespace

int scratch_address(int device, void** out)
{
    static std::mutex mu;
    static std::map<int, void*> addr;  // device ordinal -> address of `scratch` there
    std::lock_guard<std::mutex> g(mu);
    const auto it = addr.find(device);
    if (it != addr.end())
    {
        *out = it->second;
        return MCRT_OK;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
    {
        (void)hipGetLastError();
        return fail(MCRT_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    }
    addr[device] = p;
    *out = p;
    return MCRT_OK;
}

hipError_t launch_trace(const DevScene& s, const float* rays, int64_t n, void* hits, uint8_t* occ, int mode,
                        uint64_t* counters, hipStream_t stream)
{
    const bool any = mode == TRACE_ANY;
    // any-hit queries have their own compaction threshold (their rays end at
    // the first hit, so lanes free up faster), and a grid of at most the
    // waves their kernel keeps resident (waves claim rays, so workgroups
    // beyond those only start late)
    DevScene sa = s;
    if (any)
    {
        sa.refill_lanes = s.refill_lanes_any;
        if (s.lanes_any < sa.spill_lanes) sa.spill_lanes = s.lanes_any;
    }
    const int64_t kMaxLaunch = int64_t(1) << 30;  // the kernel indexes rays with 32 bits
    for (int64_t off = 0; off < n; off += kMaxLaunch)
    {
        const int64_t m = n - off < kMaxLaunch ? n - off : kMaxLaunch;
        const float* r = rays + 8 * off;
        void* h = any ? nullptr : static_cast<void*>(static_cast<int4*>(hits) + off);
        uint8_t* o = any ? occ + off : nullptr;
        hipError_t e = counters ? (any ? launch<true, true>(sa, r, m, h, o, counters, stream)
                                       : launch<false, true>(sa, r, m, h, o, counters, stream))
                                : (any ? launch<true, false>(sa, r, m, h, o, counters, stream)
                                       : launch<false, false>(sa, r, m, h, o, counters, stream));
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
}  // namespace mcrt
