// The traversal kernel template, shared by the query entry points
// (traverse.hip) and the path tracer's extend step (pt.hip).
//
// One lane per ray at a time, one wavefront per workgroup (its LDS stack
// slab is private to the wave).  Waves claim rays kClaim at a time from a
// counter in device memory (`wq`, zero at launch): a lane whose ray ends
// takes the next index of its wave's claim, and the wave claims again when
// the claim is used up, so every lane stays busy until the batch runs out
// instead of idling once its share of a static split is done.  Idle lanes
// wait until at least S.refill_lanes of the wave's 64 lanes are idle (the
// compaction threshold; 1 = refill every lane as soon as its ray ends),
// so the ray loads of a refill can be gathered into fewer, fuller rounds.
// The rays are split into kSlices contiguous slices, one counter each (64 B
// apart), and a wave starts on slice blockIdx % kSlices, i.e. on its XCD's slice as
// workgroups are dealt to the XCDs in turn; a wave whose slice is used up
// moves on to the next.  Device-scope atomics on one address serialise, so
// spreading them over eight counters and claiming kClaim rays (2 per lane)
// at a time keeps the counters off the critical path (the path tracer's
// launches hold up to 5 rays per path, most of them short).  Two ray
// sources:
//   * SEG = false: rays[0, n) with n known to the host;
//   * SEG = true : three segments whose sizes live in device memory (the
//     path tracer's compacted continuation rays [0, cnt[0]), its next-event
//     rays at their records' slots [seg1, seg1 + cnt[1]) and its compacted
//     BSDF-sampled next-event rays [seg2, seg2 + cnt[kSeg2Word])), so the
//     bounce loop needs no host round trip and can be captured in a graph.
// BATCH = true: the rays belong to several scenes (batched rendering); ray
// i is traversed in the BVH of scene rsid[i].
#pragma once
#include "trace_core.h"

namespace mcrt
{
constexpr int kWave = 64;
constexpr uint32_t kClaim = 128;
// Word of a segment-mode count array holding the third segment's size (on
// its own 128 B line: the path tracer takes it with its own atomic).
constexpr int kSeg2Word = 32;

// Rays of several scenes in one launch: ray i belongs to scene rsid[i]
// (BATCH).  ORDER (queries): unless *skip is set at launch, the k-th ray
// claimed is rays[order[k]], order a permutation of the batch.
struct TraceBatch
{
    const SceneRoot* roots = nullptr;
    const uint16_t* rsid = nullptr;
    const uint32_t* order = nullptr;
    const uint32_t* skip = nullptr;
};

// With an 8-entry stack the closest-hit kernel is compiled for eight waves
// per SIMD (32 per CU, with LDS for all of them: 64 VGPRs), the any-hit
// kernel for seven (held to eight waves' registers it spills scalar
// registers into vector lanes, and the spill code in the refill block costs
// more than the eighth wave gains).  Deeper stacks are held to fewer waves
// by LDS.
template <int D, bool ANY, bool COUNT, bool SEG, bool BATCH, bool ORDER>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(
    D <= 8 ? (ANY ? kTraceAnyWavesPerSimd : kTraceWavesPerSimd) : 4, 8))) void k_trace(DevScene S, const float4* __restrict__ rays, int64_t n,
                                                 const uint32_t* __restrict__ seg_cnt, int64_t seg1, int64_t seg2,
                                                 int4* __restrict__ hits, uint8_t* __restrict__ occ,
                                                 unsigned long long* __restrict__ counters, TraceBatch B,
                                                 uint32_t* __restrict__ wq)
{
    __shared__ uint32_t s_node[(D + 1) * kWave];  // + the discard row
    __shared__ float s_lo[(D + 1) * kWave];
    TraceStack st;
    st.node = s_node;
    st.lo = s_lo;
    st.stride = kWave;
    st.lane = threadIdx.x;
    st.depth = D;
    st.spill = S.spill;
    st.spill_lane = blockIdx.x * kWave + threadIdx.x;
    st.spill_stride = (uint32_t)S.spill_lanes;  // >= the grid's lanes (trace_blocks)
    // 32-bit ray indices: a launch covers at most 2^31 rays (launch_trace
    // splits larger batches; path-tracer chunks are at most 5 * 2^28 rays),
    // and a slice counter ends below its slice + 2 * kClaim * (waves of the grid).
    uint32_t n0 = (uint32_t)n, n01 = (uint32_t)n, total = (uint32_t)n;
    if (SEG)
    {
        n0 = seg_cnt[0];
        n01 = n0 + seg_cnt[1];
        total = n01 + seg_cnt[kSeg2Word];
    }
    TraceCount cnt;
    uint32_t nrays = 0;
    const uint32_t lane = threadIdx.x;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64u - lane));  // lanes under this one
    Tracer<ANY, COUNT> tr;
    SceneRoot root;
    root.nodes = S.nodes;
    root.det_eps = S.det_eps;
    root.node_span = S.node_span;
    uint32_t i = 0;
    const uint32_t* order = ORDER && *B.skip == 0 ? B.order : nullptr;
    uint32_t claim = 0, claim_end = 0;  // wave-uniform: indices [claim, claim_end) not yet started
    bool drained = false;               // wave-uniform: every slice is used up
    bool live = false;
    // slice s holds the indices [total * s / kSlices, total * (s + 1) / kSlices)
    auto slice_lo = [total](uint32_t s) { return (uint32_t)(((uint64_t)total * s) / kSlices); };
    uint32_t sl = blockIdx.x % kSlices, tried = 0;  // wave-uniform: current slice, slices found used up
    // The next window is requested when the current one is taken, and its
    // offset is read (lane 0's atomic result) only when that one runs out,
    // so the atomic's round trip overlaps traversal steps instead of
    // stalling the wave between short rays.  Every lane runs the refill loop
    // (its conditions are wave-uniform), so lane 0 is the first active lane
    // there and readfirstlane hands its result to the wave as a scalar: the
    // claim bookkeeping stays in scalar registers instead of being shuffled
    // through the LDS crossbar into vector registers.
    uint32_t ahead = 0;  // lane 0: offset of the requested next window in slice sl
    if (lane == 0) ahead = atomicAdd(wq + sl * kSliceStride, kClaim);
    const uint32_t refill = (uint32_t)S.refill_lanes;  // 1..64
    for (;;)
    {
        // idle lanes take the next indices of the claim, in lane order, once
        // enough of them are idle (every lane idle always qualifies)
        uint64_t idle = __ballot(!live);
        if ((uint32_t)__popcll(idle) < refill) idle = 0;
        while (idle != 0 && !drained)
        {
            if (claim == claim_end)
            {
                uint32_t c = __builtin_amdgcn_readfirstlane(ahead);
                uint32_t lo = slice_lo(sl), len = slice_lo(sl + 1) - lo;
                while (c >= len)  // slice used up: move on to the next one
                {
                    if (++tried == kSlices)
                    {
                        drained = true;
                        break;
                    }
                    sl = (sl + 1) % kSlices;
                    lo = slice_lo(sl);
                    len = slice_lo(sl + 1) - lo;
                    uint32_t t = 0;
                    if (lane == 0) t = atomicAdd(wq + sl * kSliceStride, kClaim);
                    c = __builtin_amdgcn_readfirstlane(t);
                }
                if (drained) break;
                if (lane == 0) ahead = atomicAdd(wq + sl * kSliceStride, kClaim);
                claim = lo + c;
                claim_end = len - c < kClaim ? lo + len : lo + c + kClaim;
            }
            const uint32_t avail = claim_end - claim;
            const uint32_t rank = (uint32_t)__popcll(idle & below);
            if (!live && rank < avail)
            {
                const uint32_t k = claim + rank;
                if (SEG)
                    i = k < n0 ? k : k < n01 ? (uint32_t)seg1 + (k - n0) : (uint32_t)seg2 + (k - n01);
                else
                    i = (ORDER && order != nullptr) ? order[k] : k;
                if (BATCH) root = B.roots[B.rsid[i]];
                const float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
                RayPrep rp = prep_ray(a.x, a.y, a.z, b.x, b.y, b.z, a.w);
                node_margin(rp, root.node_span);
                tr.start(rp, b.w);
                live = true;
            }
            const uint32_t nidle = (uint32_t)__popcll(idle);
            claim += nidle < avail ? nidle : avail;
            idle = __ballot(!live);
        }
        if (__ballot(live) == 0) break;  // drained and every lane done: the whole wave exits
        if (live && (root.nodes == nullptr || tr.step(root, st, cnt)))
        {
            float t, u, v;
            const int prim = tr.result(t, u, v);
            if (ANY)
                occ[i] = prim >= 0 ? 1 : 0;
            else
                hits[i] = make_int4(prim, (int)bits_of(t), (int)bits_of(u), (int)bits_of(v));
            nrays++;
            live = false;
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

// Launches k_trace with the scene's LDS stack depth on `blocks` workgroups;
// wq holds kWqWords words, the slice counters zero.
template <bool ANY, bool COUNT, bool SEG, bool BATCH = false, bool ORDER = false>
inline void launch_k_trace(const DevScene& S, int64_t blocks, hipStream_t st, const float4* rays, int64_t n,
                           const uint32_t* seg_cnt, int64_t seg1, int64_t seg2, int4* hits, uint8_t* occ,
                           unsigned long long* counters, uint32_t* wq, TraceBatch B = TraceBatch())
{
    const dim3 g((unsigned)blocks), b(kWave);
    switch (S.lds_depth)
    {
    case 8:
        hipLaunchKernelGGL((k_trace<8, ANY, COUNT, SEG, BATCH, ORDER>), g, b, 0, st, S, rays, n, seg_cnt, seg1, seg2, hits, occ,
                           counters, B, wq);
        break;
    case 12:
        hipLaunchKernelGGL((k_trace<12, ANY, COUNT, SEG, BATCH, ORDER>), g, b, 0, st, S, rays, n, seg_cnt, seg1, seg2, hits, occ,
                           counters, B, wq);
        break;
    case 24:
        hipLaunchKernelGGL((k_trace<24, ANY, COUNT, SEG, BATCH, ORDER>), g, b, 0, st, S, rays, n, seg_cnt, seg1, seg2, hits, occ,
                           counters, B, wq);
        break;
    default:
        hipLaunchKernelGGL((k_trace<16, ANY, COUNT, SEG, BATCH, ORDER>), g, b, 0, st, S, rays, n, seg_cnt, seg1, seg2, hits, occ,
                           counters, B, wq);
        break;
    }
}
}  // namespace mcrt
