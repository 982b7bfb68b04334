// Wavefront path tracer (mcrt_render, mcrt_batch_*) and its path-replay
// adjoint (mcrt_render_adjoint): the reference's rayTrace integrator
// (demo_pt.cpp:512-655) and render loop (demo_pt.cpp:657-720) as a
// sequence of kernels over SoA path state.
//
// Per chunk of P paths (path p = pixel-major, sample-minor):
//   generate             camera rays with the tent filter
//   for k = 0 .. max_depth + 1:
//     extend             one traversal launch over the compacted
//                        continuation rays AND the next-event rays of
//                        vertex k-1 (segment counts read on the device)
//     resolve            next-event contributions of vertex k-1 (MIS),
//                        added before anything of vertex k, as in the
//                        reference's accumulation order
//     shade (k <= max)   emitter hit or escape to the environment,
//                        next-event rays at diffuse vertices, BSDF
//                        continuation, Russian roulette
//   accumulate           per-pixel sum in sample order of each path's
//                        radiance (NaN -> 0, clamp >= 0) x (1/spp)
// Every kernel is grid-strided over a fixed grid and reads its work count
// from device memory, so the whole chunk is launch-only (graph-capturable).
//
// A diffuse vertex's next-event record has rs rays: the BSDF-sampled ray
// and the light-sampled shadow ray of the area lights, and, with an
// environment map (rs = 4), the shadow rays of its two environment samples
// (one drawn from the map's pdf, one from the cosine lobe; demo_pt.cpp:
// 421-499, 564-589).  The record's other rays sit at its slots; the
// BSDF-sampled ray is traced only where it can reach an emitter and then
// goes to a compacted segment of its own, its slot kept with the record.  A
// ray a scene does not need has a zero direction, which every box rejects.
//
// Batched rendering (BATCH = true): the paths of many scenes -- each with
// its own BVH, camera, film size, spp and depth -- share the chunks.  A
// chunk is a list of segments (runs of whole pixels of one scene); each
// path keeps its scene id and its random-number key, each ray its scene
// id, and the kernels read the scene's data from a device table, so one
// launch per stage serves every scene of the chunk and the uneven scenes
// fill the machine together.  The numbers of every path are those of the
// same scene rendered alone.
//
// Adjoint (path replay, Vicini, Speierer, Jakob, "Path Replay
// Backpropagation", SIGGRAPH 2021): sampling, MIS weights and the roulette
// (fixed survival probability) do not depend on albedo or emission, so
// every term of a path's radiance is a product of the albedos of the
// vertices before it and one emission.  Pass 1 runs the chunk forward and
// keeps each path's radiance L; with delta = dLoss/dfilm * [L >= 0] / spp,
// pass 2 replays the same paths (same counter-based random numbers) and
// adds delta * (L - L_before_vertex) / albedo to the vertex's triangle and
// delta * (coefficient of Le) to the emitting triangle of every emission
// term.  Memory is independent of path length.
#include <algorithm>
#include <vector>

#include "capi_internal.h"
#include "pt_core.h"
#include "trace_kernel.h"

namespace mcrt
{
namespace
{
constexpr int kTPB = 256;

struct PTState
{
    float4* T;           // throughput (xyz), w = 1 where the vertex it left was diffuse
    float4* L;           // radiance (xyz)
    float4* rays[2];     // [P continuation | rec_rays(rs) P next-event at the records' slots |
                         //  P BSDF-sampled next-event, compacted] x 2 float4 per ray
    uint32_t* act[2];    // path id per continuation ray
    float4* rec[2];      // next-event records, rec_words(rs) float4 each
    uint32_t* bsl[2];    // with an environment map: each record's BSDF-sampled ray slot or ~0
                         // (its 64 B record has no spare word)
    uint32_t* cnt[2];    // {continuation rays, next-event rays at record slots}, the
                         // compacted BSDF-sampled rays at word kSeg2Word, then the
                         // traversal's work counters from word kCntWq on (kSideWords
                         // words per side)
    int4* hits;          // (1 + rs) P
    float4* Ltot;        // adjoint: radiance of pass 1
    float4* delta;       // adjoint: dLoss/dL per path
    float* g_alb;        // adjoint: ntri x 3
    float* g_em;         // adjoint: ntri x 3
    uint16_t* sid;       // batch: scene of each path
    uint64_t* pkey;      // batch: mix64(mix64(seed) ^ pixel) of each path
    uint64_t* skey;      // batch: sample << 8 of each path
    uint16_t* rsid[2];   // batch: scene of each ray slot ((1 + rs) P)
};

// float4 words of a next-event record: {T, pdf_b}, {albedo, factor2},
// {path, light, BSDF ray slot or -1, -}; with the environment {T, pdf_b},
// {albedo, factor2}, {e1, path}, {e2, light}.  The directions of its rays
// are read back from the rays themselves (the resolve step needs them only
// where a ray hit an emitter).
__host__ __device__ constexpr int rec_words(int rs) { return rs == 4 ? 4 : 3; }
// Rays a record keeps at its own slots: the light sample's shadow ray and,
// with an environment map, its two environment shadow rays (the
// BSDF-sampled ray is compacted).
__host__ __device__ constexpr int rec_rays(int rs) { return rs - 1; }

// One scene of a render: its device data, camera and parameters.
struct ScenePT
{
    DevScene D;
    DevCamera cam;
    int32_t spp, max_depth, rr_start, nee;
    float rr_prob;
    float scale;         // (float)(1.0 / spp), FrameBuffer::scale (film.cpp:254-260)
    uint64_t seed;
    int64_t sample_offset;
    float* film;         // batch: this scene's device film
};

// A run of whole pixels of one scene inside a batched chunk.
struct Segment
{
    int32_t sid, spp;
    int64_t pix0;        // first pixel (index in the scene's film)
    int64_t p0;          // first chunk-local path
    int64_t q0;          // first chunk-local pixel
};

struct PTArgs
{
    int64_t P;              // paths in this chunk
    int64_t pix0;           // single scene: first pixel of the chunk
    int32_t nseg;           // batch: segments of this chunk
    int32_t rs;             // next-event rays per record: 2, or 4 with an environment map
    const Segment* seg;     // batch
    const ScenePT* tab;     // batch: per-scene table
    const SceneRoot* roots; // batch: per-scene traversal roots
    ScenePT one;            // single scene
};

__device__ __forceinline__ void store_ray(float4* rays, int64_t i, V3f o, V3f d, float tmin, float tmax)
{
    rays[2 * i] = make_float4(o.x, o.y, o.z, tmin);
    rays[2 * i + 1] = make_float4(d.x, d.y, d.z, tmax);
}
__device__ __forceinline__ V3f f3(float4 v) { return V3f{v.x, v.y, v.z}; }

// Slots of this thread in the workgroup's three compactions of one
// iteration: continuation rays (counter c[0], one slot each), next-event
// records of `per` rays each (counter c[1], in rays; sb is the record's
// first ray slot) and traced BSDF-sampled next-event rays (counter
// c[kSeg2Word], one slot each).  One 64-bit atomic on the two adjacent
// counters and one on the third, on a cache line of its own, per workgroup
// and iteration: device-scope atomics on one address serialise (about 12 ns
// each on MI355X), so one per wavefront and counter -- 2^19 for 2^24 paths
// -- would take over 5 ms by themselves.  Consecutive threads get
// consecutive slots, so the stores that follow coalesce.  Every thread of
// the workgroup calls it (it synchronises the workgroup).  Slot order never
// affects a result: every path's records are its own.
__device__ __forceinline__ void block_slots(uint32_t* c, bool a, bool b, bool e, uint32_t per, uint32_t& sa,
                                            uint32_t& sb, uint32_t& se)
{
    constexpr uint32_t W = kTPB / 64;
    __shared__ uint32_t s_n[3][W];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_base_e;
    const uint64_t ma = __ballot(a), mb = __ballot(b), me = __ballot(e);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    if (lane == 0)
    {
        s_n[0][w] = (uint32_t)__popcll(ma);
        s_n[1][w] = (uint32_t)__popcll(mb);
        s_n[2][w] = (uint32_t)__popcll(me);
    }
    __syncthreads();
    uint32_t pa = 0, pb = 0, pe = 0, ta = 0, tb = 0, te = 0;
    for (uint32_t q = 0; q < W; q++)
    {
        const uint32_t na = s_n[0][q], nb = s_n[1][q], ne = s_n[2][q];
        pa += q < w ? na : 0u;
        pb += q < w ? nb : 0u;
        pe += q < w ? ne : 0u;
        ta += na;
        tb += nb;
        te += ne;
    }
    // c[0] stays below 2^28 (the largest chunk), so the low word never
    // carries into c[1]
    if (threadIdx.x == 0)
    {
        if ((ta | tb) != 0)
            s_base = atomicAdd((unsigned long long*)c, (unsigned long long)ta | ((unsigned long long)(tb * per) << 32));
        if (te != 0) s_base_e = atomicAdd(c + kSeg2Word, te);
    }
    __syncthreads();
    // s_n and the bases are rewritten only after the next call's first
    // barrier, which every thread reaches after these reads.  Every lane
    // reads the bases, but a slot means something only to a lane that asked
    // for one, and then its workgroup's total is non-zero and the base was
    // written above.
    const unsigned long long base = s_base;
    const uint64_t below = (1ull << lane) - 1ull;
    sa = (uint32_t)base + pa + (uint32_t)__popcll(ma & below);
    sb = (uint32_t)(base >> 32) + (pb + (uint32_t)__popcll(mb & below)) * per;
    se = s_base_e + pe + (uint32_t)__popcll(me & below);
}
__device__ __forceinline__ float4 f4(V3f v) { return make_float4(v.x, v.y, v.z, 0.0f); }

// Gradient scatter of the adjoint kernels.  Triangles with a hot slot
// (every light, then the largest triangles: the few keys most path
// vertices share) are summed per workgroup in LDS and flushed to global
// memory once per launch; the rest go to the hardware float atomics after
// lanes of one wavefront holding the same key are merged (camera vertices
// of one pixel's samples mostly hit the same triangle).
constexpr int kHotFloats = 3 * MCRT_HOT_SLOTS;  // per gradient array

__device__ __forceinline__ void hot_clear(float* acc)
{
    for (int i = threadIdx.x; i < 2 * kHotFloats; i += blockDim.x) acc[i] = 0.0f;
    __syncthreads();
}

// acc[0 .. kHotFloats) albedo, acc[kHotFloats ..) emission.
__device__ __forceinline__ void hot_flush(const DevScene& D, const float* acc, float* g_alb, float* g_em)
{
    __syncthreads();
    const int n = 3 * D.nhot;
    for (int i = threadIdx.x; i < 2 * n; i += blockDim.x)
    {
        const int which = i >= n, r = i - which * n;
        const float v = acc[which * kHotFloats + r];
        if (v != 0.0f) unsafeAtomicAdd((which ? g_em : g_alb) + 3 * (int64_t)D.hot_tri[r / 3] + r % 3, v);
    }
}

// Wave-merged scatter-add of a 3-vector into g[3*key]: up to 4 groups of
// lanes with equal keys are summed across the wavefront; once a group
// holds a single lane (scattered keys) the remaining lanes add directly.
// Every lane of the wave must call it (the loops below iterate uniformly
// per wave).
__device__ __forceinline__ void scatter3(float* g, int key, V3f v, bool valid)
{
    uint64_t pending = __ballot(valid);
    for (int it = 0; it < 4 && pending; ++it)
    {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int k = __shfl(key, leader);
        const bool mine = valid && key == k;
        const uint64_t same = __ballot(mine);
        if (__popcll(same) == 1) break;
        float sx = mine ? v.x : 0.0f, sy = mine ? v.y : 0.0f, sz = mine ? v.z : 0.0f;
        for (int off = 32; off > 0; off >>= 1)
        {
            sx += __shfl_xor(sx, off);
            sy += __shfl_xor(sy, off);
            sz += __shfl_xor(sz, off);
        }
        if ((int)(threadIdx.x & 63) == leader)
        {
            unsafeAtomicAdd(g + 3 * (int64_t)k, sx);
            unsafeAtomicAdd(g + 3 * (int64_t)k + 1, sy);
            unsafeAtomicAdd(g + 3 * (int64_t)k + 2, sz);
        }
        valid = valid && !mine;
        pending &= ~same;
    }
    if (valid)
    {
        unsafeAtomicAdd(g + 3 * (int64_t)key, v.x);
        unsafeAtomicAdd(g + 3 * (int64_t)key + 1, v.y);
        unsafeAtomicAdd(g + 3 * (int64_t)key + 2, v.z);
    }
}

// One gradient term: LDS for hot triangles, scatter3 for the rest.
__device__ __forceinline__ void grad_add(float* acc, float* g, int key, int hot, V3f v, bool valid)
{
    if (valid && hot >= 0)
    {
        __hip_atomic_fetch_add(acc + 3 * hot, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(acc + 3 * hot + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(acc + 3 * hot + 2, v.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    scatter3(g, key, v, valid && hot < 0);
}

// Counts of one side, and its traversal work counters zeroed.
constexpr int kCntWq = 64;
static_assert(kSeg2Word >= 32 && kSeg2Word < kCntWq, "the third count sits on its own 128 B line before the work counters");
constexpr int kSideWords = kCntWq + (int)kWqWords;
__global__ void k_set_counts(uint32_t* cnt, uint32_t a, uint32_t b)
{
    if (blockIdx.x == 0 && threadIdx.x < kSlices) cnt[kCntWq + threadIdx.x * kSliceStride] = 0;
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        cnt[0] = a;
        cnt[1] = b;
        cnt[kSeg2Word] = 0;
    }
}

__global__ void k_fill(float* p, int64_t n, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}


// Last segment whose first path (first pixel with Q) is <= x.
template <bool Q>
__device__ __forceinline__ int find_seg(const Segment* seg, int nseg, int64_t x)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if ((Q ? seg[mid].q0 : seg[mid].p0) <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Camera::makeRay (camera.cpp:37-43) + tent filter (demo_pt.cpp:687-699).
template <bool BATCH>
__global__ void k_generate(PTState S, PTArgs a)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.P; p += (int64_t)gridDim.x * blockDim.x)
    {
        const ScenePT* sp = &a.one;
        int64_t pixel;
        uint64_t sample;
        if (BATCH)
        {
            const Segment g = a.seg[find_seg<false>(a.seg, a.nseg, p)];
            sp = &a.tab[g.sid];
            const int64_t lp = p - g.p0;
            pixel = g.pix0 + lp / g.spp;
            sample = (uint64_t)(sp->sample_offset + lp % g.spp);
            S.sid[p] = (uint16_t)g.sid;
            S.rsid[0][p] = (uint16_t)g.sid;
            S.pkey[p] = mix64(mix64(sp->seed) ^ (uint64_t)pixel);
            S.skey[p] = sample << 8;
        }
        else
        {
            pixel = a.pix0 + p / a.one.spp;
            sample = (uint64_t)(a.one.sample_offset + p % a.one.spp);
        }
        const DevCamera& cam = sp->cam;
        const int x = (int)(pixel % cam.w), y = (int)(pixel / cam.w);
        float dx = 0.0f, dy = 0.0f;
        if (sp->spp > 1)
        {
            const uint64_t key = rng_key(sp->seed, (uint64_t)pixel, sample, 255u);
            float r1 = 2.0f * rng_u01(key, 0);
            float r2 = 2.0f * rng_u01(key, 1);
            dx = r1 < 1 ? sqrtf(r1) - 1.0f : 1.0f - sqrtf(2.0f - r1);
            dy = r2 < 1 ? sqrtf(r2) - 1.0f : 1.0f - sqrtf(2.0f - r2);
        }
        float mx = (((float)x + 0.5f) + dx) / (float)cam.w;
        float my = (((float)y + 0.5f) + dy) / (float)cam.h;
        V3f su = (ld3(cam.fx) * (mx - 0.5f)) * cam.ww;
        V3f sv = (ld3(cam.fy) * (my - 0.5f)) * 1.0f;
        V3f sw = ld3(cam.fwd) * cam.dist;
        V3f d = nrm((su + sv) + sw);
        // camera ray p is path p, its throughput is one and no earlier
        // vertex was diffuse: the first k_shade takes these as constants,
        // so only the ray and the radiance are stored
        store_ray(S.rays[0], p, ld3(cam.o), d, kEpsRay, __builtin_inff());
        S.L[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// Area-light terms of a next-event record (demo_pt.cpp:591-643 with
// Scene::lightPdfW scene.cpp:239-267 and lightRadiance :345-366): the
// BSDF-sampled ray hitting an emitter, MIS-weighted against the light pdf,
// then the light sample if its shadow ray reaches the sampled light.
// em1/em2, hot1/hot2 and the coefficients c1/c2 of the two emissions are
// what the adjoint scatters.
// rb / rl: the BSDF-sampled ray and the light sample's shadow ray (2 float4
// each, the direction in the second; read only where the ray hit).
__device__ __forceinline__ V3f area_terms(const DevScene& D, const int4 h1, const int4 h2, V3f col,
                                          const float4* rb, const float4* rl, float pdf_b, float factor2, int li,
                                          V3f local, int& em1,
                                          int& em2, int& hot1, int& hot2, V3f& c1, V3f& c2)
{
    float pdf_l = 0.0f;
    V3f le1 = v3f(0.0f, 0.0f, 0.0f);
    if (h1.x >= 0)
    {
        const TriShade ts = D.shade[h1.x];
        const DevMaterial mt = D.mats[ts.mat];
        if (mt.type == MCRT_LGHT)
        {
            const V3f ns1 = shading_normal(ts, float_of((uint32_t)h1.z), float_of((uint32_t)h1.w));
            const V3f wb = f3(rb[1]);
            const float cosl = dot(ns1, -wb);
            if (!(cosl < 1e-6f))
            {
                const DevLight Lt = D.lights[ts.light];
                const float pdfA = (1.0f / Lt.area) * Lt.pdf;
                const float t1 = float_of((uint32_t)h1.y);
                pdf_l = pdfA * (t1 * t1) / cosl;
            }
            if (!(dot(ns1, -wb) < 0))
            {
                le1 = ld3(mt.rgb);
                em1 = h1.x;
                hot1 = ts.hot;
            }
        }
    }
    const float mis_b = pdf_b / (pdf_b + pdf_l);
    local = local + (col * le1) * mis_b;
    c1 = col * mis_b;
    V3f f = v3f(0.0f, 0.0f, 0.0f);
    if (li >= 0)
    {
        V3f le2 = v3f(0.0f, 0.0f, 0.0f);
        if (h2.x >= 0)
        {
            const TriShade ts = D.shade[h2.x];
            const DevMaterial mt = D.mats[ts.mat];
            if (mt.type == MCRT_LGHT)
            {
                const V3f ns2 = shading_normal(ts, float_of((uint32_t)h2.z), float_of((uint32_t)h2.w));
                const V3f wo = f3(rl[1]);
                if (!(dot(ns2, -wo) < 0) && ts.light == li)
                {
                    le2 = ld3(mt.rgb);
                    em2 = h2.x;
                    hot2 = ts.hot;
                }
            }
        }
        const float ip = 1.0f / kPi;
        f = v3f(ip, ip, ip) * col;  // LambertianBSDF::evaluate (bsdf.cpp:96-116)
        local = local + (f * le2) * factor2;
    }
    c2 = f * factor2;
    return local;
}

// Next-event contributions of one diffuse vertex: the environment terms
// (demo_pt.cpp:564-589), each counted when its shadow ray escapes, then the
// area-light terms, in the reference's order of accumulation.  ENV: the
// records hold the environment terms (rs = 4).
template <bool ADJ, bool BATCH, bool ENV>
__global__ void k_resolve(PTArgs a, PTState S, int cur)
{
    __shared__ float acc[ADJ ? 2 * kHotFloats : 1];
    if (ADJ) hot_clear(acc);
    const int64_t P = a.P;
    constexpr int rs = ENV ? 4 : 2, rf = rec_rays(rs);
    const uint32_t nrec = S.cnt[cur][1] / (uint32_t)rf;
    const float4* R = S.rec[cur];
    const float4* rays = S.rays[cur];
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nrec; base += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t m = base + threadIdx.x;
        const bool act = m < nrec;
        int em1 = -1, em2 = -1, hot1 = -1, hot2 = -1;
        V3f g1 = v3f(0.0f, 0.0f, 0.0f), g2 = g1;
        if (act)
        {
            const float4* Rm = R + rec_words(rs) * m;
            const float4 r0 = Rm[0], r1 = Rm[1], r2 = Rm[2], r3 = ENV ? Rm[3] : r2;
            const V3f T = f3(r0), col = f3(r1);
            const uint32_t path = bits_of(ENV ? r2.w : r2.x);
            const DevScene& D = BATCH ? a.tab[S.sid[path]].D : a.one.D;
            const int64_t h0 = P + rf * m;  // the record's first ray slot
            V3f local = v3f(0.0f, 0.0f, 0.0f);
            if (ENV)
            {
                const V3f z = v3f(0.0f, 0.0f, 0.0f);
                const V3f e1 = S.hits[h0 + 1].x < 0 ? f3(r2) : z;
                const V3f e2 = S.hits[h0 + 2].x < 0 ? f3(r3) : z;
                local = local + (e1 + e2);
            }
            if (D.nlights > 0)
            {
                const int li = (int)bits_of(ENV ? r3.w : r2.y);
                const int4 miss = make_int4(-1, 0, 0, 0);
                // the BSDF-sampled ray in the compacted segment after the
                // records' slots where it was traced, else a miss; the light
                // sample's shadow ray at the record's first slot
                const int bs = ENV ? (int)S.bsl[cur][m] : (int)bits_of(r2.z);
                const int64_t hb = P + rf * P + bs;
                V3f c1, c2;
                local = area_terms(D, bs >= 0 ? S.hits[hb] : miss, li >= 0 ? S.hits[h0] : miss, col, rays + 2 * hb,
                                   rays + 2 * h0, r0.w, r1.w, li, local, em1, em2, hot1, hot2, c1, c2);
                if (ADJ)
                {
                    const V3f dl = f3(S.delta[path]);
                    g1 = (dl * T) * c1;
                    g2 = (dl * T) * c2;
                }
            }
            // A path's radiance starts at +0 and only ever has terms added,
            // so it is never -0 and adding a zero term leaves its bits as
            // they are: records whose term is zero (shadowed light sample,
            // BSDF ray missing every emitter) skip the read and the write.
            const V3f x = T * local;
            if (!(x.x == 0.0f && x.y == 0.0f && x.z == 0.0f))
            {
                const V3f Lv = f3(S.L[path]) + x;
                S.L[path] = f4(Lv);
            }
        }
        if (ADJ)
        {
            grad_add(acc + kHotFloats, S.g_em, em1, hot1, g1, em1 >= 0);
            grad_add(acc + kHotFloats, S.g_em, em2, hot2, g2, em2 >= 0);
        }
    }
    if (ADJ) hot_flush(a.one.D, acc, S.g_alb, S.g_em);
}

// Environment next-event terms of a diffuse vertex before visibility
// (EnvmapSampler, demo_pt.cpp:421-499, with MIS :564-589): e1 from a sample
// of the map's pdf, e2 from a sample of the cosine lobe in the frame
// u = normalize((0.3, 0.4, 0.5) x n), v = n x u.  d1 / d2 are their shadow
// rays' directions, zero (rejected by every box) where the sample lies
// below the surface and the term is zero.
__device__ __forceinline__ void env_terms(const DevScene& D, uint64_t key, V3f ns, V3f col, V3f& e1, V3f& e2,
                                          V3f& d1, V3f& d2)
{
    const float ip = 1.0f / kPi;
    const V3f fl = v3f(ip, ip, ip) * col;  // LambertianBSDF::evaluate
    V3f Ld, lrad;
    float lpdf;
    env_sample(D, rng_u01(key, 8), rng_u01(key, 9), rng_u01(key, 10), rng_u01(key, 11), Ld, lrad, lpdf);
    const float g1 = fmax_ref(0.0f, dot(Ld, ns));
    if (g1 > 0)
    {
        const float epdf = env_pdf(D, env_pixel(D, Ld));
        V3f irr = ((fl * g1) * (1.0f / lpdf)) * lrad;
        if (!(irr.x == irr.x && irr.y == irr.y && irr.z == irr.z)) irr = v3f(0.0f, 0.0f, 0.0f);
        e1 = irr * (epdf / (epdf + fmax_ref(dot(Ld, ns), kEpsCos) / kPi));
        d1 = Ld;
    }
    const float x = rng_u01(key, 12) * kTwoPi;
    const float r2 = rng_u01(key, 13);
    const float r2s = sqrtf(r2);
    const V3f uu = nrm(cross(v3f(0.3f, 0.4f, 0.5f), ns));
    const V3f vv = cross(ns, uu);
    float sx, cx;
    own_sincos(x, &sx, &cx);
    const V3f wo2 = nrm(((uu * (cx * r2s)) + (vv * (sx * r2s))) + ns * sqrtf(1.0f - r2));
    const float g2 = fmax_ref(0.0f, dot(wo2, ns));
    if (g2 > 0)
    {
        const int pix = env_pixel(D, wo2);
        const float bpdf = fmax_ref(dot(wo2, ns), kEpsCos) / kPi;
        e2 = (col * env_rad(D, pix)) * (bpdf / (bpdf + env_pdf(D, pix)));
        d2 = wo2;
    }
}

// Of a vertex at hit h of ray j: the ray's direction d, the hit point, the
// shading normal and, where it has a next-event record and the scene area
// lights, the BSDF-sampled next-event direction wb and whether that ray is
// traced.  The BSDF-sampled ray adds a term only when its closest hit is an
// emitter, and the traversal accepts a triangle only after its reference's
// box passes the slab test.  That test is monotone in the box
// (round-to-nearest keeps (b - o) * inv_dir ordered in b), so a ray failing
// it against the box around every emitter reference (with tmax = inf)
// reaches none, and its terms are a miss's: it is not traced.  Rays with an
// infinite inv_dir take the reference's NaN-prone sequential test and are
// always traced.
__device__ __forceinline__ void vertex_dirs(const DevScene& D, const float4* rays, int64_t j, int4 h, const TriShade& ts,
                                            bool want_rec, uint64_t key, V3f& d, V3f& phit, V3f& ns, V3f& wb,
                                            bool& trace_wb)
{
    const float4 ra = rays[2 * j], rb = rays[2 * j + 1];
    const V3f o = f3(ra);
    d = f3(rb);
    const float t = float_of((uint32_t)h.y), u = float_of((uint32_t)h.z), v = float_of((uint32_t)h.w);
    phit = o + d * t;
    ns = shading_normal(ts, u, v);
    if (want_rec && D.nlights > 0)
    {
        wb = lambert_dir(ns, rng_u01(key, 0), rng_u01(key, 1));
        if (D.emit_any)
        {
            const RayPrep rw = prep_ray(phit.x, phit.y, phit.z, wb.x, wb.y, wb.z, kEpsRay);
            float lo;
            trace_wb = rw.exact || slab(rw, D.emit_lo[0], D.emit_hi[0], D.emit_lo[1], D.emit_hi[1], D.emit_lo[2],
                                        D.emit_hi[2], __builtin_inff(), lo);
        }
    }
}

// One vertex of every live path (demo_pt.cpp:523-651).  ENV: a scene of the
// launch has an environment map (escape radiance, rs = 4 next-event rays);
// without one the environment code is compiled out, which keeps the
// kernel's registers, and so its occupancy, those of area lights alone.
// What a vertex emits -- a next-event record, a continuation ray -- depends
// only on its material, the parameters and the roulette draw, so it is
// decided first and the workgroup takes its slots (block_slots) before
// anything is computed or stored.
template <bool ADJ, bool BATCH, bool ENV>
__global__ void k_shade(PTArgs a, PTState S, int cur, int k)
{
    __shared__ float acc[ADJ ? 2 * kHotFloats : 1];
    if (ADJ) hot_clear(acc);
    constexpr int rs = ENV ? 4 : 2, rf = rec_rays(rs);
    const int nxt = cur ^ 1;
    const uint32_t A = S.cnt[cur][0];
    const float4* rays = S.rays[cur];
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < A; base += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t j = base + threadIdx.x;
        const bool act = j < A;
        int alb_key = -1, em_key = -1, hot = -1;
        V3f g_alb = v3f(0.0f, 0.0f, 0.0f), g_em = g_alb;
        uint32_t p = 0;
        int4 h = make_int4(-1, 0, 0, 0);
        uint16_t sid = 0;
        if (act)
        {
            p = k == 0 ? (uint32_t)j : S.act[cur][j];
            h = S.hits[j];
            if (BATCH) sid = S.sid[p];
        }
        const ScenePT& sp = BATCH ? a.tab[sid] : a.one;
        const DevScene& D = sp.D;
        TriShade ts{};
        DevMaterial mt{};
        uint64_t key = 0;
        bool want_rec = false, want_cont = false;
        if (act && h.x >= 0)
        {
            ts = D.shade[h.x];
            mt = D.mats[ts.mat];
            if (mt.type != MCRT_LGHT)
            {
                if (BATCH)
                    key = mix64(S.pkey[p] ^ (S.skey[p] | (uint64_t)(uint32_t)k));
                else
                {
                    const int64_t pixel = a.pix0 + p / sp.spp;
                    const uint64_t sample = (uint64_t)(sp.sample_offset + p % sp.spp);
                    key = rng_key(sp.seed, (uint64_t)pixel, sample, (uint32_t)k);
                }
                want_rec = sp.nee && mt.type == MCRT_DIFF && (D.nlights > 0 || (ENV && D.env_w > 0));
                // no continuation after the last vertex, or when the roulette ends the path
                want_cont = !(k >= sp.max_depth) && !(k >= sp.rr_start && rng_u01(key, 7) >= sp.rr_prob);
            }
        }
        // The hit point, the shading normal and the BSDF-sampled next-event
        // direction come before the workgroup takes its slots: whether that
        // ray is traced decides whether it takes a slot of the compacted
        // segment.
        V3f d = v3f(0.0f, 0.0f, 0.0f), phit = d, ns = d, wb = d;
        bool trace_wb = false;
        if (act && h.x >= 0) vertex_dirs(D, rays, j, h, ts, want_rec, key, d, phit, ns, wb, trace_wb);
        // The path's throughput and whether its previous vertex was diffuse
        // are loaded before the slots too: their latency overlaps the slot
        // atomic.  The first vertex has unit throughput and no predecessor.
        V3f T = v3f(1.0f, 1.0f, 1.0f);
        bool no_direct = true;
        if (act && k != 0 && (h.x >= 0 || (ENV && D.env_w > 0)))
        {
            const float4 t4 = S.T[p];
            T = f3(t4);
            no_direct = !sp.nee || t4.w == 0.0f;
        }
        uint32_t cslot, rslot, bslot;
        block_slots(S.cnt[nxt], want_cont, want_rec, trace_wb, (uint32_t)rf, cslot, rslot, bslot);
        do
        {
            if (!act) break;
            if (h.x < 0)
            {
                // escape: the passive environment term (demo_pt.cpp:528-540)
                if (ENV && D.env_w > 0 && no_direct)
                {
                    const V3f Lv = f3(S.L[p]) + T * env_rad(D, env_pixel(D, f3(rays[2 * j + 1])));
                    S.L[p] = f4(Lv);
                }
                break;
            }
            const V3f col = ld3(mt.rgb);
            if (mt.type == MCRT_LGHT)
            {
                if (no_direct && dot(d, ns) < 0)
                {
                    V3f Lv = f3(S.L[p]) + T * col;
                    S.L[p] = f4(Lv);
                    if (ADJ)
                    {
                        em_key = h.x;
                        hot = ts.hot;
                        g_em = f3(S.delta[p]) * T;
                    }
                }
                break;
            }
            if (ADJ)
            {
                // every later term carries this vertex's albedo once
                const V3f rest = f3(S.Ltot[p]) - f3(S.L[p]);
                const V3f dl = f3(S.delta[p]);
                alb_key = h.x;
                hot = ts.hot;
                g_alb = v3f(col.x > 0.0f ? dl.x * rest.x / col.x : 0.0f,
                            col.y > 0.0f ? dl.y * rest.y / col.y : 0.0f,
                            col.z > 0.0f ? dl.z * rest.z / col.z : 0.0f);
            }
            if (want_rec)
            {
                const int64_t m = rslot / (uint32_t)rf;
                const int64_t r0i = a.P + (int64_t)rslot;
                float4* R = S.rec[nxt] + rec_words(rs) * m;
                float4* nr = S.rays[nxt];
                const V3f zero = v3f(0.0f, 0.0f, 0.0f);
                V3f wo = zero;
                float pdf_b = 1.0f, factor2 = 0.0f;
                int li_rec = -1;
                if (D.nlights > 0)
                {
                    pdf_b = fmax_ref(dot(wb, ns), kEpsCos) / kPi;
                    // Scene::sampleLightArea (scene.cpp:298-343)
                    const int li = sample_light(D.lights, D.nlights, rng_u01(key, 2));
                    const DevLight Lt = D.lights[li];
                    const float* tv = D.verts + 9 * (int64_t)Lt.tri;
                    const float r0 = sqrtf(rng_u01(key, 3));
                    const float lu = 1.0f - r0;
                    const float lv = rng_u01(key, 4) * r0;
                    const V3f pl = (ld3(tv) * ((1.0f - lu) - lv) + ld3(tv + 3) * lu) + ld3(tv + 6) * lv;
                    const V3f nl = shading_normal(D.shade[Lt.tri], lu, lv);
                    const float pdfA = (1.0f / Lt.area) * Lt.pdf;
                    wo = pl - phit;
                    const float dist2 = len2(wo);
                    wo = nrm(wo);
                    const float cosL = dot(nl, -wo);
                    float pdfW = pdfA * dist2 / fabsf(cosL);  // __pdfAtoW (geometry.cpp:222-225)
                    if (cosL <= 0) pdfW = 0;
                    if (pdfW > 0)
                    {
                        const float pdf_bp = fmax_ref(dot(wo, ns), kEpsCos) / kPi;
                        const float mis_l = pdfW / (pdfW + pdf_bp);
                        const float cosr = stdmax(dot(ns, wo), 0.0f);
                        factor2 = cosr / pdfW * mis_l;
                        li_rec = li;
                    }
                }
                R[0] = make_float4(T.x, T.y, T.z, pdf_b);
                R[1] = make_float4(col.x, col.y, col.z, factor2);
                const uint32_t bs = trace_wb ? bslot : ~0u;
                if (ENV)
                    S.bsl[nxt][m] = bs;
                else
                    R[2] = make_float4(float_of(p), float_of((uint32_t)li_rec), float_of(bs), 0.0f);
                // a zero direction fails every slab test (ray.cpp:77-79): the
                // rays a vertex does not need cost one root visit
                store_ray(nr, r0i, phit, li_rec >= 0 ? wo : zero, kEpsRay, __builtin_inff());
                if (trace_wb) store_ray(nr, a.P + rf * a.P + bslot, phit, wb, kEpsRay, __builtin_inff());
                if (ENV)
                {
                    V3f e1 = zero, e2 = zero, d1 = zero, d2 = zero;
                    if (D.env_w > 0) env_terms(D, key, ns, col, e1, e2, d1, d2);
                    R[2] = make_float4(e1.x, e1.y, e1.z, float_of(p));
                    R[3] = make_float4(e2.x, e2.y, e2.z, float_of((uint32_t)li_rec));
                    store_ray(nr, r0i + 1, phit, d1, kEpsRay, __builtin_inff());
                    store_ray(nr, r0i + 2, phit, d2, kEpsRay, __builtin_inff());
                }
                if (BATCH)
                {
                    for (int i = 0; i < rf; i++) S.rsid[nxt][r0i + i] = sid;
                    if (trace_wb) S.rsid[nxt][a.P + rf * a.P + bslot] = sid;
                }
            }
            if (!want_cont) break;
            // Continuation (Material::sample, demo_pt.cpp:645-650)
            const V3f wi = -d;
            V3f nd, weight;
            if (mt.type == MCRT_GLSY)
            {
                float w;
                nd = glossy_dir(wi, ns, mt.gloss, rng_u01(key, 5), rng_u01(key, 6), w);
                weight = col * w;
            }
            else
            {
                nd = lambert_dir(ns, rng_u01(key, 5), rng_u01(key, 6));
                weight = col;
            }
            T = T * weight;
            if (k >= sp.rr_start) T = T * (1.0f / sp.rr_prob);
            S.T[p] = make_float4(T.x, T.y, T.z, mt.type == MCRT_DIFF ? 1.0f : 0.0f);
            store_ray(S.rays[nxt], cslot, phit, nd, kEpsRay, __builtin_inff());
            S.act[nxt][cslot] = p;
            if (BATCH) S.rsid[nxt][cslot] = sid;
        } while (false);
        if (ADJ)
        {
            grad_add(acc, S.g_alb, alb_key, hot, g_alb, alb_key >= 0);
            grad_add(acc + kHotFloats, S.g_em, em_key, hot, g_em, em_key >= 0);
        }
    }
    if (ADJ) hot_flush(a.one.D, acc, S.g_alb, S.g_em);
}

// Adjoint seed: keep pass 1's radiance and form dLoss/dL of each path
// through the clamp (slope 1 where L >= 0), the NaN rule and the 1/spp
// scale of the film.
__global__ void k_delta(PTState S, PTArgs a, const float* dfilm, float scale)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.P; p += (int64_t)gridDim.x * blockDim.x)
    {
        const float4 L = S.L[p];
        S.Ltot[p] = L;
        const bool nan = !(L.x == L.x && L.y == L.y && L.z == L.z);
        const float* g = dfilm + 3 * (a.pix0 + p / a.one.spp);
        S.delta[p] = make_float4(!nan && L.x >= 0.0f ? g[0] * scale : 0.0f,
                                 !nan && L.y >= 0.0f ? g[1] * scale : 0.0f,
                                 !nan && L.z >= 0.0f ? g[2] * scale : 0.0f, 0.0f);
    }
}

// FrameBuffer::accumulatePixel over the samples in order, each path's
// radiance taken with demo_pt.cpp:653 (NaN -> 0) and :703 (clamp >= 0 per
// channel), then scale((float)(1.0/spp)) (demo_pt.cpp:700-710,
// film.cpp:204-260), for the npix pixels of the chunk.
template <bool BATCH>
__global__ void k_accumulate(PTState S, PTArgs a, float* film, int64_t npix)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < npix; q += (int64_t)gridDim.x * blockDim.x)
    {
        int spp;
        int64_t p0, pixel;
        float scale;
        float* out;
        if (BATCH)
        {
            const Segment g = a.seg[find_seg<true>(a.seg, a.nseg, q)];
            spp = g.spp;
            p0 = g.p0 + (q - g.q0) * spp;
            pixel = g.pix0 + (q - g.q0);
            scale = a.tab[g.sid].scale;
            out = a.tab[g.sid].film;
        }
        else
        {
            spp = a.one.spp;
            p0 = q * spp;
            pixel = a.pix0 + q;
            scale = a.one.scale;
            out = film;
        }
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        auto add = [&](float4 L) {
            if (!(L.x == L.x && L.y == L.y && L.z == L.z)) L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            gx = gx + fmax_ref(0.0f, L.x);
            gy = gy + fmax_ref(0.0f, L.y);
            gz = gz + fmax_ref(0.0f, L.z);
        };
        // A lane's samples are consecutive; loading eight (128 B) before
        // summing them lets one fetch of each line serve all eight, where
        // one load per step saw the line evicted between steps.  The sum
        // keeps the sample order.
        int s = 0;
        for (; s + 8 <= spp; s += 8)
        {
            float4 v[8];
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = S.L[p0 + s + i];
#pragma unroll
            for (int i = 0; i < 8; i++) add(v[i]);
        }
        for (; s < spp; s++) add(S.L[p0 + s]);
        float* f = out + 3 * pixel;
        f[0] = gx * scale;
        f[1] = gy * scale;
        f[2] = gz * scale;
    }
}

// Workgroups of kTPB threads of kernel K resident at once on a CU (queried
// once per kernel).
template <auto K>
int resident_per_cu()
{
    static const int n = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(K), kTPB, 0) != hipSuccess)
        {
            (void)hipGetLastError();
            b = 0;
        }
        return b < 1 ? 1 : b;
    }();
    return n;
}

// Grid of a grid-strided kernel over n items: at most the workgroups that
// are resident at once, so all of them start together and none is left
// for a second round (a grid a little larger than what fits doubles the
// time of its last workgroups).
template <auto K>
dim3 grid_of(int64_t n, int num_cus)
{
    int64_t b = (n + kTPB - 1) / kTPB;
    const int64_t cap = (int64_t)num_cus * resident_per_cu<K>();
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

// The bounce loop of one chunk (generate .. last resolve).  Dt carries the
// traversal settings (LDS depth, spill stacks) of the launch.
template <bool ADJ, bool BATCH>
void bounce_loop(const DevScene& Dt, PTState& S, const PTArgs& a, int max_depth, int cus, hipStream_t st)
{
    hipLaunchKernelGGL(k_generate<BATCH>, (grid_of<k_generate<BATCH>>(a.P, cus)), dim3(kTPB), 0, st, S, a);
    hipLaunchKernelGGL(k_set_counts, dim3(1), dim3(64), 0, st, S.cnt[0], (uint32_t)a.P, 0u);
    int cur = 0;
    for (int k = 0; k <= max_depth + 1; k++)
    {
        const int nxt = cur ^ 1;
        TraceBatch tb;
        tb.roots = a.roots;
        tb.rsid = S.rsid[cur];
        launch_k_trace<false, false, true, BATCH>(Dt, trace_blocks(Dt, (1 + a.rs) * a.P), st, S.rays[cur], (int64_t)0,
                                                  S.cnt[cur], a.P, a.P + rec_rays(a.rs) * a.P, S.hits, (uint8_t*)nullptr,
                                                  (unsigned long long*)nullptr, S.cnt[cur] + kCntWq, tb);
        if (a.rs == 4)
            hipLaunchKernelGGL((k_resolve<ADJ, BATCH, true>), (grid_of<k_resolve<ADJ, BATCH, true>>(a.P, cus)),
                               dim3(kTPB), 0, st, a, S, cur);
        else
            hipLaunchKernelGGL((k_resolve<ADJ, BATCH, false>), (grid_of<k_resolve<ADJ, BATCH, false>>(a.P, cus)),
                               dim3(kTPB), 0, st, a, S, cur);
        if (k <= max_depth)
        {
            hipLaunchKernelGGL(k_set_counts, dim3(1), dim3(64), 0, st, S.cnt[nxt], 0u, 0u);
            if (a.rs == 4)
                hipLaunchKernelGGL((k_shade<ADJ, BATCH, true>), (grid_of<k_shade<ADJ, BATCH, true>>(a.P, cus)),
                                   dim3(kTPB), 0, st, a, S, cur, k);
            else
                hipLaunchKernelGGL((k_shade<ADJ, BATCH, false>), (grid_of<k_shade<ADJ, BATCH, false>>(a.P, cus)),
                                   dim3(kTPB), 0, st, a, S, cur, k);
        }
        cur = nxt;
    }
}

PTState state_of(StreamRes* r, bool adjoint)
{
    PTWork& w = r->ptw;
    PTState S;
    S.T = w.T.as<float4>();
    S.L = w.L.as<float4>();
    for (int i = 0; i < 2; i++)
    {
        S.rays[i] = w.rays[i].as<float4>();
        S.act[i] = w.act[i].as<uint32_t>();
        S.rec[i] = w.rec[i].as<float4>();
        S.bsl[i] = w.bsl[i].as<uint32_t>();
        S.cnt[i] = w.cnt.as<uint32_t>() + kSideWords * i;
        S.rsid[i] = w.rsid[i].as<uint16_t>();
    }
    S.hits = w.hits.as<int4>();
    S.Ltot = adjoint ? w.Ltot.as<float4>() : nullptr;
    S.delta = adjoint ? w.delta.as<float4>() : nullptr;
    S.g_alb = S.g_em = nullptr;
    S.sid = w.sid.as<uint16_t>();
    S.pkey = w.pkey.as<uint64_t>();
    S.skey = w.skey.as<uint64_t>();
    return S;
}

ScenePT scene_pt(mcrt_scene* s, const DevScene& D, const mcrt_render_params* prm)
{
    ScenePT sp;
    sp.D = D;
    sp.cam = s->dev_cam;
    sp.spp = prm->spp;
    sp.max_depth = prm->max_depth;
    sp.rr_start = prm->rr_start;
    sp.nee = prm->nee;
    sp.rr_prob = prm->rr_prob;
    sp.scale = (float)(1.0 / prm->spp);
    sp.seed = prm->seed;
    sp.sample_offset = prm->sample_offset;
    sp.film = nullptr;
    return sp;
}
}  // namespace

// Workspace for up to P paths with rs next-event rays per record (batch:
// plus the per-path / per-ray scene ids and random-number keys).
int pt_workspace(StreamRes* r, int64_t P, bool adjoint, bool batch, int rs)
{
    PTWork& w = r->ptw;
    int rc;
    const size_t n = (size_t)P, nr = (size_t)(1 + rs) * n;
    if ((rc = w.T.ensure(n * 16))) return rc;
    if ((rc = w.L.ensure(n * 16))) return rc;
    for (int i = 0; i < 2; i++)
    {
        if ((rc = w.rays[i].ensure(nr * 32))) return rc;
        if ((rc = w.act[i].ensure(n * 4))) return rc;
        if ((rc = w.rec[i].ensure(n * 16 * rec_words(rs)))) return rc;
        if (rs == 4 && (rc = w.bsl[i].ensure(n * 4))) return rc;
    }
    if ((rc = w.cnt.ensure(2 * kSideWords * sizeof(uint32_t)))) return rc;
    if ((rc = w.hits.ensure(nr * 16))) return rc;
    if (adjoint)
    {
        if ((rc = w.Ltot.ensure(n * 16))) return rc;
        if ((rc = w.delta.ensure(n * 16))) return rc;
    }
    if (batch)
    {
        if ((rc = w.sid.ensure(n * 2))) return rc;
        if ((rc = w.pkey.ensure(n * 8))) return rc;
        if ((rc = w.skey.ensure(n * 8))) return rc;
        for (int i = 0; i < 2; i++)
            if ((rc = w.rsid[i].ensure(nr * 2))) return rc;
    }
    return MCRT_OK;
}

int render_scene(mcrt_scene* s, StreamRes* r, const mcrt_render_params* prm, float* film, const AdjointIO* adj,
                 hipStream_t st)
{
    DevScene D;
    int rc = render_view(s, r, D);
    if (rc) return rc;
    const DevCamera& cam = s->dev_cam;
    const int64_t npix = (int64_t)cam.w * cam.h;
    const int spp = prm->spp;
    int64_t pix_per_chunk = s->pt_chunk_paths / spp;
    if (pix_per_chunk < 1) pix_per_chunk = 1;
    if (pix_per_chunk > npix) pix_per_chunk = npix;
    const int rs = D.env_w > 0 ? 4 : 2;
    if ((rc = pt_workspace(r, pix_per_chunk * spp, adj != nullptr, false, rs))) return rc;
    PTState S = state_of(r, adj != nullptr);
    S.g_alb = adj ? adj->d_albedo : nullptr;
    S.g_em = adj ? adj->d_emission : nullptr;
    const int cus = s->ctx->num_cus;
    if (adj)
    {
        const int64_t ng = 3 * s->dev.ntri;
        hipLaunchKernelGGL(k_fill, (grid_of<k_fill>(ng, cus)), dim3(kTPB), 0, st, adj->d_albedo, ng, 0.0f);
        hipLaunchKernelGGL(k_fill, (grid_of<k_fill>(ng, cus)), dim3(kTPB), 0, st, adj->d_emission, ng, 0.0f);
    }
    PTArgs a;
    a.nseg = 0;
    a.rs = rs;
    a.seg = nullptr;
    a.tab = nullptr;
    a.roots = nullptr;
    a.one = scene_pt(s, D, prm);
    for (int64_t pix0 = 0; pix0 < npix; pix0 += pix_per_chunk)
    {
        const int64_t np = std::min(pix_per_chunk, npix - pix0);
        a.P = np * spp;
        a.pix0 = pix0;
        bounce_loop<false, false>(D, S, a, prm->max_depth, cus, st);
        if (adj)
        {
            hipLaunchKernelGGL(k_delta, (grid_of<k_delta>(a.P, cus)), dim3(kTPB), 0, st, S, a, adj->dfilm,
                               a.one.scale);
            bounce_loop<true, false>(D, S, a, prm->max_depth, cus, st);
        }
        if (film)
        {
            hipLaunchKernelGGL(k_accumulate<false>, (grid_of<k_accumulate<false>>(np, cus)), dim3(kTPB), 0, st, S, a,
                               film, np);
        }
        HIPCHK(hipGetLastError());
    }
    return MCRT_OK;
}

// Batched rendering: device tables of the scenes, chunk segments cut at
// pixel boundaries, traversal settings of the first scene with the largest
// stack bound of the batch.
int batch_init(mcrt_batch* b)
{
    const size_t n = b->scenes.size();
    mcrt_scene* s0 = b->scenes[0];
    int64_t pmax = s0->pt_chunk_paths;
    int max_sb = 0;
    b->max_depth = 0;
    b->rs = 2;
    b->env_gen.clear();
    b->det_eps.clear();
    std::vector<ScenePT> tab(n);
    std::vector<SceneRoot> roots(n);
    for (size_t i = 0; i < n; i++)
    {
        mcrt_scene* s = b->scenes[i];
        tab[i] = scene_pt(s, s->dev, &b->params[i]);
        tab[i].film = b->films[i];
        roots[i].nodes = s->dev.nodes;
        roots[i].det_eps = s->dev.det_eps;
        roots[i].node_span = s->dev.node_span;
        pmax = std::max<int64_t>(pmax, b->params[i].spp);
        max_sb = std::max(max_sb, s->dev.stack_bound);
        if (s->dev.env_w > 0) b->rs = 4;
        b->env_gen.push_back(s->env_gen);
        b->det_eps.push_back(s->dev.det_eps);
        b->max_depth = std::max(b->max_depth, b->params[i].max_depth);
    }
    // chunks of whole pixels, at most pmax paths each
    std::vector<Segment> segs;
    b->chunks.clear();
    BatchChunk c{0, 0, 0, 0};
    b->total_paths = 0;
    for (size_t i = 0; i < n; i++)
    {
        const int spp = b->params[i].spp;
        const int64_t npix = (int64_t)b->scenes[i]->dev_cam.w * b->scenes[i]->dev_cam.h;
        for (int64_t pix = 0; pix < npix;)
        {
            const int64_t room = (pmax - c.P) / spp;
            if (room == 0)
            {
                b->chunks.push_back(c);
                c = BatchChunk{0, 0, (int64_t)segs.size(), 0};
                continue;
            }
            const int64_t take = std::min(room, npix - pix);
            Segment g;
            g.sid = (int32_t)i;
            g.spp = spp;
            g.pix0 = pix;
            g.p0 = c.P;
            g.q0 = c.npix;
            segs.push_back(g);
            c.nseg++;
            c.P += take * spp;
            c.npix += take;
            pix += take;
        }
        b->total_paths += npix * spp;
    }
    if (c.nseg) b->chunks.push_back(c);
    int rc;
    if ((rc = b->d_tab.upload(tab.data(), tab.size() * sizeof(ScenePT)))) return rc;
    if ((rc = b->d_roots.upload(roots.data(), roots.size() * sizeof(SceneRoot)))) return rc;
    if ((rc = b->d_seg.upload(segs.data(), segs.size() * sizeof(Segment)))) return rc;
    int64_t pcap = 0;
    for (const BatchChunk& k : b->chunks) pcap = std::max(pcap, k.P);
    if ((rc = pt_workspace(&b->res, pcap, false, true, b->rs))) return rc;
    DevScene& Dt = b->trace_view;
    Dt = s0->dev;
    Dt.nodes = nullptr;
    Dt.stack_bound = max_sb;
    trace_geometry(Dt, b->ctx->num_cus, s0->pt_trace_waves_per_cu, s0->pt_trace_lds_depth);
    Dt.refill_lanes = s0->pt_trace_refill_lanes;
    if (Dt.spill_depth > 0)
    {
        if ((rc = b->res.spill.ensure((size_t)Dt.spill_lanes * Dt.spill_depth * sizeof(uint64_t)))) return rc;
        Dt.spill = b->res.spill.as<uint64_t>();
    }
    return MCRT_OK;
}

// Every kernel of the batch on `st` (launches only: capturable).
int batch_enqueue(mcrt_batch* b, hipStream_t st)
{
    PTState S = state_of(&b->res, false);
    const int cus = b->ctx->num_cus;
    PTArgs a;
    a.pix0 = 0;
    a.rs = b->rs;
    a.tab = b->d_tab.as<ScenePT>();
    a.roots = b->d_roots.as<SceneRoot>();
    a.one = ScenePT();
    for (const BatchChunk& c : b->chunks)
    {
        a.P = c.P;
        a.nseg = (int32_t)c.nseg;
        a.seg = b->d_seg.as<Segment>() + c.seg0;
        bounce_loop<false, true>(b->trace_view, S, a, b->max_depth, cus, st);
        hipLaunchKernelGGL(k_accumulate<true>, (grid_of<k_accumulate<true>>(c.npix, cus)), dim3(kTPB), 0, st, S, a,
                           (float*)nullptr, c.npix);
        HIPCHK(hipGetLastError());
    }
    return MCRT_OK;
}
}  // namespace mcrt
