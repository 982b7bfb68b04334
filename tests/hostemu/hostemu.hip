// Test helper: runs the product's traversal core (mcrt_amd/csrc/trace_core.h,
// the same code the GPU kernels execute) on the host over the product's
// SBVH builder and 4-wide collapse, so the whole query pipeline can be
// checked against the oracle on a machine without a GPU.  Built into
// tests/hostemu/libhostemu.so; loaded only by tests.
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "sbvh.h"
#include "trace_core.h"
#include "wide.h"

using namespace mcrt;

namespace
{
struct Emu
{
    std::vector<Tri> tris;
    WideBVH wb;
    uint64_t digest = 0;
    long long stats[8] = {0};
};
}  // namespace

extern "C" void* emu_build(const float* v9, long long n, int threads)
{
    Emu* e = new Emu();
    e->tris.resize(n);
    for (long long i = 0; i < n; i++)
        for (int k = 0; k < 3; k++)
        {
            e->tris[i].a[k] = v9[9 * i + k];
            e->tris[i].b[k] = v9[9 * i + 3 + k];
            e->tris[i].c[k] = v9[9 * i + 6 + k];
        }
    BTree bt;
    build_sbvh(e->tris, threads, bt);
    e->digest = tree_digest(bt);
    long long leaves = 0, empty = 0;
    for (const BNode& b : bt.nodes)
        if (b.leaf) { leaves++; if (b.ref_count == 0) empty++; }
    collapse_wide(bt, e->tris, e->wb);
    e->stats[0] = (long long)bt.nodes.size();
    e->stats[1] = leaves;
    e->stats[2] = (long long)bt.refs.size();
    e->stats[3] = bt.max_depth;
    e->stats[4] = empty;
    e->stats[5] = (long long)e->wb.nodes.size();
    e->stats[6] = e->wb.stack_bound;
    e->stats[7] = e->wb.depth;
    return e;
}

extern "C" unsigned long long emu_digest(void* h, long long* stats)
{
    Emu* e = (Emu*)h;
    for (int k = 0; k < 8; k++) stats[k] = e->stats[k];
    return e->digest;
}

extern "C" void emu_free(void* h) { delete (Emu*)h; }

// rays n x 8; hits n x 4 int32 {prim, t, u, v}; occ n bytes (any-hit);
// counters 3 totals (node fetches, ref fetches, MT tests).
extern "C" void emu_trace(void* h, const float* rays, long long n, int* hits, unsigned char* occ, int any,
                          int lds_depth, long long* counters, int threads)
{
    Emu* e = (Emu*)h;
    DevScene S;
    S.nodes = e->wb.nodes.empty() ? nullptr : e->wb.nodes.data();
    S.refs = e->wb.refs.data();
    S.stack_bound = e->wb.stack_bound;
    // std::thread, not OpenMP: this TU is compiled by hipcc, whose OpenMP
    // runtime differs from the libgomp that the rest of the process uses.
    const int nt = std::max(1, threads);
    std::atomic<long long> next{0};
    std::vector<long long> cnts(3 * nt, 0);
    auto work = [&](int tid) {
        std::vector<uint32_t> sn(lds_depth);
        std::vector<float> sl(lds_depth);
        int spill_depth = e->wb.stack_bound > lds_depth ? e->wb.stack_bound - lds_depth : 0;
        std::vector<uint64_t> spill(spill_depth + 1);
        TraceStack st;
        st.node = sn.data();
        st.lo = sl.data();
        st.stride = 1;
        st.lane = 0;
        st.depth = lds_depth;
        st.spill = spill.data();
        st.spill_lane = 0;
        st.spill_depth = spill_depth;
        while (true)
        {
            long long b = next.fetch_add(1024);
            if (b >= n) break;
            long long end = std::min(n, b + 1024);
            for (long long i = b; i < end; i++)
            {
                const float* r = rays + 8 * i;
                RayPrep rp = prep_ray(r[0], r[1], r[2], r[4], r[5], r[6], r[3]);
                TraceCount c;
                float t, u, v;
                int prim = any ? trace<true, true>(S, st, rp, r[7], t, u, v, c)
                               : trace<false, true>(S, st, rp, r[7], t, u, v, c);
                if (any)
                    occ[i] = prim >= 0;
                else
                {
                    hits[4 * i] = prim;
                    hits[4 * i + 1] = (int)bits_of(t);
                    hits[4 * i + 2] = (int)bits_of(u);
                    hits[4 * i + 3] = (int)bits_of(v);
                }
                cnts[3 * tid] += c.nodes;
                cnts[3 * tid + 1] += c.refs;
                cnts[3 * tid + 2] += c.tris;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++) pool.emplace_back(work, t);
    for (auto& t : pool) t.join();
    long long cn = 0, cr = 0, ct = 0;
    for (int t = 0; t < nt; t++) { cn += cnts[3 * t]; cr += cnts[3 * t + 1]; ct += cnts[3 * t + 2]; }
    if (counters) { counters[0] = cn; counters[1] = cr; counters[2] = ct; }
}
