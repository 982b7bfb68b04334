// Test helper: runs the product's traversal core (mcrt_amd/csrc/trace_core.h,
// the same code the GPU kernels execute) on the host over the product's
// SBVH builder and 4-wide collapse, so the whole query pipeline can be
// checked against the oracle on a machine without a GPU.  Built into
// tests/hostemu/libhostemu.so; loaded only by tests.
#include <algorithm>
#include <cstring>
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
    std::vector<QNode> recs;  // the device's record array: nodes, then references (layout.h)
    uint64_t digest = 0;
    long long stats[8] = {0};
    float det_eps = 1e-6f;  // Moller-Trumbore determinant cut of the traversal
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
    if (!device_records(e->wb, e->recs))
    {
        delete e;
        return nullptr;
    }
    const size_t nq = e->recs.size();
    e->recs.resize(nq + e->wb.refs.size());
    static_assert(sizeof(QNode) == sizeof(WRef), "records are 64 B");
    if (!e->wb.refs.empty()) std::memcpy(&e->recs[nq], e->wb.refs.data(), e->wb.refs.size() * sizeof(WRef));
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

extern "C" void emu_set_eps(void* h, float eps) { ((Emu*)h)->det_eps = eps; }

// Quantised nodes against the exact ones: out[0] = planes checked, out[1] =
// decoded planes that do not contain the exact child box, out[2] = child
// links or rank tables (inverse order tables) that differ, out[3] = sum over checked planes of
// (decoded - exact) outward growth relative to the node extent, x 1e6.
extern "C" void emu_check_qnodes(void* h, long long* out)
{
    Emu* e = (Emu*)h;
    long long checked = 0, bad = 0, links = 0;
    double grow = 0.0;
    for (size_t i = 0; i < e->wb.nodes.size(); i++)
    {
        const mcrt::WNode& w = e->wb.nodes[i];
        const mcrt::QNode& q = e->wb.qnodes[i];
        for (int o = 0; o < 8; o++)
            for (int k = 0; k < 4; k++)
            {
                const uint32_t ord = o < 4 ? w.order_lo : w.order_hi, rk = o < 4 ? q.rank_lo : q.rank_hi;
                const uint32_t s = (ord >> (8 * (o & 3) + 2 * k)) & 3u;
                if (((rk >> (8 * (o & 3) + 2 * s)) & 3u) != (uint32_t)k) links++;
            }
        const float* lo[3] = {w.lo_x, w.lo_y, w.lo_z};
        const float* hi[3] = {w.hi_x, w.hi_y, w.hi_z};
        const uint32_t qlo[3] = {q.q_lo_x, q.q_lo_y, q.q_lo_z};
        const uint32_t qhi[3] = {q.q_hi_x, q.q_hi_y, q.q_hi_z};
        for (int c = 0; c < 4; c++)
        {
            if (q.child[c] != w.child[c]) links++;
            if (w.child[c] == MCRT_CHILD_EMPTY) continue;
            for (int a = 0; a < 3; a++)
            {
                const float sc = mcrt::float_of(((q.exps >> (8 * a)) & 0xffu) << 23);
                const float dl = mcrt::qnode_plane(q.origin[a], (qlo[a] >> (8 * c)) & 0xffu, sc);
                const float dh = mcrt::qnode_plane(q.origin[a], (qhi[a] >> (8 * c)) & 0xffu, sc);
                checked += 2;
                if (!(dl <= lo[a][c])) bad++;
                if (!(dh >= hi[a][c])) bad++;
                float ext = 0.0f;
                for (int k = 0; k < 4; k++)
                    if (w.child[k] != MCRT_CHILD_EMPTY) ext = std::max(ext, hi[a][k] - q.origin[a]);
                if (ext > 0.0f) grow += ((double)(lo[a][c] - dl) + (double)(dh - hi[a][c])) / ext;
            }
        }
    }
    out[0] = checked;
    out[1] = bad;
    out[2] = links;
    out[3] = (long long)(grow * 1e6);
}

// rays n x 8; hits n x 4 int32 {prim, t, u, v}; occ n bytes (any-hit);
// counters 3 totals (node fetches, ref fetches, MT tests).
extern "C" void emu_trace(void* h, const float* rays, long long n, int* hits, unsigned char* occ, int any,
                          int lds_depth, long long* counters, int threads)
{
    Emu* e = (Emu*)h;
    DevScene S;
    S.nodes = e->wb.qnodes.empty() ? nullptr : e->recs.data();
    S.stack_bound = e->wb.stack_bound;
    S.node_span = qnode_span(e->wb.qnodes);
    S.det_eps = e->det_eps;
    // std::thread, not OpenMP: this TU is compiled by hipcc, whose OpenMP
    // runtime differs from the libgomp that the rest of the process uses.
    const int nt = std::max(1, threads);
    std::atomic<long long> next{0};
    std::vector<long long> cnts(3 * nt, 0);
    auto work = [&](int tid) {
        std::vector<uint32_t> sn(lds_depth + 1);  // + the discard row
        std::vector<float> sl(lds_depth + 1);
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
        st.spill_stride = 1;
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

// The node-test margin (trace_core.h: node_margin) against the reference's
// slab arithmetic on single planes: n random cases of node origin, scale
// 2^k, plane code, ray origin and direction (magnitudes from tiny to huge,
// zero direction components), each a node of span |origin| + 255 * 2^k
// (the tightest span a scene can give it).  out[0] = cases checked (rays
// the margin's bounds admit), out[1] = cases where the lower bound exceeds
// or the upper bound falls below RN(RN(plane - o) * inv), out[2] = cases
// sent to the exact path; worst = the largest share of the margin m the
// rounding used: (t_near - (t_ref - m)) / m or ((t_ref + m) - t_far) / m
// (1 or more: a bound crossed t_ref).  `wscale` multiplies the
// margin (1 for the product's; below 1 to show the check catches a margin
// that is too small).
extern "C" void emu_margin_check(long long n, unsigned long long seed, float wscale, long long* out, double* worst)
{
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&st]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    auto unif = [&]() { return (double)(next() >> 11) * 0x1p-53; };
    long long checked = 0, bad = 0, exact = 0;
    double wmax = 0.0;
    for (long long c = 0; c < n; c++)
    {
        const int k = (int)(unif() * 90.0) - 60;  // scale 2^k, k in [-60, 29]
        const float s = std::ldexp(1.0f, k);
        const float O = (float)((unif() * 2.0 - 1.0) * std::ldexp(1.0, k + (int)(unif() * 12.0)));
        const uint32_t q = (uint32_t)(next() & 0xffu);
        const float span = std::nextafter((float)(std::fabs((double)O) + 255.0 * (double)s), INFINITY);
        float o[3], d[3];
        for (int a = 0; a < 3; a++)
        {
            const double mag = std::fabs((double)O) + 255.0 * (double)s;
            const int kind = (int)(next() % 4);
            if (kind == 0)
                o[a] = (float)((unif() * 2.0 - 1.0) * mag * std::pow(10.0, unif() * 6.0 - 3.0));
            else if (kind == 1)
                o[a] = std::fma((float)(next() & 0xffu), s, O);  // on a plane of the node
            else
                o[a] = (float)((double)O + (unif() * 2.0 - 1.0) * 300.0 * (double)s);
            const int dk = (int)(next() % 8);
            const double dm = dk == 0 ? 0.0 : std::pow(10.0, unif() * 60.0 - 30.0);
            d[a] = (float)((next() & 1) ? dm : -dm);
        }
        RayPrep r = prep_ray(o[0], o[1], o[2], d[0], d[1], d[2], 1e-4f);
        node_margin(r, span);
        if (wscale != 1.0f)
        {
            // the x axis's offsets re-formed with the margin scaled
            r.w *= wscale;
            const float m = r.w * std::fabs(r.ix), p = r.ox * r.ix;
            r.cnx = -p - m;
            r.cfx = m - p;
        }
        if (r.exact)
        {
            exact++;
            continue;
        }
        const float oo = r.ox, inv = r.ix;  // the x axis carries the case
        const float P = __builtin_fmaf((float)q, s, O);
        const float tref = (P - oo) * inv;
        const float A = s * inv;
        // the kernels' node-step plane map (trace_core.h: Tracer::visit)
        const float tn = __builtin_fmaf((float)q, A, __builtin_fmaf(O, inv, r.cnx));
        const float tf = __builtin_fmaf((float)q, A, __builtin_fmaf(O, inv, r.cfx));
        checked++;
        if (!(tn <= tref && tf >= tref)) bad++;
        const double m = (double)r.w * std::fabs((double)inv);
        if (m > 0.0)
        {
            const double en = ((double)tn - ((double)tref - m)) / m, ef = (((double)tref + m) - (double)tf) / m;
            wmax = std::max(wmax, std::max(en, ef));
        }
    }
    out[0] = checked;
    out[1] = bad;
    out[2] = exact;
    *worst = wmax;
}
