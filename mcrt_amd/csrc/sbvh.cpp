// Host SBVH builder (see sbvh.h).  Decisions follow SplitBVH::build
// (splitbvh.cpp:1435-1625) exactly; the speed comes from
//   * one pass over the references that bins all three axes for both the
//     object (SAH) and the spatial splitter, chunked across threads and
//     merged exactly (integer counts, min/max box unions);
//   * the median splitter's cost from a selection (nth_element) instead of
//     a full sort whenever no tie straddles the median -- the full std::sort
//     of the reference (same comparator, same libstdc++ algorithm, hence
//     the same permutation) runs only when the order is actually needed;
//   * subtrees built as OpenMP tasks, numbered afterwards in pre-order.
#include "sbvh.h"

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <omp.h>

namespace mcrt
{
namespace
{
const int kLeafSize = 4;            // k_sbvh_leaf_node_size, splitbvh.h:27
const int kBins = 16;               // k_sbvh_num_binning_buckets, splitbvh.h:28
const float kEdgeThreshold = 1e-3f; // k_edge_split_threshold, splitbvh.h:30
const size_t kParChunk = 1 << 15;   // references per binning/partition task
const size_t kParTask = 1 << 11;    // spawn subtree tasks above this size

struct TNode
{
    Box box;
    int leaf = 0, axis = -1, depth = 0;
    TNode* l = nullptr;
    TNode* r = nullptr;
    int nref = 0;
    SRef ref[kLeafSize];
};

// Lock-light node arena: each thread carves nodes from its own block.  The
// per-thread cursor is tagged with the arena's unique id (not its address,
// which a later arena may reuse).
class Arena
{
  public:
    Arena() : id_(next_id().fetch_add(1) + 1) {}
    TNode* alloc()
    {
        thread_local Local L;
        if (L.arena_id != id_ || L.used == kBlock)
        {
            TNode* blk = new TNode[kBlock];
            {
                std::lock_guard<std::mutex> g(mu_);
                blocks_.emplace_back(blk);
            }
            L.cur = blk;
            L.used = 0;
            L.arena_id = id_;
        }
        return &L.cur[L.used++];
    }

  private:
    static const int kBlock = 4096;
    struct Local
    {
        TNode* cur = nullptr;
        int used = 0;
        uint64_t arena_id = 0;
    };
    static std::atomic<uint64_t>& next_id()
    {
        static std::atomic<uint64_t> n{0};
        return n;
    }
    const uint64_t id_;
    std::mutex mu_;
    std::vector<std::unique_ptr<TNode[]>> blocks_;
};

struct Bin
{
    Box box;
    int entries = 0, exits = 0, valid = 0;
};
struct BinSet
{
    Bin sah[3][kBins];
    Bin spa[3][kBins];
};

inline int bin_of(float v, float nmin, float delta)  // splitbvh.cpp:286-287
{
    int i = int((v - nmin) / delta);
    return i < kBins - 1 ? i : kBins - 1;
}

struct Ctx
{
    const std::vector<Tri>* tris;
    Arena arena;
    std::atomic<int> max_depth{0};
};

// Bins refs [lo, hi) for every enabled axis: SAH bins by the min corner
// (splitbvh.cpp:280-295), spatial bins by entry/exit bin with tight clips
// of every spanned bin (:720-778).  The z sweep of the reference starts its
// middle-bin loop at the entry bin (:1047), re-clipping that bin with the
// identical arguments; the repeat cannot change a box or flag and is not
// performed here.
void bin_range(const Ctx& c, const SRef* refs, size_t lo, size_t hi, const Box& nb,
               const float* delta, const bool* en, BinSet& bs)
{
    const std::vector<Tri>& tris = *c.tris;
    for (size_t i = lo; i < hi; i++)
    {
        const SRef& r = refs[i];
        const Tri& t = tris[r.tri];
        for (int ax = 0; ax < 3; ax++)
        {
            if (!en[ax]) continue;
            const float nmin = nb.lo[ax], d = delta[ax];
            int k = bin_of(r.box.lo[ax], nmin, d);
            Bin& sb = bs.sah[ax][k];
            sb.entries++;
            sb.box.grow(r.box);
            sb.valid = 1;

            int i0 = k;
            int i1 = bin_of(r.box.hi[ax], nmin, d);
            float bmin = nmin + i0 * d;
            Box cb = clip_tight(t, r.box, bmin, bmin + d, ax);
            if (cb.valid()) { Bin& b = bs.spa[ax][i0]; b.entries++; b.box.grow(cb); b.valid = 1; }
            for (int j = i0 + 1; j < i1; j++)
            {
                bmin = nmin + j * d;
                cb = clip_tight(t, r.box, bmin, bmin + d, ax);
                if (cb.valid()) { Bin& b = bs.spa[ax][j]; b.box.grow(cb); b.valid = 1; }
            }
            bmin = nmin + i1 * d;
            cb = clip_tight(t, r.box, bmin, bmin + d, ax);
            if (cb.valid()) { Bin& b = bs.spa[ax][i1]; b.exits++; b.box.grow(cb); b.valid = 1; }
        }
    }
}

void merge_bins(BinSet& dst, const BinSet& src)
{
    for (int ax = 0; ax < 3; ax++)
        for (int k = 0; k < kBins; k++)
        {
            Bin& a = dst.sah[ax][k];
            const Bin& b = src.sah[ax][k];
            a.entries += b.entries; a.exits += b.exits; a.valid |= b.valid; a.box.grow(b.box);
            Bin& p = dst.spa[ax][k];
            const Bin& q = src.spa[ax][k];
            p.entries += q.entries; p.exits += q.exits; p.valid |= q.valid; p.box.grow(q.box);
        }
}

// Plane sweep shared by both binned splitters (splitbvh.cpp:307-367,
// :791-854): skip on the left cost alone when >= best, accept on <=.
void sweep(const Bin* bins, bool spatial, int axis, float& best, int& best_axis, int& best_bin)
{
    for (int i = 1; i < kBins; i++)
    {
        Box lb;
        int cl = 0;
        for (int j = 0; j < i; j++)
            if (bins[j].valid) { cl += bins[j].entries; lb.grow(bins[j].box); }
        float cost = 0.0f;
        if (cl > 0) cost += lb.half_area() * cl;
        else cost = kInf;
        if (cost >= best) continue;
        Box rb;
        int cr = 0;
        for (int j = i; j < kBins; j++)
            if (bins[j].valid) { cr += spatial ? bins[j].exits : bins[j].entries; rb.grow(bins[j].box); }
        if (cr > 0) cost += rb.half_area() * cr;
        else cost = kInf;
        if (cost <= best) { best = cost; best_axis = axis; best_bin = i; }
    }
}

struct CenterLess
{
    int ax;
    bool operator()(const SRef& a, const SRef& b) const { return a.box.center(ax) < b.box.center(ax); }
};

size_t n_chunks(size_t n) { return (n + kParChunk - 1) / kParChunk; }

TNode* build_node(Ctx& c, std::vector<SRef>* refs_p, int depth);

// Spawns the two subtrees (tasks above kParTask references).
void build_children(Ctx& c, TNode* node, std::vector<SRef>* l, std::vector<SRef>* r, int depth)
{
    if (l->size() + r->size() > kParTask)
    {
        TNode* ln = nullptr;
#pragma omp task shared(ln, c) firstprivate(l, depth)
        ln = build_node(c, l, depth + 1);
        TNode* rn = build_node(c, r, depth + 1);
#pragma omp taskwait
        node->l = ln;
        node->r = rn;
    }
    else
    {
        node->l = build_node(c, l, depth + 1);
        node->r = build_node(c, r, depth + 1);
    }
}

TNode* build_node(Ctx& c, std::vector<SRef>* refs_p, int depth)
{
    std::unique_ptr<std::vector<SRef>> owned(refs_p);
    std::vector<SRef>& refs = *refs_p;
    const size_t n = refs.size();
    TNode* node = c.arena.alloc();
    node->depth = depth;
    {
        int md = c.max_depth.load(std::memory_order_relaxed);
        while (depth > md && !c.max_depth.compare_exchange_weak(md, depth)) {}
    }
    Box nb;
    for (const SRef& r : refs) nb.grow(r.box);
    node->box = nb;
    if ((int)n <= kLeafSize)  // splitbvh.cpp:1462-1482
    {
        node->leaf = 1;
        node->nref = (int)n;
        for (size_t i = 0; i < n; i++) node->ref[i] = refs[i];
        return node;
    }

    // --- binning for the object and spatial splitters -----------------
    F3 edge = sub(nb.hi, nb.lo);
    float delta[3];
    bool en[3];
    for (int ax = 0; ax < 3; ax++)
    {
        en[ax] = edge[ax] > kEdgeThreshold;
        delta[ax] = edge[ax] / kBins;
    }
    BinSet bins;
    if (n > kParChunk)
    {
        size_t nc = n_chunks(n);
        std::vector<BinSet> part(nc);
        for (size_t k = 0; k < nc; k++)
        {
#pragma omp task shared(c, refs, nb, delta, en, part) firstprivate(k)
            bin_range(c, refs.data(), k * kParChunk, std::min(n, (k + 1) * kParChunk), nb, delta, en,
                      part[k]);
        }
#pragma omp taskwait
        for (size_t k = 0; k < nc; k++) merge_bins(bins, part[k]);
    }
    else
        bin_range(c, refs.data(), 0, n, nb, delta, en, bins);

    float sah_cost = kInf, spa_cost = kInf;
    int sah_axis = -1, sah_bin = -1, spa_axis = -1, spa_bin = -1;
    for (int ax = 0; ax < 3; ax++)
        if (en[ax]) sweep(bins.sah[ax], false, ax, sah_cost, sah_axis, sah_bin);
    for (int ax = 0; ax < 3; ax++)
        if (en[ax]) sweep(bins.spa[ax], true, ax, spa_cost, spa_axis, spa_bin);

    // --- median splitter (splitbvh.cpp:178-236) -------------------------
    const int med_axis = nb.largest_axis();
    const int mid = (int)n / 2;
    std::vector<SRef> sorted;  // filled when the exact order is known
    float med_cost;
    {
        std::vector<float> keys(n);
        bool nan = false;
        for (size_t i = 0; i < n; i++)
        {
            keys[i] = refs[i].box.center(med_axis);
            nan |= keys[i] != keys[i];
        }
        bool decided = false;
        if (!nan)
        {
            std::vector<float> sel(keys);
            std::nth_element(sel.begin(), sel.begin() + mid, sel.end());
            const float K = sel[mid];
            size_t lt = 0;
            for (float k : keys) lt += k < K;
            if (lt == (size_t)mid)
            {
                Box lb, rb;
                for (size_t i = 0; i < n; i++)
                    if (keys[i] < K) lb.grow(refs[i].box);
                    else rb.grow(refs[i].box);
                med_cost = lb.half_area() * mid + rb.half_area() * ((int)n - mid);
                decided = true;
            }
        }
        if (!decided)
        {
            sorted = refs;
            std::sort(sorted.begin(), sorted.end(), CenterLess{med_axis});
            Box lb, rb;
            for (int i = 0; i < mid; i++) lb.grow(sorted[i].box);
            for (size_t i = mid; i < n; i++) rb.grow(sorted[i].box);
            med_cost = lb.half_area() * mid + rb.half_area() * ((int)n - mid);
        }
    }

    // --- choose (splitbvh.cpp:1520-1537): SAH first, strict > swaps -------
    int types[3] = {1, 0, 2};
    float costs[3] = {sah_cost, med_cost, spa_cost};
    if (costs[0] > costs[1]) { std::swap(costs[0], costs[1]); std::swap(types[0], types[1]); }
    if (costs[1] > costs[2]) { std::swap(costs[1], costs[2]); std::swap(types[1], types[2]); }
    if (costs[0] > costs[1]) { std::swap(costs[0], costs[1]); std::swap(types[0], types[1]); }

    auto* left = new std::vector<SRef>();
    auto* right = new std::vector<SRef>();
    if (types[0] == 0)
    {
        if (sorted.empty())
        {
            sorted = refs;
            std::sort(sorted.begin(), sorted.end(), CenterLess{med_axis});
        }
        node->axis = med_axis;
        left->assign(sorted.begin(), sorted.begin() + mid);
        right->assign(sorted.begin() + mid, sorted.end());
    }
    else
    {
        const bool spatial = types[0] == 2;
        const int ax = spatial ? spa_axis : sah_axis;
        const int bin = spatial ? spa_bin : sah_bin;
        node->axis = ax;
        const float d = delta[ax];
        const float nmin = nb.lo[ax], nmax = nb.hi[ax];
        const float clip = nmin + bin * d;  // splitbvh.cpp:1232
        const std::vector<Tri>& tris = *c.tris;
        // SAHBVHSplitter::sort (:624-662) / SplitBVHSplitter::sort (:1216-1290),
        // order preserving; large inputs are split into chunks whose outputs
        // are concatenated in chunk order.
        auto part = [&](size_t lo, size_t hi, std::vector<SRef>& L, std::vector<SRef>& R) {
            for (size_t i = lo; i < hi; i++)
            {
                const SRef& x = refs[i];
                int i0 = bin_of(x.box.lo[ax], nmin, d);
                if (!spatial)
                {
                    (i0 < bin ? L : R).push_back(x);
                    continue;
                }
                int i1 = bin_of(x.box.hi[ax], nmin, d);
                const Tri& t = tris[x.tri];
                if (i0 < bin)
                {
                    Box cb = clip_tight(t, x.box, nmin, clip, ax);
                    if (cb.valid()) L.push_back(SRef{cb, x.tri});
                }
                if (i1 >= bin)
                {
                    Box cb = clip_tight(t, x.box, clip, nmax, ax);
                    if (cb.valid()) R.push_back(SRef{cb, x.tri});
                }
            }
        };
        if (n > kParChunk)
        {
            size_t nc = n_chunks(n);
            std::vector<std::vector<SRef>> pl(nc), pr(nc);
            for (size_t k = 0; k < nc; k++)
            {
#pragma omp task shared(pl, pr, part) firstprivate(k)
                part(k * kParChunk, std::min(n, (k + 1) * kParChunk), pl[k], pr[k]);
            }
#pragma omp taskwait
            size_t sl = 0, sr = 0;
            for (size_t k = 0; k < nc; k++) { sl += pl[k].size(); sr += pr[k].size(); }
            left->reserve(sl);
            right->reserve(sr);
            for (size_t k = 0; k < nc; k++)
            {
                left->insert(left->end(), pl[k].begin(), pl[k].end());
                right->insert(right->end(), pr[k].begin(), pr[k].end());
            }
        }
        else
            part(0, n, *left, *right);
    }
    owned.reset();
    { std::vector<SRef>().swap(sorted); }
    build_children(c, node, left, right, depth);
    return node;
}

int32_t emit(const TNode* t, BTree& out)
{
    int32_t id = (int32_t)out.nodes.size();
    out.nodes.emplace_back();
    BNode bn;
    bn.box = t->box;
    bn.leaf = t->leaf;
    bn.axis = t->axis;
    bn.left = bn.right = -1;
    bn.ref_begin = bn.ref_count = 0;
    if (t->leaf)
    {
        bn.ref_begin = (int32_t)out.refs.size();
        bn.ref_count = t->nref;
        for (int i = 0; i < t->nref; i++) out.refs.push_back(t->ref[i]);
        out.nodes[id] = bn;
        return id;
    }
    bn.left = emit(t->l, out);
    bn.right = emit(t->r, out);
    out.nodes[id] = bn;
    return id;
}
}  // namespace

void build_sbvh(const std::vector<Tri>& tris, int threads, BTree& out)
{
    out.nodes.clear();
    out.refs.clear();
    out.max_depth = 0;
    const size_t n = tris.size();
    if (n == 0) return;
    Ctx c;
    c.tris = &tris;
    auto* refs = new std::vector<SRef>(n);
    for (size_t i = 0; i < n; i++) (*refs)[i] = SRef{tri_box(tris[i]), (int32_t)i};  // :1425-1430
    TNode* root = nullptr;
    int nt = threads > 0 ? threads : omp_get_num_procs();
#pragma omp parallel num_threads(nt)
    {
#pragma omp single
        root = build_node(c, refs, 0);
    }
    out.max_depth = c.max_depth.load();
    out.nodes.reserve(2 * n);
    emit(root, out);
}

namespace
{
struct Fnv
{
    uint64_t h = 1469598103934665603ull;
    void bytes(const void* p, size_t n)
    {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    }
    void f(float x) { if (x == 0.0f) x = 0.0f; bytes(&x, 4); }
    void i(int32_t x) { bytes(&x, 4); }
    void box(const Box& b) { for (int k = 0; k < 3; k++) f(b.lo[k]); for (int k = 0; k < 3; k++) f(b.hi[k]); }
};
}  // namespace

uint64_t tree_digest(const BTree& t)
{
    Fnv f;
    if (t.nodes.empty()) return f.h;
    std::vector<int32_t> st{0};
    while (!st.empty())
    {
        const BNode& n = t.nodes[st.back()];
        st.pop_back();
        f.i(n.leaf ? 1 : 0);
        f.box(n.box);
        if (n.leaf)
        {
            f.i(n.ref_count);
            for (int k = 0; k < n.ref_count; k++)
            {
                const SRef& r = t.refs[n.ref_begin + k];
                f.i(r.tri);
                f.box(r.box);
            }
            continue;
        }
        f.i(n.axis);
        st.push_back(n.right);
        st.push_back(n.left);
    }
    return f.h;
}
}  // namespace mcrt
