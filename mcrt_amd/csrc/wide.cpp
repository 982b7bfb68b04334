// Binary SBVH -> 4-wide nodes.  A wide node is a treelet of up to three
// binary branches whose frontier (up to four binary nodes) becomes its
// children.  The frontier grows by opening the largest-area branch, or by
// expanding the largest-area leaf into its references -- preferred unless
// a branch is over four times larger -- which then become children (each
// with its own clipped box, tested in the node instead of being fetched to
// be tested; when they do not all fit, the last child holds the rest).  Dropping the
// treelet's inner boxes is exact: a child box lies inside its parent's box
// (node box = union of its references' clipped boxes, and every child
// reference's box lies inside a parent reference's box), and the slab
// interval of a contained box is contained after rounding, so a child
// passes only if its parent passes.  Leaf boxes are likewise implied by
// the boxes of their references.  The visit order of the binary traversal
// (near child first iff dir[axis] > 0, splitbvh.cpp:73-134; a leaf's
// references in slot order) is stored per ray octant.
#include "wide.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace mcrt
{
namespace
{
struct TL
{
    int axis;  // split axis; -1: a leaf's references, l before r in every octant
    int l, r;  // >= 0: treelet node; < 0: slot -(x+1)
};

struct Builder
{
    const BTree& bt;
    const std::vector<Tri>& tris;
    WideBVH& out;

    void dfs_slots(const std::vector<TL>& tl, int k, std::vector<int>& seq, const int pos[3]) const
    {
        const TL& t = tl[k];
        int first = (t.axis < 0 || pos[t.axis]) ? t.l : t.r;
        int second = (t.axis < 0 || pos[t.axis]) ? t.r : t.l;
        for (int x : {first, second})
        {
            if (x >= 0) dfs_slots(tl, x, seq, pos);
            else seq.push_back(-x - 1);
        }
    }

    // Leaf link to references [k0, k0 + cnt) of leaf b (cnt = 0: all of them).
    uint32_t leaf_link(const BNode& b, int k0 = 0, int cnt = 0)
    {
        uint32_t first = (uint32_t)out.refs.size();
        const int k1 = cnt > 0 ? k0 + cnt : b.ref_count;
        for (int k = k0; k < k1; k++)
        {
            const SRef& s = bt.refs[b.ref_begin + k];
            const Tri& t = tris[s.tri];
            WRef w;
            for (int a = 0; a < 3; a++)
            {
                w.lo[a] = s.box.lo[a];
                w.hi[a] = s.box.hi[a];
                w.v0[a] = t.a[a];
                w.e1[a] = t.b[a] - t.a[a];  // v0v1 (triangle.cpp:16)
                w.e2[a] = t.c[a] - t.a[a];  // v0v2 (triangle.cpp:17)
            }
            w.prim = s.tri;
            out.refs.push_back(w);
        }
        return MCRT_CHILD_LEAF | ((uint32_t)(k1 - k0 - 1) << 29) | first;
    }

    // Returns the wide index and writes the subtree's stack need and depth.
    int32_t make(const std::vector<int32_t>& frontier_in, const std::vector<TL>& tl_in, int& need,
                 int& depth)
    {
        std::vector<int32_t> slot = frontier_in;
        // sref >= 0: the slot holds references [sref, sref + scnt) of its leaf
        std::vector<int> sref(slot.size(), -1), scnt(slot.size(), 0);
        std::vector<TL> tl = tl_in;
        auto hook = [&](int from, int tk) {
            for (size_t k = 0; k + 1 < tl.size(); k++)
            {
                if (tl[k].l == -(from + 1)) tl[k].l = tk;
                else if (tl[k].r == -(from + 1)) tl[k].r = tk;
            }
        };
        // Open the largest-area branch, or expand the largest-area leaf into
        // its references, until four children.
        while (slot.size() < 4)
        {
            int best = -1, best_leaf = -1;
            float best_a = -1.0f, best_la = -1.0f;
            for (size_t i = 0; i < slot.size(); i++)
            {
                if (sref[i] >= 0) continue;
                const BNode& b = bt.nodes[slot[i]];
                float a = b.box.half_area();
                if (b.leaf)
                {
                    if (b.ref_count >= 2 && a > best_la)
                    {
                        best_la = a;
                        best_leaf = (int)i;
                    }
                    continue;
                }
                if (a > best_a) { best_a = a; best = (int)i; }
            }
            // leaves first unless a branch is over four times larger: a
            // reference tested in the node saves a 64 B fetch and a step
            // whenever its box misses (measured 0.25..0.1 best, 1 slower)
            if (best_leaf >= 0 && (best < 0 || best_la > 0.25f * best_a))
            {
                // a leaf's references become children, in slot order; when
                // they do not all fit, the last child holds the rest
                const BNode& b = bt.nodes[slot[best_leaf]];
                const int base = (int)slot.size();
                const int items = std::min(b.ref_count, 4 - base + 1);
                sref[best_leaf] = 0;
                scnt[best_leaf] = 1;
                int prev_slot = best_leaf;
                for (int k = 1; k < items; k++)
                {
                    slot.push_back(slot[best_leaf]);
                    sref.push_back(k);
                    scnt.push_back(k + 1 < items ? 1 : b.ref_count - k);
                    const int tk = (int)tl.size();
                    tl.push_back(TL{-1, -(prev_slot + 1), -(base + k - 1 + 1)});
                    hook(prev_slot, tk);
                    prev_slot = base + k - 1;
                }
                continue;
            }
            if (best < 0) break;
            const BNode& b = bt.nodes[slot[best]];
            int ns = (int)slot.size();
            slot[best] = b.left;
            slot.push_back(b.right);
            sref.push_back(-1);
            scnt.push_back(0);
            int tk = (int)tl.size();
            tl.push_back(TL{b.axis, -(best + 1), -(ns + 1)});
            hook(best, tk);
        }
        // Renumber slots in left-first treelet order.
        std::vector<int> canon;
        {
            const int all_pos[3] = {1, 1, 1};
            dfs_slots(tl, 0, canon, all_pos);
        }
        std::vector<int> remap(slot.size());
        std::vector<int32_t> cslot(slot.size());
        std::vector<int> csref(slot.size()), cscnt(slot.size());
        for (size_t i = 0; i < canon.size(); i++)
        {
            remap[canon[i]] = (int)i;
            cslot[i] = slot[canon[i]];
            csref[i] = sref[canon[i]];
            cscnt[i] = scnt[canon[i]];
        }
        for (TL& t : tl)
        {
            if (t.l < 0) t.l = -(remap[-t.l - 1] + 1);
            if (t.r < 0) t.r = -(remap[-t.r - 1] + 1);
        }

        const int32_t idx = (int32_t)out.nodes.size();
        out.nodes.emplace_back();
        WNode w;
        std::memset(&w, 0, sizeof(w));
        uint64_t order = 0;
        for (int o = 0; o < 8; o++)
        {
            const int pos[3] = {o & 1, (o >> 1) & 1, (o >> 2) & 1};
            std::vector<int> seq;
            dfs_slots(tl, 0, seq, pos);
            for (int s = 0; s < 4; s++)
                if (std::find(seq.begin(), seq.end(), s) == seq.end()) seq.push_back(s);
            uint32_t byte = 0;
            for (int k = 0; k < 4; k++) byte |= (uint32_t)seq[k] << (2 * k);
            order |= (uint64_t)byte << (8 * o);
        }
        w.order_lo = (uint32_t)order;
        w.order_hi = (uint32_t)(order >> 32);

        int nchild = 0, child_need = 0, child_depth = 0;
        for (int s = 0; s < 4; s++)
        {
            w.child[s] = MCRT_CHILD_EMPTY;
            if (s >= (int)cslot.size()) continue;
            const BNode& b = bt.nodes[cslot[s]];
            if (b.leaf && b.ref_count == 0) continue;  // cannot produce a hit
            // references as a child: the union of their clipped boxes
            Box cb = b.box;
            if (csref[s] >= 0)
            {
                cb = bt.refs[b.ref_begin + csref[s]].box;
                for (int k = 1; k < cscnt[s]; k++) cb.grow(bt.refs[b.ref_begin + csref[s] + k].box);
            }
            w.lo_x[s] = cb.lo[0]; w.hi_x[s] = cb.hi[0];
            w.lo_y[s] = cb.lo[1]; w.hi_y[s] = cb.hi[1];
            w.lo_z[s] = cb.lo[2]; w.hi_z[s] = cb.hi[2];
            nchild++;
            if (b.leaf)
            {
                w.child[s] = csref[s] >= 0 ? leaf_link(b, csref[s], cscnt[s]) : leaf_link(b);
                continue;
            }
            int cn = 0, cd = 0;
            std::vector<TL> ctl{TL{b.axis, -1, -2}};
            w.child[s] = (uint32_t)make({b.left, b.right}, ctl, cn, cd);
            child_need = std::max(child_need, cn);
            child_depth = std::max(child_depth, cd);
        }
        w.nchild = (uint32_t)nchild;
        out.nodes[idx] = w;
        // The kernel pushes every passing child and pops one.
        need = std::max(nchild, nchild - 1 + child_need);
        depth = 1 + child_depth;
        return idx;
    }
};
}  // namespace

void collapse_wide(const BTree& bt, const std::vector<Tri>& tris, WideBVH& out)
{
    out.nodes.clear();
    out.refs.clear();
    out.stack_bound = 0;
    out.depth = 0;
    if (bt.nodes.empty()) return;
    Builder b{bt, tris, out};
    const BNode& root = bt.nodes[0];
    int need = 0, depth = 0;
    if (root.leaf)
    {
        WNode w;
        std::memset(&w, 0, sizeof(w));
        for (int s = 0; s < 4; s++) w.child[s] = MCRT_CHILD_EMPTY;
        w.order_lo = w.order_hi = 0xe4e4e4e4u;  // 0,1,2,3 for every octant
        out.nodes.push_back(w);
        if (root.ref_count > 0)
        {
            WNode& r = out.nodes[0];
            r.lo_x[0] = root.box.lo[0]; r.hi_x[0] = root.box.hi[0];
            r.lo_y[0] = root.box.lo[1]; r.hi_y[0] = root.box.hi[1];
            r.lo_z[0] = root.box.lo[2]; r.hi_z[0] = root.box.hi[2];
            uint32_t link = b.leaf_link(root);
            out.nodes[0].child[0] = link;
            out.nodes[0].nchild = 1;
        }
        out.stack_bound = 1;
        out.depth = 1;
        quantize_nodes(out);
        return;
    }
    std::vector<TL> tl{TL{root.axis, -1, -2}};
    b.make({root.left, root.right}, tl, need, depth);
    out.stack_bound = need;
    out.depth = depth;
    quantize_nodes(out);
}

namespace
{
// One axis of a node: origin = lowest child plane; scale 2^k with the
// smallest k in [-126, 127] at which every child's planes round outwards
// within 8 bits (checked with the decode expression, layout.h: qnode_plane).
bool quantize_axis(const float* lo, const float* hi, const bool* valid, float& origin, uint32_t& ebyte,
                   uint32_t& qlo, uint32_t& qhi)
{
    bool any = false;
    float o = 0.0f, top = 0.0f;
    for (int c = 0; c < 4; c++)
        if (valid[c])
        {
            o = any ? std::min(o, lo[c]) : lo[c];
            top = any ? std::max(top, hi[c]) : hi[c];
            any = true;
        }
    origin = o;
    qlo = qhi = 0;
    if (!any)
    {
        ebyte = 127;
        return true;
    }
    const double ext = (double)top - (double)o;
    int k = ext > 0.0 ? (int)std::ceil(std::log2(ext / 250.0)) : -126;
    k = std::max(k, -126);
    for (; k <= 127; k++)
    {
        const float sc = std::ldexp(1.0f, k);
        uint32_t ql = 0, qh = 0;
        bool ok = true;
        for (int c = 0; c < 4 && ok; c++)
        {
            if (!valid[c]) continue;
            double a = std::floor(((double)lo[c] - (double)o) / sc);
            uint32_t l = (uint32_t)std::min(255.0, std::max(0.0, a));
            while (l > 0 && qnode_plane(o, l, sc) > lo[c]) l--;
            double b = std::ceil(((double)hi[c] - (double)o) / sc);
            if (!(b <= 255.0))
            {
                ok = false;
                break;
            }
            uint32_t h = (uint32_t)std::max(0.0, b);
            while (h < 255 && qnode_plane(o, h, sc) < hi[c]) h++;
            if (qnode_plane(o, l, sc) > lo[c] || qnode_plane(o, h, sc) < hi[c])
            {
                ok = false;
                break;
            }
            ql |= l << (8 * c);
            qh |= h << (8 * c);
        }
        if (ok)
        {
            ebyte = (uint32_t)(k + 127);
            qlo = ql;
            qhi = qh;
            return true;
        }
    }
    return false;
}
}  // namespace

float qnode_span(const std::vector<QNode>& q)
{
    double g = 0.0;
    for (const QNode& n : q)
        for (int a = 0; a < 3 && (n.exps >> 24) != 0; a++)  // nodes with children
        {
            const double sc = std::ldexp(1.0, (int)((n.exps >> (8 * a)) & 0xffu) - 127);
            const double v = std::fabs((double)n.origin[a]) + 255.0 * sc;
            if (!(v <= (double)std::numeric_limits<float>::max())) return std::numeric_limits<float>::infinity();
            g = std::max(g, v);
        }
    // round up to a float
    float f = (float)g;
    if ((double)f < g) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return f;
}

bool device_records(const WideBVH& wb, std::vector<QNode>& nodes_out)
{
    const uint64_t nq = wb.qnodes.size();
    if (nq + wb.refs.size() > (uint64_t)MCRT_LEAF_REF_MASK + 1u) return false;
    nodes_out = wb.qnodes;
    for (QNode& q : nodes_out)
        for (int c = 0; c < 4; c++)
            if (q.child[c] != MCRT_CHILD_EMPTY && (q.child[c] & MCRT_CHILD_LEAF))
                q.child[c] += (uint32_t)nq;  // stays below bit 29: every reference index + nq fits
    return true;
}

void quantize_nodes(WideBVH& out)
{
    out.qnodes.resize(out.nodes.size());
    for (size_t i = 0; i < out.nodes.size(); i++)
    {
        const WNode& w = out.nodes[i];
        QNode& q = out.qnodes[i];
        std::memset(&q, 0, sizeof(q));
        bool valid[4];
        for (int c = 0; c < 4; c++)
        {
            valid[c] = w.child[c] != MCRT_CHILD_EMPTY;
            q.child[c] = w.child[c];
        }
        uint32_t e[3];
        const bool ok = quantize_axis(w.lo_x, w.hi_x, valid, q.origin[0], e[0], q.q_lo_x, q.q_hi_x) &&
                        quantize_axis(w.lo_y, w.hi_y, valid, q.origin[1], e[1], q.q_lo_y, q.q_hi_y) &&
                        quantize_axis(w.lo_z, w.hi_z, valid, q.origin[2], e[2], q.q_lo_z, q.q_hi_z);
        if (!ok)
        {
            // only non-finite boxes lack a conservative grid (any finite box
            // set fits at scale 2^127): planes [-FLT_MAX, +inf] on every axis
            for (int a = 0; a < 3; a++) q.origin[a] = -std::numeric_limits<float>::max();
            e[0] = e[1] = e[2] = 254;
            q.q_lo_x = q.q_lo_y = q.q_lo_z = 0;
            q.q_hi_x = q.q_hi_y = q.q_hi_z = 0xffffffffu;
        }
        q.exps = e[0] | (e[1] << 8) | (e[2] << 16) | (w.nchild << 24);
        q.rank_lo = order_to_rank(w.order_lo);
        q.rank_hi = order_to_rank(w.order_hi);
    }
}
}  // namespace mcrt
