// Host SBVH builder: the reference SplitBVH (splitbvh.cpp:178-1625)
// decision for decision, built with parallel binning and parallel
// subtrees.  Output is a binary tree in depth-first pre-order.
#pragma once
#include <cstdint>
#include <vector>

#include "geom.h"

namespace mcrt
{
struct SRef
{
    Box box;  // clipped bound (BoundedTriangle::m_clipped_bound)
    int32_t tri;
};

struct BNode
{
    Box box;
    int32_t leaf;     // 1 = leaf
    int32_t axis;     // split axis of a branch
    int32_t left;     // branch: child node indices
    int32_t right;
    int32_t ref_begin;  // leaf: refs [ref_begin, ref_begin + ref_count)
    int32_t ref_count;
};

struct BTree
{
    std::vector<BNode> nodes;  // nodes[0] = root (empty if no triangles)
    std::vector<SRef> refs;
    int32_t max_depth = 0;
};

// Builds the tree of SplitBVH(mesh) for triangles v9 (n x 9 floats).
// threads <= 0 uses every core.
void build_sbvh(const std::vector<Tri>& tris, int threads, BTree& out);

// FNV-1a digest of the canonical pre-order stream (kind, box, axis |
// count, [prim, clipped box]...), -0.0 folded to +0.0.
uint64_t tree_digest(const BTree& t);
}  // namespace mcrt
