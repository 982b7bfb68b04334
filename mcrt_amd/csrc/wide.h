// Collapse of the binary SBVH into 4-wide nodes (WNode/WRef, layout.h).
#pragma once
#include <vector>

#include "layout.h"
#include "sbvh.h"

namespace mcrt
{
struct WideBVH
{
    std::vector<WNode> nodes;  // nodes[0] = root (absent if no triangles)
    std::vector<QNode> qnodes; // nodes quantised for the kernels (same indices)
    std::vector<WRef> refs;
    int32_t stack_bound = 0;   // stack entries a traversal can ever hold
    int32_t depth = 0;         // wide levels
};

void collapse_wide(const BTree& bt, const std::vector<Tri>& tris, WideBVH& out);

// Fills out.qnodes from out.nodes (collapse_wide calls it).
void quantize_nodes(WideBVH& out);

// An upper bound, over every node and axis, of |origin| + 255 * scale (so
// of the magnitude of every decoded plane and of 255 quantisation steps):
// the scale of the rounding-error margin of the kernels' node tests
// (trace_core.h: node_margin).  +inf if a node has non-finite planes.
float qnode_span(const std::vector<QNode>& q);

// The device's record array, nodes then references (layout.h): the QNodes
// with every leaf link rebased by the node count.  Writes the nodes into
// `nodes_out` (the references follow them unchanged).  Returns false when
// the records do not fit the 29-bit link index.
bool device_records(const WideBVH& wb, std::vector<QNode>& nodes_out);
}  // namespace mcrt
