// Device-resident scene layout shared by the host flattener and the HIP
// kernels.  Everything is 16-byte aligned SoA-within-node so a lane fetches
// a node with 128-bit loads:
//
//   WNode (128 B): four child boxes as x/y/z lo/hi float4 rows, four child
//                  links, and the octant visit-order table that reproduces
//                  the binary tree's near-child-first order (host side).
//   QNode ( 64 B): the WNode with outward-quantised child boxes; what the
//                  kernels traverse.
//   WRef  ( 64 B): one leaf reference = clipped box + the triangle as
//                  v0, e1 = v1 - v0, e2 = v2 - v0 (the reference computes
//                  the same two differences in MT, so storing them is
//                  bit-identical) + primitive id.
#pragma once
#include <stdint.h>

namespace mcrt
{
// Child link: EMPTY, an interior wide node (index), or a leaf: bit 31 set,
// bits 29-30 = count - 1 (1..4 references), bits 0-28 = first WRef.
// On the device the QNodes and WRefs are one array of 64 B records, the
// nodes first (device_records, wide.h), and a leaf link's bits 0-28 index
// that array: the node count plus the first WRef's index.  A traversal step
// then finds its record at one base whichever kind it visits.
#define MCRT_CHILD_EMPTY 0xffffffffu
#define MCRT_CHILD_LEAF 0x80000000u
#define MCRT_LEAF_REF_MASK 0x1fffffffu

struct alignas(128) WNode
{
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    uint32_t child[4];
    // order[o] for octant o = (dx>0) | (dy>0)<<1 | (dz>0)<<2: byte o holds
    // the slots in visit order, 2 bits each, first visit in the low bits.
    uint32_t order_lo;  // octants 0..3
    uint32_t order_hi;  // octants 4..7
    uint32_t nchild;    // number of non-empty children
    uint32_t pad;
};
static_assert(sizeof(WNode) == 128, "WNode must be 128 bytes");

// The node the kernels read: a WNode with its child boxes quantised to 8
// bits per plane, rounded outwards, in 64 B (four 16 B rows):
//   row 0: origin x, y, z (float) | scale exponents ex | ey << 8 | ez << 16
//          | child count << 24
//   row 1: child[4] (links as in WNode)
//   row 2: lo_x, hi_x, lo_y, hi_y   (byte c of each word = child c)
//   row 3: lo_z, hi_z, rank_lo, rank_hi
// A plane decodes as origin + (float)q * 2^(e - 127) (q * 2^k is exact, one
// rounding in the addition); the host chooses q so that every decoded box
// contains the exact child box.  A contained box's slab interval is
// contained after rounding, so a quantised test passes whenever the exact
// one does; the extra subtrees it admits hold only references whose exact
// clipped boxes fail, so results are those of the exact boxes.  The visit
// order is stored inverted: byte o of rank_lo/rank_hi holds, 2 bits per
// slot, each slot's position in octant o's visit order (the inverse of
// WNode::order), so a kernel places the passing children on its stack
// with fixed-slot arithmetic instead of indexing by slot.
struct alignas(64) QNode
{
    float origin[3];
    uint32_t exps;
    uint32_t child[4];
    uint32_t q_lo_x, q_hi_x, q_lo_y, q_hi_y;
    uint32_t q_lo_z, q_hi_z;
    uint32_t rank_lo, rank_hi;
};

// The rank table of a WNode order table (either 32-bit half).
inline uint32_t order_to_rank(uint32_t order)
{
    uint32_t rank = 0;
    for (int o = 0; o < 4; o++)
        for (int k = 0; k < 4; k++)
        {
            const uint32_t s = (order >> (8 * o + 2 * k)) & 3u;
            rank |= (uint32_t)k << (8 * o + 2 * s);
        }
    return rank;
}
static_assert(sizeof(QNode) == 64, "QNode must be 64 bytes");

// Decoded plane of a QNode: the host checks with it that every decoded box
// contains the exact child box.  The kernels test the decoded box
// conservatively (trace_core.h: node_margin), passing wherever the
// reference's slab arithmetic on this plane passes.
inline float qnode_plane(float origin, uint32_t q, float scale) { return origin + (float)q * scale; }

struct alignas(64) WRef
{
    float lo[3];
    float hi[3];
    float v0[3];
    float e1[3];
    float e2[3];
    int32_t prim;
};
static_assert(sizeof(WRef) == 64, "WRef must be 64 bytes");

// Per-triangle shading data for the integrator.
struct alignas(16) TriShade
{
    float na[3];
    int32_t mat;
    float nb[3];
    int32_t light;  // light index or -1
    float nc[3];
    int32_t hot;    // gradient slot accumulated per workgroup in LDS, or -1
};

// Triangles whose gradients the adjoint accumulates in LDS first (lights,
// then the largest triangles: the ones most paths touch).
#define MCRT_HOT_SLOTS 64
static_assert(sizeof(TriShade) == 48, "TriShade must be 48 bytes");
}  // namespace mcrt
