// Device-resident scene layout shared by the host flattener and the HIP
// kernels.  Everything is 16-byte aligned SoA-within-node so a lane fetches
// a node with 128-bit loads:
//
//   WNode (128 B): four child boxes as x/y/z lo/hi float4 rows, four child
//                  links, and the octant visit-order table that reproduces
//                  the binary tree's near-child-first order.
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
