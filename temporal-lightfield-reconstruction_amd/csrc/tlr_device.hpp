// Device-side data layout of one reconstruction context (all in HBM).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tlr_math.hpp"

namespace tlr {

// Input sample buffer, SoA, indexed (y*W+x)*spp+i: the fields of
// UVTSampleBuffer the plain path reads (SampleBuffer.hpp:116-129,177-180).
struct InSoA {
    const float *x, *y, *w, *u, *v, *t, *r, *g, *b, *mvx, *mvy, *mvw;
};

// Scalars produced on the device during a build.
struct GridInfo {
    int bbmin_key[2];        // order-preserving int encodings of the bbox
    int bbmax_key[2];
    int max_hits;
    unsigned num_kept;
    unsigned num_bad;        // reprojected to a non-finite position
    float bbmin[2], bbmax[2];
    float scale;
    int bbmin_i[2];
    int gw, gh;
    unsigned num_bucketed;
    int error;               // tlr_status of the device-side checks
    int pad[3];
};

// Tree sample layout (Node::s0..s1 ranges index it): the reprojected samples
// in leaf order.  What the per-sample test reads is one 32-byte record, so a
// lane walking a leaf reads consecutive bytes (four samples per 128-byte
// line); the colour, read only for samples within R, is separate.
//   pab[2k]   = (x, y, w, t)          position at the (u,v,t) centre
//   pab[2k+1] = (mv_x, mv_y, mv_w, 0) homogeneous motion
//   pc[k]     = (r, g, b, 1)          colour
// pab holds kPointPad extra zeroed records past the last sample so that
// group loads of up to kPointPad+1 samples stay inside the allocation.
constexpr int kPointPad = 4;
struct Points {
    float4 *pab, *pc;
};
__device__ __forceinline__ float4 pt_a(const Points& P, int k) { return P.pab[2 * (size_t)k]; }
__device__ __forceinline__ float4 pt_b(const Points& P, int k) { return P.pab[2 * (size_t)k + 1]; }

// Leaf record, indexed by leaf id (ids follow the sample layout).  The first
// 16 bytes hold what ordering a leaf for a query needs (w, t, mv.w of the
// leaf's first sample: the key of ReconstructionOur.cpp:384-385) and s0.
struct __align__(16) LeafRec {
    float w0, t0, mvw0;
    int s0;
    int s1;
    int valid;
    float pad[2];
};

// Traversal node: 12 plane coefficients + children / leaf id + escape link.
// Internal nodes [0, n-1), leaf slots [n-1, 2n-1).
struct __align__(16) TravNode {
    float p[12];
    int child;               // >= 0: first child node index; < 0: -(leaf id + 1)
    int escape;              // next node when this subtree is done; -1 = end
    int child1;              // second child of an internal node
    int pad;
};

// Both children of internal node g (g < n-1), for walks that test a node's
// children from the node's own record and load a child's record only when
// they enter it: the children's 12 plane coefficients, references (>= 0:
// internal node index; < 0: -(leaf id + 1)) and, for a leaf child, the three
// terms of its key (w0, t0, mv.w0 of LeafRec; zero for an internal child).
// 128 bytes, one L2 line.
struct __align__(128) TravPair {
    float a[12];
    float b[12];
    int ra, rb;
    float ka[3], kb[3];
};

TLR_HD int order_key(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}
TLR_HD float order_val(int k) {
    int i = k >= 0 ? k : (k ^ 0x7FFFFFFF);
    return __int_as_float(i);
}

}  // namespace tlr
