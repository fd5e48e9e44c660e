// Device helpers shared by the per-output-sample kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "query.hpp"

namespace tlr {

__device__ __forceinline__ unsigned lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

struct OSample {
    float x, y, u, v, t;
};

// Output sample q = (pixel, i): pattern template hashBits(x, y) % 64 offset by
// the pixel (Reconstruction.cpp:665-690).
__device__ __forceinline__ OSample make_osample(const QueryParams& Q, unsigned q) {
    const unsigned pix = q / (unsigned)Q.M;
    const int i = (int)(q - pix * (unsigned)Q.M);
    const int px = (int)(pix % (unsigned)Q.W), py = (int)(pix / (unsigned)Q.W);
    const int pat = (int)(hash_bits((uint32_t)px, (uint32_t)py) % 64u);
    const float* tp = Q.pattern + 5 * ((size_t)pat * Q.M + i);
    OSample o;
    o.x = tp[0] + (float)px;
    o.y = tp[1] + (float)py;
    o.u = tp[2];
    o.v = tp[3];
    o.t = tp[4];
    return o;
}

// Node::intersect: Reconstruction.hpp:204-216
__device__ __forceinline__ bool node_hit(const float* p, const OSample& o, float R) {
    const float mnx = plane_eval(p + 0, o.t, o.u) - R;
    const float mny = plane_eval(p + 6, o.t, o.v) - R;
    const float mxx = plane_eval(p + 3, o.t, o.u) + R;
    const float mxy = plane_eval(p + 9, o.t, o.v) + R;
    return (o.x >= mnx && o.y >= mny && o.x <= mxx && o.y <= mxy);
}

// Key of leaf id at output time t (ReconstructionOur.cpp:384-385), from the
// first 16 bytes of its record.
__device__ __forceinline__ float leaf_key(const LeafRec* leaves, int id, float t) {
    const float4 r = *reinterpret_cast<const float4*>(leaves + id);
    return r.x + (t - r.y) * r.z;
}

}  // namespace tlr
