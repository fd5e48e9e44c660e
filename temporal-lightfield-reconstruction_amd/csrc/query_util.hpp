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
    const int px = (int)(pix % (unsigned)Q.W), py = (int)(pix / (unsigned)Q.W) + Q.y0;
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
    // all four tests, not short-circuited: a node's 64 bytes then load as one
    // request, where the first plane alone and the rest after it passes cost
    // walks two dependent loads per node they enter
    return (o.x >= mnx) & (o.y >= mny) & (o.x <= mxx) & (o.y <= mxy);
}

// Key of leaf id at output time t (ReconstructionOur.cpp:384-385), from the
// first 16 bytes of its record.
__device__ __forceinline__ float leaf_key(const LeafRec* leaves, int id, float t) {
    const float4 r = *reinterpret_cast<const float4*>(leaves + id);
    return r.x + (t - r.y) * r.z;
}

// Early rejection for the per-sample test: the same reprojection with the
// hardware reciprocal (v_rcp_f32, <= 1 ulp) instead of the two correctly
// rounded divisions.  True only when the sample lies farther from the output
// sample than R plus a bound on the difference to the exact computation
// (1e-5 relative to the magnitudes involved, about 80 ulp), so the exact test
// of ReconstructionOur.cpp:438-455 would reject it as well, and so would the
// GPU variant's (ReconstructionCudaKernels.cu:419-437, one division by w
// instead of two; both are within a few ulp of the real value).  NaN results
// are never rejected here and take the exact path.
//
// It runs for every sample a lane meets, so it is written for few
// instructions: P and w as in the exact computation; one formula for static
// samples too (with zero motion P / w is x within a few ulp, well inside the
// bound); the later steps fused, whose rounding the bound also covers; the
// bound's terms that depend only on the output sample are loop-invariant.
__device__ __forceinline__ bool surely_outside(const float4 a, const float4 b, const OSample& o, float cc0, float cc1,
                                               float R) {
    const float sx = a.x, sy = a.y, sw = a.z, st = a.w;
    const float dt = o.t - st;
    const float Px = sx * sw + dt * b.x;
    const float Py = sy * sw + dt * b.y;
    const float Pw = 1.0f * sw + dt * b.z;
    const float r = __builtin_amdgcn_rcpf(Pw);
    const float px = Px * r, py = Py * r;
    const float c0r = cc0 * r;
    const float coc = c0r + cc1;
    const float dx = __builtin_fmaf(coc, o.u, px) - o.x;
    const float dy = __builtin_fmaf(coc, o.v, py) - o.y;
    // R + 1e-5 (|px| + |py| + |c0r| + |cc1| + |o.x| + |o.y|) + 1e-6
    const float slack = __builtin_fmaf(1e-5f, fabsf(cc1) + fabsf(o.x) + fabsf(o.y), R + 1e-6f);
    const float lim = __builtin_fmaf(1e-5f, fabsf(px) + fabsf(py) + fabsf(c0r), slack);
    return __builtin_fmaf(dx, dx, dy * dy) > lim * lim;
}

// Squared screen distance of a tree sample from the output sample at the
// output's (u,v,t), with the hardware reciprocal: an estimate for features,
// not for decisions.
__device__ __forceinline__ float approx_d2(const float4 a, const float4 b, const OSample& o, float cc0, float cc1) {
    const float dt = o.t - a.w;
    const float Px = a.x * a.z + dt * b.x;
    const float Py = a.y * a.z + dt * b.y;
    const float Pw = a.z + dt * b.z;
    const float r = __builtin_amdgcn_rcpf(Pw);
    const float coc = cc0 * r + cc1;
    const float dx = __builtin_fmaf(coc, o.u, Px * r) - o.x;
    const float dy = __builtin_fmaf(coc, o.v, Py * r) - o.y;
    return __builtin_fmaf(dx, dx, dy * dy);
}

}  // namespace tlr
