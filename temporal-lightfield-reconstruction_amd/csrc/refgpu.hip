// The reference's GPU-variant method: recKernel, failKernel and nearestKernel
// (ReconstructionCudaKernels.cu:242-868), driven as in
// CudaReconstruction::reconstructGPU (ReconstructionCuda.cpp:327-465).
//
// Stages:
//   k_ref_first     one lane per output sample, 1R, leaf ids in LDS
//                   (kRefFirstCap per lane): the first surface's tent sum
//                   when its points surround the output sample, else the
//                   sample goes to the 2R or the nearest list;
//   k_ref_fail_wave one wave per output sample of the 2R list: frontier
//                   walk, rank sort, surfaces front to back with the
//                   sameSurface checks and sample tests across the lanes,
//                   lane-parallel triangle test;
//   k_ref_nearest   one lane per output sample: the nearest reprojected
//                   sample within 1, 2, 4, 8, 16 x the dispersion, radius
//                   shrinking during the walk;
//   k_ref_stage     one lane per output sample with per-lane lists in
//                   global memory, sized for the samples the first two
//                   stages could not hold (STAGE 0: 1R, STAGE 1: 2R).
// A sample whose leaves or points exceed a launch's lists goes to an
// overflow list and is run again with lists sized for it.  The reference
// instead caps lists at MAX_LEAVES_R1/R2 and MAX_POINTS (.cu:39-43) and
// reconstructs the rest on the CPU (ReconstructionCuda.cpp:412-434).
//
// Results are per output sample (qcol) and combined per pixel in sample
// order; the reference adds them atomically (.cu:200-208).  Arithmetic is
// IEEE FP32 without contraction: correctly rounded division and square root
// where the reference builds with -use_fast_math (.cu:35).
#include "query.hpp"
#include "query_util.hpp"

#include <algorithm>

namespace tlr {
namespace {

constexpr int kRefFirstBlock = 128;
constexpr int kRefStageBlock = 256;
constexpr int kEmpty = (int)0x80000000;  // leaf-list flag: no point within 1R

// Per-lane list of leaf ids (ref_collect keeps (id, key) pairs while they
// fit and compacts them to ids; keys are otherwise recomputed from the leaf
// records): element k at p[k * stride].
struct LeafList {
    int* p;
    int stride;
    __device__ __forceinline__ int get(int k) const { return p[(size_t)k * stride]; }
    __device__ __forceinline__ void put(int k, int v) const { p[(size_t)k * stride] = v; }
};
struct PointList {
    float2* p;
    int stride;
    __device__ __forceinline__ float2 get(int k) const { return p[(size_t)k * stride]; }
    __device__ __forceinline__ void put(int k, float2 v) const { p[(size_t)k * stride] = v; }
};

// Reprojection of a tree sample (records a, b) to the output sample
// (.cu:419-437).
__device__ __forceinline__ void ref_point_ab(const float4 a, const float4 b, const OSample& o, float cc0, float cc1,
                                             float& dx, float& dy, float& d2) {
    const float p0x = a.x * a.z, p0y = a.y * a.z, p0w = a.z;
    const float px = p0x + (o.t - a.w) * b.x;
    const float py = p0y + (o.t - a.w) * b.y;
    const float pw = p0w + (o.t - a.w) * b.z;
    const float oow = 1.f / pw;
    const float cocr = cc0 * oow + cc1;
    const float x = px * oow + cocr * o.u;
    const float y = py * oow + cocr * o.v;
    dx = x - o.x;
    dy = y - o.y;
    d2 = dx * dx + dy * dy;
}

// Angular extent of points around the output sample (.cu:443-445).
__device__ __forceinline__ void wedge(float* box, float dx, float dy) {
    const float k = dx / dy;
    if (dy >= 0.f) {
        box[0] = fminf(box[0], k);
        box[1] = fmaxf(box[1], k);
    } else {
        box[2] = fminf(box[2], k);
        box[3] = fmaxf(box[3], k);
    }
}

// Leaves within R, sorted by (key, leaf id) (.cu:262-330).  Returns the
// number of leaves found; only the first `cap` are stored.  SIZES: npts =
// the samples those leaves hold (a bound on the points of any surface).
template <bool SIZES>
__device__ __forceinline__ int ref_collect(const QueryParams& Q, const OSample& o, float R, const LeafList& L,
                                           int cap, int& npts) {
    int n = 0;
    npts = 0;
    // the list holds (id, key) pairs while they fit (cap / 2 leaves), ids
    // only after that, keys then recomputed from the leaf records on compare
    bool pairs = true;
    int node = Q.n_nodes > 0 ? Q.root : -1;
    while (node >= 0) {
        const float4* tn4 = reinterpret_cast<const float4*>(Q.trav + node);
        const float4 a = tn4[0], b = tn4[1], c = tn4[2], d = tn4[3];
        const float p[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
        const int child = __float_as_int(d.x), escape = __float_as_int(d.y);
        if (node_hit(p, o, R)) {
            if (child >= 0) {
                node = child;
                continue;
            }
            if (n < cap) {
                const int id = -child - 1;
                const float key = leaf_key(Q.leaves, id, o.t);
                if (pairs && n == cap / 2) {
                    for (int j = 1; j < n; j++) L.put(j, L.get(2 * j));
                    pairs = false;
                }
                int j = n;
                if (pairs) {
                    while (j > 0) {
                        const int e = L.get(2 * j - 2);
                        const float kk = __int_as_float(L.get(2 * j - 1));
                        if (!(kk > key || (!(key > kk) && e > id))) break;
                        L.put(2 * j, e);
                        L.put(2 * j + 1, __float_as_int(kk));
                        j--;
                    }
                    L.put(2 * j, id);
                    L.put(2 * j + 1, __float_as_int(key));
                } else {
                    while (j > 0) {
                        const int e = L.get(j - 1);
                        const float kk = leaf_key(Q.leaves, e, o.t);
                        if (!(kk > key || (!(key > kk) && e > id))) break;
                        L.put(j, e);
                        j--;
                    }
                    L.put(j, id);
                }
            }
            if (SIZES) {
                const int id = -child - 1;
                npts += Q.leaves[id].s1 - Q.leaves[id].s0;
            }
            n++;
        }
        node = escape;
    }
    if (pairs)
        for (int j = 1; j < min(n, cap); j++) L.put(j, L.get(2 * j));
    return n;
}

__device__ __forceinline__ void sig_of(const QueryParams& Q, int id, const OSample& o, SurfSig& s) {
    const float RANGE = 1.0f / sqrtf(128.0f);
    surf_sig(Q.leaf_b[id].p, o.u - RANGE, o.v - RANGE, o.t - 0.5f * RANGE, o.u + RANGE, o.v + RANGE,
             o.t + 0.5f * RANGE, s);
}

enum { R_COLOUR = 0, R_FAIL = 1, R_NEAREST = 2, R_OVERFLOW = 3 };

// recKernel after the gather (.cu:332-473) over the n sorted 1R leaves.
// tested / fetched: samples reprojected, colours read.
__device__ __forceinline__ int ref_first_surface(const QueryParams& Q, const OSample& o, const LeafList& L, int n,
                                                 float4& out, unsigned& tested, unsigned& fetched) {
    const float R = dispersion_for_spp(Q.spp_pow2);
    const float R2 = R * R;
    const float4* __restrict__ pab = Q.pts.pab;
    float wb[4] = {kFltMax, -kFltMax, kFltMax, -kFltMax};
    float cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
    bool multi = false;
    for (int li = 0; li < n; li++) {
        const int id = L.get(li);
        const int k0 = Q.leaves[id].s0, k1 = Q.leaves[id].s1;
        // first sample strictly within R; four records per load, samples
        // that surely lie outside R skip the exact test (pab is padded past
        // its last record for the group loads)
        int i0 = k1;
        for (int kg = k0; kg < k1 && i0 == k1; kg += 4) {
            const float4* g = pab + 2 * (size_t)kg;
            float4 ga[4], gb[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                ga[e] = g[2 * e];
                gb[e] = g[2 * e + 1];
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if (kg + e >= k1) break;
                tested++;
                if (surely_outside(ga[e], gb[e], o, Q.cc0, Q.cc1, R)) continue;
                float dx, dy, d2;
                ref_point_ab(ga[e], gb[e], o, Q.cc0, Q.cc1, dx, dy, d2);
                if (d2 < R2) {
                    i0 = kg + e;
                    break;
                }
            }
        }
        const bool found = i0 < k1;
        if (!found) {
            L.put(li, id | kEmpty);
        } else {
            SurfSig cur, prev;
            sig_of(Q, id, o, cur);
            for (int j = 0; j < li; j++) {
                const int pj = L.get(j);
                if (pj & kEmpty) continue;
                sig_of(Q, pj, o, prev);
                if (!same_surface_sig(cur, prev)) {
                    multi = true;
                    break;
                }
            }
        }
        if (multi) break;
        // the rest of the leaf from that sample on, in order
        for (int kg = i0; kg < k1; kg += 4) {
            const float4* g = pab + 2 * (size_t)kg;
            float4 ga[4], gb[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                ga[e] = g[2 * e];
                gb[e] = g[2 * e + 1];
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int k = kg + e;
                if (k >= k1) break;
                tested++;
                if (surely_outside(ga[e], gb[e], o, Q.cc0, Q.cc1, R)) continue;
                float dx, dy, d2;
                ref_point_ab(ga[e], gb[e], o, Q.cc0, Q.cc1, dx, dy, d2);
                if (d2 <= R2) {
                    wedge(wb, dx, dy);
                    const float w = fmaxf(0.f, 1.f - sqrtf(d2) / R);
                    const float4 c = Q.pts.pc[k];
                    fetched++;
                    cr += w * c.x;
                    cg += w * c.y;
                    cb += w * c.z;
                    ca += w * 1.0f;
                }
            }
        }
    }
    if (ca > 0.f && wb[1] > wb[2] && wb[0] < wb[3]) {
        const float s = 1.f / ca;
        out = make_float4(cr * s, cg * s, cb * s, ca * s);
        return R_COLOUR;
    }
    return multi ? R_FAIL : R_NEAREST;
}

// checkTriangles (.cu:586-632): is there a triangle of the surface's points,
// all edges shorter than 2 x dispersion, around the output sample?  For each
// point p0, a later point p1 on the non-negative side of p0 and a later point
// p2 on the negative side; the reference sorts the later points into the two
// sides in a buffer first, which changes nothing in the answer.
__device__ __forceinline__ bool ref_triangles(const PointList& xy, int n, float disp) {
    if (n < 3) return false;
    const float dcomp = 4.f * disp * disp;
    for (int i = 0; i < n - 2; i++) {
        const float2 p0 = xy.get(i);
        for (int j = i + 1; j < n; j++) {
            const float2 p1 = xy.get(j);
            const float ex = p0.x - p1.x, ey = p0.y - p1.y;
            if (!(ex * ex + ey * ey < dcomp)) continue;
            if (!(p0.x * p1.y - p1.x * p0.y >= 0.f)) continue;
            for (int k = i + 1; k < n; k++) {
                const float2 p2 = xy.get(k);
                const float fx = p0.x - p2.x, fy = p0.y - p2.y;
                if (!(fx * fx + fy * fy < dcomp)) continue;
                if (p0.x * p2.y - p2.x * p0.y >= 0.f) continue;
                const float gx = p1.x - p2.x, gy = p1.y - p2.y;
                if (p1.x * p2.y >= p2.x * p1.y && gx * gx + gy * gy < dcomp) return true;
            }
        }
    }
    return false;
}

// failKernel after the gather (.cu:706-867) over the n sorted 2R leaves.
// R_OVERFLOW: a surface held more points than xy_cap and needed the
// triangle test.
__device__ __forceinline__ int ref_fail(const QueryParams& Q, const OSample& o, const LeafList& L, int n,
                                        const PointList& xy, int xy_cap, float4& out) {
    const float disp = dispersion_for_spp(Q.spp_pow2);
    const float R = disp * 2.f;
    const float R2 = R * R;
    const float4* __restrict__ pab = Q.pts.pab;
    float cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
    int lidx = 0;
    do {
        const int sfirst = lidx;
        int np = 0;
        cr = cg = cb = ca = 0.f;
        float ab[4] = {kFltMax, -kFltMax, kFltMax, -kFltMax};
        float wb[4] = {kFltMax, -kFltMax, kFltMax, -kFltMax};
        do {
            const int id = L.get(lidx);
            if (np >= 4) {
                SurfSig cur, prev;
                sig_of(Q, id, o, cur);
                bool end = false;
                for (int j = sfirst; j < lidx; j++) {
                    sig_of(Q, L.get(j), o, prev);
                    if (!same_surface_sig(cur, prev)) {
                        end = true;
                        break;
                    }
                }
                if (end) break;
            }
            const int k0 = Q.leaves[id].s0, k1 = Q.leaves[id].s1;
            for (int kg = k0; kg < k1; kg += 4) {
                const float4* g = pab + 2 * (size_t)kg;
                float4 ga[4], gb[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    ga[e] = g[2 * e];
                    gb[e] = g[2 * e + 1];
                }
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int i = kg + e;
                    if (i >= k1) break;
                    if (surely_outside(ga[e], gb[e], o, Q.cc0, Q.cc1, R)) continue;
                    float dx, dy, d2;
                    ref_point_ab(ga[e], gb[e], o, Q.cc0, Q.cc1, dx, dy, d2);
                    if (d2 <= R2) {
                        if (np < xy_cap) xy.put(np, make_float2(dx, dy));
                        np++;
                        const float w = 1.f - sqrtf(d2) / disp;
                        if (w <= 0.f) {
                            wedge(ab, dx, dy);
                        } else {
                            wedge(wb, dx, dy);
                            const float4 c = Q.pts.pc[i];
                            cr += w * c.x;
                            cg += w * c.y;
                            cb += w * c.z;
                            ca += w * 1.0f;
                        }
                    }
                }
            }
            lidx++;
        } while (lidx < n);
        if (lidx < n) {
            ab[0] = fminf(ab[0], wb[0]);
            ab[1] = fmaxf(ab[1], wb[1]);
            ab[2] = fminf(ab[2], wb[2]);
            ab[3] = fmaxf(ab[3], wb[3]);
            if (ab[1] > ab[2] && ab[0] < ab[3]) {
                if (wb[1] > wb[2] && wb[0] < wb[3]) break;
                if (np > xy_cap) return R_OVERFLOW;
                if (ref_triangles(xy, np, disp)) break;
            }
        }
    } while (lidx < n);
    if (ca > 0.f) {
        const float s = 1.f / ca;
        out = make_float4(cr * s, cg * s, cb * s, ca * s);
        return R_COLOUR;
    }
    return R_NEAREST;
}

// nearestKernel (.cu:476-579).  Zero colour when no sample lies within
// 16 x dispersion.
__device__ __forceinline__ float4 ref_nearest(const QueryParams& Q, const OSample& o) {
    const float disp = dispersion_for_spp(Q.spp_pow2);
    const float4* __restrict__ pab = Q.pts.pab;
    for (float rmul = 1.f; rmul <= 16.f; rmul *= 2.f) {
        float radius = rmul * disp;
        int best = -1;
        int node = Q.n_nodes > 0 ? Q.root : -1;
        while (node >= 0) {
            const float4* tn4 = reinterpret_cast<const float4*>(Q.trav + node);
            const float4 a = tn4[0], b = tn4[1], c = tn4[2], d = tn4[3];
            const float p[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
            const int child = __float_as_int(d.x), escape = __float_as_int(d.y);
            if (node_hit(p, o, radius)) {
                if (child >= 0) {
                    node = child;
                    continue;
                }
                const int id = -child - 1;
                const int k0 = Q.leaves[id].s0, k1 = Q.leaves[id].s1;
                for (int kg = k0; kg < k1; kg += 4) {
                    const float4* g = pab + 2 * (size_t)kg;
                    float4 ga[4], gb[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        ga[e] = g[2 * e];
                        gb[e] = g[2 * e + 1];
                    }
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int i = kg + e;
                        if (i >= k1) break;
                        if (surely_outside(ga[e], gb[e], o, Q.cc0, Q.cc1, radius)) continue;
                        float dx, dy, d2;
                        ref_point_ab(ga[e], gb[e], o, Q.cc0, Q.cc1, dx, dy, d2);
                        if (d2 <= radius * radius) {
                            radius = sqrtf(d2);
                            best = i;
                        }
                    }
                }
            }
            node = escape;
        }
        if (best >= 0) {
            const float4 c = Q.pts.pc[best];
            return make_float4(c.x, c.y, c.z, 1.0f);
        }
    }
    return make_float4(0.f, 0.f, 0.f, 0.f);
}

// Wave-aggregated append of v to list where pred holds (one atomic per wave).
// Every lane of the wave calls it.
__device__ __forceinline__ void wave_append(bool pred, unsigned v, unsigned* list, unsigned* count) {
    const uint64_t m = __ballot(pred);
    if (!m) return;
    unsigned base = 0;
    if ((threadIdx.x & 63) == (unsigned)(__builtin_ctzll(m))) base = atomicAdd(count, (unsigned)__popcll(m));
    base = (unsigned)__shfl((int)base, __builtin_ctzll(m), 64);
    if (pred) list[base + lanes_below(m)] = v;
}

__device__ __forceinline__ void finish(const RefLaunch& R, bool active, unsigned q, int r, float4 col,
                                       int stage, int need_leaves, int need_points) {
    if (active && r == R_COLOUR) {
        R.q.qcol[q] = col;
        if (R.stage) R.stage[q] = stage;
    }
    wave_append(active && r == R_FAIL, q, R.to_fail, R.n_fail);
    wave_append(active && r == R_NEAREST, q, R.to_near, R.n_near);
    wave_append(active && r == R_OVERFLOW, q, R.to_over, R.n_over);
    if (active && r == R_OVERFLOW) {
        atomicMax(&R.need[0], (unsigned)need_leaves);
        atomicMax(&R.need[1], (unsigned)need_points);
    }
}

// With q.collect_stats: leaves, samples tested and colours read by this
// launch, summed into q.stats (ST_LEAF_NODES, ST_SAMPLES_IN_LEAVES,
// ST_WITHIN_1R).
template <bool STATS>
__global__ __launch_bounds__(kRefFirstBlock) __attribute__((amdgpu_waves_per_eu(7, 7))) void k_ref_first(RefLaunch R) {
    __shared__ int lst[kRefFirstCap][kRefFirstBlock];
    const unsigned q = blockIdx.x * kRefFirstBlock + threadIdx.x;
    const bool active = q < R.q.num_queries;
    int r = R_COLOUR, n = 0;
    unsigned tested = 0, fetched = 0;
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) {
        const OSample o = make_osample(R.q, q);
        const LeafList L{&lst[0][threadIdx.x], kRefFirstBlock};
        int np = 0;
        n = ref_collect<false>(R.q, o, dispersion_for_spp(R.q.spp_pow2), L, R.leaf_cap, np);
        if (n == 0)
            r = R_FAIL;
        else if (n > R.leaf_cap)
            r = R_OVERFLOW;
        else
            r = ref_first_surface(R.q, o, L, n, col, tested, fetched);
    }
    finish(R, active, q, r, col, 0, n, 0);
    if (STATS) {
        unsigned long long v[3] = {r == R_OVERFLOW ? 0ull : (unsigned long long)n, tested, fetched};
#pragma unroll
        for (int k = 0; k < 3; k++) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) v[k] += __shfl_xor(v[k], d, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&R.q.stats[ST_LEAF_NODES], v[0]);
            atomicAdd(&R.q.stats[ST_SAMPLES_IN_LEAVES], v[1]);
            atomicAdd(&R.q.stats[ST_WITHIN_1R], v[2]);
        }
    }
}

// Persistent waves over list[0 .. *count).  STAGE 0: 1R; STAGE 1: 2R.
template <int STAGE>
__global__ __launch_bounds__(kRefStageBlock) void k_ref_stage(RefLaunch R) {
    const unsigned lane = threadIdx.x & 63;
    const size_t wave = ((size_t)blockIdx.x * kRefStageBlock + threadIdx.x) >> 6;
    const LeafList L{R.lists + wave * (size_t)R.leaf_cap * 64 + lane, 64};
    const PointList XY{R.xy + wave * (size_t)R.point_cap * 64 + lane, 64};
    const unsigned total = *R.count;
    for (;;) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(R.next, 64u);
        base = (unsigned)__builtin_amdgcn_readfirstlane((int)base);
        if (base >= total) break;
        const unsigned slot = base + lane;
        const bool active = slot < total;
        const unsigned q = active ? R.list[slot] : 0u;
        int r = R_COLOUR, n = 0, np = 0;
        float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
        if (active) {
            const OSample o = make_osample(R.q, q);
            const float disp = dispersion_for_spp(R.q.spp_pow2);
            n = ref_collect<STAGE == 1>(R.q, o, STAGE == 0 ? disp : disp * 2.f, L, R.leaf_cap, np);
            if (n == 0) {
                r = STAGE == 0 ? R_FAIL : R_NEAREST;
            } else if (n > R.leaf_cap) {
                r = R_OVERFLOW;
            } else if (STAGE == 0) {
                unsigned tested = 0, fetched = 0;
                r = ref_first_surface(R.q, o, L, n, col, tested, fetched);
            } else {
                r = ref_fail(R.q, o, L, n, XY, R.point_cap, col);
            }
        }
        finish(R, active, q, r, col, STAGE, n, np);
    }
}

__global__ __launch_bounds__(kRefStageBlock) void k_ref_nearest(RefLaunch R) {
    const unsigned lane = threadIdx.x & 63;
    const unsigned total = *R.count;
    for (;;) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(R.next, 64u);
        base = (unsigned)__builtin_amdgcn_readfirstlane((int)base);
        if (base >= total) break;
        const unsigned slot = base + lane;
        if (slot < total) {
            const unsigned q = R.list[slot];
            R.q.qcol[q] = ref_nearest(R.q, make_osample(R.q, q));
            if (R.stage) R.stage[q] = 2;
        }
    }
}

// ---------------------------------------------------------------------------
// failKernel with one wave per output sample (the 2R stage): the wave walks
// the hierarchy as a frontier, sorts the leaves by (key, leaf id) by rank,
// then follows the surfaces front to back with one leaf's samples across the
// lanes: the sameSurface check against the surface's earlier leaves and the
// per-sample tests run lane-parallel, the wedge boxes are lane reductions
// (min and max do not depend on order) and the tent sum is accumulated in
// sample order.  Per-wave lists live in global scratch (RefWaveMem).
// ---------------------------------------------------------------------------
struct RefWaveMem {
    int* stack;
    float* lkey;
    int* lid;
    int* sid;
    float* px;
    float* py;
    int* pos;
    int* neg;
};
__device__ __forceinline__ RefWaveMem ref_wave_mem(float* base) {
    RefWaveMem m;
    m.stack = reinterpret_cast<int*>(base);
    m.lkey = base + kRefWaveStack;
    m.lid = reinterpret_cast<int*>(m.lkey + kRefWaveLeaves);
    m.sid = m.lid + kRefWaveLeaves;
    m.px = reinterpret_cast<float*>(m.sid + kRefWaveLeaves);
    m.py = m.px + kRefWavePoints;
    m.pos = reinterpret_cast<int*>(m.py + kRefWavePoints);
    m.neg = m.pos + kRefWavePoints;
    return m;
}

__device__ __forceinline__ float wave_fmin(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ float wave_fmax(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ int wuni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float wunif(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
// Orders the wave's global scratch writes before its later reads.
__device__ __forceinline__ void wave_mem_sync() { __threadfence_block(); }

// checkTriangles (.cu:586-632) over the wave's np points: for each p0, the
// later points near it on its non-negative and on its negative side are
// listed, then every (positive, negative) pair is tested, positives across
// the lanes.
__device__ bool ref_triangles_wave(const RefWaveMem& m, int np, float disp) {
    const int lane = threadIdx.x & 63;
    if (np < 3) return false;
    const float dcomp = 4.f * disp * disp;
    for (int i = 0; i < np - 2; i++) {
        const float p0x = wunif(m.px[i]), p0y = wunif(m.py[i]);
        int npos = 0, nneg = 0;
        for (int j0 = i + 1; j0 < np; j0 += 64) {
            const int j = j0 + lane;
            bool near = false, posv = false;
            if (j < np) {
                const float p1x = m.px[j], p1y = m.py[j];
                const float ex = p0x - p1x, ey = p0y - p1y;
                near = ex * ex + ey * ey < dcomp;
                posv = near && (p0x * p1y - p1x * p0y >= 0.f);
            }
            const bool negv = near && !posv;
            const uint64_t pm = __ballot(posv), nm = __ballot(negv);
            if (posv) m.pos[npos + (int)lanes_below(pm)] = j;
            if (negv) m.neg[nneg + (int)lanes_below(nm)] = j;
            npos += __popcll(pm);
            nneg += __popcll(nm);
        }
        wave_mem_sync();
        if (npos == 0 || nneg == 0) continue;
        for (int a0 = 0; a0 < npos; a0 += 64) {
            const int a = a0 + lane;
            float p1x = 0.f, p1y = 0.f;
            if (a < npos) {
                const int ia = m.pos[a];
                p1x = m.px[ia];
                p1y = m.py[ia];
            }
            bool any = false;
            for (int b = 0; b < nneg; b++) {
                const int kb = wuni(m.neg[b]);
                const float p2x = wunif(m.px[kb]), p2y = wunif(m.py[kb]);
                const float fx = p1x - p2x, fy = p1y - p2y;
                if (a < npos && p1x * p2y >= p2x * p1y && fx * fx + fy * fy < dcomp) any = true;
            }
            if (__ballot(any)) return true;
        }
    }
    return false;
}

// One output sample through failKernel's body; R_OVERFLOW when its leaves or
// (where the triangle test needs them) its points exceed the wave's lists.
__device__ int ref_fail_wave(const QueryParams& Q, const RefWaveMem& m, const OSample& o, int leaf_cap,
                             int point_cap, float4& out) {
    const int lane = threadIdx.x & 63;
    const float disp = dispersion_for_spp(Q.spp_pow2);
    const float R = disp * 2.f;
    const float R2 = R * R;
    // --- leaves within 2R (.cu:654-689): frontier walk, any visiting order ---
    int nl = 0;
    if (Q.n_nodes > 0) {
        int sp = 1;
        if (lane == 0) m.stack[0] = Q.root;
        wave_mem_sync();
        while (sp > 0) {
            const int take = sp < 64 ? sp : 64;
            const int base = sp - take;
            int node = -1;
            if (lane < take) node = m.stack[base + lane];
            bool hit = false, internal = false;
            int c0 = 0, c1 = 0, lidv = 0;
            float key = 0.0f;
            if (node >= 0) {
                const float4* tn4 = reinterpret_cast<const float4*>(Q.trav + node);
                const float4 a = tn4[0], b = tn4[1], c = tn4[2], d = tn4[3];
                const float p[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
                // links taken without a branch: the node loads as one request
                const int child = __float_as_int(d.x);
                hit = node_hit(p, o, R);
                internal = child >= 0;
                c0 = child;
                c1 = __float_as_int(d.z);
                lidv = -child - 1;
                if (hit && !internal) key = leaf_key(Q.leaves, lidv, o.t);
            }
            const uint64_t im = __ballot(hit && internal);
            const uint64_t lm = __ballot(hit && !internal);
            const int ni = __popcll(im), nlh = __popcll(lm);
            if (base + 2 * ni > kRefWaveStack || nl + nlh > leaf_cap) return R_OVERFLOW;
            if (hit && internal) {
                const int r = (int)lanes_below(im);
                m.stack[base + 2 * r] = c0;
                m.stack[base + 2 * r + 1] = c1;
            }
            if (hit && !internal) {
                const int r = nl + (int)lanes_below(lm);
                m.lkey[r] = key;
                m.lid[r] = lidv;
            }
            sp = base + 2 * ni;
            nl += nlh;
            wave_mem_sync();
        }
    }
    if (nl == 0) return R_NEAREST;
    // --- front to back by (key, leaf id): rank of each entry ---
    for (int j = lane; j < nl; j += 64) {
        const float kj = m.lkey[j];
        const int ij = m.lid[j];
        int rank = 0;
        for (int i = 0; i < nl; i++) {
            const float ki = m.lkey[i];
            const int ii = m.lid[i];
            rank += (ki < kj || (!(kj < ki) && ii < ij)) ? 1 : 0;
        }
        m.sid[rank] = ij;
    }
    wave_mem_sync();
    // --- surfaces (.cu:706-867) ---
    const float RANGE = 1.0f / sqrtf(128.0f);
    const float u0 = o.u - RANGE, v0 = o.v - RANGE, t0 = o.t - 0.5f * RANGE;
    const float u1 = o.u + RANGE, v1 = o.v + RANGE, t1 = o.t + 0.5f * RANGE;
    const float inf = __builtin_inff();
    const float4* __restrict__ pab = Q.pts.pab;
    float cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
    int lidx = 0;
    do {
        const int sfirst = lidx;
        int np = 0;
        cr = cg = cb = ca = 0.f;
        float ab0 = kFltMax, ab1 = -kFltMax, ab2 = kFltMax, ab3 = -kFltMax;
        float wb0 = kFltMax, wb1 = -kFltMax, wb2 = kFltMax, wb3 = -kFltMax;
        do {
            const int id = wuni(m.sid[lidx]);
            if (np >= 4) {
                bool end = false;
                const float* nb = Q.leaf_b[id].p;
                for (int j0 = sfirst; j0 < lidx && !end; j0 += 64) {
                    const int j = j0 + lane;
                    bool ok = true;
                    if (j < lidx) ok = same_surface_flat(nb, Q.leaf_b[m.sid[j]].p, u0, v0, t0, u1, v1, t1);
                    end = __ballot(!ok) != 0;
                }
                if (end) break;
            }
            const int s0 = wuni(Q.leaves[id].s0), s1 = wuni(Q.leaves[id].s1);
            for (int c = s0; c < s1; c += 64) {
                const int k = c + lane;
                bool kept = false;
                float dx = 0.f, dy = 0.f, d2 = 0.f;
                if (k < s1) {
                    ref_point_ab(pab[2 * (size_t)k], pab[2 * (size_t)k + 1], o, Q.cc0, Q.cc1, dx, dy, d2);
                    kept = d2 <= R2;
                }
                const uint64_t km = __ballot(kept);
                if (!km) continue;
                const int pos = np + (int)lanes_below(km);
                if (kept && pos < point_cap) {
                    m.px[pos] = dx;
                    m.py[pos] = dy;
                }
                np += __popcll(km);
                const float w = kept ? 1.f - sqrtf(d2) / disp : 0.f;
                const float kk = kept ? dx / dy : 0.f;
                const bool inA = kept && w <= 0.f, inW = kept && !(w <= 0.f);
                const bool up = dy >= 0.f;
                ab0 = fminf(ab0, wave_fmin(inA && up ? kk : inf));
                ab1 = fmaxf(ab1, wave_fmax(inA && up ? kk : -inf));
                ab2 = fminf(ab2, wave_fmin(inA && !up ? kk : inf));
                ab3 = fmaxf(ab3, wave_fmax(inA && !up ? kk : -inf));
                wb0 = fminf(wb0, wave_fmin(inW && up ? kk : inf));
                wb1 = fmaxf(wb1, wave_fmax(inW && up ? kk : -inf));
                wb2 = fminf(wb2, wave_fmin(inW && !up ? kk : inf));
                wb3 = fmaxf(wb3, wave_fmax(inW && !up ? kk : -inf));
                // tent sum in sample order over lane-held weights and colours
                const float4 col = inW ? Q.pts.pc[k] : make_float4(0.f, 0.f, 0.f, 0.f);
                for (uint64_t cm = __ballot(inW); cm; cm &= cm - 1) {
                    const int j = __builtin_ctzll(cm);
                    const float wj = __shfl(w, j, 64);
                    cr += wj * __shfl(col.x, j, 64);
                    cg += wj * __shfl(col.y, j, 64);
                    cb += wj * __shfl(col.z, j, 64);
                    ca += wj * 1.0f;
                }
            }
            lidx++;
        } while (lidx < nl);
        if (lidx < nl) {
            ab0 = fminf(ab0, wb0);
            ab1 = fmaxf(ab1, wb1);
            ab2 = fminf(ab2, wb2);
            ab3 = fmaxf(ab3, wb3);
            if (ab1 > ab2 && ab0 < ab3) {
                if (wb1 > wb2 && wb0 < wb3) break;
                if (np > point_cap) return R_OVERFLOW;
                wave_mem_sync();
                if (ref_triangles_wave(m, np, disp)) break;
            }
        }
    } while (lidx < nl);
    if (ca > 0.f) {
        const float s = 1.f / ca;
        out = make_float4(cr * s, cg * s, cb * s, ca * s);
        return R_COLOUR;
    }
    return R_NEAREST;
}

// Persistent waves over list[0 .. *count): one output sample per wave.
__global__ __launch_bounds__(kRefStageBlock) void k_ref_fail_wave(RefLaunch R) {
    const int lane = threadIdx.x & 63;
    // wave-uniform (so the scratch pointers derived from it live in scalar registers)
    const size_t wave = (size_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * kRefStageBlock + threadIdx.x) >> 6));
    const RefWaveMem m = ref_wave_mem(R.wscratch + wave * kRefWaveScratchWords);
    const unsigned total = *R.count;
    // four consecutive entries per claim (atomics on the one counter serialise)
    constexpr unsigned kChunk = 4;
    unsigned e = 0, e_end = 0;
    for (;;) {
        if (e == e_end) {
            unsigned e0 = 0;
            if (lane == 0) e0 = atomicAdd(R.next, kChunk);
            e = (unsigned)wuni((int)e0);
            if (e >= total) break;
            e_end = total - e < kChunk ? total : e + kChunk;
        }
        const unsigned q = (unsigned)wuni((int)R.list[e]);
        e++;
        const OSample o = make_osample(R.q, q);
        float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
        const int r = ref_fail_wave(R.q, m, o, R.leaf_cap, R.point_cap, col);
        if (lane == 0) {
            if (r == R_COLOUR) {
                R.q.qcol[q] = col;
                if (R.stage) R.stage[q] = 1;
            } else if (r == R_NEAREST) {
                R.to_near[atomicAdd(R.n_near, 1u)] = q;
            } else {
                // lists for a rerun: this sample's leaves and their samples
                int npts = 0;
                const LeafList none{nullptr, 0};
                const int n = ref_collect<true>(R.q, o, dispersion_for_spp(R.q.spp_pow2) * 2.f, none, 0, npts);
                atomicMax(&R.need[0], (unsigned)n);
                atomicMax(&R.need[1], (unsigned)npts);
                R.to_over[atomicAdd(R.n_over, 1u)] = q;
            }
        }
        wave_mem_sync();
    }
}

}  // namespace

int ref_stage_blocks(int cus, int waves_per_cu) { return std::max(1, cus * waves_per_cu / (kRefStageBlock / 64)); }

hipError_t launch_ref_first(const RefLaunch& r, hipStream_t s) {
    if (r.q.num_queries == 0) return hipSuccess;
    const unsigned blocks = (r.q.num_queries + kRefFirstBlock - 1) / kRefFirstBlock;
    if (r.q.collect_stats)
        k_ref_first<true><<<blocks, kRefFirstBlock, 0, s>>>(r);
    else
        k_ref_first<false><<<blocks, kRefFirstBlock, 0, s>>>(r);
    return hipGetLastError();
}

hipError_t launch_ref_stage(const RefLaunch& r, int stage, int blocks, hipStream_t s) {
    if (stage == 0)
        k_ref_stage<0><<<blocks, kRefStageBlock, 0, s>>>(r);
    else
        k_ref_stage<1><<<blocks, kRefStageBlock, 0, s>>>(r);
    return hipGetLastError();
}

hipError_t launch_ref_fail_wave(const RefLaunch& r, int blocks, hipStream_t s) {
    k_ref_fail_wave<<<blocks, kRefStageBlock, 0, s>>>(r);
    return hipGetLastError();
}

hipError_t launch_ref_nearest(const RefLaunch& r, int blocks, hipStream_t s) {
    k_ref_nearest<<<blocks, kRefStageBlock, 0, s>>>(r);
    return hipGetLastError();
}

}  // namespace tlr
