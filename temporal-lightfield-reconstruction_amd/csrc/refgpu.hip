// The reference's GPU-variant method: recKernel, failKernel and nearestKernel
// (ReconstructionCudaKernels.cu:242-868), driven as in
// CudaReconstruction::reconstructGPU (ReconstructionCuda.cpp:327-465).
//
// One lane per output sample in every stage:
//   k_ref_first   1R, leaf list in LDS (kRefFirstCap entries per lane);
//                 the first surface's tent sum when its points surround the
//                 output sample, else the sample goes to the 2R or the
//                 nearest list;
//   k_ref_stage   persistent waves over a list: 1R for the samples whose
//                 leaves did not fit the LDS list (STAGE 0), or 2R surface by
//                 surface with the triangle test (STAGE 1); per-lane leaf
//                 and point lists in global memory, lane-interleaved per wave;
//   k_ref_nearest the nearest reprojected sample within 1, 2, 4, 8, 16 x
//                 the dispersion, radius shrinking during the walk.
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

// Per-lane list of (key bits, leaf id) pairs: element k at p[k * stride].
struct LeafList {
    int2* p;
    int stride;
    __device__ __forceinline__ int2 get(int k) const { return p[(size_t)k * stride]; }
    __device__ __forceinline__ void put(int k, int2 v) const { p[(size_t)k * stride] = v; }
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
                int j = n;
                while (j > 0) {
                    const int2 e = L.get(j - 1);
                    const float kk = __int_as_float(e.x);
                    if (!(kk > key || (!(key > kk) && e.y > id))) break;
                    L.put(j, e);
                    j--;
                }
                L.put(j, make_int2(__float_as_int(key), id));
            }
            if (SIZES) {
                const int id = -child - 1;
                npts += Q.leaves[id].s1 - Q.leaves[id].s0;
            }
            n++;
        }
        node = escape;
    }
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
        const int2 e = L.get(li);
        const int id = e.y;
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
            L.put(li, make_int2(e.x, id | kEmpty));
        } else {
            SurfSig cur, prev;
            sig_of(Q, id, o, cur);
            for (int j = 0; j < li; j++) {
                const int pj = L.get(j).y;
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
            const int id = L.get(lidx).y;
            if (np >= 4) {
                SurfSig cur, prev;
                sig_of(Q, id, o, cur);
                bool end = false;
                for (int j = sfirst; j < lidx; j++) {
                    sig_of(Q, L.get(j).y, o, prev);
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
__global__ __launch_bounds__(kRefFirstBlock) void k_ref_first(RefLaunch R) {
    __shared__ int2 lst[kRefFirstCap][kRefFirstBlock];
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

hipError_t launch_ref_nearest(const RefLaunch& r, int blocks, hipStream_t s) {
    k_ref_nearest<<<blocks, kRefStageBlock, 0, s>>>(r);
    return hipGetLastError();
}

}  // namespace tlr
