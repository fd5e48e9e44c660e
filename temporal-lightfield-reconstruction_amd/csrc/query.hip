// Output-sample reconstruction: one wavefront per output sample (x,y,u,v,t).
//
// Follows Filterer::reconstruct / collectInputSamples2
// (ReconstructionOur.cpp:50-511) and FilterTask::process
// (Reconstruction.cpp:642-720).  All decisions use the same FP32 values as
// the sequential reference; lanes only split independent work:
//   - hierarchy traversal: a wave-wide frontier over the tree (the set of
//     leaves found does not depend on visiting order);
//   - leaf ordering: rank sort by (key, leaf id);
//   - per-leaf sample tests: one sample per lane, kept samples compacted in
//     sample order with ballot + mbcnt;
//   - sameSurface against earlier leaves: one leaf per lane, AND-reduced;
//   - triangle coverage: one (k,l) vertex pair per lane for each m.
// Sums that the reference accumulates sequentially (tent filter) are
// accumulated in the same order by every lane.
#include "query.hpp"
#include "query_util.hpp"

namespace tlr {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
enum { XPOS = 1, YPOS = 2, XNEG = 4, YNEG = 8 };

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
// Orders global-memory scratch writes of the wave before later reads.
__device__ __forceinline__ void wave_sync_mem() { __threadfence_block(); }

// The wave's scratch area: one base pointer, each array at a constant
// offset from it (computed where it is used, so that only the base stays
// live in registers).
struct WaveMem {
    float* base;
    template <class T>
    __device__ __forceinline__ T* at(size_t words) const {
        return reinterpret_cast<T*>(base + words);
    }
    static constexpr size_t kStack = 0;
    static constexpr size_t kLkey = kStack + kStackCap;
    static constexpr size_t kLid = kLkey + kLeafCap;
    static constexpr size_t kSid = kLid + kLeafCap;
    static constexpr size_t kSx = kSid + kLeafCap;
    static constexpr size_t kSy = kSx + kSampleCap;
    static constexpr size_t kSd = kSy + kSampleCap;
    static constexpr size_t kSidx = kSd + kSampleCap;
    static constexpr size_t kSrpos = kSidx + kSampleCap;
    static constexpr size_t kFStart = kSrpos + kSampleCap;
    static constexpr size_t kFSize = kFStart + kLeafCap;
    static constexpr size_t kFQuad = kFSize + kLeafCap;
    static constexpr size_t kFMini = kFQuad + kLeafCap;
    static constexpr size_t kFMin = kFMini + kLeafCap;
    static constexpr size_t kEnd = kFMin + kLeafCap;
    __device__ __forceinline__ int* stack() const { return at<int>(kStack); }
    __device__ __forceinline__ float* lkey() const { return at<float>(kLkey); }
    __device__ __forceinline__ int* lid() const { return at<int>(kLid); }
    __device__ __forceinline__ int* sid() const { return at<int>(kSid); }
    __device__ __forceinline__ float* sx() const { return at<float>(kSx); }
    __device__ __forceinline__ float* sy() const { return at<float>(kSy); }
    __device__ __forceinline__ float* sd() const { return at<float>(kSd); }
    __device__ __forceinline__ int* sidx() const { return at<int>(kSidx); }
    __device__ __forceinline__ int* srpos() const { return at<int>(kSrpos); }
    __device__ __forceinline__ int* f_start() const { return at<int>(kFStart); }
    __device__ __forceinline__ int* f_size() const { return at<int>(kFSize); }
    __device__ __forceinline__ int* f_quad() const { return at<int>(kFQuad); }
    __device__ __forceinline__ int* f_mini() const { return at<int>(kFMini); }
    __device__ __forceinline__ float* f_min() const { return at<float>(kFMin); }
};
static_assert(WaveMem::kEnd == kWaveScratchWords, "the wave's arrays fill its scratch words");

__device__ __forceinline__ WaveMem wave_mem(float* base) { return WaveMem{base}; }

// Reprojection of a tree sample (pa, pb records) to the output's (u,v,t):
// ReconstructionOur.cpp:438-451
__device__ __forceinline__ void reproject_point(const float4 a, const float4 b, const OSample& o, float cc0,
                                                float cc1, float& x, float& y) {
    const float sx = a.x, sy = a.y, sw = a.z, st = a.w;
    const float mvx = b.x, mvy = b.y, mvw = b.z;
    if (mvx != 0.0f || mvy != 0.0f || mvw != 0.0f) {
        const float dt = o.t - st;
        const float Px = sx * sw + dt * mvx;
        const float Py = sy * sw + dt * mvy;
        const float Pw = 1.0f * sw + dt * mvw;
        const float c = rcp(Pw);
        const float coc = cc0 / Pw + cc1;
        x = Px * c + coc * o.u;
        y = Py * c + coc * o.v;
    } else {
        const float coc = cc0 / sw + cc1;
        x = sx + coc * o.u;
        y = sy + coc * o.v;
    }
}

__device__ __forceinline__ void reproject_to(const Points& P, int k, const OSample& o, float cc0, float cc1,
                                             float& x, float& y) {
    reproject_point(pt_a(P, k), pt_b(P, k), o, cc0, cc1, x, y);
}

// Per-wave LDS area, shared by the gather (traversal frontier, leaf keys,
// ids and sample ranges of the first kStage leaves) and the triangle search
// that runs after it (samples of one surface).
// (Both capacities can be set at compile time; a test build sets them tiny
// so that every query takes the paths past them.)
#ifndef TLR_GATHER_STAGE
#define TLR_GATHER_STAGE 128
#endif
#ifndef TLR_GATHER_LDS_STACK
#define TLR_GATHER_LDS_STACK 256
#endif
constexpr int kStage = TLR_GATHER_STAGE;
constexpr int kLdsStack = TLR_GATHER_LDS_STACK;   // traversal frontier entries in LDS
static_assert(kStage >= 1 && kLdsStack >= 64, "the depth-6 start puts up to 64 entries on the stack");
constexpr int kTriCap = 200;
constexpr int kGatherLdsWords = 5 * kStage + kLdsStack;
constexpr int kTriLdsWords = 3 * kTriCap + kTriCap / 2;   // tx, ty, tr (words), tl (shorts)
// (a multiple of 4 words: the leaf keys and ids are read 16 bytes at a time)
constexpr int kWaveLdsWords = ((kGatherLdsWords > kTriLdsWords ? kGatherLdsWords : kTriLdsWords) + 3) & ~3;
static_assert(kStage % 4 == 0, "leaf ids start 16-byte aligned after the keys");
__device__ __forceinline__ float* wave_lds() {
    __shared__ __align__(16) float area[kWavesPerBlock][kWaveLdsWords];
    return area[threadIdx.x >> 6];
}

struct Gather {
    int total;     // gathered samples
    int nsurf;     // surfaces
    int used;      // sample slots in use (>= total after merges)
    int overflow;
    int nl;        // leaves found (sorted ids: first kStage in LDS, the rest in m.sid())
};

// collectInputSamples2 for one output sample at radius R (wave-uniform).
// STATS: count the reference's Stats (the overflow count is kept always).
template <bool STATS>
__device__ Gather gather(const QueryParams& Q, const WaveMem& m, const OSample& o, float R,
                         unsigned long long* st) {
    const int lane = threadIdx.x & 63;
    Gather g{0, 0, 0, 0, 0};
    int nl = 0;
    // leaves found (traversal order) and sorted, the first kStage in LDS,
    // the rest in the wave's scratch: key, id; sorted id, s0, s1
    float* const Lkey = wave_lds();
    int* const Lid = reinterpret_cast<int*>(Lkey + kStage);
    int* const Sid = Lid + kStage;
    int* const S0 = Sid + kStage;
    int* const S1 = S0 + kStage;
    // --- traversal: wave-wide frontier stack, in LDS until it outgrows
    // kLdsStack entries, then in the wave's scratch (a generic pointer) ---
    int* stk = S1 + kStage;
    bool lds_stack = true;
    if (Q.n_nodes > 0) {
        int sp = 1;
        const int nv = (Q.n_nodes + 1) / 2;
        if (nv >= 128) {
            // Start at depth 6: lane l takes the node whose path from the
            // root follows the bits of l (internal node of leaf range
            // [lo, hi) is (lo + hi) / 2 - 1), and keeps it when it and its
            // six ancestors are hit -- the nodes a walk from the root would
            // reach at that depth, tested in one step instead of six.
            int lo = 0, hi = nv;
            bool ok = true;
#pragma unroll
            for (int d = 0; d < 6; d++) {
                const int mid = (lo + hi) / 2;
                const TravNode& an = Q.trav[mid - 1];
                float p[12];
#pragma unroll
                for (int q = 0; q < 12; q++) p[q] = an.p[q];
                ok = ok && node_hit(p, o, R);
                if ((lane >> (5 - d)) & 1)
                    lo = mid;
                else
                    hi = mid;
            }
            const uint64_t km = __ballot(ok);
            if (ok) stk[lanes_below(km)] = (lo + hi) / 2 - 1;
            sp = __popcll(km);
        } else if (lane == 0) {
            stk[0] = Q.root;
        }
        wave_sync_mem();
        while (sp > 0) {
            const int take = sp < 64 ? sp : 64;
            const int base = sp - take;
            int node = -1;
            if (lane < take) node = stk[base + lane];
            bool hit = false, internal = false;
            int c0 = 0, c1 = 0, lidv = 0;
            float key = 0.0f;
            if (node >= 0) {
                // the whole 64-byte node at once: links are taken without a
                // branch, so they do not load after the bounds test
                const float4* tn4 = reinterpret_cast<const float4*>(Q.trav + node);
                const float4 a = tn4[0], b = tn4[1], c = tn4[2], d = tn4[3];
                const float p[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
                const int child = __float_as_int(d.x);
                hit = node_hit(p, o, R);
                internal = child >= 0;
                c0 = child;
                c1 = __float_as_int(d.z);
                lidv = -child - 1;
                if (hit && !internal) {
                    const LeafRec& lr = Q.leaves[lidv];
                    key = lr.w0 + (o.t - lr.t0) * lr.mvw0;   // ReconstructionOur.cpp:384-385
                }
            }
            const uint64_t im = __ballot(hit && internal);
            const uint64_t lm = __ballot(hit && !internal);
            const int ni = __popcll(im), nlh = __popcll(lm);
            if (base + 2 * ni > kStackCap || nl + nlh > kLeafCap) {
                g.overflow = 1;
                break;
            }
            if (lds_stack && base + 2 * ni > kLdsStack) {
                // the entries still pending move to scratch
                for (int e = lane; e < base; e += 64) m.stack()[e] = stk[e];
                stk = m.stack();
                lds_stack = false;
            }
            if (hit && internal) {
                const int r = (int)lanes_below(im);
                stk[base + 2 * r] = c0;
                stk[base + 2 * r + 1] = c1;
            }
            if (hit && !internal) {
                const int r = nl + (int)lanes_below(lm);
                if (r < kStage) {
                    Lkey[r] = key;
                    Lid[r] = lidv;
                } else {
                    m.lkey()[r] = key;
                    m.lid()[r] = lidv;
                }
            }
            sp = base + 2 * ni;
            nl += nlh;
            wave_sync_mem();
        }
    }
    if (g.overflow) return g;
    if (STATS) st[ST_LEAF_NODES] += (unsigned long long)nl;
    // --- front-to-back order: rank by (key, leaf id) ---
    const int nls = nl < kStage ? nl : kStage;
    for (int j = lane; j < nl; j += 64) {
        const float kj = j < kStage ? Lkey[j] : m.lkey()[j];
        const int ij = j < kStage ? Lid[j] : m.lid()[j];
        const LeafRec& lr = Q.leaves[ij];
        const int s0 = lr.s0, s1 = lr.s1;
        int rank = 0;
        // the staged entries four at a time (16-byte LDS reads), then the rest
        int i = 0;
        for (; i + 4 <= nls; i += 4) {
            const float4 k4 = *reinterpret_cast<const float4*>(Lkey + i);
            const int4 i4 = *reinterpret_cast<const int4*>(Lid + i);
            rank += (k4.x < kj || (!(kj < k4.x) && i4.x < ij)) ? 1 : 0;
            rank += (k4.y < kj || (!(kj < k4.y) && i4.y < ij)) ? 1 : 0;
            rank += (k4.z < kj || (!(kj < k4.z) && i4.z < ij)) ? 1 : 0;
            rank += (k4.w < kj || (!(kj < k4.w) && i4.w < ij)) ? 1 : 0;
        }
        for (; i < nls; i++) rank += (Lkey[i] < kj || (!(kj < Lkey[i]) && Lid[i] < ij)) ? 1 : 0;
        for (i = kStage; i < nl; i++) {
            const float ki = m.lkey()[i];
            const int ii = m.lid()[i];
            rank += (ki < kj || (!(kj < ki) && ii < ij)) ? 1 : 0;
        }
        if (rank < kStage) {
            Sid[rank] = ij;
            S0[rank] = s0;
            S1[rank] = s1;
        } else {
            m.sid()[rank] = ij;
        }
    }
    wave_sync_mem();
    // --- per-leaf sample tests and surface assignment ---
    const float RANGE = 1.0f / sqrtf(128.0f);
    const float u0 = o.u - RANGE, v0 = o.v - RANGE, t0 = o.t - 0.5f * RANGE;
    const float u1 = o.u + RANGE, v1 = o.v + RANGE, t1 = o.t + 0.5f * RANGE;
    const float R2 = R * R;
    int total = 0, nsurf = 0, sstart = 0;
    int c_start = 0, c_size = 0, c_quad = 0, c_mini = -1;
    float c_min = kFltMax;
    unsigned long long in_leaves = 0;
    for (int li = 0; li < nl; li++) {
        int lid, s0, s1;
        if (li < kStage) {
            lid = uni(Sid[li]);
            s0 = uni(S0[li]);
            s1 = uni(S1[li]);
        } else {
            lid = uni(m.sid()[li]);
            const LeafRec lr = Q.leaves[lid];
            s0 = uni(lr.s0);
            s1 = uni(lr.s1);
        }
        in_leaves += (unsigned long long)(s1 - s0);
        bool first = true;
        for (int c = s0; c < s1; c += 64) {
            const int k = c + lane;
            bool kept = false;
            float x = 0, y = 0, d2 = 0, dx = 0, dy = 0;
            if (k < s1) {
                reproject_to(Q.pts, k, o, Q.cc0, Q.cc1, x, y);
                dx = x - o.x;
                dy = y - o.y;
                d2 = dx * dx + dy * dy;
                kept = d2 <= R2;
            }
            const uint64_t km = __ballot(kept);
            if (km == 0) continue;
            if (first) {
                first = false;
                bool same = true;
                // both leaves' 12 coefficients as three 16-byte loads each,
                // issued together ahead of the plane-by-plane test
                const float4* np4 = reinterpret_cast<const float4*>(Q.leaf_b[lid].p);
                const float4 na = np4[0], nb = np4[1], nc = np4[2];
                const float np[12] = {na.x, na.y, na.z, na.w, nb.x, nb.y, nb.z, nb.w, nc.x, nc.y, nc.z, nc.w};
                // with no surface yet a new one opens whatever the test says
                for (int j0 = sstart; nsurf > 0 && j0 < li && same; j0 += 64) {
                    const int j = j0 + lane;
                    bool ok = true;
                    if (j < li) {
                        const float4* b4 = reinterpret_cast<const float4*>(Q.leaf_b[j < kStage ? Sid[j] : m.sid()[j]].p);
                        const float4 ba = b4[0], bb = b4[1], bc = b4[2];
                        const float bp[12] = {ba.x, ba.y, ba.z, ba.w, bb.x, bb.y, bb.z, bb.w, bc.x, bc.y, bc.z, bc.w};
                        ok = same_surface(np, bp, u0, v0, t0, u1, v1, t1);
                    }
                    same = __ballot(!ok) == 0;
                }
                if (!same || nsurf == 0) {
                    if (nsurf > 0 && lane == 0) {
                        m.f_start()[nsurf - 1] = c_start;
                        m.f_size()[nsurf - 1] = c_size;
                        m.f_quad()[nsurf - 1] = c_quad;
                        m.f_min()[nsurf - 1] = c_min;
                        m.f_mini()[nsurf - 1] = c_mini;
                    }
                    sstart = li;
                    nsurf++;
                    c_start = total;
                    c_size = 0;
                    c_quad = 0;
                    c_min = kFltMax;
                    c_mini = -1;
                }
            }
            const int cnt = __popcll(km);
            if (total + cnt > kSampleCap / 2) {
                g.overflow = 1;
                return g;
            }
            const int rpos = (dx >= 0 ? XPOS : XNEG) | (dy >= 0 ? YPOS : YNEG);
            const float dist = kept ? sqrtf(d2) : kFltMax;
            if (kept) {
                const int pos = total + (int)lanes_below(km);
                m.sx()[pos] = x;
                m.sy()[pos] = y;
                m.sd()[pos] = dist;
                m.sidx()[pos] = k;
                m.srpos()[pos] = rpos;
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (__ballot(kept && (rpos & (XPOS | YPOS)) == q)) c_quad |= 1 << q;
            // first occurrence of the chunk minimum (sequential strict-< rule)
            float dm = dist;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                const float od = __shfl_xor(dm, d, 64);
                dm = od < dm ? od : dm;
            }
            const uint64_t am = __ballot(kept && dist == dm);
            if (dm < c_min) {
                const int first_lane = __ffsll((unsigned long long)am) - 1;
                const int pos0 = total + (int)__popcll(km & ((1ull << first_lane) - 1ull));
                c_min = dm;
                c_mini = pos0 - c_start;
            }
            total += cnt;
            c_size += cnt;
        }
    }
    if (nsurf > 0 && lane == 0) {
        m.f_start()[nsurf - 1] = c_start;
        m.f_size()[nsurf - 1] = c_size;
        m.f_quad()[nsurf - 1] = c_quad;
        m.f_min()[nsurf - 1] = c_min;
        m.f_mini()[nsurf - 1] = c_mini;
    }
    wave_sync_mem();
    if (STATS) {
        st[ST_SAMPLES_IN_LEAVES] += in_leaves;
        st[ST_FETCHED] += (unsigned long long)total;
    }
    g.total = total;
    g.nsurf = nsurf;
    g.used = total;
    g.nl = nl;
    return g;
}

// Pair index p -> (a, b), a < b, pairs ordered b-major: (0,1),(0,2),(1,2),...
__device__ __forceinline__ void pair_of(int p, int& a, int& b) {
    int bb = (int)floorf((1.0f + sqrtf(1.0f + 8.0f * (float)p)) * 0.5f);
    while (bb * (bb - 1) / 2 > p) bb--;
    while ((bb + 1) * bb / 2 <= p) bb++;
    b = bb;
    a = p - bb * (bb - 1) / 2;
}

// Does any triangle of the surface's samples cover o?  The reference loops
// over index triples m < k < l (ReconstructionOur.cpp:221-259); the answer is
// a plain "any", so the visiting order does not matter, and every triple that
// is tested is tested exactly as there (same vertex order, same FP32 edge
// lengths and distances, FP64 edge functions).  The surface's samples
// (n < 200, :219) are staged in LDS with their quadrant bits and whether they
// lie within the dispersion.  Triples that cannot pass are skipped by exact
// tests only:
//   - no triangle can have all four quadrants unless the surface holds two
//     samples in opposite quadrants (XPOS|YPOS with XNEG|YNEG, or XPOS|YNEG
//     with XNEG|YPOS);
//   - for each m, k and l are taken from the samples j > m whose edge to m
//     is at most 2 * dispersion (the e01 / e02 tests of :236-238).
__device__ bool triangle_search(const WaveMem& m, int start, int n, const OSample& o, float disp) {
    float* const tx = wave_lds();
    float* const ty = tx + kTriCap;
    int* const tr = reinterpret_cast<int*>(ty + kTriCap);
    short* const tl = reinterpret_cast<short*>(tr + kTriCap);
    const int lane = threadIdx.x & 63;
    constexpr int kNear = 16;
    constexpr int kAll = XPOS | XNEG | YPOS | YNEG;
    int cls = 0;  // bit q: some sample has quadrant q (q = rpos & (XPOS|YPOS))
    for (int j = lane; j < n; j += 64) {
        const float x = m.sx()[start + j], y = m.sy()[start + j];
        const float d = sqrtf((x - o.x) * (x - o.x) + (y - o.y) * (y - o.y));
        const int rp = m.srpos()[start + j];
        tx[j] = x;
        ty[j] = y;
        tr[j] = rp | (!(d > disp) ? kNear : 0);
        cls |= 1 << (rp & (XPOS | YPOS));
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) cls |= __shfl_xor(cls, d, 64);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // quadrant q = (x >= 0 ? XPOS : 0) | (y >= 0 ? YPOS : 0); opposite pairs
    // are (XPOS|YPOS, 0) = (3, 0) and (XPOS, YPOS) = (1, 2)
    const bool opposite = ((cls & (1 << 3)) && (cls & (1 << 0))) || ((cls & (1 << 1)) && (cls & (1 << 2)));
    bool any = false;
    for (int mm = 0; opposite && mm < n - 2 && !any; mm++) {
        const float r0x = tx[mm], r0y = ty[mm];
        const int p0 = tr[mm];
        // candidates j > mm with |r0 - rj| <= 2 disp, in index order
        int s = 0;
        for (int jb = mm + 1; jb < n; jb += 64) {
            const int j = jb + lane;
            bool c = false;
            if (j < n) {
                const float rx = tx[j], ry = ty[j];
                const float e = sqrtf((r0x - rx) * (r0x - rx) + (r0y - ry) * (r0y - ry));
                c = !(e > 2 * disp);
            }
            const uint64_t cm = __ballot(c);
            if (c) tl[s + (int)lanes_below(cm)] = (short)j;
            s += __popcll(cm);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int np = s * (s - 1) / 2;
        for (int pb = 0; pb < np; pb += 64) {
            const int p = pb + lane;
            bool cov = false;
            if (p < np) {
                int a, b;
                pair_of(p, a, b);
                const int k = tl[a], l = tl[b];
                const int p1 = tr[k], p2 = tr[l];
                if (((p0 | p1 | p2) & kAll) == kAll && ((p0 | p1 | p2) & kNear)) {
                    const float r1x = tx[k], r1y = ty[k];
                    const float r2x = tx[l], r2y = ty[l];
                    const float e12 = sqrtf((r1x - r2x) * (r1x - r2x) + (r1y - r2y) * (r1y - r2y));
                    if (!(e12 > 2 * disp)) cov = tri_covers(r0x, r0y, r1x, r1y, r2x, r2y, o.x, o.y);
                }
            }
            if (__ballot(cov)) {
                any = true;
                break;
            }
        }
        // tl is rewritten for the next m
        __builtin_amdgcn_wave_barrier();
    }
    // the LDS area is rewritten by the next call
    __builtin_amdgcn_wave_barrier();
    return any;
}

// One output sample of the hand-off list, the whole algorithm
// (Filterer::reconstruct, ReconstructionOur.cpp:50-358), wave-uniform.
// DEBUG also writes the debug-image values (a separate instantiation: the
// extra live values cost the plain one registers).
// w of tree sample k at output time t: the key collectInputSamples2 gives a
// gathered sample (ReconstructionOur.cpp:438-451)
__device__ __forceinline__ float sample_w_at(const Points& P, int k, float t) {
    const float4 a = pt_a(P, k), b = pt_b(P, k);
    if (b.x != 0.0f || b.y != 0.0f || b.z != 0.0f) return 1.0f * a.z + (t - a.w) * b.z;
    return a.z;
}

// Does a gathered sample at offset (dx, dy) = o.xy - its xy, with w = key at
// the output's t and w-gradient (wgx, wgy), light a shadow query at w = ow?
// ReconstructionOur.cpp:289 (SHADOW_BIAS 0.01, :46); FW's dot sums from zero.
__device__ __forceinline__ bool shadow_lit(float ow, float dx, float dy, float key, float wgx, float wgy) {
    float d = 0.0f;
    d += dx * wgx;
    d += dy * wgy;
    return ow - 0.01f <= key + d;
}

template <bool DEBUG, int MODE = 0, bool STATS = true>
__device__ __forceinline__ void general_query(const QueryParams& Q, const WaveMem& m, unsigned qi, bool skip1r,
                                              unsigned long long* st) {
    const int lane = threadIdx.x & 63;
    OSample o;
    float ow = 0.0f, density = 1.0f;
    if (MODE != 0 && Q.qin) {
        const float4 a = Q.qin[2 * (size_t)qi], b = Q.qin[2 * (size_t)qi + 1];
        o.x = unif(a.x);
        o.y = unif(a.y);
        o.u = unif(a.z);
        o.v = unif(a.w);
        o.t = unif(b.x);
        ow = unif(b.y);
        density = unif(b.z);
    } else {
        const int M = Q.M;
        const unsigned pix = qi / (unsigned)M;
        const int i = (int)(qi - pix * (unsigned)M);
        const int px = (int)(pix % (unsigned)Q.W), py = (int)(pix / (unsigned)Q.W) + Q.y0;
        const int pat = (int)(hash_bits((uint32_t)px, (uint32_t)py) % 64u);
        const float* tp = Q.pattern + 5 * ((size_t)pat * M + i);
        o.x = tp[0] + (float)px;
        o.y = tp[1] + (float)py;
        o.u = tp[2];
        o.v = tp[3];
        o.t = tp[4];
    }
    if (STATS) st[ST_COUNT] += 1;

    // Filterer::reconstruct: ReconstructionOur.cpp:50-159
    float disp = dispersion_for_spp(Q.spp_pow2);
    // non-uniform view sample density in light space: ReconstructionOur.cpp:78-80
    if (MODE == 2) disp *= sqrtf(density);
    Gather g{0, 0, 0, 0, 0};
    int overflow = 0;
    int n1 = 0;
    bool f2 = true;
    // One gather call site (the gather is large; every inlined copy
    // costs registers): pass 0 = 1R, 1 = 2R, 2..4 = emergency passes
    // (dispersion doubled before each, gather at 2x dispersion).
    int pass = skip1r ? 1 : 0;
    for (;;) {
        g = gather<STATS>(Q, m, o, pass == 0 ? disp : disp * 2, st);
        overflow = g.overflow;
        if (overflow) break;
        if (pass == 0) {
            n1 = g.total;
            const int quad0 = g.nsurf > 0 ? uni(m.f_quad()[0]) : 0;
            f2 = g.nsurf == 0 || (quad0 != 0xF && g.nsurf > 1);
            if (!f2) break;
            pass = 1;
            continue;
        }
        if (g.total != 0 || pass == 4) break;
        if (STATS && pass == 1) st[ST_AT_LEAST_ONE] += 1;
        disp *= 2.f;
        pass++;
    }
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!overflow) {
        if (STATS) {
            st[ST_2R] += f2 ? 1 : 0;
            st[ST_WITHIN_1R] += (unsigned long long)n1;
        }
        const int nsurf = g.nsurf;
        int numSurfaces = nsurf;
        int used = g.used;
        // Tiny-surface merge: ReconstructionOur.cpp:137-156.  dst's list
        // becomes dst ++ src (materialised at the end of the sample area);
        // minIndex is copied as is (it then indexes the merged list).
        if (f2) {
            for (int si = 0; si < nsurf - 1; si++) {
                const int ssize = uni(m.f_size()[si]);
                if (ssize < 4) {
                    const int sst = uni(m.f_start()[si]);
                    const int dst = uni(m.f_start()[si + 1]);
                    const int dsize = uni(m.f_size()[si + 1]);
                    if (used + dsize + ssize > kSampleCap) {
                        overflow = 1;
                        break;
                    }
                    for (int j = lane; j < dsize + ssize; j += 64) {
                        const int from = j < dsize ? dst + j : sst + (j - dsize);
                        m.sx()[used + j] = m.sx()[from];
                        m.sy()[used + j] = m.sy()[from];
                        m.sd()[used + j] = m.sd()[from];
                        m.sidx()[used + j] = m.sidx()[from];
                        m.srpos()[used + j] = m.srpos()[from];
                    }
                    const float smin = unif(m.f_min()[si]);
                    const float dmin = unif(m.f_min()[si + 1]);
                    const int smini = uni(m.f_mini()[si]);
                    wave_sync_mem();
                    if (lane == 0) {
                        m.f_start()[si + 1] = used;
                        m.f_size()[si + 1] = dsize + ssize;
                        m.f_size()[si] = 0;
                        if (smin < dmin) {
                            m.f_min()[si + 1] = smin;
                            m.f_mini()[si + 1] = smini;
                        }
                    }
                    wave_sync_mem();
                    used += dsize + ssize;
                    numSurfaces--;
                }
            }
        }
        if (STATS) {
            st[ST_SURFACES] += (unsigned long long)numSurfaces;
            st[ST_MERGED] += (numSurfaces != nsurf) ? 1 : 0;
        }

        // Coverage and filtering, front to back: ReconstructionOur.cpp:165-302
        bool found = false;
        float cr = 0, cg = 0, cb = 0, ca = 0;
        float nearest = kFltMax;
        int nI0 = -1, nI1 = 1;
        for (int si = 0; si < nsurf && !found && !overflow; si++) {
            const int fs = uni(m.f_start()[si]);
            const int n = uni(m.f_size()[si]);
            if (n == 0) continue;
            const float fmin = unif(m.f_min()[si]);
            const int fmini = uni(m.f_mini()[si]);
            if (fmin < nearest) {
                nearest = fmin;
                nI0 = si;
                nI1 = fmini;
            }
            if (!f2 || fmin < 0.f || si == nsurf - 1) found = true;
            if (n >= 200) found = true;
            if (!found) found = triangle_search(m, fs, n, o, disp);
            if (found) {
                cr = cg = cb = ca = 0;
                // weights and colours of 64 samples load in parallel; the
                // sum then runs in sample order over lane-held values
                for (int c0 = 0; c0 < n; c0 += 64) {
                    const int k = c0 + lane;
                    float wgt = 0.0f;
                    float4 c = make_float4(0, 0, 0, 0);
                    if (k < n) {
                        const float rx = m.sx()[fs + k], ry = m.sy()[fs + k];
                        const float dx = o.x - rx, dy = o.y - ry;
                        wgt = 1 - sqrtf(dx * dx + dy * dy) / disp;
                        if (wgt > 0) {
                            const int sk = m.sidx()[fs + k];
                            c = Q.pts.pc[sk];
                            if (MODE == 2) {
                                // per-input-sample thresholding: ReconstructionOur.cpp:283-291
                                const float4 e = Q.pext[sk];
                                if (!shadow_lit(ow, dx, dy, sample_w_at(Q.pts, sk, o.t), e.x, e.y))
                                    c.x = c.y = c.z = 0.0f;
                            }
                        }
                    }
                    const int cnt = min(64, n - c0);
                    for (int j = 0; j < cnt; j++) {
                        const float wj = __shfl(wgt, j, 64);
                        if (wj <= 0) continue;
                        cr += wj * __shfl(c.x, j, 64);
                        cg += wj * __shfl(c.y, j, 64);
                        cb += wj * __shfl(c.z, j, 64);
                        ca += wj * 1.0f;
                    }
                }
                const float oow = rcp(ca);
                cr *= oow; cg *= oow; cb *= oow; ca *= oow;
                nI0 = si;
                nI1 = fmini;
            }
        }
        if (MODE != 0 && lane == 0 && Q.qext) {
            // Filterer::Result's xy, w, wg and density from the nearest
            // sample: ReconstructionOur.cpp:325-331 (zero without one)
            float4 e0 = make_float4(0.f, 0.f, 0.f, 0.f), e1 = e0;
            if (!overflow && nI0 != -1) {
                const int fs = m.f_start()[nI0];
                const int sk = m.sidx()[fs + nI1];
                const float4 pe = Q.pext[sk];
                e0 = make_float4(m.sx()[fs + nI1], m.sy()[fs + nI1], sample_w_at(Q.pts, sk, o.t), pe.z);
                e1 = make_float4(pe.x, pe.y, 0.f, 0.f);
            }
            Q.qext[2 * (size_t)qi] = e0;
            Q.qext[2 * (size_t)qi + 1] = e1;
        }
        if (!overflow && nI0 != -1) {
            if (ca == 0) {
                const int fs = uni(m.f_start()[nI0]);
                const int sk = uni(m.sidx()[fs + nI1]);
                const float4 c = Q.pts.pc[sk];
                cr = c.x;
                cg = c.y;
                cb = c.z;
                ca = 1.0f;
                if (MODE == 2) {
                    // ReconstructionOur.cpp:341-347
                    const float4 pe = Q.pext[sk];
                    const float nx = unif(m.sx()[fs + nI1]), ny = unif(m.sy()[fs + nI1]);
                    if (!shadow_lit(ow, o.x - nx, o.y - ny, sample_w_at(Q.pts, sk, o.t), pe.x, pe.y))
                        cr = cg = cb = 0.0f;
                }
            }
            if (MODE != 2 && Q.lens_gauss) {
                // Gaussian(x, 0.5, 0): Reconstruction.hpp:378
                const float e = 2.718281828f;
                const float gu = powf(e, -((o.u - 0) * (o.u - 0)) / (2 * (0.5f * 0.5f)));
                const float gv = powf(e, -((o.v - 0) * (o.v - 0)) / (2 * (0.5f * 0.5f)));
                const float gg = gu * gv;
                cr *= gg; cg *= gg; cb *= gg; ca *= gg;
            }
            if (MODE != 2 && Q.time_gauss) {
                const float e = 2.718281828f;
                const float gg = powf(e, -((o.t - 0.5f) * (o.t - 0.5f)) / (2 * (0.25f * 0.25f)));
                cr *= gg; cg *= gg; cb *= gg; ca *= gg;
            }
            col = make_float4(cr, cg, cb, ca);
        }
    }
    if (overflow) st[ST_OVERFLOW] += 1;
    if (lane == 0) {
        Q.qcol[qi] = col;
        // debug image values (Reconstruction.hpp:46, ReconstructionOur.cpp:312-314):
        // surfaces of the last gather, samples within 1R (known to the fast
        // kernel when it handed the 2R fetch over)
        if (DEBUG) {
            Q.qdebug[qi].x = (float)g.nsurf;
            if (!skip1r) Q.qdebug[qi].y = (float)n1;
        }
    }
    wave_sync_mem();
}

__device__ __forceinline__ void flush_stats(unsigned long long* dst, const unsigned long long* st, bool all) {
    if ((threadIdx.x & 63) != 0) return;
    if (all) {
#pragma unroll
        for (int k = 0; k < kNumStats; k++)
            if (st[k]) atomicAdd(&dst[k], st[k]);
    } else if (st[ST_OVERFLOW]) {
        atomicAdd(&dst[ST_OVERFLOW], st[ST_OVERFLOW]);
    }
}

// 8 waves per SIMD (the register limit this sets spills; the kernel waits
// on memory and gains more from resident waves: 7 and 6 were 6 % and 15 %
// slower).  Persistent: waves take queries from an atomic counter until the
// list is done.
constexpr unsigned kGeneralChunk = 4;

template <bool DEBUG, int MODE, bool STATS>
__global__ __launch_bounds__(kBlock, 8) void k_query_general(QueryParams Q) {
    const int lane = threadIdx.x & 63;
    // wave-uniform (so the scratch pointers derived from it live in scalar registers)
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    const WaveMem m = wave_mem(Q.scratch + (size_t)wave * kWaveScratchWords);
    unsigned long long st[kNumStats];
#pragma unroll
    for (int k = 0; k < kNumStats; k++) st[k] = 0;
    const unsigned nq = Q.qlist ? *Q.qcount : Q.num_queries;
    for (;;) {
        // kGeneralChunk consecutive queries per claim (atomics on the one
        // counter serialise)
        unsigned q0 = 0;
        if (lane == 0) q0 = atomicAdd(Q.next_query, kGeneralChunk);
        q0 = (unsigned)uni((int)q0);
        if (q0 >= nq) break;
        const unsigned q1 = nq - q0 < kGeneralChunk ? nq : q0 + kGeneralChunk;
        for (unsigned e = q0; e < q1; e++) {
            unsigned qi = e;
            bool skip1r = false;
            if (Q.qlist) {
                qi = (unsigned)uni((int)Q.qlist[e]);
                if (qi == kSkipEntry) continue;  // accepted by the learned gate
                // top bit: the fast kernel already found that 2R is needed
                skip1r = (qi & kNeeds2R) != 0;
                qi &= ~kNeeds2R;
            }
            general_query<DEBUG, MODE, STATS>(Q, m, qi, skip1r, st);
        }
    }
    flush_stats(Q.stats, st, STATS && Q.collect_stats != 0);
}

// Batched form: one list of hand-offs from every buffer of a batch, with the
// buffer of each entry beside it.  A wave claims kBatchChunk entries at a
// time and reads a buffer's parameters once per run of entries from it (the
// fast kernel's waves append their hand-offs contiguously): inside a run they
// stay in scalar registers as kernel arguments do.
constexpr unsigned kBatchChunk = 8;

__global__ __launch_bounds__(kBlock, 7) void k_query_general_batch(BatchQuery B) {
    const int lane = threadIdx.x & 63;
    // wave-uniform (so the scratch pointers derived from it live in scalar registers)
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    const WaveMem m = wave_mem(B.scratch + (size_t)wave * kWaveScratchWords);
    unsigned long long st[kNumStats];
#pragma unroll
    for (int k = 0; k < kNumStats; k++) st[k] = 0;
    const unsigned nq = *B.list_count;
    for (;;) {
        unsigned e = 0;
        if (lane == 0) e = atomicAdd(B.list_next, kBatchChunk);
        e = (unsigned)uni((int)e);
        if (e >= nq) break;
        const unsigned end = min(nq, e + kBatchChunk);
        while (e < end) {
            const int item = uni((int)B.list_item[e]);
            // word by word into scalar registers (a plain struct copy lands
            // in scratch memory)
            QueryParams Q;
            {
                const int* src = reinterpret_cast<const int*>(B.items + item);
                int* dst = reinterpret_cast<int*>(&Q);
#pragma unroll
                for (int k = 0; k < (int)(sizeof(QueryParams) / 4); k++)
                    dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
            }
            do {
                const unsigned ent = (unsigned)uni((int)B.list[e]);
                if (ent != kSkipEntry)
                    general_query<false, 0, false>(Q, m, ent & ~kNeeds2R, (ent & kNeeds2R) != 0, st);
                e++;
            } while (e < end && uni((int)B.list_item[e]) == item);
        }
    }
    flush_stats(B.stats, st, false);
}

// ---------------------------------------------------------------------------
// Fast path: one lane per output sample at 1R (collectInputSamples2 with
// R = dispersion, ReconstructionOur.cpp:94).  The lane walks the tree
// stacklessly (escape links) and keeps the leaves it finds sorted by
// (key, leaf id) in LDS.  Then, in three passes over them: which leaves hold a
// sample within R; the surfaces over those leaves, front to back, exactly as
// the general kernel forms them; surface 0's quadrant mask, nearest sample
// and tent-weighted sum in sample order.  The query is finished here when the
// reference would not fetch 2R (ReconstructionOur.cpp:99): one surface, or a
// second surface behind a first one that covers all four quadrants; then the
// found surface is surface 0 (:204) and its filtered colour is the result
// (:268-297).  Every other query (no surface, a 2R fetch, or more than
// kFastLeafCap leaves) is appended to fail_list for the general kernel.
// STATS visits every leaf and sample to count the reference's Stats;
// otherwise a leaf's scan stops at its first sample within R and the surface
// pass stops when the second surface starts.
// ---------------------------------------------------------------------------
// Gate features of a hand-off entry (kGateFeatures floats), from what the
// fast kernel's 1R passes found; `prov` false (no first-surface colour: no
// surface at 1R, or more leaves than the lane's list holds) marks the entry
// with f[0] = -1.  Ratios that can be unbounded pass through a signed log;
// everything is clamped to [-16, 16] and NaN becomes 0.
//   0 log1p(leaves at 1R)  1 log1p(leaves with a sample within R)
//   2 log1p(such leaves in the first surface)  3 quadrants of the first
//   surface / 4  4..7 its quadrant bits  8 log1p(its samples within R)
//   9 their mean tent weight  10 nearest distance / R  11 log1p(leaves in
//   front of the first surface with no sample within R)  12 slog(first-surface
//   key)  13 slog(relative key gap to the second surface)  14,15 centroid
//   offset (x, y) of the first surface's samples / R  16 nearest sample of
//   the leaves in front / R (16 if none)  17..19 output (u, v, t)
//   20 slog(CoC at the first surface / 16)  21 fraction of kept leaves in
//   the first surface  22 log1p(samples tested in its leaves)
//   23 log1p(kept leaves behind it)
struct GateAcc {
    int cnt0 = 0, tested0 = 0;
    float sdx = 0.f, sdy = 0.f;
};

__device__ __forceinline__ float slog(float v) { return copysignf(log1pf(fabsf(v)), v); }

__device__ __forceinline__ void write_gate_features(float* out, bool prov, const GateAcc& g, int n, int nkept,
                                                    int nkept0, int quad0, float aw, float min0, int nsurf,
                                                    float key0, float key1, int front, float front_d2,
                                                    const OSample& o, float R, float cc0, float cc1) {
    float f[kGateFeatures];
#pragma unroll
    for (int k = 0; k < kGateFeatures; k++) f[k] = 0.0f;
    if (prov) {
        const float c = (float)(g.cnt0 > 0 ? g.cnt0 : 1);
        f[0] = log1pf((float)n);
        f[1] = log1pf((float)nkept);
        f[2] = log1pf((float)nkept0);
        f[3] = 0.25f * (float)__popc((unsigned)quad0);
#pragma unroll
        for (int q = 0; q < 4; q++) f[4 + q] = (quad0 >> q) & 1 ? 1.0f : 0.0f;
        f[8] = log1pf((float)g.cnt0);
        f[9] = aw / c;
        f[10] = min0 / R;
        f[11] = log1pf((float)front);
        f[12] = slog(key0);
        f[13] = nsurf > 1 ? slog((key1 - key0) / fmaxf(fabsf(key0), 1e-6f)) : 0.0f;
        f[14] = g.sdx / c / R;
        f[15] = g.sdy / c / R;
        f[16] = front ? sqrtf(front_d2) / R : 16.0f;
        f[17] = o.u;
        f[18] = o.v;
        f[19] = o.t;
        f[20] = slog((cc0 / key0 + cc1) / 16.0f);
        f[21] = nkept ? (float)nkept0 / (float)nkept : 0.0f;
        f[22] = log1pf((float)g.tested0);
        f[23] = log1pf((float)(nkept - nkept0));
#pragma unroll
        for (int k = 0; k < kGateFeatures; k++) {
            const float v = f[k];
            f[k] = v == v ? fminf(fmaxf(v, -16.0f), 16.0f) : 0.0f;
        }
    } else {
        f[0] = -1.0f;
    }
    float4* o4 = reinterpret_cast<float4*>(out);
#pragma unroll
    for (int k = 0; k < kGateFeatures / 4; k++) o4[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
}

// One block of the fast path: `slot` is this lane's query index (or list
// position).  lid: the block's LDS leaf lists, CAP entries per lane.
// GATE: hand-offs that have a first surface at 1R still get its filtered
// colour in qcol (a provisional result the learned gate may keep), and every
// hand-off entry gets its gate features in Q.features (row = list position).
// RELAY (list mode over a hand-off list): entries already marked kNeeds2R
// are appended to fail_list unchanged, the others are run.
// MODE (as the general kernel's): 1 = also the extended result of each
// query it completes (qext); 2 = shadow queries (radius scaled by the
// query's density, every contributing sample and the nearest-sample fallback
// kept only where they light the query, no lens/time weighting) with
// extended results.  In modes 1 and 2 the queries come from qin when it is
// set (2 float4 each: (x, y, u, v), (t, w, density, 0)), else from the
// output pattern.
template <bool STATS, int CAP, int BLOCK, bool GATE = false, bool RELAY = false, int MODE = 0>
__device__ __forceinline__ void fast_query(const QueryParams& Q, unsigned slot, unsigned item, int (*lid)[BLOCK]) {
    const int tid = threadIdx.x;
    const int M = Q.M;
    const bool listed = slot < (Q.fast_list ? *Q.qcount : Q.num_queries);
    unsigned q = listed ? (Q.fast_list ? Q.qlist[slot] : slot) : 0u;
    const bool relay = RELAY && listed && (q & kNeeds2R) != 0u;
    if (RELAY) q &= ~kNeeds2R;
    const bool active = listed && !relay;
    const unsigned pix = q / (unsigned)M;
    const int i = (int)(q - pix * (unsigned)M);
    const int px = (int)(pix % (unsigned)Q.W), py = (int)(pix / (unsigned)Q.W) + Q.y0;
    bool fail = false;
    bool needs2r = false;  // failed because the reference fetches 2R
    unsigned long long s_leaves = 0, s_in_leaves = 0, s_kept = 0, s_surf = 0;
    // gate features (GATE only)
    bool prov = false;
    GateAcc gacc;
    int g_front = 0;            // leaves in front of the first with a sample within R
    float g_front_d2 = kFltMax; // their nearest sample, squared distance
    int g_n = 0, g_nkept = 0, g_nkept0 = 0, g_quad0 = 0, g_nsurf = 0;
    float g_aw = 0.f, g_min0 = 0.f, g_key0 = 0.f, g_key1 = 0.f;
    OSample o{0.f, 0.f, 0.f, 0.f, 0.f};
    float ow = 0.0f, density = 1.0f;   // modes 1, 2 with explicit queries
    if (active) {
        if (MODE != 0 && Q.qin) {
            const float4 a = Q.qin[2 * (size_t)q], b = Q.qin[2 * (size_t)q + 1];
            o.x = a.x;
            o.y = a.y;
            o.u = a.z;
            o.v = a.w;
            o.t = b.x;
            ow = b.y;
            density = b.z;
        } else {
            const int pat = (int)(hash_bits((uint32_t)px, (uint32_t)py) % 64u);
            const float* tp = Q.pattern + 5 * ((size_t)pat * M + i);
            o.x = tp[0] + (float)px;
            o.y = tp[1] + (float)py;
            o.u = tp[2];
            o.v = tp[3];
            o.t = tp[4];
        }
    }
    float disp = dispersion_for_spp(Q.spp_pow2);
    // non-uniform view sample density in light space: ReconstructionOur.cpp:78-80
    if (MODE == 2) disp *= sqrtf(density);
    const float R = disp;
    // --- wave-wide descent: the deepest node whose subtree holds every leaf
    // any lane's walk can reach.  From a node, the walk enters a child only
    // if the lane hits the node and the child; when no lane would enter one
    // of the two children, the wave moves to the other, keeping the lanes
    // that hit the node (`alive`: hit every ancestor of the current node).
    // Node loads are uniform.  Each lane then walks the subtree of `start`,
    // finding exactly the leaves a walk from the root would.
    bool alive = active;
    int start = -1;
    if (Q.n_nodes > 0) {
        int node = Q.root;
        for (;;) {
            const TravNode tn = Q.trav[node];
            if (tn.child < 0) break;
            const TravNode ta = Q.trav[tn.child], tb = Q.trav[tn.child1];
            const bool hn = alive && node_hit(tn.p, o, R);
            const uint64_t ba = __ballot(hn && node_hit(ta.p, o, R));
            const uint64_t bb = __ballot(hn && node_hit(tb.p, o, R));
            if (ba && bb) break;
            if (!ba && !bb) {
                alive = false;
                break;
            }
            node = uni(ba ? tn.child : tn.child1);
            alive = hn;
        }
        start = node;
    }
    if (active) {
        // --- depth-first walk of start's subtree over child-pair records:
        // a node's record holds both children's bounds, so the lane tests
        // them there and loads a child's record only when it enters it;
        // a second hit child waits on a short stack held in registers (a
        // query that would overflow it is handed over).  Leaves found are
        // appended to the LDS column, (id, key) pairs while it has room for
        // them (CAP / 2 leaves), ids only after that, and sorted by
        // (key, id) once the walk is done: sorting inside the walk would make
        // every lane of the wave wait for the insertions of whichever lane
        // just found a leaf, at almost every step ---
        int n = 0;
        bool pairs = true;
        int stk[kWalkStack];
#pragma unroll
        for (int k = 0; k < kWalkStack; k++) stk[k] = -1;
        int sp = 0;
        int node = -1;
        // leaves to append after the current step, with their key terms
        int fl[2] = {-1, -1};
        float fk[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        if (alive) {
            // the walk's first node, tested from its own record
            const float4* tn4 = reinterpret_cast<const float4*>(Q.trav + start);
            const float4 a = tn4[0], b = tn4[1], c = tn4[2], d = tn4[3];
            const float p[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
            const int child = __float_as_int(d.x);
            if (node_hit(p, o, R)) {
                if (child >= 0) {
                    node = start;
                } else {
                    fl[0] = -child - 1;
                    const LeafRec& lr = Q.leaves[fl[0]];
                    fk[0][0] = lr.w0;
                    fk[0][1] = lr.t0;
                    fk[0][2] = lr.mvw0;
                }
            }
        }
        for (;;) {
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int id = fl[c];
                if (id < 0 || fail) continue;
                if (n == CAP) {
                    fail = true;
                    continue;
                }
                if (pairs && n == CAP / 2) {
                    // the pairs fill the column: keep the ids only
                    for (int j = 1; j < n; j++) lid[j][tid] = lid[2 * j][tid];
                    pairs = false;
                }
                if (pairs) {
                    lid[2 * n][tid] = id;
                    // ReconstructionOur.cpp:384-385
                    lid[2 * n + 1][tid] = __float_as_int(fk[c][0] + (o.t - fk[c][1]) * fk[c][2]);
                } else {
                    lid[n][tid] = id;
                }
                n++;
            }
            if (fail || node < 0) break;
            const float4* pr4 = reinterpret_cast<const float4*>(Q.pairs + node);
            const float4 a0 = pr4[0], a1 = pr4[1], a2 = pr4[2];
            const float4 b0 = pr4[3], b1 = pr4[4], b2 = pr4[5];
            const float4 r0 = pr4[6], r1 = pr4[7];
            const float pa[12] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, a2.x, a2.y, a2.z, a2.w};
            const float pb[12] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w, b2.x, b2.y, b2.z, b2.w};
            const int ra = __float_as_int(r0.x), rb = __float_as_int(r0.y);
            const bool ha = node_hit(pa, o, R), hb = node_hit(pb, o, R);
            fl[0] = (ha && ra < 0) ? -ra - 1 : -1;
            fl[1] = (hb && rb < 0) ? -rb - 1 : -1;
            fk[0][0] = r0.z;
            fk[0][1] = r0.w;
            fk[0][2] = r1.x;
            fk[1][0] = r1.y;
            fk[1][1] = r1.z;
            fk[1][2] = r1.w;
            const bool ea = ha && ra >= 0, eb = hb && rb >= 0;
            if (ea && eb) {
                if (sp == kWalkStack) {
                    fail = true;
                    break;
                }
#pragma unroll
                for (int k = kWalkStack - 1; k > 0; k--) stk[k] = stk[k - 1];
                stk[0] = rb;
                sp++;
                node = ra;
            } else if (ea) {
                node = ra;
            } else if (eb) {
                node = rb;
            } else if (sp > 0) {
                node = stk[0];
#pragma unroll
                for (int k = 0; k < kWalkStack - 1; k++) stk[k] = stk[k + 1];
                sp--;
            } else {
                node = -1;
            }
        }
        if (!fail) {
            // sort by (key, id), leaving the ids in lid[0 .. n)
            if (pairs) {
                for (int i = 1; i < n; i++) {
                    const int id = lid[2 * i][tid];
                    const float key = __int_as_float(lid[2 * i + 1][tid]);
                    int j = i;
                    while (j > 0) {
                        const int ii = lid[2 * j - 2][tid];
                        const float kk = __int_as_float(lid[2 * j - 1][tid]);
                        if (!(kk > key || (!(key > kk) && ii > id))) break;
                        lid[2 * j][tid] = ii;
                        lid[2 * j + 1][tid] = __float_as_int(kk);
                        j--;
                    }
                    lid[2 * j][tid] = id;
                    lid[2 * j + 1][tid] = __float_as_int(key);
                }
                for (int j = 1; j < n; j++) lid[j][tid] = lid[2 * j][tid];
            } else {
                for (int i = 1; i < n; i++) {
                    const int id = lid[i][tid];
                    const float key = leaf_key(Q.leaves, id, o.t);
                    int j = i;
                    while (j > 0) {
                        const int ii = lid[j - 1][tid];
                        const float kk = leaf_key(Q.leaves, ii, o.t);
                        if (!(kk > key || (!(key > kk) && ii > id))) break;
                        lid[j][tid] = ii;
                        j--;
                    }
                    lid[j][tid] = id;
                }
            }
        }
        if (!fail) {
            s_leaves = (unsigned long long)n;
            const float RANGE = 1.0f / sqrtf(128.0f);
            const float u0 = o.u - RANGE, v0 = o.v - RANGE, t0 = o.t - 0.5f * RANGE;
            const float u1 = o.u + RANGE, v1 = o.v + RANGE, t1 = o.t + 0.5f * RANGE;
            const float R2 = R * R;
            const float4* __restrict__ pab = Q.pts.pab;
            const float4* __restrict__ pc = Q.pts.pc;
            // Three passes, so that the per-leaf sameSurface work does not
            // sit inside the sample loop, where some lane of the wave meets
            // the first kept sample of some leaf at almost every step:
            //  1. which leaves hold a sample within R (a leaf's scan stops at
            //     its first such sample unless every sample is counted);
            //  2. surfaces over those leaves, front to back
            //     (ReconstructionOur.cpp:466-484);
            //  3. surface 0's samples in order: quadrant mask, nearest sample
            //     and tent-weighted sum (:486-499, :268-297).
            // One loop per lane over (leaf, group of 4 samples) steps rather
            // than a loop over leaves around a loop over a leaf's groups: a
            // wave then runs as many steps as its busiest lane, not the sum
            // over leaf positions of the lanes' largest leaf.  The next
            // leaf's sample range is loaded a leaf ahead.
            uint64_t kmask = 0;
            // offset from the leaf's first sample of the first sample within
            // R, for the first four leaves that hold one, a byte each (0 when
            // it does not fit); every sample before it is farther than R
            uint32_t kst = 0;
            {
                int li = 0, k0 = 0, kg = 0, k1 = 0, nk0 = 0, nk1 = 0;
                if (n > 0) {
                    const LeafRec& r = Q.leaves[lid[0][tid]];
                    kg = k0 = r.s0;
                    k1 = r.s1;
                }
                if (n > 1) {
                    const LeafRec& r = Q.leaves[lid[1][tid]];
                    nk0 = r.s0;
                    nk1 = r.s1;
                }
                s_in_leaves += (unsigned long long)(k1 - kg);
                // GATE: nearest sample (hardware-reciprocal estimate) of the
                // leaves in front of the first one with a sample within R
                float leaf_min = kFltMax;
                while (li < n) {
                    const float4* g = pab + 2 * (size_t)kg;
                    float4 ga[4], gb[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        ga[e] = g[2 * e];
                        gb[e] = g[2 * e + 1];
                    }
                    bool hit = false;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        if (kg + e >= k1) break;
                        if (GATE && kmask == 0) leaf_min = fminf(leaf_min, approx_d2(ga[e], gb[e], o, Q.cc0, Q.cc1));
                        if (surely_outside(ga[e], gb[e], o, Q.cc0, Q.cc1, R)) continue;
                        float x, y;
                        reproject_point(ga[e], gb[e], o, Q.cc0, Q.cc1, x, y);
                        const float dx = x - o.x, dy = y - o.y;
                        if (!(dx * dx + dy * dy <= R2)) continue;
                        if (!((kmask >> li) & 1ull)) {
                            const int j = __popcll(kmask);
                            const int off = kg + e - k0;
                            if (j < 4 && off < 256) kst |= (uint32_t)off << (8 * j);
                        }
                        kmask |= 1ull << li;
                        hit = true;
                        s_kept++;
                        if (!STATS) break;
                    }
                    kg += 4;
                    // a leaf's scan stops at its first sample within R
                    // unless every sample is counted
                    if ((!STATS && hit) || kg >= k1) {
                        if (GATE && kmask == 0) {
                            g_front++;
                            g_front_d2 = fminf(g_front_d2, leaf_min);
                        }
                        leaf_min = kFltMax;
                        li++;
                        kg = k0 = nk0;
                        k1 = nk1;
                        if (li < n) s_in_leaves += (unsigned long long)(k1 - kg);
                        if (li + 1 < n) {
                            const LeafRec& r = Q.leaves[lid[li + 1][tid]];
                            nk0 = r.s0;
                            nk1 = r.s1;
                        }
                    }
                }
            }
            // 2. surfaces; surface 0 = kept leaves in [first0, end0)
            int nsurf = 0, sstart = 0, first0 = n, end0 = n;
            for (uint64_t km = kmask; km; km &= km - 1) {
                const int li = __builtin_ctzll(km);
                if (nsurf == 0) {
                    nsurf = 1;
                    sstart = li;
                    first0 = li;
                    continue;
                }
                bool same = true;
                SurfSig sa, sb;
                surf_sig(Q.leaf_b[lid[li][tid]].p, u0, v0, t0, u1, v1, t1, sa);
                Bounds nb = Q.leaf_b[lid[sstart][tid]];
                for (int j = sstart; j < li && same; ++j) {
                    const Bounds cb = nb;
                    if (j + 1 < li) nb = Q.leaf_b[lid[j + 1][tid]];
                    surf_sig(cb.p, u0, v0, t0, u1, v1, t1, sb);
                    same = same_surface_sig(sa, sb);
                }
                if (!same) {
                    if (nsurf == 1) end0 = li;
                    nsurf++;
                    sstart = li;
                    if (!STATS) break;  // surface 0 is complete
                }
            }
            // 3. surface 0
            int quad0 = 0;
            float ar = 0, ag = 0, ab = 0, aw = 0;
            float min0 = kFltMax;
            float4 near = make_float4(0, 0, 0, 0);
            // modes 1, 2: the nearest sample's index and reprojected position
            int kn = 0;
            float nx = 0.0f, ny = 0.0f;
            {
                // one loop over (leaf, group) steps, as in pass 1; surface 0's
                // leaves are the first kept ones, so the first four start at
                // their first sample within R
                uint64_t km = nsurf > 0 ? (kmask & ((end0 < 64 ? (1ull << end0) : 0ull) - 1ull)) : 0ull;
                int kg = 0, k1 = 0, nk0 = 0, nk1 = 0;
                if (km) {
                    const LeafRec& r = Q.leaves[lid[__builtin_ctzll(km)][tid]];
                    if (GATE) gacc.tested0 += r.s1 - r.s0;
                    kg = r.s0 + (int)(kst & 255u);
                    k1 = r.s1;
                    km &= km - 1;
                }
                if (km) {
                    const LeafRec& r = Q.leaves[lid[__builtin_ctzll(km)][tid]];
                    nk0 = r.s0;
                    nk1 = r.s1;
                }
                while (kg < k1) {
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
                        if (surely_outside(ga[e], gb[e], o, Q.cc0, Q.cc1, R)) continue;
                        float x, y;
                        reproject_point(ga[e], gb[e], o, Q.cc0, Q.cc1, x, y);
                        const float dx = x - o.x, dy = y - o.y;
                        const float d2 = dx * dx + dy * dy;
                        if (!(d2 <= R2)) continue;
                        quad0 |= 1 << ((dx >= 0 ? XPOS : 0) | (dy >= 0 ? YPOS : 0));
                        const float dist = sqrtf(d2);
                        if (GATE) {
                            gacc.cnt0++;
                            gacc.sdx += dx;
                            gacc.sdy += dy;
                        }
                        const float4 c0 = pc[k];
                        if (dist < min0) {
                            min0 = dist;
                            near = c0;
                            if (MODE != 0) {
                                kn = k;
                                nx = x;
                                ny = y;
                            }
                        }
                        const float wgt = 1 - dist / disp;
                        if (wgt > 0) {
                            float4 col = c0;
                            if (MODE == 2) {
                                // per-input-sample thresholding: ReconstructionOur.cpp:283-291
                                const float4 pe = Q.pext[k];
                                const float key = (gb[e].x != 0.0f || gb[e].y != 0.0f || gb[e].z != 0.0f)
                                                      ? 1.0f * ga[e].z + (o.t - ga[e].w) * gb[e].z
                                                      : ga[e].z;
                                if (!shadow_lit(ow, o.x - x, o.y - y, key, pe.x, pe.y)) col.x = col.y = col.z = 0.0f;
                            }
                            ar += wgt * col.x;
                            ag += wgt * col.y;
                            ab += wgt * col.z;
                            aw += wgt * 1.0f;
                        }
                    }
                    kg += 4;
                    if (kg >= k1) {
                        // next kept leaf of surface 0, if any (leaves hold
                        // at least one sample, so kg < k1 again)
                        if (km) {
                            km &= km - 1;
                            kg = nk0;
                            k1 = nk1;
                            if (GATE) gacc.tested0 += k1 - kg;
                            kst >>= 8;   // the next kept leaf's offset (0 past four)
                            kg += (int)(kst & 255u);
                            if (km) {
                                const LeafRec& r = Q.leaves[lid[__builtin_ctzll(km)][tid]];
                                nk0 = r.s0;
                                nk1 = r.s1;
                            }
                        }
                    }
                }
            }
            const bool stop = nsurf > 1;
            const bool f2 = nsurf == 0 || (quad0 != 0xF && (nsurf > 1 || stop));
            if (f2) {
                fail = true;
                needs2r = true;
            }
            if (GATE && f2 && nsurf > 0) {
                prov = true;
                g_n = n;
                g_nkept = __popcll(kmask);
                g_nkept0 = __popcll(kmask & ((end0 < 64 ? (1ull << end0) : 0ull) - 1ull));
                g_quad0 = quad0;
                g_aw = aw;
                g_min0 = min0;
                g_nsurf = nsurf;
                g_key0 = leaf_key(Q.leaves, lid[first0][tid], o.t);
                g_key1 = nsurf > 1 ? leaf_key(Q.leaves, lid[end0][tid], o.t) : 0.0f;
            }
            if (!fail || prov) {
                if (!fail) s_surf = (unsigned long long)nsurf;
                const float oow = rcp(aw);
                float cr = ar * oow, cg = ag * oow, cb = ab * oow, ca = aw * oow;
                if (ca == 0) {
                    cr = near.x; cg = near.y; cb = near.z; ca = 1.0f;
                    if (MODE == 2 && !fail) {
                        // ReconstructionOur.cpp:341-347
                        const float4 pe = Q.pext[kn];
                        if (!shadow_lit(ow, o.x - nx, o.y - ny, sample_w_at(Q.pts, kn, o.t), pe.x, pe.y))
                            cr = cg = cb = 0.0f;
                    }
                }
                if (MODE != 0 && !fail && Q.qext) {
                    // Filterer::Result's xy, w, wg and density from the found
                    // surface's nearest sample: ReconstructionOur.cpp:325-331
                    const float4 pe = Q.pext[kn];
                    Q.qext[2 * (size_t)q] = make_float4(nx, ny, sample_w_at(Q.pts, kn, o.t), pe.z);
                    Q.qext[2 * (size_t)q + 1] = make_float4(pe.x, pe.y, 0.0f, 0.0f);
                }
                if (MODE != 2 && Q.lens_gauss) {
                    const float e = 2.718281828f;
                    const float gu = powf(e, -((o.u - 0) * (o.u - 0)) / (2 * (0.5f * 0.5f)));
                    const float gv = powf(e, -((o.v - 0) * (o.v - 0)) / (2 * (0.5f * 0.5f)));
                    const float gg = gu * gv;
                    cr *= gg; cg *= gg; cb *= gg; ca *= gg;
                }
                if (MODE != 2 && Q.time_gauss) {
                    const float e = 2.718281828f;
                    const float gg = powf(e, -((o.t - 0.5f) * (o.t - 0.5f)) / (2 * (0.25f * 0.25f)));
                    cr *= gg; cg *= gg; cb *= gg; ca *= gg;
                }
                Q.qcol[q] = make_float4(cr, cg, cb, ca);
            }
        }
    }
    // failures go to the general kernel (one atomic per wave)
    if (RELAY && relay) needs2r = true;
    const bool hand = (active && fail) || relay;
    const uint64_t fm = __ballot(hand);
    if (fm) {
        const int lane = tid & 63;
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(Q.fail_count, (unsigned)__popcll(fm));
        base = (unsigned)__builtin_amdgcn_readfirstlane((int)base);
        if (hand) {
            const unsigned e = base + lanes_below(fm);
            Q.fail_list[e] = q | (needs2r ? kNeeds2R : 0u);
            if (Q.fail_item) Q.fail_item[e] = item;
            if (GATE)
                write_gate_features(Q.features + (size_t)e * kGateFeatures, prov, gacc, g_n, g_nkept, g_nkept0,
                                    g_quad0, g_aw, g_min0, g_nsurf, g_key0, g_key1, g_front, g_front_d2, o, R,
                                    Q.cc0, Q.cc1);
        }
    }
    if (STATS) {
        // finished queries report everything; queries handed over for a 2R
        // fetch report their (complete) 1R pass, the general kernel the rest
        const bool done = active && !fail;
        const bool one = done || (active && needs2r);
        if (Q.qdebug && one) {
            if (done) Q.qdebug[q] = make_float2((float)s_surf, (float)s_kept);
            else Q.qdebug[q].y = (float)s_kept;
        }
        unsigned long long v[6] = {one ? s_leaves : 0ull, one ? s_in_leaves : 0ull, one ? s_kept : 0ull,
                                   done ? s_surf : 0ull, done ? 1ull : 0ull, one ? 1ull : 0ull};
#pragma unroll
        for (int k = 0; k < 6; k++) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) v[k] += __shfl_xor(v[k], d, 64);
        }
        if ((tid & 63) == 0 && v[5]) {
            atomicAdd(&Q.stats[ST_LEAF_NODES], v[0]);
            atomicAdd(&Q.stats[ST_SAMPLES_IN_LEAVES], v[1]);
            atomicAdd(&Q.stats[ST_WITHIN_1R], v[2]);
            atomicAdd(&Q.stats[ST_FETCHED], v[2]);
            atomicAdd(&Q.stats[ST_SURFACES], v[3]);
            atomicAdd(&Q.stats[ST_COUNT], v[4]);
            atomicAdd(&Q.stats[RELAY ? ST_WIDE_LEAVES : ST_FAST_LEAVES], v[0]);
            atomicAdd(&Q.stats[RELAY ? ST_WIDE_IN_LEAVES : ST_FAST_IN_LEAVES], v[1]);
            atomicAdd(&Q.stats[RELAY ? ST_WIDE_KEPT : ST_FAST_KEPT], v[2]);
        }
    }
}

// Query index of fast-kernel lane `slot` (0xFFFFFFFF: no query).  With tile
// order each wave takes one output-sample index over an 8x8 pixel tile, so
// its lanes' (u,v,t) come from one time stratum of the patterns; lanes of
// tiles that overhang the image get no query.  Lists keep their own order.
// Tile order 2 keeps the whole device on one band of tile rows (the band's
// samples stay in the last-level cache) while each XCD (workgroup b runs on
// XCD b % 8) works on its own eighth of the band's columns, so that its L2
// holds a window an eighth of the band's width.  The eighth moves by one
// segment per tile row: a fixed segment per XCD would leave the XCDs with
// the cheap parts of the image idle while the others finish theirs (XCDs
// take their workgroups independently).  Measured at 1920x1080 x 16: the
// L2's fetches fall from 310 to about 110 GB per launch; fixed segments
// are 22% slower than tile order 1, rotated ones 2% faster.
template <int BLOCK>
__device__ __forceinline__ unsigned tile_slot(const QueryParams& Q, unsigned b, unsigned t) {
    const unsigned slot = b * BLOCK + t;
    if (!Q.tile_order || Q.fast_list) return slot;
    const unsigned lane = t & 63u, M = (unsigned)Q.M;
    unsigned tx, ty, i;
    if (Q.tile_order == 2) {
        constexpr unsigned kWpb = BLOCK / 64;
        static_assert(kWpb >= 1 && 16 % kWpb == 0, "a segment's waves are whole workgroups");
        const unsigned seg_tiles = ((((unsigned)Q.W + 7u) / 8u + 15u) & ~15u) / 8u;
        const unsigned seg_blocks = seg_tiles * M / kWpb;
        const unsigned x = b & 7u, k = b >> 3;
        ty = k / seg_blocks;
        const unsigned ws = (k - ty * seg_blocks) * kWpb + (t >> 6);
        tx = ((x + ty) & 7u) * seg_tiles + ws / M;
        i = ws % M;
    } else {
        const unsigned grp = slot >> 6;
        const unsigned tiles_x = ((unsigned)Q.W + 7u) / 8u;
        i = grp % M;
        const unsigned tile = grp / M;
        tx = tile % tiles_x;
        ty = tile / tiles_x;
    }
    const unsigned x = tx * 8u + (lane & 7u), y = ty * 8u + (lane >> 3);
    return (x < (unsigned)Q.W && y < (unsigned)Q.H) ? (y * (unsigned)Q.W + x) * M + i : 0xFFFFFFFFu;
}

template <bool STATS>
__global__ __launch_bounds__(kFastBlock) void k_query_fast(QueryParams Q) {
    // per lane: leaf ids sorted by (key, id); keys are recomputed on compare
    __shared__ int lid[kFastLeafCap][kFastBlock];
    // Queries in scan order (pixel-major): the whole device sweeps one band
    // of rows, whose samples stay resident in the last-level cache.
    fast_query<STATS, kFastLeafCap, kFastBlock>(Q, tile_slot<kFastBlock>(Q, blockIdx.x, threadIdx.x), 0u, lid);
}

// 1R lane kernel for the general kernel's modes 1 and 2 (extended results,
// shadow queries), over the pattern, explicit queries (qin) or a list.
template <bool STATS, int MODE>
__global__ __launch_bounds__(kFastBlock) void k_query_fast_mode(QueryParams Q) {
    __shared__ int lid[kFastLeafCap][kFastBlock];
    fast_query<STATS, kFastLeafCap, kFastBlock, false, false, MODE>(
        Q, tile_slot<kFastBlock>(Q, blockIdx.x, threadIdx.x), 0u, lid);
}

// Second 1R pass over the first pass's hand-off list, for the output samples
// whose 1R leaves did not fit the first pass's lists: the same lane kernel
// with CAP leaves per lane in 64-lane workgroups, persistent (each wave takes
// the next 64 entries until the list is done); entries the first pass marked
// for 2R are passed on unchanged.  Its hand-offs (2R fetches and output
// samples with more than CAP leaves) go to fail_list for the general kernel.
template <bool STATS, int CAP, int MODE>
__global__ __launch_bounds__(64) void k_query_fast_wide(QueryParams Q) {
    __shared__ int lid[CAP][64];
    const unsigned n = *Q.qcount;
    for (;;) {
        unsigned base = 0;
        if (threadIdx.x == 0) base = atomicAdd(Q.next_query, 64u);
        base = (unsigned)uni((int)base);
        if (base >= n) break;
        fast_query<STATS, CAP, 64, false, true, MODE>(Q, base + threadIdx.x, 0u, lid);
    }
}

// Gated fast path (single buffer): as k_query_fast<false>, plus provisional
// colours and gate features for the hand-offs.
__global__ __launch_bounds__(kFastBlock) void k_query_fast_gate(QueryParams Q) {
    __shared__ int lid[kFastLeafCap][kFastBlock];
    fast_query<false, kFastLeafCap, kFastBlock, true>(Q, tile_slot<kFastBlock>(Q, blockIdx.x, threadIdx.x), 0u,
                                                      lid);
}

// Batched fast path: one launch over every buffer of a batch.  Block b of the
// grid is logical block b of the concatenated buffers (each buffer's lanes in
// tile order, padded to whole blocks); its buffer is found by binary search
// over the buffers' first blocks.
template <int CAP, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_query_fast_batch(BatchQuery B) {
    __shared__ int lid[CAP][BLOCK];
    const unsigned b = blockIdx.x;
    int lo = 0, hi = B.n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.block_prefix[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const QueryParams Q = B.items[lo];
    fast_query<false, CAP, BLOCK>(Q, tile_slot<BLOCK>(Q, b - B.block_prefix[lo], threadIdx.x), (unsigned)lo, lid);
}

// Learned-gate decision over the fast kernel's hand-off list: entry e keeps
// its provisional first-surface colour when the gate's logit reaches the
// threshold and the entry has one (feature 0 >= 0); it is then replaced by
// kSkipEntry in place, so the general kernel walks the list in its original
// (spatially coherent) order and skips it.  A NaN logit keeps the entry.
// *accepted counts the skipped entries.
__global__ __launch_bounds__(256) void k_gate_accept(const float* __restrict__ logits,
                                                     const float* __restrict__ features,
                                                     unsigned* __restrict__ entries,
                                                     const unsigned* __restrict__ n_entries, float thr,
                                                     unsigned* __restrict__ accepted) {
    const unsigned n = *n_entries;
    unsigned mine = 0;
    for (unsigned e0 = blockIdx.x * 256; e0 < n; e0 += gridDim.x * 256) {
        const unsigned e = e0 + threadIdx.x;
        bool acc = false;
        if (e < n) {
            // the feature row (one cache line per entry) only where the logit passes
            const float z = logits[e];
            if (z >= thr) acc = features[(size_t)e * kGateFeatures] >= 0.0f;
        }
        if (acc) {
            entries[e] = kSkipEntry;
            mine++;
        }
    }
    // one atomic per block on the shared count (atomics on one address serialise)
    __shared__ unsigned wsum[4];
    unsigned v = mine;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t) atomicAdd(accepted, t);
    }
}

// Provisional colour of every hand-off entry (before the general kernel
// overwrites qcol).
__global__ __launch_bounds__(256) void k_gate_prov(const float4* __restrict__ qcol,
                                                   const unsigned* __restrict__ entries,
                                                   const unsigned* __restrict__ n_entries,
                                                   float4* __restrict__ prov) {
    const unsigned n = *n_entries;
    for (unsigned e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256)
        prov[e] = qcol[entries[e] & ~kNeeds2R];
}

// Training label of every hand-off entry: 1 when it has a provisional colour
// and the plain method's final colour differs from it by at most tau in
// every channel.
__global__ __launch_bounds__(256) void k_gate_label(const float4* __restrict__ qcol,
                                                    const unsigned* __restrict__ entries,
                                                    const unsigned* __restrict__ n_entries,
                                                    const float* __restrict__ features,
                                                    const float4* __restrict__ prov, float tau,
                                                    float* __restrict__ labels) {
    const unsigned n = *n_entries;
    for (unsigned e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
        const float4 a = qcol[entries[e] & ~kNeeds2R], b = prov[e];
        const float d = fmaxf(fmaxf(fabsf(a.x - b.x), fabsf(a.y - b.y)), fmaxf(fabsf(a.z - b.z), fabsf(a.w - b.w)));
        labels[e] = (features[(size_t)e * kGateFeatures] >= 0.0f && d <= tau) ? 1.0f : 0.0f;
    }
}

// debugImage of reconstructDofMotion (Reconstruction.cpp:276): a pixel holds
// its last output sample's (#surfaces, #samples within 1R, 0, 0)
// (ReconstructionOur.cpp:312-314; every output sample overwrites it).
__global__ __launch_bounds__(256) void k_debug_image(const float2* __restrict__ qdebug, int W, int H, int M,
                                                     float4* __restrict__ out) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= W * H) return;
    const float2 d = qdebug[(size_t)p * M + (M - 1)];
    out[p] = make_float4(d.x, d.y, 0.0f, 0.0f);
}

// The gather of listed output samples (collectInputSamples2 at radius R,
// ReconstructionOur.cpp:412-511), as the general kernel runs it: for output
// sample j, its leaves front to back (leaf_cap per row), and its gathered
// samples (tree layout index, surface) in gather order (sample_cap per row).
// Counts are -1 when the gather exceeded the kernel's capacities.
__global__ __launch_bounds__(kBlock) void k_query_gather(QueryParams Q, const float* __restrict__ o5, int n, float R,
                                                         int leaf_cap, int* __restrict__ leaves,
                                                         int* __restrict__ n_leaves, int sample_cap,
                                                         int* __restrict__ samples, int* __restrict__ surface_of,
                                                         int* __restrict__ n_samples) {
    const int lane = threadIdx.x & 63;
    // wave-uniform (so the scratch pointers derived from it live in scalar registers)
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    const int waves = gridDim.x * kWavesPerBlock;
    const WaveMem m = wave_mem(Q.scratch + (size_t)wave * kWaveScratchWords);
    unsigned long long st[kNumStats];
#pragma unroll
    for (int k = 0; k < kNumStats; k++) st[k] = 0;
    for (int j = wave; j < n; j += waves) {
        OSample o;
        o.x = o5[5 * j]; o.y = o5[5 * j + 1]; o.u = o5[5 * j + 2]; o.v = o5[5 * j + 3]; o.t = o5[5 * j + 4];
        const Gather g = gather<false>(Q, m, o, R, st);
        if (g.overflow) {
            if (lane == 0) n_leaves[j] = n_samples[j] = -1;
            continue;
        }
        const int* Sid = reinterpret_cast<const int*>(wave_lds()) + 2 * kStage;
        for (int li = lane; li < g.nl && li < leaf_cap; li += 64)
            leaves[(size_t)j * leaf_cap + li] = li < kStage ? Sid[li] : m.sid()[li];
        for (int si = 0; si < g.nsurf; si++) {
            const int fs = uni(m.f_start()[si]), fn = uni(m.f_size()[si]);
            for (int k = lane; k < fn; k += 64)
                if (fs + k < sample_cap) {
                    samples[(size_t)j * sample_cap + fs + k] = m.sidx()[fs + k];
                    surface_of[(size_t)j * sample_cap + fs + k] = si;
                }
        }
        if (lane == 0) {
            n_leaves[j] = g.nl;
            n_samples[j] = g.total;
        }
        // the next gather rewrites the wave's LDS area
        __builtin_amdgcn_wave_barrier();
        wave_sync_mem();
    }
}

// Pixel = sum of its output samples in order, normalised by alpha; red if
// nothing was found (Reconstruction.cpp:665-719).
// Per pixel: the sum of its M output samples in sample order, normalised by
// the weight (red where it is 0).  When M divides 64, one thread per output
// sample: a wave loads 64 consecutive output samples (64 / M pixels) in one
// coalesced access and every lane sums its pixel's samples, in order, from
// the lanes that hold them; otherwise one thread per pixel.
__global__ __launch_bounds__(kBlock) void k_combine(const float4* __restrict__ qcol, int W, int H, int M,
                                                    float* __restrict__ out) {
    const int npix = W * H;
    float r = 0, g = 0, b = 0, a = 0;
    int p;
    if (M <= 64 && (64 % M) == 0) {
        const size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x;
        const bool in = k < (size_t)npix * (size_t)M;
        const float4 c = in ? qcol[k] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int lane = threadIdx.x & 63, first = lane & ~(M - 1);
        for (int i = 0; i < M; i++) {
            r += __shfl(c.x, first + i, 64);
            g += __shfl(c.y, first + i, 64);
            b += __shfl(c.z, first + i, 64);
            a += __shfl(c.w, first + i, 64);
        }
        if (!in || lane != first) return;
        p = (int)(k / (size_t)M);
    } else {
        p = blockIdx.x * kBlock + threadIdx.x;
        if (p >= npix) return;
        for (int i = 0; i < M; i++) {
            const float4 c = qcol[(size_t)p * M + i];
            r += c.x; g += c.y; b += c.z; a += c.w;
        }
    }
    if (a == 0) {
        r = 1; g = 0; b = 0; a = 1;
    }
    const float rw = rcp(a);
    float4 o = make_float4(r * rw, g * rw, b * rw, a * rw);
    reinterpret_cast<float4*>(out)[p] = o;
}

// Counter calibration kernels: a streaming copy and gathers of 32-byte and
// 64-byte records in a fixed scattered order (i * odd mod n, a bijection for
// n a power of two).
__global__ __launch_bounds__(256) void k_calib_stream(const float4* __restrict__ src, float4* __restrict__ dst,
                                                      size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <int REC>
__global__ __launch_bounds__(256) void k_calib_gather(const float4* __restrict__ src, float4* __restrict__ dst,
                                                      size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t j = (i * (size_t)0x9E3779B1u) & (n - 1);
    const float4* r = src + (size_t)REC * j;
    float4 acc = r[0];
#pragma unroll
    for (int e = 1; e < REC; e++) {
        const float4 v = r[e];
        acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
    }
    dst[i] = acc;
}

// Graph-safe clear of small counters (no memset nodes in captured graphs).
__global__ void k_fill_u32(unsigned* __restrict__ p, int n, unsigned v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// dst[r] = src[r * stride], rows of `width` 32-bit words.
__global__ void k_take_rows(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t rows, int width,
                            int stride) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= rows * (size_t)width) return;
    const size_t r = i / (size_t)width, c = i - r * (size_t)width;
    dst[i] = src[r * (size_t)stride * width + c];
}

}  // namespace

int general_waves_per_cu() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_query_general<false, 0, false>, kBlock, 0) != hipSuccess ||
        blocks < 1)
        blocks = 1;
    return blocks * kWavesPerBlock;
}

hipError_t launch_query_general(const QueryParams& p, hipStream_t s) {
    if (p.num_queries == 0) return hipSuccess;
    const int blocks = p.num_waves / kWavesPerBlock;
    if (p.mode == 1 && p.collect_stats)
        k_query_general<false, 1, true><<<blocks, kBlock, 0, s>>>(p);
    else if (p.mode == 1)
        k_query_general<false, 1, false><<<blocks, kBlock, 0, s>>>(p);
    else if (p.mode == 2 && p.collect_stats)
        k_query_general<false, 2, true><<<blocks, kBlock, 0, s>>>(p);
    else if (p.mode == 2)
        k_query_general<false, 2, false><<<blocks, kBlock, 0, s>>>(p);
    else if (p.qdebug)
        k_query_general<true, 0, true><<<blocks, kBlock, 0, s>>>(p);
    else if (p.collect_stats)
        k_query_general<false, 0, true><<<blocks, kBlock, 0, s>>>(p);
    else
        k_query_general<false, 0, false><<<blocks, kBlock, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t launch_query_fast(const QueryParams& p, bool stats, hipStream_t s) {
    if (p.num_queries == 0) return hipSuccess;
    // list mode launches for the worst case; surplus lanes exit at once
    const size_t slots = p.fast_list ? p.num_queries : fast_slots(p);
    const unsigned blocks = (unsigned)((slots + kFastBlock - 1) / kFastBlock);
    if (stats)
        k_query_fast<true><<<blocks, kFastBlock, 0, s>>>(p);
    else
        k_query_fast<false><<<blocks, kFastBlock, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t launch_query_fast_mode(const QueryParams& p, bool stats, hipStream_t s) {
    if (p.num_queries == 0) return hipSuccess;
    const size_t slots = p.fast_list ? p.num_queries : fast_slots(p);
    const unsigned blocks = (unsigned)((slots + kFastBlock - 1) / kFastBlock);
    if (p.mode == 2) {
        if (stats)
            k_query_fast_mode<true, 2><<<blocks, kFastBlock, 0, s>>>(p);
        else
            k_query_fast_mode<false, 2><<<blocks, kFastBlock, 0, s>>>(p);
    } else {
        if (stats)
            k_query_fast_mode<true, 1><<<blocks, kFastBlock, 0, s>>>(p);
        else
            k_query_fast_mode<false, 1><<<blocks, kFastBlock, 0, s>>>(p);
    }
    return hipGetLastError();
}

hipError_t launch_query_fast_wide(const QueryParams& p, bool stats, int blocks, hipStream_t s) {
    if (p.num_queries == 0) return hipSuccess;
    switch (p.mode * 2 + (stats ? 1 : 0)) {
        case 0: k_query_fast_wide<false, kWideLeafCap, 0><<<blocks, 64, 0, s>>>(p); break;
        case 1: k_query_fast_wide<true, kWideLeafCap, 0><<<blocks, 64, 0, s>>>(p); break;
        case 2: k_query_fast_wide<false, kWideLeafCap, 1><<<blocks, 64, 0, s>>>(p); break;
        case 3: k_query_fast_wide<true, kWideLeafCap, 1><<<blocks, 64, 0, s>>>(p); break;
        case 4: k_query_fast_wide<false, kWideLeafCap, 2><<<blocks, 64, 0, s>>>(p); break;
        default: k_query_fast_wide<true, kWideLeafCap, 2><<<blocks, 64, 0, s>>>(p); break;
    }
    return hipGetLastError();
}

int fast_wide_blocks_per_cu() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_query_fast_wide<false, kWideLeafCap, 0>, 64, 0) !=
            hipSuccess ||
        b <= 0)
        b = 1;
    return b;
}

hipError_t launch_query_fast_gate(const QueryParams& p, hipStream_t s) {
    if (p.num_queries == 0) return hipSuccess;
    const size_t slots = fast_slots(p);
    k_query_fast_gate<<<(unsigned)((slots + kFastBlock - 1) / kFastBlock), kFastBlock, 0, s>>>(p);
    return hipGetLastError();
}

// Launches over a device-side entry count use a grid that covers max_n
// entries, capped; the kernels stride over the rest.
static unsigned entry_grid(size_t max_n) {
    const size_t b = (max_n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

hipError_t launch_gate_accept(const float* logits, const float* features, unsigned* entries,
                              const unsigned* n_entries, size_t max_n, float thr, unsigned* accepted,
                              hipStream_t s) {
    if (max_n == 0) return hipSuccess;
    k_gate_accept<<<entry_grid(max_n), 256, 0, s>>>(logits, features, entries, n_entries, thr, accepted);
    return hipGetLastError();
}

hipError_t launch_gate_prov(const float4* qcol, const unsigned* entries, const unsigned* n_entries, size_t max_n,
                            float4* prov, hipStream_t s) {
    if (max_n == 0) return hipSuccess;
    k_gate_prov<<<entry_grid(max_n), 256, 0, s>>>(qcol, entries, n_entries, prov);
    return hipGetLastError();
}

hipError_t launch_gate_label(const float4* qcol, const unsigned* entries, const unsigned* n_entries, size_t max_n,
                             const float* features, const float4* prov, float tau, float* labels, hipStream_t s) {
    if (max_n == 0) return hipSuccess;
    k_gate_label<<<entry_grid(max_n), 256, 0, s>>>(qcol, entries, n_entries, features, prov, tau, labels);
    return hipGetLastError();
}

hipError_t launch_debug_image(const float2* qdebug, int W, int H, int M, float* out, hipStream_t s) {
    const int n = W * H;
    if (n == 0) return hipSuccess;
    k_debug_image<<<(n + 255) / 256, 256, 0, s>>>(qdebug, W, H, M, reinterpret_cast<float4*>(out));
    return hipGetLastError();
}

hipError_t launch_query_gather(const QueryParams& p, const float* o5, int n, float R, int leaf_cap, int* leaves,
                               int* n_leaves, int sample_cap, int* samples, int* surface_of, int* n_samples,
                               hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int blocks = p.num_waves / kWavesPerBlock;
    k_query_gather<<<blocks, kBlock, 0, s>>>(p, o5, n, R, leaf_cap, leaves, n_leaves, sample_cap, samples,
                                             surface_of, n_samples);
    return hipGetLastError();
}

hipError_t launch_combine(const float4* qcol, int W, int H, int M, float* out, hipStream_t s) {
    const int n = W * H;
    if (n == 0) return hipSuccess;
    const size_t threads = (M <= 64 && (64 % M) == 0) ? (size_t)n * M : (size_t)n;
    k_combine<<<(unsigned)((threads + kBlock - 1) / kBlock), kBlock, 0, s>>>(qcol, W, H, M, out);
    return hipGetLastError();
}


__global__ void k_stamp(unsigned long long* dst) {
    if (threadIdx.x == 0) *dst = wall_clock64();
}

hipError_t launch_stamp(unsigned long long* dst, hipStream_t s) {
    k_stamp<<<1, 64, 0, s>>>(dst);
    return hipGetLastError();
}

hipError_t launch_calibrate(int kind, const float4* src, float4* dst, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (kind == 0)
        k_calib_stream<<<blocks, 256, 0, s>>>(src, dst, n);
    else if (kind == 1)
        k_calib_gather<2><<<blocks, 256, 0, s>>>(src, dst, n);
    else
        k_calib_gather<4><<<blocks, 256, 0, s>>>(src, dst, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned* p, int n, unsigned v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_fill_u32<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

// Fast-kernel variants the batch autotuner chooses from: (leaf cap, block).
// A cap of 22 lets the lists of 7 waves per SIMD fit the CU's 160 KB of LDS.
static const int kFastVariants[kNumFastVariants][2] = {{22, 128}, {16, 128}, {22, 64}, {16, 256}};

int fast_variant_cap(int v) { return kFastVariants[v][0]; }
int fast_variant_block(int v) { return kFastVariants[v][1]; }

hipError_t launch_query_fast_batch(const BatchQuery& b, int variant, hipStream_t s) {
    const unsigned grid = b.total_blocks;
    if (grid == 0) return hipSuccess;
    switch (variant) {
        case 0: k_query_fast_batch<22, 128><<<grid, 128, 0, s>>>(b); break;
        case 1: k_query_fast_batch<16, 128><<<grid, 128, 0, s>>>(b); break;
        case 2: k_query_fast_batch<22, 64><<<grid, 64, 0, s>>>(b); break;
        default: k_query_fast_batch<16, 256><<<grid, 256, 0, s>>>(b); break;
    }
    return hipGetLastError();
}

hipError_t launch_query_general_batch(const BatchQuery& b, int blocks, hipStream_t s) {
    k_query_general_batch<<<blocks, kBlock, 0, s>>>(b);
    return hipGetLastError();
}

hipError_t launch_take_rows(const void* src, void* dst, size_t rows, int width_words, int stride, hipStream_t s) {
    const size_t n = rows * (size_t)width_words;
    if (n == 0) return hipSuccess;
    k_take_rows<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(static_cast<const uint32_t*>(src),
                                                          static_cast<uint32_t*>(dst), rows, width_words, stride);
    return hipGetLastError();
}

}  // namespace tlr
