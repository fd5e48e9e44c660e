// Tree build kernels: reprojection to the (u,v,t) centre, bucketing grid,
// initial kd split, per-cell sort / bounds / surface scan, leaves, hierarchy.
//
// Stage <-> reference:
//   k_reproject, k_hits_max, k_grid_setup, k_cell_keys
//       TreeGather::reprojectToUVTCenter  Reconstruction.cpp:514-640
//       Sample::reprojectToUVTCenter      Reconstruction.hpp:93-121
//   k_sat_rows/cols, k_kd_descend, k_emit_initial, k_place
//       buildInitialRecursive             ReconstructionTreeBuilder.cpp:34-96
//       candidate collection              ReconstructionTreeBuilder.cpp:133-140
//   k_leaf_sort                           ReconstructionTreeBuilder.cpp:153 (FW::sort,
//                                          src/framework/base/Sort.cpp:63-211)
//   k_points                              ReconstructionTreeBuilder.cpp:170-213
//   k_surface_scan                        ReconstructionTreeBuilder.cpp:222-340
//   k_leaf_emit                           ReconstructionTreeBuilder.cpp:345-378
//   k_kd_*, k_tree_trav                   median-split hierarchy; replaces emitNodes
//                                          (:409-451) and the task merge
//                                          (Reconstruction.cpp:110-203)
#include "build.hpp"
#include "prims.hpp"
#include <algorithm>
#include <utility>

namespace tlr {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ unsigned lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        T o = __shfl_xor(v, d, 64);
        v = o < v ? o : v;
    }
    return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        T o = __shfl_xor(v, d, 64);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ unsigned wave_sum(unsigned v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// ---------------------------------------------------------------------------
// Reprojection: grid-stride over the input samples; the bbox and counters are
// reduced per block so each block issues one set of global atomics.
// ---------------------------------------------------------------------------
constexpr int kReduceBlocks = 2048;

// Besides the per-field arrays, every kept sample gets one packed 48-byte
// record (x, y, w, t | mv_x, mv_y, mv_w, r | g, b, 0, 0) that the tree-order
// gathers (k_place, k_points) read in one or two cache lines.
__global__ __launch_bounds__(kBlock) void k_reproject(InSoA in, int n, int W, int H, float C0, float C1,
                                                      float* __restrict__ rx, float* __restrict__ ry,
                                                      float* __restrict__ rw, float* __restrict__ rt,
                                                      float4* __restrict__ packed, int* __restrict__ hits,
                                                      GridInfo* gi) {
    int kmin_x = 0x7FFFFFFF, kmin_y = 0x7FFFFFFF;
    int kmax_x = (int)0x80000000, kmax_y = (int)0x80000000;
    unsigned nok = 0, nbad = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        bool ok = false;
        const float x = in.x[i], y = in.y[i], w = in.w[i], u = in.u[i], v = in.v[i], t = in.t[i];
        const float mvx = in.mvx[i], mvy = in.mvy[i], mvw = in.mvw[i];
        const float coc = C0 / w + C1;
        const float px = x - coc * u;
        const float py = y - coc * v;
        const float Px = px * w, Py = py * w, Pw = 1.0f * w;
        const float ax = Px - t * mvx, ay = Py - t * mvy, aw = Pw - t * mvw;
        float nt = 0.5f;
        float X = ax + nt * mvx, Y = ay + nt * mvy, Wt = aw + nt * mvw;
        if (Wt <= 0.f) {
            nt = 0.0f;
            X = ax + nt * mvx; Y = ay + nt * mvy; Wt = aw + nt * mvw;
        }
        if (Wt <= 0.f) {
            nt = 1.0f;
            X = ax + nt * mvx; Y = ay + nt * mvy; Wt = aw + nt * mvw;
        }
        if (Wt > 0.f) {
            const float c = rcp(Wt);
            const float ox = X * c;
            const float oy = Y * c;
            if (isfinite(ox) && isfinite(oy)) {
                ok = true;
                rx[i] = ox; ry[i] = oy; rw[i] = Wt; rt[i] = nt;
                packed[3 * (size_t)i] = make_float4(ox, oy, Wt, nt);
                packed[3 * (size_t)i + 1] = make_float4(mvx, mvy, mvw, in.r[i]);
                packed[3 * (size_t)i + 2] = make_float4(in.g[i], in.b[i], 0.0f, 0.0f);
                const int sx = floor_i(ox), sy = floor_i(oy);
                if (sx >= 0 && sy >= 0 && sx < W && sy < H) atomicAdd(&hits[sy * W + sx], 1);
                kmin_x = min(kmin_x, order_key(ox));
                kmin_y = min(kmin_y, order_key(oy));
                kmax_x = max(kmax_x, order_key(ox));
                kmax_y = max(kmax_y, order_key(oy));
                nok++;
            } else {
                nbad++;
            }
        }
        if (!ok) rw[i] = 0.0f;  // marks a discarded sample
    }
    kmin_x = wave_min(kmin_x); kmin_y = wave_min(kmin_y);
    kmax_x = wave_max(kmax_x); kmax_y = wave_max(kmax_y);
    nok = wave_sum(nok);
    nbad = wave_sum(nbad);
    __shared__ int red[kBlock / 64][6];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = kmin_x; red[wv][1] = kmin_y; red[wv][2] = kmax_x; red[wv][3] = kmax_y;
        red[wv][4] = (int)nok; red[wv][5] = (int)nbad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; k++) {
            kmin_x = min(kmin_x, red[k][0]); kmin_y = min(kmin_y, red[k][1]);
            kmax_x = max(kmax_x, red[k][2]); kmax_y = max(kmax_y, red[k][3]);
            nok += (unsigned)red[k][4]; nbad += (unsigned)red[k][5];
        }
        if (nok) {
            atomicMin(&gi->bbmin_key[0], kmin_x);
            atomicMin(&gi->bbmin_key[1], kmin_y);
            atomicMax(&gi->bbmax_key[0], kmax_x);
            atomicMax(&gi->bbmax_key[1], kmax_y);
            atomicAdd(&gi->num_kept, nok);
        }
        if (nbad) atomicAdd(&gi->num_bad, nbad);
    }
}

// Largest per-pixel hit count: grid-stride, reduced per block, one atomic per
// block (atomics on one address serialise).
__global__ __launch_bounds__(kBlock) void k_hits_max(const int* __restrict__ hits, int n, GridInfo* gi) {
    int m = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) m = max(m, hits[i]);
    m = wave_max(m);
    __shared__ int red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; k++) m = max(m, red[k]);
        if (m > 0) atomicMax(&gi->max_hits, m);
    }
}

// Grid sizing: Reconstruction.cpp:587-603
__global__ void k_grid_setup(GridInfo* gi) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    gi->error = 0;
    if (gi->num_bad) {
        gi->error = 1;  // TLR_E_INVALID: non-finite reprojected position
        return;
    }
    const int mh = gi->max_hits;
    if (mh == 0) {
        gi->error = 3;  // TLR_E_ALL_DISCARDED
        return;
    }
    const float bx0 = order_val(gi->bbmin_key[0]), by0 = order_val(gi->bbmin_key[1]);
    const float bx1 = order_val(gi->bbmax_key[0]), by1 = order_val(gi->bbmax_key[1]);
    gi->bbmin[0] = bx0; gi->bbmin[1] = by0;
    gi->bbmax[0] = bx1; gi->bbmax[1] = by1;
    float scale = sqrtf(1.f * mh / 48);
    const float ext = fmx(bx1 - bx0, by1 - by0);
    if (scale * ext > 4096) scale = 4096 / ext;
    gi->scale = scale;
    const int ix0 = floor_i(bx0 * scale), iy0 = floor_i(by0 * scale);
    const int ix1 = floor_i(bx1 * scale) + 1, iy1 = floor_i(by1 * scale) + 1;
    gi->bbmin_i[0] = ix0;
    gi->bbmin_i[1] = iy0;
    gi->gw = ix1 - ix0;
    gi->gh = iy1 - iy0;
}

// Bucket index of each kept sample (grid-stride, one num_bucketed atomic
// per block); the cells' counts come from the sorted keys (k_cell_runs).
__global__ __launch_bounds__(kBlock) void k_cell_keys(const float* __restrict__ rx, const float* __restrict__ ry,
                                                      const float* __restrict__ rw, int n, GridInfo* gi,
                                                      uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const float scale = gi->scale;
    const int bx = gi->bbmin_i[0], by = gi->bbmin_i[1], gw = gi->gw, gh = gi->gh;
    unsigned nin = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        uint32_t key = 0xFFFFFFFFu;
        if (rw[i] > 0.0f) {
            const int sx = floor_i(scale * rx[i]) - bx;
            const int sy = floor_i(scale * ry[i]) - by;
            if (!(sx < 0 || sy < 0 || sx >= gw || sy >= gh)) {
                key = (uint32_t)(sy * gw + sx);
                nin++;
            }
        }
        keys[i] = key;
        vals[i] = (uint32_t)i;
    }
    nin = wave_sum(nin);
    __shared__ unsigned red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nin;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; k++) nin += red[k];
        if (nin) atomicAdd(&gi->num_bucketed, nin);
    }
}

// Each non-empty cell's run in the sorted keys: start[cell] = its first
// position, end[cell] = one past its last (both zero for an empty cell, from
// the memsets before).  Keys of samples outside the grid sort after every
// cell and are skipped.
__global__ __launch_bounds__(kBlock) void k_cell_runs(const uint32_t* __restrict__ skeys, int n, uint32_t cells,
                                                      uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = skeys[i];
    if (k >= cells) return;
    if (i == 0 || skeys[i - 1] != k) start[k] = (uint32_t)i;
    if (i == n - 1 || skeys[i + 1] != k) end[k] = (uint32_t)i + 1;
}

// Summed-area table over the cell counts (end - start of each cell's run),
// (gh+1) x (gw+1), zero first row/col.  Rows: one block per grid row.
__global__ __launch_bounds__(kBlock) void k_sat_rows(const uint32_t* __restrict__ cstart,
                                                     const uint32_t* __restrict__ cend, int gw, int gh,
                                                     uint32_t* __restrict__ sat) {
    const int y = blockIdx.x;
    __shared__ uint32_t carry;
    __shared__ uint32_t lds4[4];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    uint32_t* row = sat + (size_t)(y + 1) * (gw + 1);
    if (threadIdx.x == 0) row[0] = 0;
    for (int base = 0; base < gw; base += kBlock) {
        const int x = base + threadIdx.x;
        uint32_t v = (x < gw) ? cend[(size_t)y * gw + x] - cstart[(size_t)y * gw + x] : 0u;
        // inclusive block scan
        uint32_t inc = v;
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) lds4[w] = inc;
        __syncthreads();
        uint32_t off = carry;
        for (int k = 0; k < w; k++) off += lds4[k];
        if (x < gw) row[x + 1] = off + inc;
        __syncthreads();
        if (threadIdx.x == kBlock - 1) carry = off + inc;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_sat_cols(uint32_t* __restrict__ sat, int gw, int gh) {
    const int x = blockIdx.x * kBlock + threadIdx.x;
    if (x > gw) return;
    sat[x] = 0;
    uint32_t run = 0;
    for (int y = 1; y <= gh; y++) {
        uint32_t* p = sat + (size_t)y * (gw + 1) + x;
        run += *p;
        *p = run;
    }
}

__device__ __forceinline__ uint32_t rect_count(const uint32_t* __restrict__ sat, int gw, int x0, int y0, int x1,
                                               int y1) {
    const size_t s = (size_t)gw + 1;
    return sat[(size_t)y1 * s + x1] - sat[(size_t)y0 * s + x1] - sat[(size_t)y1 * s + x0] +
           sat[(size_t)y0 * s + x0];
}

// Each cell walks the kd split (buildInitialRecursive) from the root down to
// the initial leaf containing it, accumulating the sample offset of the leaf
// in depth-first order.  The leaf's top-left cell owns it.
__global__ __launch_bounds__(kBlock) void k_kd_descend(const uint32_t* __restrict__ sat, int gw, int gh,
                                                       int4* __restrict__ owner, uint32_t* __restrict__ own_flag,
                                                       int* __restrict__ own_y1) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= gw * gh) return;
    const int cx = c % gw, cy = c / gw;
    int x0 = 0, x1 = gw, y0 = 0, y1 = gh;
    uint32_t off = 0;
    for (;;) {
        const int dx = x1 - x0, dy = y1 - y0;
        bool leaf = max(dx, dy) == 1;
        if (!leaf && dx * dy <= 64) leaf = rect_count(sat, gw, x0, y0, x1, y1) <= 48u;
        if (leaf) break;
        if (dx > dy) {
            const int xm = (x1 + x0) / 2;
            if (cx < xm) {
                x1 = xm;
            } else {
                off += rect_count(sat, gw, x0, y0, xm, y1);
                x0 = xm;
            }
        } else {
            const int ym = (y1 + y0) / 2;
            if (cy < ym) {
                y1 = ym;
            } else {
                off += rect_count(sat, gw, x0, y0, x1, ym);
                y0 = ym;
            }
        }
    }
    owner[c] = make_int4(x0, y0, x1, (int)off);
    const bool own = (cx == x0 && cy == y0);
    own_flag[c] = own ? 1u : 0u;
    if (own) own_y1[c] = y1;
}

// Initial leaf records from owner cells: x0,y0,x1,y1 | offset,count.
__global__ __launch_bounds__(kBlock) void k_emit_initial(const int4* __restrict__ owner,
                                                         const uint32_t* __restrict__ own_flag,
                                                         const int* __restrict__ own_y1,
                                                         const uint32_t* __restrict__ idx,
                                                         const uint32_t* __restrict__ sat, int gw, int gh,
                                                         int4* __restrict__ rect, int2* __restrict__ offcnt) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= gw * gh) return;
    if (!own_flag[c]) return;
    const int4 o = owner[c];
    const int y1 = own_y1[c];
    const uint32_t k = idx[c];
    rect[k] = make_int4(o.x, o.y, o.z, y1);
    offcnt[k] = make_int2(o.w, (int)rect_count(sat, gw, o.x, o.y, o.z, y1));
}

// Place each bucketed sample at its candidate slot: initial-leaf offset +
// samples in the leaf's cells before it in row-major order + rank in cell.
__global__ __launch_bounds__(kBlock) void k_place(const uint32_t* __restrict__ skeys,
                                                  const uint32_t* __restrict__ svals, int nb, int gw,
                                                  const int4* __restrict__ owner,
                                                  const uint32_t* __restrict__ sat,
                                                  const uint32_t* __restrict__ cell_start,
                                                  const float4* __restrict__ packed,
                                                  uint32_t* __restrict__ cand, float* __restrict__ ct,
                                                  float* __restrict__ cw) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= nb) return;
    const uint32_t key = skeys[s];
    const uint32_t v = svals[s];
    const int cx = (int)(key % (uint32_t)gw), cy = (int)(key / (uint32_t)gw);
    const int4 o = owner[key];
    uint32_t pos = (uint32_t)o.w;
    pos += rect_count(sat, gw, o.x, o.y, o.z, cy);   // full rows above
    pos += rect_count(sat, gw, o.x, cy, cx, cy + 1);  // cells to the left
    pos += (uint32_t)s - cell_start[key];
    const float4 a = packed[3 * (size_t)v];   // x, y, w, t
    cand[pos] = v;
    ct[pos] = a.w;
    cw[pos] = a.z;
}

// ---------------------------------------------------------------------------
// FW::sort replica (src/framework/base/Sort.cpp:63-211) on (t, w) keys of one
// initial leaf; sampleCompareFuncInc at ReconstructionTreeBuilder.cpp:99-117.
// ---------------------------------------------------------------------------
struct CandView {
    uint32_t* idx;
    float* t;
    float* w;
    __device__ __forceinline__ int cmp(int a, int b) const {
        const float ta = t[a], tb = t[b];
        if (ta < tb) return -1;
        if (ta > tb) return 1;
        const float wa = w[a], wb = w[b];
        if (wa < wb) return -1;
        if (wa > wb) return 1;
        return 0;
    }
    __device__ __forceinline__ void swp(int a, int b) const {
        uint32_t i0 = idx[a]; idx[a] = idx[b]; idx[b] = i0;
        float t0 = t[a]; t[a] = t[b]; t[b] = t0;
        float w0 = w[a]; w[a] = w[b]; w[b] = w0;
    }
};

// The same order on one 64-bit key per candidate: t's bits above w's.  For
// t in [0, 1] and w > 0 (a kept sample's reprojection time and depth) the
// bit patterns order as the values do, so comparing keys is the (t, w)
// comparison above, with half the loads and one array less to swap.
struct CandKey64 {
    uint32_t* idx;
    uint64_t* key;
    __device__ __forceinline__ int cmp(int a, int b) const {
        const uint64_t ka = key[a], kb = key[b];
        return ka < kb ? -1 : (ka > kb ? 1 : 0);
    }
    __device__ __forceinline__ void swp(int a, int b) const {
        uint32_t i0 = idx[a]; idx[a] = idx[b]; idx[b] = i0;
        uint64_t k0 = key[a]; key[a] = key[b]; key[b] = k0;
    }
};

template <class V>
__device__ void fwsort_insertion(const V& v, int start, int size) {
    for (int i = 1; i < size; i++) {
        int j = start + i - 1;
        while (j >= start && v.cmp(j, j + 1) > 0) {
            v.swp(j, j + 1);
            j--;
        }
    }
}

template <class V>
__device__ int fwsort_partition(const V& v, int low, int high) {
    int l = low, c = (low + high) >> 1, h = high - 2;
    if (v.cmp(l, h) > 0) {
        int tmp = l; l = h; h = tmp;
    }
    if (v.cmp(l, c) > 0) c = l;
    const int piv = (v.cmp(c, h) > 0) ? h : c;
    v.swp(piv, high - 1);
    int i = low - 1, j = high - 1;
    for (;;) {
        do i++; while (v.cmp(i, high - 1) < 0);
        do j--; while (v.cmp(j, high - 1) > 0);
        if (i >= j) break;
        v.swp(i, j);
    }
    v.swp(i, high - 1);
    return i;
}

template <class V>
__device__ void fwsort(const V& v, int start, int end) {
    if (end - start < 2) return;
    int low = start, high = end;
    int stack[32];
    int sp = 0;
    stack[sp++] = high;
    while (sp) {
        high = stack[--sp];
        if (high - low < 16 || sp + 2 > 32) {
            fwsort_insertion(v, low, high - low);
            low = high + 1;
            continue;
        }
        const int i = fwsort_partition(v, low, high);
        if (high - i > 2) stack[sp++] = high;
        if (i - low > 1) stack[sp++] = i;
        else low = i + 1;
    }
}

// One wave of 64 initial leaves (one thread each).  The wave copies its
// leaves into LDS one leaf at a time (each copy a contiguous, coalesced
// range; (t, w) as one 64-bit key), every thread runs the sort on its own
// leaf there, and the sorted input indices go back the same way; the sorted
// keys are not needed afterwards.  A wave whose leaves hold more than
// kSortLds candidates sorts in global memory instead.  Either way the
// permutation is FW::sort's.
constexpr int kSortBlock = 64;
constexpr int kSortLds = 2304;

__global__ __launch_bounds__(kSortBlock) void k_leaf_sort(const int2* __restrict__ offcnt, int nil, uint32_t* cand,
                                                          float* ct, float* cw) {
    __shared__ uint64_t s_key[kSortLds];
    __shared__ uint32_t s_idx[kSortLds];
    const int lane = threadIdx.x;
    const int k = blockIdx.x * kSortBlock + lane;
    const int2 oc = k < nil ? offcnt[k] : make_int2(0, 0);
    // exclusive prefix of the leaf sizes across the wave
    int incl = oc.y;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    const int base = incl - oc.y;
    const int total = __shfl(incl, 63, 64);
    if (total > kSortLds) {
        if (k < nil) fwsort(CandView{cand + oc.x, ct + oc.x, cw + oc.x}, 0, oc.y);
        return;
    }
    for (int j = 0; j < kSortBlock; j++) {
        const int off = __shfl(oc.x, j, 64), n = __shfl(oc.y, j, 64), b = __shfl(base, j, 64);
        for (int e = lane; e < n; e += 64) {
            s_idx[b + e] = cand[off + e];
            s_key[b + e] = ((uint64_t)__float_as_uint(ct[off + e]) << 32) | (uint64_t)__float_as_uint(cw[off + e]);
        }
    }
    __syncthreads();
    if (k < nil) fwsort(CandKey64{s_idx + base, s_key + base}, 0, oc.y);
    __syncthreads();
    for (int j = 0; j < kSortBlock; j++) {
        const int off = __shfl(oc.x, j, 64), n = __shfl(oc.y, j, 64), b = __shfl(base, j, 64);
        for (int e = lane; e < n; e += 64) cand[off + e] = s_idx[b + e];
    }
}

// ---------------------------------------------------------------------------
// Per-sample tree data in layout order: points and t=0/t=1 boxes
// (ReconstructionTreeBuilder.cpp:170-213).  Boxes of samples with an empty
// time span are zero.  A sample's bounds: TimeLensBounds from xywt0 =
// (xy*w, w) - t*mv and mv (:174-175).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sample_bounds(const float4 a, const float4 m, float W, float H, float cc0, float cc1,
                                              Bounds& b) {
    const float x = a.x, y = a.y, w = a.z, t = a.w;
    const float px = x * w - t * m.x, py = y * w - t * m.y, pw = w - t * m.z;
    tux_bounds(px, py, pw, m.x, m.y, m.z, W, H, cc0, cc1, b);
}

__global__ __launch_bounds__(kBlock) void k_points(const uint32_t* __restrict__ cand, int nb,
                                                   const float4* __restrict__ packed, float W, float H, float cc0,
                                                   float cc1, Points pts, Bounds* __restrict__ bounds,
                                                   float4* __restrict__ boxes) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nb) return;
    const uint32_t i = cand[k];
    const float4 r0 = packed[3 * (size_t)i], r1 = packed[3 * (size_t)i + 1], r2 = packed[3 * (size_t)i + 2];
    const float x = r0.x, y = r0.y, w = r0.z, t = r0.w;
    const float mvx = r1.x, mvy = r1.y, mvw = r1.z;
    const float4 pa = make_float4(x, y, w, t), pm = make_float4(mvx, mvy, mvw, 0.0f);
    pts.pab[2 * (size_t)k] = pa;
    pts.pab[2 * (size_t)k + 1] = pm;
    pts.pc[k] = make_float4(r1.w, r2.x, r2.y, 1.0f);

    Bounds b;
    sample_bounds(pa, pm, W, H, cc0, cc1, b);
    float4 b0 = make_float4(0, 0, 0, 0), b1 = make_float4(0, 0, 0, 0);
    if (bounds_valid(b)) {
        const float ta = b.ts0, tb = b.ts1;
        const float ax = x * w + (ta - t) * mvx, ay = y * w + (ta - t) * mvy, aw = w + (ta - t) * mvw;
        const float bx = x * w + (tb - t) * mvx, by = y * w + (tb - t) * mvy, bw = w + (tb - t) * mvw;
        const float ca = rcp(aw), cb = rcp(bw);
        const float xta = ax * ca, yta = ay * ca;
        const float xtb = bx * cb, ytb = by * cb;
        const float oodt = 1.0f / (tb - ta);
        const float dxdt = (xtb - xta) * oodt;
        const float dydt = (ytb - yta) * oodt;
        const float x0 = xta - ta * dxdt, y0 = yta - ta * dydt;
        const float x1 = xtb + (1.0f - tb) * dxdt, y1 = ytb + (1.0f - tb) * dydt;
        const float cta = cc0 / aw + cc1;
        const float ctb = cc0 / bw + cc1;
        b0 = make_float4(x0 + cta, y0 + cta, x0 - cta, y0 - cta);   // boxa, boxb (t=0)
        b1 = make_float4(x1 + ctb, y1 + ctb, x1 - ctb, y1 - ctb);   // boxaT1, boxbT1
    }
    bounds[k] = b;
    boxes[2 * (size_t)k] = b0;
    boxes[2 * (size_t)k + 1] = b1;
}

// Sequential surface scan of one initial leaf (ReconstructionTreeBuilder.cpp:
// 222-340, keeping the cocboundst[0] typo at :334,336).  Marks leaf starts.
__global__ __launch_bounds__(kBlock) void k_surface_scan(const int2* __restrict__ offcnt, int nil,
                                                         const float4* __restrict__ pab,
                                                         const float4* __restrict__ boxes,
                                                         uint32_t* __restrict__ start_flag) {
    const int li = blockIdx.x * kBlock + threadIdx.x;
    if (li >= nil) return;
    const int2 oc = offcnt[li];
    const int base = oc.x, n = oc.y;
    if (n <= 0) return;
    const float4* B = boxes + 2 * (size_t)base;
    const float4* T = pab + 2 * (size_t)base;  // (x, y, w, t) at T[2k]
    start_flag[base] = 1u;
    int sb = 0;
    bool reset = true;
    float cb00 = 0, cb01 = 0, cb10 = 0, cb11 = 0, mn0 = 0, mn1 = 0, mx0 = 0, mx1 = 0, st = 0;
    const float TH = 1 * 2.0f;
    // samples in groups of four, each group's records loaded together: the
    // wave's threads stream 64 leaves at once, and a line read one sample at
    // a time would leave the cache before its thread came back for the next
    for (int k0 = 1; k0 < n; k0 += 4) {
        float4 ga[4], gc[4];
        float gt[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + q < n ? k0 + q : n - 1;
            ga[q] = B[2 * k];
            gc[q] = B[2 * k + 1];
            gt[q] = T[2 * k].w;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + q;
            if (k >= n) break;
            if (reset) {
                const float4 a = B[2 * sb], c = B[2 * sb + 1];
                cb00 = cb01 = 0.5f * (a.x - a.z);
                cb10 = cb11 = 0.5f * (c.x - c.z);
                mn0 = 0.5f * (c.x + c.z) - 0.5f * (a.x + a.z);
                mn1 = 0.5f * (c.y + c.w) - 0.5f * (a.y + a.w);
                mx0 = mn0; mx1 = mn1;
                st = T[2 * sb].w;
                reset = false;
            }
            if (st != gt[q]) {
                sb = k;
                start_flag[base + k] = 1u;
                reset = true;
                continue;
            }
            const float4 a = ga[q], c = gc[q];
            const float c0 = 0.5f * (a.x - a.z);
            const float c1 = 0.5f * (c.x - c.z);
            const float m0 = 0.5f * (c.x + c.z) - 0.5f * (a.x + a.z);
            const float m1 = 0.5f * (c.y + c.w) - 0.5f * (a.y + a.w);
            const float dmv = fmx(fmx(fabsf(mn0 - m0), fabsf(mn1 - m1)), fmx(fabsf(mx0 - m0), fabsf(mx1 - m1)));
            const bool early = fmx(fabsf(c0 - cb00), fabsf(c0 - cb01)) < 0.5f * TH &&
                               fmx(fabsf(c1 - cb10), fabsf(c1 - cb11)) < 0.5f * TH && dmv < TH;
            bool conflict = false;
            if (!early) {
                for (int j = sb; j < k; ++j) {
                    const float4 ja = B[2 * j], jc = B[2 * j + 1];
                    const bool bX = tern_eq4(tern_cmp(ja.x, a.x, TH), tern_cmp(ja.z, a.z, TH),
                                             tern_cmp(jc.x, c.x, TH), tern_cmp(jc.z, c.z, TH));
                    const bool bY = tern_eq4(tern_cmp(ja.y, a.y, TH), tern_cmp(ja.w, a.w, TH),
                                             tern_cmp(jc.y, c.y, TH), tern_cmp(jc.w, c.w, TH));
                    if (!(bX && bY)) {
                        sb = k;
                        start_flag[base + k] = 1u;
                        conflict = true;
                        reset = true;
                        break;
                    }
                }
            }
            if (!conflict) {
                cb00 = fmn(c0, cb00);
                cb01 = fmx(c0, cb00);
                cb10 = fmn(c1, cb10);
                cb11 = fmx(c1, cb10);
                mn0 = fmn(mn0, m0); mn1 = fmn(mn1, m1);
                mx0 = fmx(mx0, m0); mx1 = fmx(mx1, m1);
            }
        }
    }
}

// Start sample of every leaf (leaf id = the scanned start flag), and nb
// after the last.
__global__ __launch_bounds__(kBlock) void k_leaf_starts(const uint32_t* __restrict__ start_flag,
                                                        const uint32_t* __restrict__ leaf_pos, int nb, int L,
                                                        int* __restrict__ leaf_start) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < nb && start_flag[k]) leaf_start[leaf_pos[k]] = k;
    if (k == 0) leaf_start[L] = nb;
}

// One thread per leaf: bounds folded in sample order with node.tlb =
// TimeLensBounds(bounds[k], node.tlb) (ReconstructionTreeBuilder.cpp:
// 345-378).  The 64 leaves of a block cover one contiguous sample range; the
// block copies its samples' bounds into LDS (coalesced) and each thread folds
// its leaf there.  A block whose range exceeds kEmitLds samples folds from
// global memory.
constexpr int kEmitBlock = 64;
constexpr int kEmitLds = 512;

__global__ __launch_bounds__(kEmitBlock) void k_leaf_emit(const int* __restrict__ leaf_start, int L,
                                                          const Bounds* __restrict__ bounds, Points pts,
                                                          LeafRec* __restrict__ leaves, Bounds* __restrict__ leaf_b,
                                                          uint32_t* __restrict__ valid_flag) {
    __shared__ float4 sb[kEmitLds * 4];
    const int lane = threadIdx.x;
    const int l0 = blockIdx.x * kEmitBlock;
    const int id = l0 + lane;
    const int lend = min(l0 + kEmitBlock, L);
    const int r0 = leaf_start[l0], r1 = leaf_start[lend];
    const bool staged = r1 - r0 <= kEmitLds;
    if (staged) {
        const float4* src = reinterpret_cast<const float4*>(bounds + r0);
        for (int e = lane; e < 4 * (r1 - r0); e += kEmitBlock) sb[e] = src[e];
    }
    __syncthreads();
    if (id >= L) return;
    const int k = leaf_start[id], e = leaf_start[id + 1];
    Bounds acc;
    bounds_zero(acc);
    for (int j = k; j < e; j++) {
        Bounds bj, nb2;
        if (staged) {
            const float4* q = sb + 4 * (j - r0);
            float* d = reinterpret_cast<float*>(&bj);
            const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
            d[0] = q0.x; d[1] = q0.y; d[2] = q0.z; d[3] = q0.w;
            d[4] = q1.x; d[5] = q1.y; d[6] = q1.z; d[7] = q1.w;
            d[8] = q2.x; d[9] = q2.y; d[10] = q2.z; d[11] = q2.w;
            d[12] = q3.x; d[13] = q3.y; d[14] = q3.z; d[15] = q3.w;
        } else {
            bj = bounds[j];
        }
        bounds_merge(bj, acc, nb2);
        acc = nb2;
    }
    LeafRec r;
    r.s0 = k;
    r.s1 = e;
    const float4 a = pt_a(pts, k);
    r.w0 = a.z;
    r.t0 = a.w;
    r.mvw0 = pt_b(pts, k).z;
    r.valid = bounds_valid(acc) ? 1 : 0;
    r.pad[0] = r.pad[1] = 0.0f;
    leaves[id] = r;
    leaf_b[id] = acc;
    valid_flag[id] = r.valid ? 1u : 0u;
}

// ---------------------------------------------------------------------------
// Hierarchy over the valid leaves: balanced binary tree of recursive median
// splits.  The split features are a leaf's centre in x at the four (t, u)
// corners of the shutter-lens square, (0,-1), (1,-1), (1,1), (0,1), and its
// centre in y at the same (t, v) corners; every node splits its leaves at the
// median of the feature of widest extent.  Leaves close at every corner stay
// close wherever a query's (t, u, v) falls, so the merged planes stay tight.
// Against splitting on the centre at (t = 0.5, u = v = 0), the CoC slope and
// the t-slopes: 9% fewer node tests per query at 480x270, and at 1920x1080 x
// 16 spp the 1R lane passes 62.1 -> 59.6 ms for 0.6 ms more build.
// Built level by level: per-node extents, then a stable radix sort of
// (node at this level, feature quantised to kd_qbits(level) bits), then the
// halving split.  Internal node g is the node whose split lies between sorted
// positions g and g+1.  Replaces emitNodes (ReconstructionTreeBuilder.cpp:
// 409-451); which leaves a query finds does not depend on the hierarchy.
// ---------------------------------------------------------------------------
constexpr int kKdFeat = 8;

// x lower/upper and y lower/upper bound of a leaf at the four corners.
__device__ __forceinline__ void kd_corners(const Bounds& b, float xl[4], float xh[4], float yl[4], float yh[4]) {
    const float ct[4] = {0.0f, 1.0f, 1.0f, 0.0f};
    const float cu[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        xl[c] = plane_eval(b.p + 0, ct[c], cu[c]);
        xh[c] = plane_eval(b.p + 3, ct[c], cu[c]);
        yl[c] = plane_eval(b.p + 6, ct[c], cu[c]);
        yh[c] = plane_eval(b.p + 9, ct[c], cu[c]);
    }
}

__device__ __forceinline__ int kd_qbits(int level) { return level < 16 ? 16 : 32 - level; }

__device__ __forceinline__ uint32_t kd_quant(float v, float lo, float hi, int bits) {
    if (!(hi > lo)) return 0;
    const uint32_t m = (1u << bits) - 1u;
    const float sc = (float)m / (hi - lo);
    const float f = (v - lo) * sc;
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)m) return m;
    return (uint32_t)f;
}

// Node (index at `level`, virtual below early leaves) containing position p.
__device__ __forceinline__ int kd_node_at(int n, int level, int p, int& lo, int& hi) {
    lo = 0;
    hi = n;
    int node = 0;
    for (int k = 0; k < level; k++) {
        if (hi - lo < 2) return node << (level - k);
        const int mid = (lo + hi) >> 1;
        if (p < mid) {
            hi = mid;
            node = node * 2;
        } else {
            lo = mid;
            node = node * 2 + 1;
        }
    }
    return node;
}

// Range of node `node` at `level` (empty for virtual nodes).
__device__ __forceinline__ void kd_range(int n, int level, int node, int& lo, int& hi) {
    lo = 0;
    hi = n;
    for (int k = level - 1; k >= 0; k--) {
        if (hi - lo < 2) {
            if (node & ((1 << (k + 1)) - 1)) lo = hi;
            return;
        }
        const int mid = (lo + hi) >> 1;
        if ((node >> k) & 1) lo = mid;
        else hi = mid;
    }
}

__global__ __launch_bounds__(kBlock) void k_kd_features(const uint32_t* __restrict__ valid_flag,
                                                        const uint32_t* __restrict__ valid_pos, int L,
                                                        const Bounds* __restrict__ leaf_b, float* __restrict__ feat,
                                                        int nv, uint32_t* __restrict__ vid_leaf,
                                                        uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= L || !valid_flag[i]) return;
    const uint32_t v = valid_pos[i];
    float xl[4], xh[4], yl[4], yh[4];
    kd_corners(leaf_b[i], xl, xh, yl, yh);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        feat[c * (size_t)nv + v] = 0.5f * (xl[c] + xh[c]);
        feat[(4 + c) * (size_t)nv + v] = 0.5f * (yl[c] + yh[c]);
    }
    vid_leaf[v] = (uint32_t)i;
    vals[v] = v;
}

// Per-node split dimension and its range: info[node] = (dim, lo, hi).
__device__ __forceinline__ void kd_pick(const float mn[kKdFeat], const float mx[kKdFeat], float4& out) {
    int best = 0;
    float bext = mx[0] - mn[0];
#pragma unroll
    for (int d = 1; d < kKdFeat; d++)
        if (mx[d] - mn[d] > bext) {
            bext = mx[d] - mn[d];
            best = d;
        }
    out = make_float4(__int_as_float(best), mn[best], mx[best], 0.0f);
}

// Order-preserving float <-> uint32 map for atomic min/max.
__device__ __forceinline__ uint32_t f2ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__global__ __launch_bounds__(kBlock) void k_kd_ext_init(int nodes, uint32_t* __restrict__ emn,
                                                        uint32_t* __restrict__ emx) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= kKdFeat * nodes) return;
    emn[i] = 0xFFFFFFFFu;
    emx[i] = 0u;
}

// Min/max of the features over positions [a, b) of node `node`, reduced
// over the block, then one atomic per feature.  Called by the whole block.
__device__ void kd_block_extent(const float* __restrict__ feat, int nv, const uint32_t* __restrict__ vals, int a,
                                int b, int node, uint32_t* __restrict__ emn, uint32_t* __restrict__ emx) {
    __shared__ float smn[kKdFeat][kBlock / 64], smx[kKdFeat][kBlock / 64];
    float mn[kKdFeat], mx[kKdFeat];
#pragma unroll
    for (int d = 0; d < kKdFeat; d++) {
        mn[d] = kFltMax;
        mx[d] = -kFltMax;
    }
    for (int p = a + threadIdx.x; p < b; p += kBlock) {
        const uint32_t v = vals[p];
#pragma unroll
        for (int d = 0; d < kKdFeat; d++) {
            const float f = feat[d * (size_t)nv + v];
            mn[d] = fminf(mn[d], f);
            mx[d] = fmaxf(mx[d], f);
        }
    }
#pragma unroll
    for (int d = 0; d < kKdFeat; d++) {
        mn[d] = wave_min(mn[d]);
        mx[d] = wave_max(mx[d]);
        if ((threadIdx.x & 63) == 0) {
            smn[d][threadIdx.x >> 6] = mn[d];
            smx[d][threadIdx.x >> 6] = mx[d];
        }
    }
    __syncthreads();
    if (threadIdx.x < kKdFeat && b > a) {
        const int d = threadIdx.x;
        float lo = smn[d][0], hi = smx[d][0];
        for (int w = 1; w < kBlock / 64; w++) {
            lo = fminf(lo, smn[d][w]);
            hi = fmaxf(hi, smx[d][w]);
        }
        atomicMin(&emn[kKdFeat * node + d], f2ord(lo));
        atomicMax(&emx[kKdFeat * node + d], f2ord(hi));
    }
    __syncthreads();
}

// Nodes of at least kKdChunk leaves: one block per chunk of kKdChunk
// positions; a chunk meets at most two nodes (node sizes at one level differ
// by at most one).
constexpr int kKdChunk = 2048;
__global__ __launch_bounds__(kBlock) void k_kd_ext_atomic(const float* __restrict__ feat, int nv,
                                                          const uint32_t* __restrict__ vals, int level,
                                                          uint32_t* __restrict__ emn, uint32_t* __restrict__ emx) {
    const int p0 = blockIdx.x * kKdChunk;
    const int p1 = min(nv, p0 + kKdChunk);
    int lo, hi;
    const int n0 = kd_node_at(nv, level, p0, lo, hi);
    const int cut = min(hi, p1);
    kd_block_extent(feat, nv, vals, p0, cut, n0, emn, emx);
    if (cut < p1) {
        const int n1 = kd_node_at(nv, level, cut, lo, hi);
        kd_block_extent(feat, nv, vals, cut, p1, n1, emn, emx);
    }
}

// Mid-size nodes: one wave per node.
__global__ __launch_bounds__(kBlock) void k_kd_extent_wave(const float* __restrict__ feat, int nv,
                                                           const uint32_t* __restrict__ vals, int level, int nodes,
                                                           float4* __restrict__ info) {
    const int node = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (node >= nodes) return;
    int lo, hi;
    kd_range(nv, level, node, lo, hi);
    if (hi - lo < 2) return;
    float mn[kKdFeat], mx[kKdFeat];
#pragma unroll
    for (int d = 0; d < kKdFeat; d++) {
        mn[d] = kFltMax;
        mx[d] = -kFltMax;
    }
    for (int p = lo + (threadIdx.x & 63); p < hi; p += 64) {
        const uint32_t v = vals[p];
#pragma unroll
        for (int d = 0; d < kKdFeat; d++) {
            const float f = feat[d * (size_t)nv + v];
            mn[d] = fminf(mn[d], f);
            mx[d] = fmaxf(mx[d], f);
        }
    }
#pragma unroll
    for (int d = 0; d < kKdFeat; d++) {
        mn[d] = wave_min(mn[d]);
        mx[d] = wave_max(mx[d]);
    }
    if ((threadIdx.x & 63) == 0) kd_pick(mn, mx, info[node]);
}

__global__ __launch_bounds__(kBlock) void k_kd_ext_final(int nv, int level, int nodes,
                                                         const uint32_t* __restrict__ emn,
                                                         const uint32_t* __restrict__ emx,
                                                         float4* __restrict__ info) {
    const int node = blockIdx.x * kBlock + threadIdx.x;
    if (node >= nodes) return;
    int lo, hi;
    kd_range(nv, level, node, lo, hi);
    if (hi - lo < 2) return;
    float mn[kKdFeat], mx[kKdFeat];
#pragma unroll
    for (int d = 0; d < kKdFeat; d++) {
        mn[d] = ord2f(emn[kKdFeat * node + d]);
        mx[d] = ord2f(emx[kKdFeat * node + d]);
    }
    kd_pick(mn, mx, info[node]);
}

// Small nodes: one thread per node.
__global__ __launch_bounds__(kBlock) void k_kd_extent_thread(const float* __restrict__ feat, int nv,
                                                             const uint32_t* __restrict__ vals, int level,
                                                             int nodes, float4* __restrict__ info) {
    const int node = blockIdx.x * kBlock + threadIdx.x;
    if (node >= nodes) return;
    int lo, hi;
    kd_range(nv, level, node, lo, hi);
    if (hi - lo < 2) return;
    float mn[kKdFeat], mx[kKdFeat];
#pragma unroll
    for (int d = 0; d < kKdFeat; d++) {
        mn[d] = kFltMax;
        mx[d] = -kFltMax;
    }
    for (int p = lo; p < hi; p++) {
        const uint32_t v = vals[p];
#pragma unroll
        for (int d = 0; d < kKdFeat; d++) {
            const float f = feat[d * (size_t)nv + v];
            mn[d] = fminf(mn[d], f);
            mx[d] = fmaxf(mx[d], f);
        }
    }
    kd_pick(mn, mx, info[node]);
}

__global__ __launch_bounds__(kBlock) void k_kd_keys(const float* __restrict__ feat, int nv,
                                                    const uint32_t* __restrict__ vals, int level,
                                                    const float4* __restrict__ info, uint32_t* __restrict__ keys) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= nv) return;
    int lo, hi;
    const int node = kd_node_at(nv, level, p, lo, hi);
    const int bits = kd_qbits(level);
    uint32_t q = 0;
    if (hi - lo >= 2) {
        const float4 in = info[node];
        const int d = __float_as_int(in.x);
        q = kd_quant(feat[d * (size_t)nv + vals[p]], in.y, in.z, bits);
    }
    keys[p] = ((uint32_t)node << bits) | q;
}

// Leaf slot -> leaf id after the last level.
// Stable sort of one level's keys within chunks of whole nodes, for levels
// whose nodes hold at most kSegSort leaves: chunk c = nodes [c*K, c*K + K)
// (contiguous positions, as the previous level left them grouped by node),
// sorted in LDS as (key, position in chunk) by a bitonic network.  Equals
// the stable radix sort of the whole array: keys start with the node index.
// Chunks of about kSegSortTarget keep the network short and every CU busy;
// nodes of up to kSegSort leaves take a chunk of their own.
constexpr int kSegSort = 2048;
constexpr int kSegSortTarget = 512;
constexpr int kSegSortBlock = 256;
__global__ __launch_bounds__(kSegSortBlock) void k_kd_seg_sort(const uint32_t* __restrict__ keys_in,
                                                               const uint32_t* __restrict__ vals_in,
                                                               uint32_t* __restrict__ keys_out,
                                                               uint32_t* __restrict__ vals_out, int nv, int level,
                                                               int nodes, int K) {
    __shared__ unsigned long long sk[kSegSort];
    const int n0 = blockIdx.x * K;
    const int n1 = min(nodes, n0 + K) - 1;
    int lo0, hi0, lo1, hi1;
    kd_range(nv, level, n0, lo0, hi0);
    kd_range(nv, level, n1, lo1, hi1);
    const int base = lo0;
    const int len = hi1 - lo0;
    int P = 1;
    while (P < len) P <<= 1;
    for (int i = threadIdx.x; i < P; i += kSegSortBlock)
        sk[i] = i < len ? (((unsigned long long)keys_in[base + i] << 12) | (unsigned long long)i) : ~0ull;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < P; i += kSegSortBlock) {
                const int j = i ^ stride;
                if (j > i) {
                    const unsigned long long a = sk[i], b = sk[j];
                    const bool up = (i & size) == 0;
                    if ((a > b) == up) {
                        sk[i] = b;
                        sk[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < len; i += kSegSortBlock) {
        const unsigned long long e = sk[i];
        keys_out[base + i] = (uint32_t)(e >> 12);
        vals_out[base + i] = vals_in[base + (int)(e & 0xFFFull)];
    }
}

__global__ __launch_bounds__(kBlock) void k_kd_order(const uint32_t* __restrict__ vals,
                                                     const uint32_t* __restrict__ vid_leaf, int nv,
                                                     uint32_t* __restrict__ ids) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < nv) ids[p] = vid_leaf[vals[p]];
}

// Internal node g: children, parent links, range, split owner.
__global__ __launch_bounds__(kBlock) void k_kd_nodes(int n, int* __restrict__ child0, int* __restrict__ child1,
                                                     int* __restrict__ parent, int2* __restrict__ range,
                                                     int* __restrict__ split_owner) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n - 1) return;
    int lo = 0, hi = n, mid = n >> 1;
    for (;;) {
        mid = (lo + hi) >> 1;
        if (g == mid - 1) break;
        if (g < mid - 1) hi = mid;
        else lo = mid;
    }
    const int c0 = (mid - lo == 1) ? (n - 1 + lo) : ((lo + mid) >> 1) - 1;
    const int c1 = (hi - mid == 1) ? (n - 1 + mid) : ((mid + hi) >> 1) - 1;
    child0[g] = c0;
    child1[g] = c1;
    parent[c0] = g;
    parent[c1] = g;
    range[g] = make_int2(lo, hi - 1);
    split_owner[g] = g;
}

// Bounds of the median-split hierarchy, one launch per level from the
// deepest: node `node` at `level` merges its two children (child0, child1),
// which are leaves or nodes of the level below.
__global__ __launch_bounds__(kBlock) void k_kd_bounds_level(const uint32_t* __restrict__ ids, int n, int level,
                                                            const Bounds* __restrict__ leaf_b,
                                                            const int* __restrict__ child0,
                                                            const int* __restrict__ child1,
                                                            Bounds* __restrict__ node_b) {
    const int node = blockIdx.x * kBlock + threadIdx.x;
    if (node >= (1 << level)) return;
    int lo, hi;
    kd_range(n, level, node, lo, hi);
    if (hi - lo < 2) return;
    const int g = ((lo + hi) >> 1) - 1;
    const int c0 = child0[g], c1 = child1[g];
    const Bounds a = (c0 >= n - 1) ? leaf_b[ids[c0 - (n - 1)]] : node_b[c0];
    const Bounds b = (c1 >= n - 1) ? leaf_b[ids[c1 - (n - 1)]] : node_b[c1];
    Bounds r;
    bounds_merge(a, b, r);
    node_b[g] = r;
}

// Child-pair records of the internal nodes, from the traversal records.
__global__ __launch_bounds__(kBlock) void k_tree_pairs(const TravNode* __restrict__ trav, int n,
                                                       const LeafRec* __restrict__ leaves,
                                                       TravPair* __restrict__ pairs) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n - 1) return;
    const TravNode& t = trav[g];
    const TravNode& ta = trav[t.child];
    const TravNode& tb = trav[t.child1];
    TravPair r;
#pragma unroll
    for (int q = 0; q < 12; q++) {
        r.a[q] = ta.p[q];
        r.b[q] = tb.p[q];
    }
    // a leaf slot's record holds -(leaf id + 1); an internal node is its index
    r.ra = ta.child < 0 ? ta.child : t.child;
    r.rb = tb.child < 0 ? tb.child : t.child1;
    r.ka[0] = r.ka[1] = r.ka[2] = r.kb[0] = r.kb[1] = r.kb[2] = 0.0f;
    if (r.ra < 0) {
        const LeafRec& l = leaves[-r.ra - 1];
        r.ka[0] = l.w0;
        r.ka[1] = l.t0;
        r.ka[2] = l.mvw0;
    }
    if (r.rb < 0) {
        const LeafRec& l = leaves[-r.rb - 1];
        r.kb[0] = l.w0;
        r.kb[1] = l.t0;
        r.kb[2] = l.mvw0;
    }
    pairs[g] = r;
}

// Traversal records with escape links (next node after a subtree).
__global__ __launch_bounds__(kBlock) void k_tree_trav(const uint32_t* __restrict__ ids, int n,
                                                      const Bounds* __restrict__ leaf_b,
                                                      const Bounds* __restrict__ node_b,
                                                      const int* __restrict__ child0,
                                                      const int* __restrict__ child1,
                                                      const int2* __restrict__ range,
                                                      const int* __restrict__ split_owner,
                                                      TravNode* __restrict__ trav) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= 2 * n - 1) return;
    TravNode t;
    int hi;
    const Bounds* b;
    t.child1 = -1;
    t.pad = 0;
    if (k < n - 1) {
        b = &node_b[k];
        t.child = child0[k];
        t.child1 = child1[k];
        hi = range[k].y;
    } else {
        const int slot = k - (n - 1);
        const uint32_t lid = ids[slot];
        b = &leaf_b[lid];
        t.child = -(int)lid - 1;
        hi = slot;
    }
#pragma unroll
    for (int q = 0; q < 12; q++) t.p[q] = b->p[q];
    t.escape = (hi == n - 1) ? -1 : child1[split_owner[hi]];
    trav[k] = t;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
#define TLR_GRID(n) dim3((unsigned)(((n) + kBlock - 1) / kBlock)), dim3(kBlock)

void launch_reproject(const InSoA& in, int n, int W, int H, float C0, float C1, float* rx, float* ry, float* rw,
                      float* rt, float4* packed, int* hits, GridInfo* gi, hipStream_t s) {
    const unsigned rb = (unsigned)std::min<long long>(kReduceBlocks, ((long long)n + kBlock - 1) / kBlock);
    if (n > 0) k_reproject<<<rb, kBlock, 0, s>>>(in, n, W, H, C0, C1, rx, ry, rw, rt, packed, hits, gi);
    const int nh = W * H;
    unsigned blocks = (unsigned)std::min(1024, (nh + kBlock - 1) / kBlock);
    if (blocks == 0) blocks = 1;
    k_hits_max<<<blocks, kBlock, 0, s>>>(hits, nh, gi);
    k_grid_setup<<<1, 64, 0, s>>>(gi);
}

void launch_cell_keys(const float* rx, const float* ry, const float* rw, int n, GridInfo* gi, uint32_t* keys,
                      uint32_t* vals, hipStream_t s) {
    const unsigned cb = (unsigned)std::min<long long>(kReduceBlocks, ((long long)n + kBlock - 1) / kBlock);
    if (n > 0) k_cell_keys<<<cb, kBlock, 0, s>>>(rx, ry, rw, n, gi, keys, vals);
}

void launch_cell_runs(const uint32_t* skeys, int n, uint32_t cells, uint32_t* start, uint32_t* end, hipStream_t s) {
    if (n > 0) k_cell_runs<<<TLR_GRID(n), 0, s>>>(skeys, n, cells, start, end);
}

void launch_sat(const uint32_t* cstart, const uint32_t* cend, int gw, int gh, uint32_t* sat, hipStream_t s) {
    k_sat_rows<<<gh, kBlock, 0, s>>>(cstart, cend, gw, gh, sat);
    k_sat_cols<<<TLR_GRID(gw + 1), 0, s>>>(sat, gw, gh);
}

void launch_kd(const uint32_t* sat, int gw, int gh, int4* owner, uint32_t* own_flag, int* own_y1, hipStream_t s) {
    k_kd_descend<<<TLR_GRID(gw * gh), 0, s>>>(sat, gw, gh, owner, own_flag, own_y1);
}

void launch_emit_initial(const int4* owner, const uint32_t* own_flag, const int* own_y1, const uint32_t* idx,
                         const uint32_t* sat, int gw, int gh, int4* rect, int2* offcnt, hipStream_t s) {
    k_emit_initial<<<TLR_GRID(gw * gh), 0, s>>>(owner, own_flag, own_y1, idx, sat, gw, gh, rect, offcnt);
}

void launch_place(const uint32_t* skeys, const uint32_t* svals, int nb, int gw, const int4* owner,
                  const uint32_t* sat, const uint32_t* cell_start, const float4* packed, uint32_t* cand, float* ct,
                  float* cw, hipStream_t s) {
    if (nb > 0) k_place<<<TLR_GRID(nb), 0, s>>>(skeys, svals, nb, gw, owner, sat, cell_start, packed, cand, ct, cw);
}

void launch_leaf_sort(const int2* offcnt, int nil, uint32_t* cand, float* ct, float* cw, hipStream_t s) {
    if (nil > 0) k_leaf_sort<<<(nil + kSortBlock - 1) / kSortBlock, kSortBlock, 0, s>>>(offcnt, nil, cand, ct, cw);
}

void launch_points(const uint32_t* cand, int nb, const float4* packed, float W, float H, float cc0, float cc1,
                   const Points& pts, Bounds* bounds, float4* boxes, hipStream_t s) {
    if (nb > 0) k_points<<<TLR_GRID(nb), 0, s>>>(cand, nb, packed, W, H, cc0, cc1, pts, bounds, boxes);
}

void launch_surface_scan(const int2* offcnt, int nil, const float4* pab, const float4* boxes, uint32_t* flags,
                         hipStream_t s) {
    if (nil > 0) k_surface_scan<<<TLR_GRID(nil), 0, s>>>(offcnt, nil, pab, boxes, flags);
}

void launch_leaf_emit(const uint32_t* flags, const uint32_t* leaf_pos, int nb, int L, int* leaf_start,
                      const Bounds* bounds, const Points& pts, LeafRec* leaves, Bounds* leaf_b,
                      uint32_t* valid_flag, hipStream_t s) {
    if (nb <= 0 || L <= 0) return;
    k_leaf_starts<<<TLR_GRID(nb), 0, s>>>(flags, leaf_pos, nb, L, leaf_start);
    k_leaf_emit<<<(L + kEmitBlock - 1) / kEmitBlock, kEmitBlock, 0, s>>>(leaf_start, L, bounds, pts, leaves, leaf_b,
                                                                         valid_flag);
}

hipError_t launch_kdtree(const uint32_t* valid_flag, const uint32_t* valid_pos, int L, const Bounds* leaf_b,
                         int nv, float* feat, uint32_t* vid_leaf, uint32_t* k0, uint32_t* v0, uint32_t* k1,
                         uint32_t* v1, float4* info, uint32_t* emn, uint32_t* emx, uint32_t* hist, uint32_t* tmp, uint32_t* ids, int* c0, int* c1,
                         int* parent, int2* range, int* split_owner, Bounds* node_b, hipStream_t s) {
    if (nv <= 0) return hipSuccess;
    k_kd_features<<<TLR_GRID(L), 0, s>>>(valid_flag, valid_pos, L, leaf_b, feat, nv, vid_leaf, v0);
    uint32_t* vals = v0;
    uint32_t* keys = k0;
    uint32_t* ko = k1;
    uint32_t* vo = v1;
    int levels = 0;
    while ((1 << levels) < nv) levels++;
    for (int level = 0; level < levels; level++) {
        const int nodes = 1 << level;
        // average node size decides how extents are reduced
        if (nv / nodes >= kKdChunk) {
            k_kd_ext_init<<<TLR_GRID(kKdFeat * nodes), 0, s>>>(nodes, emn, emx);
            k_kd_ext_atomic<<<(nv + kKdChunk - 1) / kKdChunk, kBlock, 0, s>>>(feat, nv, vals, level, emn, emx);
            k_kd_ext_final<<<TLR_GRID(nodes), 0, s>>>(nv, level, nodes, emn, emx, info);
        } else if (nv / nodes > 64) {
            k_kd_extent_wave<<<(nodes + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, s>>>(feat, nv, vals, level,
                                                                                        nodes, info);
        } else {
            k_kd_extent_thread<<<TLR_GRID(nodes), 0, s>>>(feat, nv, vals, level, nodes, info);
        }
        k_kd_keys<<<TLR_GRID(nv), 0, s>>>(feat, nv, vals, level, info, keys);
        const int smax = (nv + nodes - 1) / nodes;   // largest node at this level
        if (smax <= kSegSort) {
            const int K = smax <= kSegSortTarget ? kSegSortTarget / smax : 1;
            const int chunks = (nodes + K - 1) / K;
            k_kd_seg_sort<<<chunks, kSegSortBlock, 0, s>>>(keys, vals, ko, vo, nv, level, nodes, K);
            std::swap(keys, ko);
            std::swap(vals, vo);
            continue;
        }
        const int bits = level + (level < 16 ? 16 : 32 - level);
        bool first = true;
        hipError_t e = radix_sort_pairs(keys, vals, ko, vo, (size_t)nv, bits, hist, tmp, &first, s);
        if (e != hipSuccess) return e;
        if (!first) {
            std::swap(keys, ko);
            std::swap(vals, vo);
        }
    }
    k_kd_order<<<TLR_GRID(nv), 0, s>>>(vals, vid_leaf, nv, ids);
    if (nv > 1) k_kd_nodes<<<TLR_GRID(nv - 1), 0, s>>>(nv, c0, c1, parent, range, split_owner);
    for (int level = levels - 1; level >= 0; level--)
        k_kd_bounds_level<<<TLR_GRID(1 << level), 0, s>>>(ids, nv, level, leaf_b, c0, c1, node_b);
    return hipGetLastError();
}

void launch_tree_trav(const uint32_t* ids, int n, const Bounds* leaf_b, const Bounds* node_b, const int* c0,
                      const int* c1, const int2* range, const int* split_owner, const LeafRec* leaves,
                      TravNode* trav, TravPair* pairs, hipStream_t s) {
    if (n > 0) k_tree_trav<<<TLR_GRID(2 * n - 1), 0, s>>>(ids, n, leaf_b, node_b, c0, c1, range, split_owner, trav);
    if (n > 1) k_tree_pairs<<<TLR_GRID(n - 1), 0, s>>>(trav, n, leaves, pairs);
}

}  // namespace tlr
