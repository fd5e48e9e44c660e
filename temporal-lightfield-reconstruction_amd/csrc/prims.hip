// Exclusive scan and stable LSD radix sort (wave64, gfx950).
#include "prims.hpp"

namespace tlr {

hipError_t DeviceArena::reserve(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (base) {
        hipError_t e = hipFree(base);
        if (e != hipSuccess) return e;
        base = nullptr;
        cap = 0;
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), bytes);
    if (e != hipSuccess) return e;
    cap = bytes;
    top = 0;
    return hipSuccess;
}

void DeviceArena::release() {
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = 0;
    top = 0;
}

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds4, uint32_t& total) {
    uint32_t inc = wave_incl_scan(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int k = 0; k < w; k++) off += lds4[k];
    total = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return off + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_block_sums(const uint32_t* __restrict__ in, size_t n,
                                                             uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds4[4];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        size_t i = base + (size_t)j * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    (void)block_excl_scan(s, lds4, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Scans each 4096-element tile and adds the tile's offset (offs may be null).
__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const uint32_t* in, uint32_t* out, size_t n,
                                                            const uint32_t* __restrict__ offs,
                                                            uint32_t* total) {
    __shared__ uint32_t tile[kScanTile];
    __shared__ uint32_t lds4[4];
    const size_t base = (size_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        size_t i = base + (size_t)j * kScanThreads + threadIdx.x;
        tile[j * kScanThreads + threadIdx.x] = (i < n) ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        v[j] = tile[threadIdx.x * kScanItems + j];
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(s, lds4, tot);
    if (offs) run += offs[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        tile[threadIdx.x * kScanItems + j] = run;
        run += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        size_t i = base + (size_t)j * kScanThreads + threadIdx.x;
        if (i < n) out[i] = tile[j * kScanThreads + threadIdx.x];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) {
        // run is the inclusive sum of this thread's items = grand total
        *total = run;
    }
}

// LSD radix sort, 8-bit digits.  A tile of 4096 keys per 256-thread block;
// each of its four waves takes a contiguous quarter, so input order inside a
// tile is (wave, round, lane).
constexpr int kRadixThreads = 256;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixWaves = kRadixThreads / 64;
constexpr int kRadixWaveItems = kRadixTile / kRadixWaves;

// Digit counts per tile, stored digit-major (hist[d * nb + tile]) so that one
// exclusive scan gives every (digit, tile) run its output offset.
__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(const uint32_t* __restrict__ keys, size_t n,
                                                              int shift, uint32_t nb, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[256];
    const int tid = threadIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRadixTile;
#pragma unroll
    for (int j = 0; j < kRadixItems; j++) {
        const size_t i = base + (size_t)j * kRadixThreads + tid;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)tid * nb + blockIdx.x] = cnt[tid];
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Stable scatter of one tile: every key is ranked inside the tile (digit
// start in the tile + earlier waves' keys of its digit + its rank in its
// wave), the tile is staged sorted in LDS and written out as one contiguous
// run per digit at the run's offset from the scanned histogram.
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                 const uint32_t* __restrict__ vin,
                                                                 uint32_t* __restrict__ kout,
                                                                 uint32_t* __restrict__ vout, size_t n, int shift,
                                                                 uint32_t nb, const uint32_t* __restrict__ hist) {
    __shared__ uint32_t sk[kRadixTile];
    __shared__ uint32_t sv[kRadixTile];
    __shared__ uint32_t wcnt[kRadixWaves][256];
    __shared__ uint32_t dstart[256];
    __shared__ uint32_t goff[256];
    __shared__ uint32_t wtot[kRadixWaves];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kRadixTile;
    // this tile's count of digit tid: the next (digit, tile) offset minus its own
    uint32_t cnt;
    {
        const size_t e = (size_t)tid * nb + tile;
        const uint32_t g = hist[e];
        const uint32_t g1 = (e + 1 < (size_t)256 * nb) ? hist[e + 1] : (uint32_t)n;
        goff[tid] = g;
        cnt = g1 - g;
    }
    uint32_t inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wtot[w] = inc;
#pragma unroll
    for (int k = 0; k < kRadixWaves; k++) wcnt[k][tid] = 0;
    __syncthreads();
    {
        uint32_t off = 0;
        for (int k = 0; k < w; k++) off += wtot[k];
        dstart[tid] = off + inc - cnt;
    }
    // rank pass: each wave over its quarter, in input order
    uint32_t key[kRadixItems], val[kRadixItems], loc[kRadixItems];
    const size_t wbase = base + (size_t)w * kRadixWaveItems;
#pragma unroll
    for (int j = 0; j < kRadixItems; j++) {
        const size_t i = wbase + (size_t)j * 64 + lane;
        const bool act = i < n;
        key[j] = act ? kin[i] : 0u;
        val[j] = act ? vin[i] : 0u;
        const uint32_t d = (key[j] >> shift) & 255u;
        uint64_t m = __ballot(act);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(act && bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t before = wcnt[w][d];
        loc[j] = before + lanes_below(m);
        __builtin_amdgcn_wave_barrier();
        // the group's highest lane advances the wave's count of the digit
        if (act && (m >> lane) == 1ull) wcnt[w][d] = before + (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: tile start plus the counts of the earlier waves
    {
        uint32_t acc = dstart[tid];
#pragma unroll
        for (int k = 0; k < kRadixWaves; k++) {
            const uint32_t c = wcnt[k][tid];
            wcnt[k][tid] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRadixItems; j++) {
        const size_t i = wbase + (size_t)j * 64 + lane;
        if (i < n) {
            const uint32_t p = wcnt[w][(key[j] >> shift) & 255u] + loc[j];
            sk[p] = key[j];
            sv[p] = val[j];
        }
    }
    __syncthreads();
    const uint32_t tn = (uint32_t)((n - base) < (size_t)kRadixTile ? (n - base) : (size_t)kRadixTile);
    for (uint32_t p = tid; p < tn; p += kRadixThreads) {
        const uint32_t k = sk[p];
        const uint32_t d = (k >> shift) & 255u;
        const uint32_t dst = goff[d] + (p - dstart[d]);
        kout[dst] = k;
        vout[dst] = sv[p];
    }
}

}  // namespace

size_t scan_tmp_words(size_t n) {
    if (n <= (size_t)kScanTile) return 0;
    size_t nb = (n + kScanTile - 1) / kScanTile;
    return nb + 64 + scan_tmp_words(nb);
}

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, size_t n, uint32_t* tmp,
                              hipStream_t s) {
    if (n == 0) {
        if (total) return hipMemsetAsync(total, 0, sizeof(uint32_t), s);
        return hipSuccess;
    }
    if (n <= (size_t)kScanTile) {
        k_tile_scan<<<1, kScanThreads, 0, s>>>(in, out, n, nullptr, total);
        return hipGetLastError();
    }
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    uint32_t* sums = tmp;
    k_block_sums<<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = exclusive_scan_u32(sums, sums, total, nb, tmp + nb + 64, s);
    if (e != hipSuccess) return e;
    k_tile_scan<<<(unsigned)nb, kScanThreads, 0, s>>>(in, out, n, sums, nullptr);
    return hipGetLastError();
}

size_t radix_hist_words(size_t n) {
    size_t nb = (n + kRadixTile - 1) / kRadixTile;
    return 256 * (nb ? nb : 1);
}

hipError_t radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                            uint32_t* hist, uint32_t* tmp, bool* in_first, hipStream_t s) {
    *in_first = true;
    if (n == 0 || bits <= 0) return hipSuccess;
    const uint32_t nb = (uint32_t)((n + kRadixTile - 1) / kRadixTile);
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (int shift = 0; shift < bits; shift += 8) {
        k_radix_hist<<<nb, kRadixThreads, 0, s>>>(ki, n, shift, nb, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = exclusive_scan_u32(hist, hist, nullptr, (size_t)256 * nb, tmp, s);
        if (e != hipSuccess) return e;
        k_radix_scatter<<<nb, kRadixThreads, 0, s>>>(ki, vi, ko, vo, n, shift, nb, hist);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        std::swap(ki, ko);
        std::swap(vi, vo);
        *in_first = !*in_first;
    }
    return hipSuccess;
}

}  // namespace tlr
