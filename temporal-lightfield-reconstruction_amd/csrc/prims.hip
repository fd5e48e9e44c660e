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

constexpr int kRadixItems = 16;
constexpr int kRadixTile = 64 * kRadixItems;  // one wave per block

__global__ __launch_bounds__(64) void k_radix_hist(const uint32_t* __restrict__ keys, size_t n, int shift,
                                                   uint32_t nb, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[256];
    const int lane = threadIdx.x;
    for (int d = lane; d < 256; d += 64) cnt[d] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRadixTile;
#pragma unroll
    for (int j = 0; j < kRadixItems; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    for (int d = lane; d < 256; d += 64) hist[(size_t)d * nb + blockIdx.x] = cnt[d];
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__global__ __launch_bounds__(64) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      size_t n, int shift, uint32_t nb,
                                                      const uint32_t* __restrict__ hist) {
    __shared__ uint32_t run[256];
    const int lane = threadIdx.x;
    for (int d = lane; d < 256; d += 64) run[d] = hist[(size_t)d * nb + blockIdx.x];
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRadixTile;
    for (int j = 0; j < kRadixItems; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        const bool act = i < n;
        uint32_t key = act ? kin[i] : 0u;
        uint32_t val = act ? vin[i] : 0u;
        uint32_t d = (key >> shift) & 255u;
        uint64_t m = __ballot(act);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(act && bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = lanes_below(m);
        uint32_t dst = 0;
        if (act) dst = run[d] + rank;
        __builtin_amdgcn_wave_barrier();
        if (act) {
            kout[dst] = key;
            vout[dst] = val;
            if ((m >> lane) == 1ull) run[d] = dst + 1;
        }
        __builtin_amdgcn_wave_barrier();
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
        k_radix_hist<<<nb, 64, 0, s>>>(ki, n, shift, nb, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = exclusive_scan_u32(hist, hist, nullptr, (size_t)256 * nb, tmp, s);
        if (e != hipSuccess) return e;
        k_radix_scatter<<<nb, 64, 0, s>>>(ki, vi, ko, vo, n, shift, nb, hist);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        std::swap(ki, ko);
        std::swap(vi, vo);
        *in_first = !*in_first;
    }
    return hipSuccess;
}

}  // namespace tlr
