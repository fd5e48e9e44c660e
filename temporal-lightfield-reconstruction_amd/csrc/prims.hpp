// Device-wide primitives used by the build: exclusive scan and a stable LSD
// radix sort of (u32 key, u32 value) pairs.  Hand-written for wave64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace tlr {

// Scratch allocator owned by a context: bump allocation inside one device
// buffer that grows on demand between builds (never during one).  A failed
// allocation is sticky: `failed` stays set until the next reset(), so a
// caller can check once after a group of allocations (a later, smaller one
// succeeding does not hide an earlier failure).
struct DeviceArena {
    char* base = nullptr;
    size_t cap = 0;
    size_t top = 0;
    bool failed = false;
    hipError_t reserve(size_t bytes);
    void reset() {
        top = 0;
        failed = false;
    }
    template <class T>
    T* alloc(size_t n) {
        size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
        if (failed || top + bytes > cap) {
            failed = true;
            return nullptr;
        }
        T* p = reinterpret_cast<T*>(base + top);
        top += bytes;
        return p;
    }
    void release();
};

// out[i] = sum_{j<i} in[i]; if total != nullptr, *total (device) = sum of all.
// `in` and `out` may alias.  tmp must hold scan_tmp_words(n) uint32.
size_t scan_tmp_words(size_t n);
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, size_t n, uint32_t* tmp,
                              hipStream_t s);

// Stable sort of n pairs by the low `bits` bits of key.  Ping-pongs between
// (k0,v0) and (k1,v1); returns in *result_in_first whether the sorted data
// ends in (k0,v0).  hist must hold radix_hist_words(n) uint32 and tmp
// scan_tmp_words(radix_hist_words(n)).
size_t radix_hist_words(size_t n);
hipError_t radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                            uint32_t* hist, uint32_t* tmp, bool* result_in_first, hipStream_t s);

}  // namespace tlr
