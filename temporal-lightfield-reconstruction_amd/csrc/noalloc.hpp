// Fixed device scratch of the allocation-free reconstruction entry point
// (tlr_reconstruct_noalloc).  One array per GPU, in device memory for the
// life of the process; every call on a GPU shares it.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "query.hpp"

namespace tlr {

// Output samples per band: the image is reconstructed in row bands of at
// most this many output samples, so any image size fits the fixed array.
constexpr size_t kNoAllocBandQueries = (size_t)1 << 20;
// Resident waves of the general kernel (each holds kWaveScratchWords).
constexpr int kNoAllocWaves = 64;
// Largest output spp (the pattern area holds 64 * this many templates).
constexpr int kNoAllocMaxM = 256;

// Byte offsets inside the array.
constexpr size_t kNoAllocCounters = 0;                                  // next, fail count, pad, stats
constexpr size_t kNoAllocPattern = 256;                                 // 64 * M * 5 floats
constexpr size_t kNoAllocQcol = kNoAllocPattern + (((size_t)64 * kNoAllocMaxM * 5 * 4 + 255) & ~(size_t)255);
constexpr size_t kNoAllocFail = kNoAllocQcol + kNoAllocBandQueries * 16;
constexpr size_t kNoAllocWave = kNoAllocFail + kNoAllocBandQueries * 4;
constexpr size_t kNoAllocBytes = kNoAllocWave + (size_t)kNoAllocWaves * kWaveScratchWords * 4;
static_assert(kNoAllocWave % 256 == 0, "wave scratch is 256-byte aligned");
static_assert(8 + 2 * kNumStats <= 64, "counters and stats fit the first 256 bytes");

// Device address of this GPU's array (device must be current).  The first
// call per device resolves the symbol; later calls return the kept address.
hipError_t noalloc_scratch(int device, char** ptr);

}  // namespace tlr
