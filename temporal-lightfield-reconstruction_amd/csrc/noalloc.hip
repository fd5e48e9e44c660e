// The allocation-free entry point's device scratch: a file-scope array in
// device memory.  In host code the array's name is a host-side stand-in, not
// a device address, so its address is resolved once per device with
// hipGetSymbolAddress and only that address is ever handed to kernels or
// runtime calls.
#include <mutex>
#include <vector>

#include "noalloc.hpp"

namespace tlr {

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

hipError_t noalloc_scratch(int device, char** ptr) {
    static std::mutex mu;
    static std::vector<char*> addr;   // per device ordinal
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return hipErrorInvalidDevice;
    if ((size_t)device >= addr.size()) addr.resize((size_t)device + 1, nullptr);
    if (!addr[device]) {
        char* p = nullptr;
        hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

}  // namespace tlr
