// Device scratch of the allocation-free multiply (lfg_spmv_noalloc): the solver state
// slot, the completion counters and the per-block partials of the multiply's in-kernel
// p.q reduction, in one file-scope array per device instead of buffers the library
// allocates. Each device holds its own copy at its own address, which
// hipGetSymbolAddress gives once per device (with that device current) when a context
// is created there; host code uses only that address, never the array's name.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lfg_kernels.hpp"

namespace {

static __device__ __align__(256) unsigned char scratch[lfg::kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

}  // namespace

namespace lfg {

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess) return err;
    g_scratch[device] = ptr;
    return hipSuccess;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch[device];
}

}  // namespace lfg
