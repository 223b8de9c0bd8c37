// noalloc.hip -- device scratch for the entry point that allocates nothing.
//
// bfm_run_simulation_on_gpu_noalloc stages its copy-back (the unsorted,
// unwrapped positions the unwrap kernel writes) in this file-scope array
// instead of a buffer the updater allocates on its first copy-back.  Each GPU
// that loads the library holds its own copy until the process ends.  In host
// code the array's name stands for no device address: its address on a
// device comes from hipGetSymbolAddress with that device current, taken once
// per device and kept.
#include "noalloc.hpp"

#include <mutex>
#include <stdexcept>

#include "updater.hpp"

namespace pscbfm {
namespace {

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void *g_scratch[kMaxDevices] = {};

}  // namespace

void *noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices)
        throw std::invalid_argument("no-allocation scratch: device id out of range");
    std::lock_guard<std::mutex> guard(g_lock);
    if (!g_scratch[device]) {
        int current = -1;
        hip_check(hipGetDevice(&current), "hipGetDevice");
        if (current != device)
            hip_check(hipSetDevice(device), "hipSetDevice");
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (current != device)
            hipSetDevice(current);
        hip_check(e, "hipGetSymbolAddress (no-allocation scratch)");
        if (!p)
            throw std::runtime_error("hipGetSymbolAddress returned no address for the no-allocation scratch");
        g_scratch[device] = p;
    }
    return g_scratch[device];
}

}  // namespace pscbfm
