// Device memory of the allocation-free entry point (pf_peaks_noalloc).
//
// One file-scope device array per GPU, created when the code object is
// loaded on that GPU and held until the process ends.  Host code never uses
// the array's name as an address (on the host it names a stand-in, not
// device memory): pf_ctx_create resolves the device address once per device
// with hipGetSymbolAddress, with that device current, and every later use
// goes through the resolved pointer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "pf_kernels.hpp"

static __device__ __align__(256) unsigned char scratch[pf::kNoAllocScratchBytes];

namespace pf {

namespace {

struct ScratchSlot {
    bool resolved = false;
    hipError_t rc = hipSuccess;
    void* ptr = nullptr;
};

std::mutex g_scratch_mu;
std::vector<ScratchSlot> g_scratch;  // indexed by device ordinal

}  // namespace

hipError_t noalloc_scratch(int device, void** ptr) {
    *ptr = nullptr;
    if (device < 0) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (static_cast<size_t>(device) >= g_scratch.size()) g_scratch.resize(device + 1);
    ScratchSlot& s = g_scratch[device];
    if (!s.resolved) {
        void* p = nullptr;
        s.rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        s.ptr = (s.rc == hipSuccess) ? p : nullptr;
        s.resolved = true;
    }
    *ptr = s.ptr;
    return s.rc;
}

}  // namespace pf
