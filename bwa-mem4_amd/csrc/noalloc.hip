// Device scratch of the allocation-free entry point (bwg_extend_batch_noalloc).
//
// One file-scope device array per GPU, sized kNoAllocScratchBytes, holds everything a call needs
// beyond the caller's inputs and outputs: the staged arena, job descriptors and results, the bin
// order arrays and queue heads, the compacted queues of the ungapped pre-pass, and the plain
// kernel's per-job column state.  It lives as long as the process, so the entry point never calls
// hipMalloc or hipFree.
//
// In host code `scratch` names a host-side stand-in, not device memory: its device address is
// looked up once per device with hipGetSymbolAddress (noalloc_setup, called from bwg_create with
// that device current) and kept per device; host code uses only that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "bwg_runtime.h"

namespace bwg {
namespace {

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_setup_mu;
uint8_t *g_ptr[kMaxDevices] = {};
std::mutex g_call_mu[kMaxDevices];

}  // namespace

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return BWG_ENODEV;
    std::lock_guard<std::mutex> g(g_setup_mu);
    if (g_ptr[device]) return BWG_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the caller made `device` current
    if (e != hipSuccess || !p) return BWG_EIO;
    g_ptr[device] = (uint8_t *)p;
    return BWG_OK;
}

uint8_t *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_setup_mu);
    return g_ptr[device];
}

std::mutex &noalloc_call_mutex(int device) { return g_call_mu[device]; }

}  // namespace bwg
