// fs_noalloc.hip -- the fixed device workspace of the allocation-free entry
// point (fs_solve_pressure_system_noalloc, include/fluidsim_noalloc.h).
#include <hip/hip_runtime.h>

#include "../../include/fluidsim_noalloc.h"
#include "fs_noalloc.hpp"

namespace fsk {

// One copy per GPU, held until the process ends and shared by every call on
// that GPU.  Its name is a host-side stand-in: host code uses only the device
// address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[FS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void* noalloc_scratch(int device, hipError_t* err) {
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    if (!g_scratch_ptr[device]) {
        // first call on this device (which is current): its copy's address
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = e != hipSuccess ? e : hipErrorInvalidSymbol;
            return nullptr;
        }
        g_scratch_ptr[device] = p;
    }
    *err = hipSuccess;
    return g_scratch_ptr[device];
}

}  // namespace fsk
