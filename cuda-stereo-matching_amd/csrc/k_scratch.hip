// k_scratch.hip -- the device scratch of the entry point that allocates
// nothing (elas_hip_process_device_noalloc).
//
// One file-scope array per GPU, in the code object's zero-filled data: every
// GPU the library runs on holds its own copy, at its own address, from the
// moment the library's code is loaded there until the process ends.  Host
// code never passes `scratch` itself to a kernel or a copy (on the host the
// name is only a stand-in for the symbol): elas_scratch_setup() looks up each
// device's address once, with that device current, and every use goes
// through that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "elas_launch.h"

static __device__ __align__(256) unsigned char scratch[ELAS_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_m;
void* g_scratch_ptr[ELAS_SCRATCH_DEVICES] = {};

}  // namespace

int elas_scratch_setup(int device, void** ptr) {
    if (device < 0 || device >= ELAS_SCRATCH_DEVICES || !ptr) return -1;
    std::lock_guard<std::mutex> g(g_scratch_m);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) {
            (void)hipGetLastError();
            return -1;
        }
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return -1;
        }
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return 0;
}
