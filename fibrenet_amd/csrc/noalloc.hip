// noalloc.hip - a fixed device workspace for callers that cannot allocate.
//
// fn_group_force_noalloc runs fn_group_force's reduction with its index list,
// network offsets, partial sums and result in `scratch`, a file-scope device
// array of kScratchBytes, instead of buffers the engine allocates.  Each GPU
// the library runs on holds its own copy of the array until the process
// ends.  Host code never passes the array's name to a kernel or a copy: its
// device address, looked up once per device with that device current, is
// what the launches and copies take.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.h"

namespace {

constexpr size_t kIdxBytes = sizeof(int) * (size_t)fnk::kScratchGroup;
constexpr size_t kOffBytes = (sizeof(int) * (size_t)(fnk::kScratchNets + 1) + 255) / 256 * 256;
constexpr size_t kPartBytes = sizeof(double) * (size_t)fnk::kSumBlocks * fnk::kScratchNets;
constexpr size_t kOutBytes = sizeof(double) * (size_t)fnk::kScratchNets;
constexpr size_t kScratchBytes = kIdxBytes + kOffBytes + kPartBytes + kOutBytes;

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};

}  // namespace

// One copy per GPU, held until the process ends.  It is initialised data
// (one nonzero byte), so the code object carries it as file bytes and its
// loaded size equals its size in the file.
static __device__ __align__(256) unsigned char scratch[kScratchBytes] = {1};

namespace fnk {

size_t scratch_bytes() { return kScratchBytes; }

hipError_t scratch_workspace(int device, ScratchWorkspace *w)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    unsigned char *p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_ptr[device]) {
            int cur = -1;
            hipError_t e = hipGetDevice(&cur);
            if (e != hipSuccess) return e;
            if (cur != device) return hipErrorInvalidDevice;   // the caller makes the device current
            void *q = nullptr;
            e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return e;
            g_ptr[device] = static_cast<unsigned char *>(q);
        }
        p = g_ptr[device];
    }
    w->idx = reinterpret_cast<int *>(p);
    w->off = reinterpret_cast<int *>(p + kIdxBytes);
    w->part = reinterpret_cast<double *>(p + kIdxBytes + kOffBytes);
    w->out = reinterpret_cast<double *>(p + kIdxBytes + kOffBytes + kPartBytes);
    return hipSuccess;
}

}  // namespace fnk
