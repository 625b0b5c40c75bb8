// fpcc_allpairs_device_noalloc: fpcc_allpairs_device with its workspace in a
// fixed device array instead of a caller's buffer, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops).
#include <stdio.h>
#include <map>
#include <mutex>
#include "../../include/fastpcc.h"
#include "engine.hpp"

using namespace fpcc;

namespace {

// The workspace: one array per device (every GPU the library runs on holds its
// own copy, at its own address, until the process ends).  In host code `scratch`
// names a host-side stand-in, not device memory: host code uses only the address
// hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchBytes = FPCC_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device, looked up on the first call
// made on that device and kept.
void *scratch_on_current_device()
{
    static std::mutex mu;
    static std::map<int, void *> addr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        fprintf(stderr, "fastpcc: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev, hipGetErrorString(e));
        return nullptr;
    }
    addr[dev] = ptr;
    return ptr;
}

void tables_message()
{
    fprintf(stderr, "fastpcc: fpcc_allpairs_device_noalloc: this job's FFT plans or tables are not built on this "
                    "device yet (one fpcc_allpairs_device call of a job of the same shape builds them)\n");
}

}  // namespace

extern "C" int fpcc_allpairs_device_noalloc(const fpcc_job *job, const float *d_x, float *d_y, void *stream)
{
    NoBuildScope nobuild;  // constant tables are looked up, never built, in this call
    const size_t need = fpcc_workspace_size(job);
    if (need == 0) {
        if (tables_missing()) {
            tables_message();
            return -5;
        }
        // an invalid job: fpcc_allpairs_device returns its code (-1, -3) from the
        // job check, before it reads the workspace or launches anything
        return fpcc_allpairs_device(job, d_x, d_y, nullptr, 0, stream);
    }
    if (need > kScratchBytes) {
        fprintf(stderr,
                "fastpcc: fpcc_allpairs_device_noalloc: the job needs %zu bytes of workspace, more than the "
                "%zu-byte scratch\n",
                need, kScratchBytes);
        return -6;
    }
    if (!d_x || !d_y) return -1;
    void *ptr = scratch_on_current_device();
    if (!ptr) return -1;
    const int r = fpcc_allpairs_device(job, d_x, d_y, ptr, kScratchBytes, stream);
    if (r && tables_missing()) {
        tables_message();
        return -5;
    }
    return r;
}
