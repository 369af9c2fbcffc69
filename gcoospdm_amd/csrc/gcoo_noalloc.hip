// gcoo_noalloc.hip -- the FP32 multiply for callers that cannot allocate.
//
// gcoo_spdm_f32_async keeps the scratch of its kernels in the handle,
// allocated on first use: the panel lists of the LDS-staged multiply (built
// once per conversion) and the long-row list and counter.  Callers that
// cannot call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid it) use gcoo_spdm_f32_noalloc_async instead: the
// same kernels and the same bits, with that scratch carved out of one
// file-scope device array per GPU, `scratch`, of kNoAllocScratchBytes.
//
// Host code never passes `scratch` itself to a kernel (in host code the name
// is a host-side stand-in, not a device address): the array's address on
// each device is looked up once, with that device current, through
// hipGetSymbolAddress and kept per device; only that address is used.
// Every call on a GPU shares that GPU's array and rebuilds what it needs in
// it (the kernels clear their own counters), so calls that overlap on one
// GPU -- concurrent streams or host threads -- race and must be serialised
// by the caller.
#include "gcoo_internal.h"

#include <mutex>

namespace {

constexpr size_t kNoAllocScratchBytes = 16u << 20;  // 16 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per
// device).
int scratch_address(void **out)
{
    int dev = 0;
    GCOO_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return gcoo::set_error(GCOO_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            return gcoo::set_error(GCOO_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                                   hipGetErrorString(e));
        }
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return GCOO_OK;
}

}  // namespace

using namespace gcoo;

extern "C" {

int gcoo_spdm_f32_noalloc_async(gcoo_handle h, const float *D, int K, int N, float *C,
                                gcoo_stream s)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    if (!h->converted) return set_error(GCOO_ERR_STATE, "matrix not converted");
    if (K != h->K)
        return set_error(GCOO_ERR_INVALID, "dense operand has %d rows, sparse has %d columns",
                         K, h->K);
    if (N < 0) return set_error(GCOO_ERR_INVALID, "N = %d", N);
    if (N > 0 && h->M > 0 && ((!D && h->K > 0) || !C))
        return set_error(GCOO_ERR_INVALID, "null operand");
    if ((long long)h->M * N > (long long)INT32_MAX * 64)
        return set_error(GCOO_ERR_OVERFLOW, "output too large");
    int dev = 0;
    GCOO_HIP_TRY(hipGetDevice(&dev));
    if (dev != h->device)
        return set_error(GCOO_ERR_INVALID, "handle lives on device %d, current is %d",
                         h->device, dev);
    void *p = nullptr;
    int rc = scratch_address(&p);
    if (rc) return rc;
    const ExtScratch xs{static_cast<char *>(p), kNoAllocScratchBytes};
    return launch_spdm_f32_ext(h, D, N, C, &xs, reinterpret_cast<hipStream_t>(s));
}

}  // extern "C"
