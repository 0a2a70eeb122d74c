// pd_noalloc.hip -- the fixed-tree sum of pd_reduce_sum without allocation,
// for callers that cannot call hipMalloc (code under HIP graph capture,
// serving loops, harnesses that forbid allocation).
//
// The block partials live in a file-scope device array, `scratch`, that the
// library holds on every GPU it runs on until the process ends.  Its device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) and kept; kernels get that address, never the array's name, which
// in host code is only a host-side stand-in.  Every call on one GPU shares
// the array: calls that overlap on one GPU (two streams at once) race.
#include "pd_internal.h"

#include <mutex>

namespace {

constexpr int kMaxBlocks = 4096;  // block partials the scratch holds: lanes up to 4096 * 256
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kMaxBlocks];

std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// The scratch array's address on `device` (looked up on the first call).
hipError_t scratch_ptr(int device, double **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        e = hipSetDevice(device);
        if (e != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        hipError_t e2 = hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (e2 != hipSuccess) return e2;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[device] = static_cast<double *>(p);
    }
    *out = g_ptr[device];
    return hipSuccess;
}

} // namespace

extern "C" int pd_reduce_sum_noalloc(int device, const double *values, int64_t n, int lanes, double *out,
                                     void *stream)
{
    if (!out || n < 0 || (n > 0 && !values) || lanes < 64 || (lanes & (lanes - 1)) != 0) return PD_E_INVALID;
    const int threads = lanes >= 256 ? 256 : lanes;
    if (lanes / threads > kMaxBlocks) return PD_E_CAPACITY;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) return PD_E_HIP;
    if (device < 0 || device >= ndev || device >= kMaxDevices) return PD_E_INVALID;
    double *partials = nullptr;
    if (scratch_ptr(device, &partials) != hipSuccess) return PD_E_HIP;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return PD_E_HIP;
    if (cur != device && hipSetDevice(device) != hipSuccess) return PD_E_HIP;
    hipError_t e = launch_field_sum_to(static_cast<hipStream_t>(stream), lanes / threads, threads, values, n,
                                       partials, out);
    if (cur != device) (void)hipSetDevice(cur);
    return e == hipSuccess ? PD_OK : PD_E_HIP;
}
