// Allocation-free calcDtNew for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It launches the same two kernels as pennant_hip_calc_dt_new,
// but keeps their per-block partials in a file-scope device array instead of
// a caller workspace.
//
// The array is per device: each GPU the library runs on holds its own copy
// until the process ends, at its own address.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the device
// symbol); it passes the address hipGetSymbolAddress returned for the device
// that is current, looked up once per device on the first call there.
//
// Every call on one device shares that device's array: calls that overlap on
// one GPU (different streams, or threads) race on it and must be ordered by
// the caller.  Each launch writes every partial it then reads, so nothing in
// the array needs clearing between calls.
#include <mutex>

#include "common.h"
#include "reduce.h"

namespace {

// One partial per block of the reduction grid, which never exceeds
// PN_RED_BLOCKS blocks whatever numz is (the kernel is grid-stride).
constexpr int kScratchParts = PN_RED_BLOCKS;
static __device__ __align__(256) double scratch[kScratchParts];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
double* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int device_scratch(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return PENNANT_EINVAL;
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return PENNANT_EINVAL;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" int pennant_hip_calc_dt_new_noalloc(const pennant_mesh* m, const double* zdl,
                                               const double* zdu, const double* zss, double cfl,
                                               double* dtnew, hipStream_t st) {
    if (!m || !dtnew || m->numz < 0) return PENNANT_EINVAL;
    const int nb = pn::nred_blocks(m->numz);
    if (nb > kScratchParts) return PENNANT_EINVAL;   // cannot happen: the grid is capped
    double* part = nullptr;
    const int rc = device_scratch(&part);
    if (rc) return rc;
    hipLaunchKernelGGL(k_calc_dt_new, dim3(nb), dim3(PN_BLOCK), 0, st, m->numz, zdl, zdu, zss,
                       cfl, part);
    hipLaunchKernelGGL(k_min_parts, dim3(1), dim3(PN_BLOCK), 0, st, part, nb, dtnew);
    return (int)hipGetLastError();
}
