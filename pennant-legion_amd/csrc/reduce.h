// The dt and dvol reductions (calcDtNew, calcDvol) as two launches each: a
// bounded grid of at most PN_RED_BLOCKS blocks writing one partial per block
// (wave64 shuffles, then the block), and a one-block fold of the partials.
// Included by every source file that launches them; each file gets its own
// copy of the kernels.
#pragma once

#include "common.h"

namespace {

using pn::block_max;
using pn::block_min;
using pn::smax;

// calcDtNewTask, src/Hydro.cc:1591-1619: per-block partial minima
__global__ void __launch_bounds__(PN_BLOCK) k_calc_dt_new(int numz, const double* zdl,
                                                          const double* zdu, const double* zss,
                                                          double cfl, double* part) {
    const double fuzz = 1.e-99;
    double acc = 1.e99;
    for (int z = blockIdx.x * PN_BLOCK + threadIdx.x; z < numz; z += gridDim.x * PN_BLOCK) {
        const double cdu = smax(zdu[z], smax(zss[z], fuzz));
        const double zdthyd = zdl[z] * cfl / cdu;
        acc = (zdthyd < acc) ? zdthyd : acc;
    }
    acc = block_min(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(PN_BLOCK) k_min_parts(const double* part, int nparts,
                                                        double* out) {
    double acc = 1.e99;
    for (int i = threadIdx.x; i < nparts; i += PN_BLOCK) acc = (part[i] < acc) ? part[i] : acc;
    acc = block_min(acc);
    if (threadIdx.x == 0) *out = acc;
}

// calcDvolTask, src/Hydro.cc:1622-1644
__global__ void __launch_bounds__(PN_BLOCK) k_calc_dvol(int numz, const double* zvol,
                                                        const double* zvol0, double* part) {
    double acc = 1.e-99;
    for (int z = blockIdx.x * PN_BLOCK + threadIdx.x; z < numz; z += gridDim.x * PN_BLOCK) {
        const double zdvov = fabs((zvol[z] - zvol0[z]) / zvol0[z]);
        acc = (zdvov > acc) ? zdvov : acc;
    }
    acc = block_max(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(PN_BLOCK) k_max_parts(const double* part, int nparts,
                                                        double* out) {
    double acc = 1.e-99;
    for (int i = threadIdx.x; i < nparts; i += PN_BLOCK) acc = (part[i] > acc) ? part[i] : acc;
    acc = block_max(acc);
    if (threadIdx.x == 0) *out = acc;
}

}  // namespace
