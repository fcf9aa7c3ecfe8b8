// pffrg_chi.h -- correlation kernels (compute_χ!): arguments shared by the host and chi.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pffrg_device.h"

namespace pffrg {

constexpr int CHI_MAX_REGIONS = 8192;  // cubature regions per (site, frequency, box) task

struct ChiArgs {
    Sys y;
    const double *mesh;  // σ | Ωs | νs | Ωt | νt, MAXN each
    ActV<double> A;      // the action (reference layout)
    const int *exch;
    const double *chi;   // χ frequencies [nchi]
    double L, atol, rtol;
    int nchi, cap;
    // cubature regions of every task, [task][cap]
    double *rc0, *rc1, *rh0, *rh1, *rval, *rerr;
    int *rsplit;
    double *box;         // [task] integral of the task's box
    int *nreg;           // [task] final region count
    double *prop;        // [nchi] propagator term of site 0
};

// tasks = nchi * ns * 9 (task = (w * ns + site) * 9 + box)
hipError_t launch_chi(const ChiArgs &Q, hipStream_t stream);

}  // namespace pffrg
