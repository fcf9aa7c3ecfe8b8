// pffrg_chi.h -- correlation kernels (compute_χ!): arguments shared by the host and chi.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pffrg_device.h"

namespace pffrg {

constexpr long CHI_MAX_EVALS = 10000000;                       // maxevals of integrate_χ_boxes
constexpr long CHI_MAX_REGIONS = 1 + (CHI_MAX_EVALS - 17) / 34;  // regions that budget allows
constexpr int CHI_T = 256;                                       // threads per cubature task
constexpr int CHI_RPB = 7;                                       // selected regions per split block
constexpr int CHI_SPLIT_T = 256;                                 // >= 34 CHI_RPB threads

// one problem of a χ computation (a context's action, meshes and χ frequencies)
struct ChiProb {
    Sys y;
    const double *mesh;  // σ | Ωs | νs | Ωt | νt, MAXN each
    ActV<double> A;      // the action (reference layout)
    const int *exch;
    const double *chi;   // χ frequencies [nchi]
    double L;
    int nchi;
    int task0;           // first global task: local task (w * ns + site) * 9 + box
    int prop0;           // first global χ frequency
    int comp;            // U(1)-DM: χxx (0), χzz (1), χxy (2); the SU(2) kernel ignores it
};

struct ChiArgs {
    const ChiProb *probs;   // [problems]
    const int *task_prob;   // [global task] problem
    const int *prop_prob;   // [global frequency] problem
    int nfreq;              // global χ frequencies (k_chi_prop blocks)
    double atol, rtol;
    long cap;            // region capacity of one task slot
    int nslots;          // task slots of this pass
    const int *tasks;    // [nslots] global task of each slot
    const int *active;   // [unfinished slots] slot of each k_chi_round block
    const int *split_tab;  // [2 * blocks] (slot, first of CHI_RPB selected regions) of each k_chi_split block
    // per slot: regions so far, regions selected for splitting this round, finished flag
    long *rn, *rk;
    int *rdone;
    // cubature regions of every slot, [slot][cap]
    double *rc0, *rc1, *rh0, *rh1, *rval, *rerr;
    int *rsplit, *rsel;
    double *box;         // [global task] integral of the task's box
    int *nreg;           // [global task] final region count
    int *status;         // [global task] 0 done, 1 the slot's capacity was too small
    double *prop;        // [global frequency] propagator term of site 0
};

// first regions of the nslots slots; one round of n_active unfinished slots; the splits of a
// round over n_blocks blocks of CHI_RPB selected regions; the site-0 propagator term
hipError_t launch_chi_init(const ChiArgs &Q, hipStream_t stream);
hipError_t launch_chi_round(const ChiArgs &Q, int n_active, hipStream_t stream);
hipError_t launch_chi_split(const ChiArgs &Q, int n_blocks, hipStream_t stream);
hipError_t launch_chi_prop(const ChiArgs &Q, hipStream_t stream);

}  // namespace pffrg
