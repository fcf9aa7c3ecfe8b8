// pffrg_rescan.h -- mesh rescanning kernels (resample_from_to, Action.jl:483-557): the
// arguments shared by the host and rescan.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pffrg_device.h"

namespace pffrg {

constexpr int SCAN_BLOCKS = 64;  // partial-argmax blocks per (component, channel)
constexpr int SCAN_ROWS = 18;    // (component, channel) pairs of the largest action: 6 components x 3

// mesh-rescanning scratch: argmax partials (values, positions), SCAN_ROWS x 4 positions,
// SCAN_ROWS cut sets of n_Omega + 3 n_nu <= 4 MAXN values, the five new meshes
constexpr size_t SCAN_OFF_PIDX = SCAN_ROWS * SCAN_BLOCKS * 8;
constexpr size_t SCAN_OFF_IDX = SCAN_OFF_PIDX + SCAN_ROWS * SCAN_BLOCKS * 8;
constexpr size_t SCAN_OFF_CUTS = SCAN_OFF_IDX + 512;
constexpr size_t SCAN_OFF_MESH = SCAN_OFF_CUTS + SCAN_ROWS * 4 * MAXN * 8;
constexpr size_t SCAN_MEM_BYTES = SCAN_OFF_MESH + 5 * MAXN * 8;
// what one scan_cuts call uses of it (everything before the new meshes)
constexpr size_t SCAN_CUTS_BYTES = SCAN_OFF_MESH;

struct ScanArgs {
    Sys y;
    const double *gamma;  // the slot's vertex (reference layout)
    double *pval;         // [3 nc][SCAN_BLOCKS] partial maxima of |q3|
    long long *pidx;      // [3 nc][SCAN_BLOCKS] their column-major positions
    int *idx;             // [3 nc][4] (site, w, v, v') of the first maximum
    double *cuts;         // [3 nc][nO + 3 nn]
};

struct ResampleArgs {
    Sys y;
    const double *mold;  // old meshes σ | Ωs | νs | Ωt | νt, then their inverse spacings, MAXN each
    const double *mnew;  // new meshes σ | Ωs | νs | Ωt | νt, MAXN each
    const double *src_gamma, *src_sigma;
    double *dst_gamma, *dst_sigma;
};

hipError_t launch_scan_cuts(const ScanArgs &A, hipStream_t stream);

// Device address of this GPU's static scan_cuts scratch (SCAN_CUTS_BYTES, one per device for the
// life of the process), looked up once per device with that device current.
hipError_t scan_scratch_address(int device, void **ptr);
hipError_t launch_resample(const ResampleArgs &A, hipStream_t stream);

}  // namespace pffrg
