// pffrg_probe.h -- the gather-ceiling probe (gather_probe.hip): random trilinear q3 lookups
// on a slot's gather copy with the vertex pass's lane layout.  Measurement only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pffrg_device.h"

namespace pffrg {

// n_blocks waves of 64 lanes, n_iter lookups per group with ν′ < planes (1..n_nu-1); out
// receives n_blocks * 64 floats.  lds_bytes of dynamic LDS per wave only caps how many waves
// share a CU.
hipError_t launch_gather_probe(const void *g, bool f32, const Sys &y, int planes, int n_blocks, size_t lds_bytes,
                               int n_iter, uint32_t seed, float *out, hipStream_t stream);

}  // namespace pffrg
