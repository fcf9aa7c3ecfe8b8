// Practical ceiling of the vertex pass's gathers: random trilinear q3 lookups on a slot's
// gather copy, with nothing else in the wave.  Measurement only (no reference counterpart).
//
// Every wave has the vertex pass's lane layout: G groups of GS = n_sites lanes, one site per
// lane.  Per iteration a group draws one random (Ω, ν, ν′) cell of the q3 block, away from the
// upper mesh edge so that all 8 corners exist, with ν′ below `planes` (all of them, or a few
// to confine the lookups to a footprint that stays in L2), and every lane loads the 8 corners of its site in
// all three channels (24 loads of one (spin, density) pair each, issued before any is used), as
// a vertex pass does for one vertex slot.  The loads are independent of each other and of the
// sum they feed, so with enough waves in flight the rate is what the L2 / Infinity Cache and
// HBM deliver for this access pattern.  The grid's occupancy can be capped with dynamic LDS so
// that the ceiling can also be read at the vertex pass's own occupancy.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pffrg_device.h"
#include "pffrg_probe.h"

namespace pffrg {
namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    // murmur3 finaliser
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

template <typename TV>
struct Pair;
template <>
struct Pair<float> {
    using T = float2;
};
template <>
struct Pair<double> {
    using T = double2;
};

// at most 4 waves per SIMD: room for all 24 loads of a lookup in flight at once
template <typename TV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_gather_probe(const TV *__restrict__ g, Sys y, int planes, int n_iter, uint32_t seed,
                                                     float *__restrict__ out) {
    using P2 = typename Pair<TV>::T;
    extern __shared__ char occupancy_cap[];  // sized by the launch only to limit waves per CU
    (void)occupancy_cap;
    const int lane = threadIdx.x;
    const int grp = lane / y.GS, site = lane - grp * y.GS;
    const bool on = grp < y.G && site < y.ns;
    const long s1 = y.ns, s2 = (long)y.ns * y.nO, s3 = (long)y.ns * y.nO * y.nn;
    const P2 *p = reinterpret_cast<const P2 *>(g);
    const uint32_t stream = mix32(seed ^ (blockIdx.x * 64u + (uint32_t)grp) * 0x9e3779b9u);
    float acc = 0.f;
    if (on) {
        for (int it = 0; it < n_iter; ++it) {
            const uint32_t h0 = mix32(stream + 2u * (uint32_t)it), h1 = mix32(stream + 2u * (uint32_t)it + 1u);
            const int w = (int)((h0 & 0xffffu) % (uint32_t)(y.nO - 1));
            const int v = (int)((h0 >> 16) % (uint32_t)(y.nn - 1));
            const int vp = (int)((h1 & 0xffffu) % (uint32_t)planes);
            const long cell = y.q3 + site + s1 * w + s2 * v + s3 * vp;
            P2 c[3][8];
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const long b = ch * y.chlen + cell;
#pragma unroll
                for (int k = 0; k < 8; ++k) c[ch][k] = p[b + ((k & 1) ? s1 : 0) + ((k & 2) ? s2 : 0) + ((k & 4) ? s3 : 0)];
            }
            // pairwise sums, so that no load has to wait for an earlier one's use
            TV s[3];
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                TV q[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    q[k] = (c[ch][2 * k].x + c[ch][2 * k].y) + (c[ch][2 * k + 1].x + c[ch][2 * k + 1].y);
                s[ch] = (q[0] + q[1]) + (q[2] + q[3]);
            }
            acc += (float)((s[0] + s[1]) + s[2]);
        }
    }
    out[blockIdx.x * 64 + lane] = acc;
}

}  // namespace

hipError_t launch_gather_probe(const void *g, bool f32, const Sys &y, int planes, int n_blocks, size_t lds_bytes,
                               int n_iter, uint32_t seed, float *out, hipStream_t stream) {
    if (planes < 1 || planes > y.nn - 1) return hipErrorInvalidValue;
    if (f32)
        hipLaunchKernelGGL(k_gather_probe<float>, dim3(n_blocks), dim3(64), lds_bytes, stream, (const float *)g, y,
                           planes, n_iter, seed, out);
    else
        hipLaunchKernelGGL(k_gather_probe<double>, dim3(n_blocks), dim3(64), lds_bytes, stream, (const double *)g, y,
                           planes, n_iter, seed, out);
    return hipGetLastError();
}

}  // namespace pffrg
