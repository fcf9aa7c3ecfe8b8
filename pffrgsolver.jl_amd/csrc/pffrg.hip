// pffrg.hip -- gfx950 kernels and C ABI of the pf-FRG flow right-hand side.
//
// One launch per pass evaluates every output element (channel, kernel, Ω, ν, ν')
// of the vertex flow: one 64-lane wavefront (= one workgroup) owns an element and
// splits its lanes into groups of n_sites lanes; each group runs the reference's
// adaptive Simpson + Richardson rule on one loop-frequency sub-interval at a time
// (lane = irreducible site), so the sub-intervals of an element are integrated
// concurrently and summed in the reference's order.  Meshes, Σ, the site exchange
// map and the t-channel overlap rows are staged in LDS; the vertex is gathered
// from a site-fastest structure-of-arrays buffer (FP64, or an FP32 copy in the
// mixed mode), so every trilinear corner is one coalesced n_sites-wide segment.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <array>
#include <vector>

#include "pffrg.h"
#include "pffrg_chi.h"
#include "pffrg_device.h"
#include "pffrg_pass.h"
#include "pffrg_probe.h"
#include "pffrg_rescan.h"
#include "pffrg_shared.h"

using namespace pffrg;

namespace {

// ======================================================================= tail corrections
struct CorrArgs {
    Sys y;
    const double *mesh;
    const double *sigA, *sigD;
    double *corrs;
    double L, atol, rtol;
    int kat;
};

// compute_corrs_kat! / compute_corrs! (Flow/bubbles.jl:99-220): one workgroup per (Ω, channel, side)
__global__ __launch_bounds__(64) void k_corrs(CorrArgs C) {
    __shared__ double s_mesh[5 * MAXN];
    __shared__ double s_sA[MAXN], s_sD[MAXN];
    __shared__ Seg segs[GK_MAX_SEGS];
    __shared__ double fv[64];
    __shared__ int ctl[4];
    __shared__ double ctld[2];
    const Sys &y = C.y;
    const int lane = threadIdx.x;
    for (int k = lane; k < 5 * MAXN; k += 64) s_mesh[k] = C.mesh[k];
    for (int k = lane; k < y.nsig; k += 64) {
        s_sA[k] = C.sigA[k];
        s_sD[k] = C.kat ? C.sigD[k] : 0.0;
    }
    __syncthreads();
    const int side = blockIdx.x & 1;
    const int ch = (blockIdx.x >> 1) % 3;
    const int wi = (blockIdx.x >> 1) / 3;
    const double W = (ch == 1) ? s_mesh[3 * MAXN + wi] : s_mesh[MAXN + wi];
    const double val = (ch == 1) ? y.valT : y.valS;
    const double *sm = s_mesh;
    const int nsig = y.nsig;
    const double L = C.L;
    const int kat = C.kat;
    auto f = [&](double v) -> double {
        double x1, x2;
        bubble_args(ch, W, v, x1, x2);
        return kat ? bubble_kat(L, x1, x2, sm, nsig, s_sA, s_sD) : bubble_plain(L, x1, x2, sm, nsig, s_sA);
    };
    const double bnd = f(side ? val : -val);
    double out = 0.0;
    if (fabs(bnd) > 1e-8) {
        const double bp[2] = {0.0, 1.0};
        const double I = gk_block(f, side ? 1 : 2, side ? val : -val, bp, 2, C.atol, C.rtol, segs, fv, ctl, ctld);
        out = I / bnd;
    }
    if (lane == 0) C.corrs[side + 2 * (ch + 3 * wi)] = out;
}

// ======================================================================= self energy
enum SigmaKind { SK_DSIGMA = 0, SK_CORR1 = 1, SK_CORR2 = 2, SK_SDE = 3 };

struct SigmaArgs {
    Sys y;
    const double *mesh;
    const int *exch;
    const double *mult;
    ActV<double> A;   // flowing action (propagator; vertex for DSIGMA/CORR2)
    ActV<double> V;   // vertex source: da_Σ (CORR1), a2 (SDE)
    const double *sigV;  // Σ of da_Σ (CORR2)
    double *out;      // Σ output
    const double *add;   // CORR2: da_Σ.Σ added first
    double L, atol, rtol;
};

// compute_dΣ! / compute_dΣ_corr! / compute_Σ! (Flow/flow_self.jl, Flow/SDE.jl:168-174):
// one workgroup per σ point w_i, i >= 1; the integrand restates compute_dΣ_kernel(_corr1,
// _corr2) (flow_su2_self.jl, flow_u1_dm_self.jl) and compute_Σ_kernel (SDE_su2.jl:45-77,
// SDE_u1_dm.jl:45-91).
// Threads of a σ-point block: the rule nodes' site terms are spread over them (a bisection's 30
// nodes on 16 sites in one round).
constexpr int SIGMA_T = 512;

template <int KIND>
__device__ __forceinline__ void sigma_point(const SigmaArgs &Q, int si) {
    __shared__ double s_mesh[5 * MAXN];
    __shared__ double s_sA[MAXN], s_sV[MAXN];
    __shared__ int s_exch[64];
    __shared__ double s_mult[64];
    __shared__ Seg segs[GK_MAX_SEGS];
    __shared__ double fv[64], ts[64];
    __shared__ double s_terms[60 * 65];  // per rule node: the local term, then each site's term
    __shared__ int ctl[4];
    __shared__ double ctld[2];
    const Sys &y = Q.y;
    const int lane = threadIdx.x;
    for (int k = lane; k < 5 * MAXN; k += blockDim.x) s_mesh[k] = Q.mesh[k];
    for (int k = lane; k < y.nsig; k += blockDim.x) {
        s_sA[k] = Q.A.Sigma[k];
        s_sV[k] = (KIND == SK_CORR2) ? Q.sigV[k] : 0.0;
    }
    if (lane < y.ns) {
        s_exch[lane] = Q.exch[lane];
        s_mult[lane] = Q.mult[lane];
    }
    __syncthreads();
    const Meshes<> M{s_mesh, s_mesh + MAXN, s_mesh + 2 * MAXN, s_mesh + 3 * MAXN, s_mesh + 4 * MAXN};
    const double w = s_mesh[si];
    const double L = Q.L;
    // The integrand at loop frequency v is ((local - Σ_j term_j) · factor): piece(v, -1) is the local
    // term of site 0, piece(v, j) site j's term with its coefficient, factor(v) the propagators.
    // SU(2): local = 3 Γs + Γd, term_j = 2 m_j S² Γd_j (flow_su2_self.jl:2-101, SDE_su2.jl:45-77);
    // U(1)-DM: local = 2 Γxx + Γzz + Γdd, term_j = 2 m_j Γdd_j (flow_u1_dm_self.jl:22-28,
    // SDE_u1_dm.jl:68-76).
    auto piece = [&](double v, int j) -> double {
        if (y.nc == 6) {
            double l6[6];
            if (KIND == SK_SDE) {
                const double INF = __builtin_huge_val();
                const Buf<> e = buf_empty<double>();
                if (j < 0) {
                    const Buf<> b2s = buf_s(M, y, v + w, INF, 0.5 * (-v + w));
                    gamma_site_u1<true, false, false>(Q.V, b2s, e, e, 0, s_exch, y, l6);
                    return 2.0 * l6[U1_XX] + l6[U1_ZZ] + l6[U1_DD];
                }
                const Buf<> b1s = buf_s(M, y, v + w, INF, 0.5 * (v - w));
                gamma_site_u1<true, false, false>(Q.V, b1s, e, e, j, s_exch, y, l6);
                return 2.0 * s_mult[j] * l6[U1_DD];
            }
            if (j < 0) {
                const Buf<> b2s = buf_s(M, y, v + w, 0.5 * (-v + w), 0.5 * (-v + w));
                const Buf<> b2t = buf_t(M, y, v - w, 0.5 * (v + w), 0.5 * (v + w));
                const Buf<> b2u = buf_u(M, y, 0.0, w, v);
                if (KIND == SK_CORR1) gamma_site_u1<true, true, false>(Q.V, b2s, b2t, b2u, 0, s_exch, y, l6);
                else gamma_site_u1<true, true, true>(Q.A, b2s, b2t, b2u, 0, s_exch, y, l6);
                return 2.0 * l6[U1_XX] + l6[U1_ZZ] + l6[U1_DD];
            }
            const Buf<> b1s = buf_s(M, y, v + w, 0.5 * (-v + w), 0.5 * (v - w));
            const Buf<> b1t = buf_t(M, y, 0.0, w, v);
            const Buf<> b1u = buf_u(M, y, -v + w, 0.5 * (v + w), 0.5 * (v + w));
            if (KIND == SK_CORR1) gamma_site_u1<true, false, true>(Q.V, b1s, b1t, b1u, j, s_exch, y, l6);
            else gamma_site_u1<true, true, true>(Q.A, b1s, b1t, b1u, j, s_exch, y, l6);
            return 2.0 * s_mult[j] * l6[U1_DD];
        }
        double sj, dj;
        if (KIND == SK_SDE) {
            const double INF = __builtin_huge_val();
            const Buf<> e = buf_empty<double>();
            if (j < 0) {
                const Buf<> b2s = buf_s(M, y, v + w, INF, 0.5 * (-v + w));
                gamma_site<double, true, false, false>(Q.V, b2s, e, e, 0, s_exch, y, sj, dj);
                return 3.0 * sj + dj;
            }
            const Buf<> b1s = buf_s(M, y, v + w, INF, 0.5 * (v - w));
            gamma_site<double, true, false, false>(Q.V, b1s, e, e, j, s_exch, y, sj, dj);
            return 2.0 * s_mult[j] * y.S2 * dj;
        }
        if (j < 0) {
            const Buf<> b2s = buf_s(M, y, v + w, 0.5 * (-v + w), 0.5 * (-v + w));
            const Buf<> b2t = buf_t(M, y, v - w, 0.5 * (v + w), 0.5 * (v + w));
            const Buf<> b2u = buf_u(M, y, 0.0, w, v);
            if (KIND == SK_CORR1) gamma_site<double, true, true, false>(Q.V, b2s, b2t, b2u, 0, s_exch, y, sj, dj);
            else gamma_site<double, true, true, true>(Q.A, b2s, b2t, b2u, 0, s_exch, y, sj, dj);
            return 3.0 * sj + dj;
        }
        const Buf<> b1s = buf_s(M, y, v + w, 0.5 * (-v + w), 0.5 * (v - w));
        const Buf<> b1t = buf_t(M, y, 0.0, w, v);
        const Buf<> b1u = buf_u(M, y, -v + w, 0.5 * (v + w), 0.5 * (v + w));
        if (KIND == SK_CORR1) gamma_site<double, true, false, true>(Q.V, b1s, b1t, b1u, j, s_exch, y, sj, dj);
        else gamma_site<double, true, true, true>(Q.A, b1s, b1t, b1u, j, s_exch, y, sj, dj);
        return 2.0 * s_mult[j] * y.S2 * dj;
    };
    auto factor = [&](double v) -> double {
        if (KIND == SK_DSIGMA) return S_single(L, v, s_mesh, y.nsig, s_sA) / (2.0 * PI_D);
        if (KIND == SK_CORR1 || KIND == SK_SDE) return -G_full(L, v, s_mesh, y.nsig, s_sA) / (2.0 * PI_D);
        const double G = G_full(L, v, s_mesh, y.nsig, s_sA);
        return G * G * sigma_at(v, s_mesh, y.nsig, s_sV) / (2.0 * PI_D);
    };
    // the rule nodes' pieces over all threads, then each node's sum in site order by one thread,
    // mapped from t ∈ (-1, 1) to v = t/(1-t²) with the weight (1+t²)/(1-t²)²
    const int per = y.ns + 1;
    auto ev = [&](int n, const double *tq, double *fo) {
        // the n ns site terms first, then the n local terms, so that a wave's lanes take one path
        const int nsite = n * y.ns;
        for (int it = lane; it < nsite + n; it += blockDim.x) {
            const bool site = it < nsite;
            const int q = site ? it / y.ns : it - nsite, j = site ? it - q * y.ns : -1;
            const double t = tq[q];
            const double den = 1.0 / (1.0 - t * t);
            s_terms[q * per + 1 + j] = piece(t * den, j);
        }
        __syncthreads();
        if (lane < n) {
            const double t = tq[lane];
            const double t2 = t * t;
            const double den = 1.0 / (1.0 - t2);
            const double v = t * den;
            double val = s_terms[lane * per];
            for (int jj = 0; jj < y.ns; ++jj) val -= s_terms[lane * per + 1 + jj];
            val *= factor(v);
            fo[lane] = val * (1.0 + t2) * den * den;
        }
    };
    // breakpoints (-Inf, -2Λ, 0, 2Λ, Inf) in the t/(1-t²) variable
    double bp[5];
    bp[0] = -1.0;
    const double inner[3] = {-2.0 * L, 0.0, 2.0 * L};
    for (int k = 0; k < 3; ++k) bp[k + 1] = 2.0 * inner[k] / (1.0 + hypot(1.0, 2.0 * inner[k]));
    bp[4] = 1.0;
    const double I = gk_block_coop(ev, bp, 5, Q.atol, Q.rtol, segs, fv, ts, ctl, ctld);
    if (lane == 0) {
        if (KIND == SK_CORR2) {
            double o = Q.out[si];
            o += Q.add[si];
            o += I;
            Q.out[si] = o;
        } else {
            Q.out[si] = I;
        }
    }
}

template <int KIND>
__global__ __launch_bounds__(SIGMA_T) void k_sigma(SigmaArgs Q) {
    sigma_point<KIND>(Q, blockIdx.x + 1);
}

// The σ points of several problems in one launch: block b takes point pts[b] & 0xffff of problem
// pts[b] >> 16, whose arguments are tab[problem]; each point is computed exactly as k_sigma does.
template <int KIND>
__global__ __launch_bounds__(SIGMA_T) void k_sigma_batch(const SigmaArgs *__restrict__ tab, const uint32_t *__restrict__ pts) {
    const uint32_t pt = pts[blockIdx.x];
    sigma_point<KIND>(tab[pt >> 16], (int)(pt & 0xffffu));
}

// ======================================================================= elementwise action kernels
// The ν <-> ν' image of component comp in channel ch reads component sym_src with sign
// sym_sign: the identity for SU(2); for U(1)-DM DM flips and zd, dz swap with a sign in the s
// and t channels, zd and dz flip in the u channel (action_u1_dm.jl:170-339)
__device__ __forceinline__ int sym_src(int nc, int ch, int comp) {
    if (nc == 6 && ch != 2 && comp == U1_ZD) return U1_DZ;
    if (nc == 6 && ch != 2 && comp == U1_DZ) return U1_ZD;
    return comp;
}
__device__ __forceinline__ bool sym_neg(int nc, int ch, int comp) {
    if (nc != 6) return false;
    if (ch != 2) return comp == U1_DM || comp == U1_ZD || comp == U1_DZ;
    return comp == U1_ZD || comp == U1_DZ;
}

// symmetrize! (action_su2.jl:141-173, action_u1_dm.jl:170-215): q3[i,w,v,v'] = ±q3'[x(i),w,v',v]
// for v < v' (the upper triangle is written from lower-triangle entries only)
__global__ void k_symmetrize(double *g, Sys y, const int *exch) {
    const long per = (long)y.ns * y.nO * y.nn * y.nn;
    const long total = 3L * y.nc * per;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int cc = (int)(t / per);  // comp*3 + ch
        long r = t % per;
        const int i = (int)(r % y.ns); r /= y.ns;
        const int w = (int)(r % y.nO); r /= y.nO;
        const int v = (int)(r % y.nn);
        const int vp = (int)(r / y.nn);
        if (v >= vp) continue;
        const int ch = cc % 3, comp = cc / 3;
        double *q3 = g + comp * y.complen + ch * y.chlen + y.q3;
        const double *q3s = g + sym_src(y.nc, ch, comp) * y.complen + ch * y.chlen + y.q3;
        const int src = (ch == 2) ? i : exch[i];
        const double x = q3s[src + (long)y.ns * (w + (long)y.nO * (vp + (long)y.nn * v))];
        q3[i + (long)y.ns * (w + (long)y.nO * (v + (long)y.nn * vp))] = sym_neg(y.nc, ch, comp) ? -x : x;
    }
}

// symmetrize_add_to! (action_su2.jl:176-250, action_u1_dm.jl:218-339): a += a_l ± image of
// a_l under ν <-> ν'
__global__ void k_symmetrize_add(const double *__restrict__ l, double *__restrict__ a, Sys y, const int *exch) {
    const long total = y.nc * y.complen;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long cc = t / y.chlen;
        const int ch = (int)(cc % 3), comp = (int)(cc / 3);
        long r = t % y.chlen;
        const long cbase = (long)sym_src(y.nc, ch, comp) * y.complen + (long)ch * y.chlen;
        const int i = (int)(r % y.ns);
        const int ix = (ch == 2) ? i : exch[i];
        double img;
        if (r < y.q21) {
            img = l[cbase + ix + (r - i)];
        } else if (r < y.q22) {
            const long k = r - y.q21;
            img = l[cbase + y.q22 + ix + (k - i)];
        } else if (r < y.q3) {
            const long k = r - y.q22;
            img = l[cbase + y.q21 + ix + (k - i)];
        } else {
            long k = r - y.q3;
            k /= y.ns;
            const int w = (int)(k % y.nO); k /= y.nO;
            const int v = (int)(k % y.nn);
            const int vp = (int)(k / y.nn);
            img = l[cbase + y.q3 + ix + (long)y.ns * (w + (long)y.nO * (vp + (long)y.nn * v))];
        }
        a[t] += sym_neg(y.nc, ch, comp) ? l[t] - img : l[t] + img;
    }
}

__global__ void k_axpy(const double *__restrict__ x, double fac, double *__restrict__ y, long n) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        y[t] += fac * x[t];
}

__global__ void k_scale(double *x, double fac, long n) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) x[t] *= fac;
}

// gather copy: y[nc t + comp] = x[comp * complen + t]
template <typename T>
__global__ void k_interleave(const double *__restrict__ x, T *__restrict__ y, long complen, int nc) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < complen; t += (long)gridDim.x * blockDim.x) {
        if (nc == 2) {
            y[2 * t] = (T)x[t];
            y[2 * t + 1] = (T)x[complen + t];
        } else {
            for (int k = 0; k < nc; ++k) y[(long)nc * t + k] = (T)x[k * complen + t];
        }
    }
}

// max |q3| over the q3 blocks of every component and channel (channel.jl:406-413): block max
// then 64-bit atomic max on the IEEE bit pattern, which orders like the value for non-negative
// doubles.
__global__ void k_absmax_q3(const double *g, Sys y, unsigned long long *out) {
    __shared__ double red[256];
    const long per = (long)y.ns * y.nO * y.nn * y.nn;
    double m = 0.0;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < 3L * y.nc * per; t += (long)gridDim.x * blockDim.x) {
        const long cc = t / per;
        const double v = fabs(g[(cc / 3) * y.complen + (cc % 3) * y.chlen + y.q3 + (t % per)]);
        m = (v > m || v != v) ? v : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

}  // namespace

// ======================================================================= host side
// One action on the device: Σ, bare and the vertex in the reference's layout, plus the
// gather copies read by the vertex passes, built on demand: the same vertex with the two
// SU(2) components of each entry stored next to each other, in FP64 and in FP32.
struct Slot {
    double *Sigma = nullptr, *bare = nullptr, *gamma = nullptr;
    double *gat64 = nullptr;
    float *gat32 = nullptr;
    bool gat64_valid = false, gat32_valid = false;
};

struct pffrg_ctx {
    pffrg_dims dims{};
    Sys sys{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<Slot> slots;
    double *d_mesh = nullptr;
    bool have_mesh = false, have_lattice = false;
    int *d_exch = nullptr, *d_ov_off = nullptr;
    double *d_mult = nullptr;
    uint32_t *d_ov_rows = nullptr;
    int n_ov = 0;
    double *d_corrs = nullptr;
    unsigned long long *d_count = nullptr;
    unsigned long long *d_max = nullptr;
    uint32_t *d_el_half = nullptr, *d_el_full = nullptr, *d_el_red = nullptr;
    int n_half = 0, n_full = 0, n_red = 0;
    // shared-node pass (FP32-mixed): work units per element-list mode, tile shape, partial sums
    uint32_t *d_units[3] = {nullptr, nullptr, nullptr};
    int n_units[3] = {0, 0, 0};
    // dispatch order of the units per (pass, element-list mode): the list the last launch ran and
    // each unit's run time there (a prior before the first launch); the next launch of that pass
    // and mode dispatches its units longest first, from the other list
    uint32_t *d_ord[9][2] = {};
    uint32_t *d_cost[9] = {};
    int ord_cur[9] = {};
    int sh_nw = 1, sh_spt = 1, sh_kb = 8, sh_kmax = SH_KMAX;
    bool sh_shaped = false;  // the shape above is set (defaults or the autotuner's choice)
    double *d_part = nullptr;
    size_t part_bytes = 0;
    int precision = PFFRG_FP64;
    long long stat_elems = 0, stat_launches = 0;
    bool timing = false;
    double time_ms = 0.0;
    long long time_launches = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // recording of per-sub-interval node counts of the next vertex pass (diagnostics)
    int32_t *h_dbg_nodes = nullptr;
    double *h_dbg_res = nullptr;
    int *d_dbg_nodes = nullptr;
    double *d_dbg_res = nullptr;
    long dbg_cap = 0;
    // correlation cubature scratch (regions of every task), grown on demand
    void *chi_mem = nullptr;
    // mesh-rescanning scratch (argmax partials, cuts, new meshes), SCAN_MEM_BYTES
    char *scan_mem = nullptr;
    // this device's static scan_cuts scratch (pffrg_scan_cuts_noalloc), looked up at creation
    char *scan_scratch = nullptr;
    size_t chi_bytes = 0;
    // host copies of the meshes and the log-domain breakpoint table
    std::vector<double> h_Os, h_Ot;
    double *d_logpts = nullptr;
    long logpts_cap = 0;
    std::vector<double> h_logpts;
    bool logpts_valid = false;
    double logpts_L = 0.0;
    int logpts_eval = 0;
};

// Several contexts on one device whose 1-loop right-hand sides run as one vertex-pass launch.
struct pffrg_batch {
    std::vector<pffrg_ctx *> ctx;  // in launch order: larger problems first
    std::vector<int> order;        // order[k] = caller's index of ctx[k]
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint32_t *d_codes = nullptr;   // half-pass elements of all problems, problem index in bits 26-31
    int n_codes = 0;
    std::vector<size_t> layout_lds;  // per-problem LDS (launch order) the launch layout was made for
    int n_codes_full = 0;            // leading codes of the problems whose blocks fit 12 per CU
    void *d_tab = nullptr;         // PassArgs of every problem
    size_t tab_bytes = 0;
    void *chi_mem = nullptr;       // correlation cubature scratch of batched χ
    size_t chi_bytes = 0;
    // shared-node pass (FP32-mixed): the work units of all problems, problem index in bits 26-31
    uint32_t *d_sh_units = nullptr;
    int n_sh_units = 0;
    uint32_t *d_sh_units2 = nullptr;  // the other unit list: launches alternate, dispatching longest first
    uint32_t *d_sh_cost = nullptr;    // each unit's run time at its place in the list last run
    void *d_sh_tab = nullptr;
    std::vector<int> sh_spts;                 // sites per wave of each problem when the list was made
    std::vector<std::array<int, 3>> sh_groups;  // (sites per wave, first unit, units) per launch
    // dΣ of all problems in one launch: the σ points (problem << 16 | point) and the arguments
    uint32_t *d_sig_pts = nullptr;
    int n_sig_pts = 0;
    void *d_sig_tab = nullptr;
};

namespace {

int fail(pffrg_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                          \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) return fail((c), PFFRG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

int check_slot(pffrg_ctx *c, int s) {
    if (!c) return PFFRG_ERR_ARG;
    if (s < 0 || s >= (int)c->slots.size()) return fail(c, PFFRG_ERR_ARG, "slot index out of range: " + std::to_string(s));
    return 0;
}

int ready(pffrg_ctx *c) {
    if (!c) return PFFRG_ERR_ARG;
    if (!c->have_mesh) return fail(c, PFFRG_ERR_STATE, "meshes not set (pffrg_set_mesh)");
    if (!c->have_lattice) return fail(c, PFFRG_ERR_STATE, "lattice not set (pffrg_set_lattice)");
    return 0;
}

uint32_t enc(int ch, int kernel, int w1, int w2, int w3) {
    return (uint32_t)ch | ((uint32_t)kernel << 2) | ((uint32_t)w1 << 5) | ((uint32_t)(w2 < 0 ? 0 : w2) << 12) |
           ((uint32_t)(w3 < 0 ? 0 : w3) << 19);
}

// element lists of compute_Γ!/compute_dΓ_*! (compute.jl) and compute_Σ! (SDE.jl:156-166),
// channel-major then Ω so that each XCD slice covers neighbouring bosonic frequencies
std::vector<uint32_t> element_list(int nO, int nn, int mode) {
    std::vector<uint32_t> e;
    if (mode == 2) {
        for (int w1 = 0; w1 < nO; ++w1) {
            e.push_back(enc(0, 1, w1, -1, -1));
            for (int w3 = 0; w3 < nn; ++w3) e.push_back(enc(0, 3, w1, -1, w3));
        }
        return e;
    }
    for (int ch = 0; ch < 3; ++ch)
        for (int w1 = 0; w1 < nO; ++w1) {
            e.push_back(enc(ch, 1, w1, -1, -1));
            for (int w3 = 0; w3 < nn; ++w3) {
                e.push_back(enc(ch, 2, w1, w3, -1));
                e.push_back(enc(ch, 3, w1, -1, w3));
            }
            for (int w3 = 0; w3 < nn; ++w3)
                for (int w2 = (mode == 1 ? 0 : w3); w2 < nn; ++w2) e.push_back(enc(ch, 4, w1, w2, w3));
        }
    return e;
}

int grid_for(long n) {
    long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

int ensure_gather(pffrg_ctx *c, int s, bool f32) {
    Slot &sl = c->slots[s];
    const long n = c->sys.complen;
    const int nc = c->sys.nc;
    if (f32) {
        if (!sl.gat32) HIPCHK(c, hipMalloc(&sl.gat32, nc * n * sizeof(float)));
        if (!sl.gat32_valid) {
            hipLaunchKernelGGL(k_interleave<float>, dim3(grid_for(n)), dim3(256), 0, c->stream, sl.gamma, sl.gat32, n,
                               nc);
            HIPCHK(c, hipGetLastError());
            sl.gat32_valid = true;
        }
    } else {
        if (!sl.gat64) HIPCHK(c, hipMalloc(&sl.gat64, nc * n * sizeof(double)));
        if (!sl.gat64_valid) {
            hipLaunchKernelGGL(k_interleave<double>, dim3(grid_for(n)), dim3(256), 0, c->stream, sl.gamma, sl.gat64, n,
                               nc);
            HIPCHK(c, hipGetLastError());
            sl.gat64_valid = true;
        }
    }
    return 0;
}

void invalidate(pffrg_ctx *c, int s) { c->slots[s].gat64_valid = c->slots[s].gat32_valid = false; }

// Kernel arguments of one pass of one context (also rebuilds the log-domain breakpoints when
// Λ, eval or the meshes changed).
template <int PASS, typename TV>
int pass_args(pffrg_ctx *c, double L, int a, int d, int out, int eval, const double *tol, int mode, PassArgs<TV> &P,
              const double *dbg_v = nullptr, double *dbg_out = nullptr, uint32_t dbg_elem = 0) {
    P.dbg_v = dbg_v;
    P.dbg_out = dbg_out;
    P.dbg_elem = dbg_elem;
    P.dbg_sub_nodes = c->d_dbg_nodes;
    P.dbg_sub_res = c->d_dbg_res;
    P.dbg_cap = c->dbg_cap;
    P.y = c->sys;
    P.mesh = c->d_mesh;
    P.exch = c->d_exch;
    P.ov_off = c->d_ov_off;
    P.ov_rows = c->d_ov_rows;
    const Slot &SA = c->slots[a], &SD = c->slots[d];
    if (std::is_same<TV, float>::value) {
        P.A = ActV<TV>{(const TV *)SA.gat32, SA.bare, SA.Sigma};
        P.D = ActV<TV>{(const TV *)SD.gat32, SD.bare, SD.Sigma};
    } else {
        P.A = ActV<TV>{(const TV *)SA.gat64, SA.bare, SA.Sigma};
        P.D = ActV<TV>{(const TV *)SD.gat64, SD.bare, SD.Sigma};
    }
    P.out = c->slots[out].gamma;
    P.corrs = c->d_corrs;
    if (!(c->logpts_valid && c->logpts_L == L && c->logpts_eval == eval)) {
        // breakpoints ξ^k·2ref of the log domains (quadrature.jl:113-139), both mesh families;
        // rebuilt only when Λ, eval or the meshes change, after the stream has drained
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const int nO = c->sys.nO;
        const long need = 2L * nO * (eval + 1);
        if (need > c->logpts_cap) {
            HIPCHK(c, hipFree(c->d_logpts));
            c->d_logpts = nullptr;
            HIPCHK(c, hipMalloc(&c->d_logpts, need * sizeof(double)));
            c->logpts_cap = need;
        }
        c->h_logpts.assign(need, 0.0);
        for (int fam = 0; fam < 2; ++fam) {
            const std::vector<double> &Om = fam ? c->h_Ot : c->h_Os;
            const double val = fam ? c->sys.valT : c->sys.valS;
            for (int w1 = 0; w1 < nO; ++w1) {
                double *pt = c->h_logpts.data() + ((long)fam * nO + w1) * (eval + 1);
                const double lo = 2.0 * (L + 0.5 * Om[w1]), hi = 1.0 * val;
                if (!(hi > lo)) {
                    pt[0] = -1.0;
                    continue;
                }
                const double xi = std::pow(hi / lo, 1.0 / eval);
                for (int k = 0; k <= eval; ++k) pt[k] = std::pow(xi, k) * lo;
            }
        }
        HIPCHK(c, hipMemcpyAsync(c->d_logpts, c->h_logpts.data(), need * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        c->logpts_valid = true;
        c->logpts_L = L;
        c->logpts_eval = eval;
    }
    P.logpts = c->d_logpts;
    P.elems = (mode == 0) ? c->d_el_half : (mode == 1) ? c->d_el_full : c->d_el_red;
    P.n_elems = (mode == 0) ? c->n_half : (mode == 1) ? c->n_full : c->n_red;
    P.L = L;
    P.eval = eval;
    P.atol = tol[0];
    P.rtol = tol[1];
    P.node_count = c->d_count;
    return 0;
}

template <int PASS, typename TV>
size_t pass_lds_bytes(const pffrg_ctx *c) {
    return (size_t)pass_lds(c->sys, c->n_ov, (int)sizeof(DescT<TV>),
                            std::is_same<typename PassScalar<TV>::T, float>::value, PASS == P_KAT).total;
}

template <int PASS, typename TV>
int launch_pass_t(pffrg_ctx *c, double L, int a, int d, int out, int eval, const double *tol, int mode,
                  const double *dbg_v = nullptr, double *dbg_out = nullptr, int dbg_n = 0, uint32_t dbg_elem = 0) {
    PassArgs<TV> P;
    int rc;
    if ((rc = pass_args<PASS, TV>(c, L, a, d, out, eval, tol, mode, P, dbg_v, dbg_out, dbg_elem))) return rc;
    const size_t lds = pass_lds_bytes<PASS, TV>(c);
    const int grid = dbg_v ? dbg_n : P.n_elems;
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, (launch_vertex_pass<PASS, TV>(P, grid, lds, c->stream)));
    if (c->timing) {
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->time_ms += ms;
        c->time_launches += 1;
    }
    c->stat_elems += P.n_elems;
    c->stat_launches += 1;
    invalidate(c, out);
    if (c->d_dbg_nodes && !dbg_v) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(c->h_dbg_nodes, c->d_dbg_nodes, c->dbg_cap * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(c->h_dbg_res, c->d_dbg_res, c->dbg_cap * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipFree(c->d_dbg_nodes));
        HIPCHK(c, hipFree(c->d_dbg_res));
        c->d_dbg_nodes = nullptr;
        c->d_dbg_res = nullptr;
        c->dbg_cap = 0;
    }
    return 0;
}

// Launch shape of the shared-node pass: 8 waves, each holding the 16 x 16 tile on
// ceil(n_sites / 8) <= SH_SPT sites (with fewer than 8 sites some waves hold none and only help
// with the descriptors and interpolations).  Every problem has the same block, so a batched
// launch runs each problem exactly as its own launch does.  Problems with more sites use the
// per-element pass.
// the shared-node pass serves the SU(2) action up to 8 SH_SPT sites; U(1)-DM runs the
// per-element pass in FP32 in that mode
// the shape (waves per block, batch capacity, most nodes per batch) is feasible for the problem:
// at most SH_SPT sites per wave and the pass's LDS (tables, overlap rows, node states, the batch)
// within the 160 KB of a CU
bool shared_shape_ok(const pffrg_ctx *c, int nw, int kb, int kmax) {
    // kb >= 4: a batch of the initial Simpson rule pads its nodes to a group of four
    if (nw < 1 || nw > 8 || kb < 4 || kb > 8 || kmax < kb || kmax > SH_KMAX) return false;
    const int spt = (c->sys.ns + nw - 1) / nw;
    if (spt > SH_SPT) return false;
    return shared_lds(c->sys, c->n_ov, nw, true, spt, kb).total + SH_STATIC_LDS <= 160 * 1024;
}

bool shared_fits(const pffrg_ctx *c) {
    if (c->sys.nc != 2 || c->sys.ns > 8 * SH_SPT) return false;
    return shared_shape_ok(c, 8, shared_batch_nodes((c->sys.ns + 7) / 8), SH_KMAX);
}

void shared_set_shape(pffrg_ctx *c, int nw, int kb, int kmax) {
    c->sh_nw = nw;
    c->sh_spt = (c->sys.ns + nw - 1) / nw;
    c->sh_kb = kb;
    c->sh_kmax = kmax;
    c->sh_shaped = true;
}

void shared_shape(pffrg_ctx *c) {
    if (!c->sh_shaped || !shared_shape_ok(c, c->sh_nw, c->sh_kb, c->sh_kmax))
        shared_set_shape(c, 8, shared_batch_nodes((c->sys.ns + 7) / 8), SH_KMAX);
}

// Work units of the shared-node pass for element-list mode m: every (channel, Ω, 16 x 16 tile)
// holding an element of the pass, once per loop-frequency domain.  The two logarithmic domains
// come first (they hold most nodes), and consecutive units differ in Ω and channel, so that the
// XCDs, which take consecutive blocks in turn, receive even shares.
std::vector<uint32_t> shared_units(const pffrg_ctx *c, int m) {
    const int nO = c->sys.nO, nn = c->sys.nn;
    const int NX2 = (m == 2) ? 1 : nn + 1, NX3 = nn + 1;
    const int n2 = (NX2 + SH_T - 1) / SH_T, n3 = (NX3 + SH_T - 1) / SH_T;
    const int nch = (m == 2) ? 1 : 3;
    std::vector<uint32_t> u;
    const int doms[4] = {0, 3, 1, 2};
    for (int d : doms)
        for (int t2 = 0; t2 < n2; ++t2)
            for (int t3 = 0; t3 < n3; ++t3)
                for (int w1 = 0; w1 < nO; ++w1)
                    for (int ch = 0; ch < nch; ++ch) {
                        bool any = false;
                        for (int i2 = t2 * SH_T; i2 < std::min(NX2, t2 * SH_T + SH_T) && !any; ++i2)
                            for (int i3 = t3 * SH_T; i3 < std::min(NX3, t3 * SH_T + SH_T) && !any; ++i3)
                                any = shared_in_pass(m, i2, i3, nn);
                        if (any) u.push_back(shared_unit(ch, w1, t2, t3, d));
                    }
    return u;
}

// Run-time prior of a unit before its first measured launch (clock ticks): t-channel tiles carry
// the overlap-row contraction and the two inner loop-frequency domains hold most nodes.
uint32_t shared_unit_prior(uint32_t u) {
    const int ch = u & 3, dom = (u >> 17) & 3;
    return 100000u * (ch == 1 ? 3u : 1u) * ((dom == 1 || dom == 2) ? 2u : 1u);
}

// Order key of a unit's run time: 1/8 octave of clock ticks, the longest bucket first.
__device__ __forceinline__ int unit_bucket(uint32_t ticks) {
    const int b = (int)(8.0f * log2f((float)ticks + 1.0f));
    return 255 - (b < 0 ? 0 : (b > 255 ? 255 : b));
}

// out = the units of in, longest run time first (a counting sort by unit_bucket; the order
// inside a bucket is the order the atomics serve, which no result depends on: every partial sum
// has one unit that owns it).  One block.
__global__ __launch_bounds__(1024) void k_order_units(const uint32_t *__restrict__ in,
                                                      const uint32_t *__restrict__ cost, uint32_t *__restrict__ out,
                                                      int n) {
    __shared__ int s_off[256];
    const int t = threadIdx.x;
    if (t < 256) s_off[t] = 0;
    __syncthreads();
    for (int i = t; i < n; i += blockDim.x) atomicAdd(&s_off[unit_bucket(cost[i])], 1);
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int b = 0; b < 256; ++b) {
            const int h = s_off[b];
            s_off[b] = acc;
            acc += h;
        }
    }
    __syncthreads();
    for (int i = t; i < n; i += blockDim.x) out[atomicAdd(&s_off[unit_bucket(cost[i])], 1)] = in[i];
}

hipError_t launch_order_units(const uint32_t *in, const uint32_t *cost, uint32_t *out, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_order_units, dim3(1), dim3(1024), 0, stream, in, cost, out, n);
    return hipGetLastError();
}

// Arguments of the shared-node pass of one context: the pass's tables, the work units of the
// element-list mode (built on first use) and the partial-sum buffer (grown on demand).
template <int PASS>
int shared_args(pffrg_ctx *c, double L, int a, int d, int out, int eval, const double *tol, int mode, SharedArgs &A) {
    int rc;
    if ((rc = pass_args<PASS, float>(c, L, a, d, out, eval, tol, mode, A.P))) return rc;
    shared_shape(c);
    if (!c->d_units[mode]) {
        std::vector<uint32_t> u = shared_units(c, mode);
        HIPCHK(c, hipMalloc(&c->d_units[mode], std::max<size_t>(1, u.size()) * sizeof(uint32_t)));
        if (!u.empty())
            HIPCHK(c, hipMemcpy(c->d_units[mode], u.data(), u.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        c->n_units[mode] = (int)u.size();
    }
    const int nn = c->sys.nn;
    A.units = c->d_units[mode];
    A.n_units = c->n_units[mode];
    A.mode = mode;
    A.NX2 = (mode == 2) ? 1 : nn + 1;
    A.NX3 = nn + 1;
    A.nw = c->sh_nw;
    A.spt = c->sh_spt;
    A.kb = c->sh_kb;
    A.kmax = c->sh_kmax;
    A.count = c->d_count;
    A.cost = nullptr;
    const size_t need = (size_t)SH_NPART * 3 * c->sys.nO * (size_t)A.NX2 * A.NX3 * 2 * c->sys.ns * sizeof(double);
    if (need > c->part_bytes) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(c->d_part));
        c->d_part = nullptr;
        c->part_bytes = 0;
        HIPCHK(c, hipMalloc(&c->d_part, need));
        c->part_bytes = need;
    }
    A.part = c->d_part;
    return 0;
}

template <int PASS>
size_t shared_lds_bytes(const pffrg_ctx *c, const SharedArgs &A) {
    return (size_t)shared_lds(c->sys, c->n_ov, A.nw, PASS == P_KAT, A.spt, A.kb).total;
}

template <int PASS>
int launch_shared_t(pffrg_ctx *c, double L, int a, int d, int out, int eval, const double *tol, int mode) {
    SharedArgs A;
    int rc;
    if ((rc = shared_args<PASS>(c, L, a, d, out, eval, tol, mode, A))) return rc;
    const size_t lds = shared_lds_bytes<PASS>(c, A);
    // dispatch order: the units of the last launch of this pass and mode, longest first
    const int slot = (PASS == P_KAT ? 0 : PASS == P_LEFT ? 1 : 2) * 3 + mode;
    if (!c->d_cost[slot] && A.n_units > 0) {
        const std::vector<uint32_t> u = shared_units(c, mode);
        std::vector<uint32_t> prior(u.size());
        for (size_t k = 0; k < u.size(); ++k) prior[k] = shared_unit_prior(u[k]);
        const size_t bytes = u.size() * sizeof(uint32_t);
        HIPCHK(c, hipMalloc(&c->d_ord[slot][0], bytes));
        HIPCHK(c, hipMalloc(&c->d_ord[slot][1], bytes));
        HIPCHK(c, hipMalloc(&c->d_cost[slot], bytes));
        HIPCHK(c, hipMemcpy(c->d_ord[slot][0], u.data(), bytes, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_cost[slot], prior.data(), bytes, hipMemcpyHostToDevice));
        c->ord_cur[slot] = 0;
    }
    if (c->d_cost[slot]) {
        const int cur = c->ord_cur[slot];
        HIPCHK(c, launch_order_units(c->d_ord[slot][cur], c->d_cost[slot], c->d_ord[slot][1 - cur], A.n_units, c->stream));
        c->ord_cur[slot] = 1 - cur;
        A.units = c->d_ord[slot][1 - cur];
        A.cost = c->d_cost[slot];
    }
    // the timing covers k_shared_pass alone (the pass's dominant kernel), not the finish
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_shared_pass<PASS>(A, lds, c->stream));
    if (c->timing) {
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->time_ms += ms;
        c->time_launches += 1;
    }
    HIPCHK(c, launch_shared_finish<PASS>(A, c->stream));
    c->stat_elems += A.P.n_elems;
    c->stat_launches += 1;
    invalidate(c, out);
    return 0;
}

template <int PASS>
int launch_pass(pffrg_ctx *c, double L, int a, int d, int out, int eval, const double *tol, int mode) {
    if (eval < 1) return fail(c, PFFRG_ERR_ARG, "eval must be >= 1");
    const bool f32 = c->precision != PFFRG_FP64;
    int rc;
    if ((rc = ensure_gather(c, a, f32))) return rc;
    if ((PASS == P_LEFT || PASS == P_CENTRAL) && (rc = ensure_gather(c, d, f32))) return rc;
    if (c->precision == PFFRG_FP32_MIXED && shared_fits(c)) return launch_shared_t<PASS>(c, L, a, d, out, eval, tol, mode);
    if (f32) return launch_pass_t<PASS, float>(c, L, a, d, out, eval, tol, mode);
    return launch_pass_t<PASS, double>(c, L, a, d, out, eval, tol, mode);
}

int launch_corrs(pffrg_ctx *c, double L, int a, int da, bool kat, const double *tol) {
    CorrArgs C;
    C.y = c->sys;
    C.mesh = c->d_mesh;
    C.sigA = c->slots[a].Sigma;
    C.sigD = kat ? c->slots[da].Sigma : c->slots[a].Sigma;
    C.corrs = c->d_corrs;
    C.L = L;
    C.atol = tol[0];
    C.rtol = tol[1];
    C.kat = kat ? 1 : 0;
    hipLaunchKernelGGL(k_corrs, dim3(c->sys.nO * 3 * 2), dim3(64), 0, c->stream, C);
    HIPCHK(c, hipGetLastError());
    return 0;
}

SigmaArgs sigma_args(const pffrg_ctx *c, double L, int a, int v, double *out, const double *sigV, const double *add,
                     const double *tol) {
    SigmaArgs Q;
    Q.y = c->sys;
    Q.mesh = c->d_mesh;
    Q.exch = c->d_exch;
    Q.mult = c->d_mult;
    Q.A = ActV<double>{c->slots[a].gamma, c->slots[a].bare, c->slots[a].Sigma};
    Q.V = ActV<double>{c->slots[v].gamma, c->slots[v].bare, c->slots[v].Sigma};
    Q.sigV = sigV;
    Q.out = out;
    Q.add = add;
    Q.L = L;
    Q.atol = tol[0];
    Q.rtol = tol[1];
    return Q;
}

template <int KIND>
int launch_sigma(pffrg_ctx *c, double L, int a, int v, double *out, const double *sigV, const double *add,
                 const double *tol) {
    const SigmaArgs Q = sigma_args(c, L, a, v, out, sigV, add, tol);
    if (c->sys.nsig > 1) {
        hipLaunchKernelGGL((k_sigma<KIND>), dim3(c->sys.nsig - 1), dim3(SIGMA_T), 0, c->stream, Q);
        HIPCHK(c, hipGetLastError());
    }
    return 0;
}

int do_symmetrize(pffrg_ctx *c, int s) {
    const long n = 3L * c->sys.nc * c->sys.ns * c->sys.nO * c->sys.nn * c->sys.nn;
    hipLaunchKernelGGL(k_symmetrize, dim3(grid_for(n)), dim3(256), 0, c->stream, c->slots[s].gamma, c->sys, c->d_exch);
    HIPCHK(c, hipGetLastError());
    invalidate(c, s);
    return 0;
}

int do_symmetrize_add(pffrg_ctx *c, int l, int a) {
    const long n = c->sys.nc * c->sys.complen;
    hipLaunchKernelGGL(k_symmetrize_add, dim3(grid_for(n)), dim3(256), 0, c->stream, c->slots[l].gamma,
                       c->slots[a].gamma, c->sys, c->d_exch);
    HIPCHK(c, hipGetLastError());
    invalidate(c, a);
    return 0;
}

int do_axpy(pffrg_ctx *c, int src, double fac, int dst, bool vertex_only) {
    const long n = c->sys.nc * c->sys.complen;
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, c->stream, c->slots[src].gamma, fac, c->slots[dst].gamma, n);
    HIPCHK(c, hipGetLastError());
    if (!vertex_only) {
        hipLaunchKernelGGL(k_axpy, dim3(1), dim3(256), 0, c->stream, c->slots[src].Sigma, fac, c->slots[dst].Sigma,
                           (long)c->sys.nsig);
        HIPCHK(c, hipGetLastError());
    }
    invalidate(c, dst);
    return 0;
}

int do_scale(pffrg_ctx *c, int s, double fac, bool vertex_only) {
    const long n = c->sys.nc * c->sys.complen;
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c->stream, c->slots[s].gamma, fac, n);
    HIPCHK(c, hipGetLastError());
    if (!vertex_only) {
        hipLaunchKernelGGL(k_scale, dim3(1), dim3(256), 0, c->stream, c->slots[s].Sigma, fac, (long)c->sys.nsig);
        HIPCHK(c, hipGetLastError());
    }
    invalidate(c, s);
    return 0;
}

// Bracket hints of one mesh of n >= 2 strictly increasing points from 0 (layout next to
// HINT_CELLS): c0 puts m[1] in cell 1 and every smaller value in cell 0; the last cell starts
// above m[n-1] unless that takes more than HINT_CELLS cells.  first[c] = first index whose
// point is >= the cell's lower bound, so the answers of cell c lie in [first[c], first[c+1]].
void mesh_hint(const double *m, int n, int &c0, int &nc, uint8_t *tab) {
    c0 = hint_key(m[1]) - 1;
    const int top = hint_key(m[n - 1]) - c0 + 1;
    nc = std::min(top + 1, HINT_CELLS);
    std::vector<int> first(nc + 1);
    first[0] = 0;
    for (int cc = 1; cc < nc; ++cc) {
        const long long bits = (long long)(cc + c0) << 49;
        double lb;
        std::memcpy(&lb, &bits, sizeof lb);
        int a = 0;
        while (a < n && m[a] < lb) ++a;
        first[cc] = std::min(a, n - 1);
    }
    first[nc] = n - 1;  // the last cell also holds every larger value, whose answer is n - 1
    for (int cc = 0; cc < nc; ++cc)
        tab[cc] = (uint8_t)(first[cc] | (first[cc + 1] - first[cc] > 3 ? 128 : 0));
}

// the device mesh table: five meshes, then the inverse spacings 1/(m[j+1]-m[j]) of each (the
// reciprocal get_param computes, param.jl:73-76, evaluated once with the same IEEE division),
// then the bracket hints of the five meshes
int mesh_table(pffrg_ctx *c, const double *sigma, const double *Os, const double *ns, const double *Ot,
               const double *nt, std::vector<double> &h) {
    const Sys &y = c->sys;
    h.assign(MESH_TABLE, 0.0);
    auto put = [&](int k, const double *m, int n) {
        for (int j = 0; j < n; ++j) h[k * MAXN + j] = m[j];
        for (int j = 0; j + 1 < n; ++j) h[(5 + k) * MAXN + j] = 1.0 / (m[j + 1] - m[j]);
    };
    put(0, sigma, y.nsig);
    put(1, Os, y.nO);
    put(2, ns, y.nn);
    put(3, Ot, y.nO);
    put(4, nt, y.nn);
    for (int k = 0; k < 5; ++k) {
        const int n = (k == 0) ? y.nsig : ((k == 1 || k == 3) ? y.nO : y.nn);
        if (h[k * MAXN] != 0.0) return fail(c, PFFRG_ERR_ARG, "meshes must start at 0");
        for (int j = 1; j < n; ++j)
            if (!(h[k * MAXN + j] > h[k * MAXN + j - 1])) return fail(c, PFFRG_ERR_ARG, "meshes must be strictly increasing");
    }
    int16_t hdr[10];
    std::vector<uint8_t> tab(5 * HINT_CELLS, 0);
    for (int k = 0; k < 5; ++k) {
        const int n = (k == 0) ? y.nsig : ((k == 1 || k == 3) ? y.nO : y.nn);
        int c0, nc;
        mesh_hint(&h[k * MAXN], n, c0, nc, &tab[k * HINT_CELLS]);
        hdr[k] = (int16_t)c0;  // keys of positive doubles are < 2^14
        hdr[5 + k] = (int16_t)nc;
    }
    unsigned char *dst = reinterpret_cast<unsigned char *>(&h[10 * MAXN]);
    std::memcpy(dst, hdr, sizeof hdr);
    std::memcpy(dst + sizeof hdr, tab.data(), tab.size());
    return 0;
}

// make a checked mesh table the context's (waits for the work queued on the old meshes)
int apply_mesh(pffrg_ctx *c, const std::vector<double> &h) {
    Sys &y = c->sys;
    y.valS = h[1 * MAXN + y.nO - 1] + h[2 * MAXN + y.nn - 1];
    y.valT = h[3 * MAXN + y.nO - 1] + h[4 * MAXN + y.nn - 1];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->h_Os.assign(h.begin() + 1 * MAXN, h.begin() + 1 * MAXN + y.nO);
    c->h_Ot.assign(h.begin() + 3 * MAXN, h.begin() + 3 * MAXN + y.nO);
    c->logpts_valid = false;
    HIPCHK(c, hipMemcpy(c->d_mesh, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    c->have_mesh = true;
    return 0;
}

// small scratch of the rescanning kernels, allocated on first use
int ensure_scan_mem(pffrg_ctx *c) {
    if (c->scan_mem) return 0;
    HIPCHK(c, hipMalloc(&c->scan_mem, SCAN_MEM_BYTES));
    return 0;
}

int do_copy(pffrg_ctx *c, int dst, int src, bool vertex_only) {
    HIPCHK(c, hipMemcpyAsync(c->slots[dst].gamma, c->slots[src].gamma, c->sys.nc * c->sys.complen * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
    if (!vertex_only)
        HIPCHK(c, hipMemcpyAsync(c->slots[dst].Sigma, c->slots[src].Sigma, c->sys.nsig * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
    invalidate(c, dst);
    return 0;
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

int pffrg_create(const pffrg_dims *d, pffrg_ctx **out) { return pffrg_create_symmetry(d, PFFRG_SU2, out); }

int pffrg_create_symmetry(const pffrg_dims *d, int32_t symmetry, pffrg_ctx **out) {
    if (!d || !out) return PFFRG_ERR_ARG;
    *out = nullptr;
    if (symmetry != PFFRG_SU2 && symmetry != PFFRG_U1_DM) return PFFRG_ERR_ARG;
    if (d->n_sites < 1 || d->n_sites > 64 || d->n_sigma < 2 || d->n_sigma > MAXN || d->n_Omega < 2 ||
        d->n_Omega > MAXN || d->n_nu < 2 || d->n_nu > MAXN || d->n_slots < 1 || d->S <= 0.0)
        return PFFRG_ERR_LIMIT;
    pffrg_ctx *c = new pffrg_ctx();
    c->dims = *d;
    c->device = d->device;
    if (hipSetDevice(d->device) != hipSuccess) { delete c; return PFFRG_ERR_HIP; }
    Sys &y = c->sys;
    y.ns = d->n_sites;
    y.nO = d->n_Omega;
    y.nn = d->n_nu;
    y.nsig = d->n_sigma;
    y.GS = y.ns;
    y.G = (y.ns <= 32) ? WAVE / y.ns : 1;
    if (y.G > MAX_GROUPS) y.G = MAX_GROUPS;
    y.S2 = 2.0 * d->S;
    y.nc = (symmetry == PFFRG_U1_DM) ? 6 : 2;
    y.q21 = (long)y.ns * y.nO;
    y.q22 = y.q21 + (long)y.ns * y.nO * y.nn;
    y.q3 = y.q22 + (long)y.ns * y.nO * y.nn;
    y.chlen = y.q3 + (long)y.ns * y.nO * y.nn * y.nn;
    y.complen = 3 * y.chlen;
    void *scratch_ptr = nullptr;
    bool ok = scan_scratch_address(d->device, &scratch_ptr) == hipSuccess;
    c->scan_scratch = static_cast<char *>(scratch_ptr);
    ok = ok && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&c->d_mesh, MESH_TABLE * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->d_corrs, 2 * 3 * y.nO * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->d_count, N_COUNTERS_ALLOC * sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipMalloc(&c->d_max, sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipMemset(c->d_corrs, 0, 2 * 3 * y.nO * sizeof(double)) == hipSuccess;
    ok = ok && hipMemset(c->d_count, 0, N_COUNTERS_ALLOC * sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess;
    c->slots.resize(d->n_slots);
    for (auto &s : c->slots) {
        ok = ok && hipMalloc(&s.Sigma, y.nsig * sizeof(double)) == hipSuccess;
        ok = ok && hipMalloc(&s.bare, y.nc * y.ns * sizeof(double)) == hipSuccess;
        ok = ok && hipMalloc(&s.gamma, y.nc * y.complen * sizeof(double)) == hipSuccess;
        ok = ok && hipMemset(s.Sigma, 0, y.nsig * sizeof(double)) == hipSuccess;
        ok = ok && hipMemset(s.bare, 0, y.nc * y.ns * sizeof(double)) == hipSuccess;
        ok = ok && hipMemset(s.gamma, 0, y.nc * y.complen * sizeof(double)) == hipSuccess;
    }
    auto upload_list = [&](uint32_t **dst, int *n, int mode) {
        std::vector<uint32_t> e = element_list(y.nO, y.nn, mode);
        *n = (int)e.size();
        ok = ok && hipMalloc(dst, e.size() * sizeof(uint32_t)) == hipSuccess;
        ok = ok && hipMemcpy(*dst, e.data(), e.size() * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess;
    };
    upload_list(&c->d_el_half, &c->n_half, 0);
    upload_list(&c->d_el_full, &c->n_full, 1);
    upload_list(&c->d_el_red, &c->n_red, 2);
    if (!ok) {
        pffrg_destroy(c);
        return PFFRG_ERR_HIP;
    }
    *out = c;
    return 0;
}

void pffrg_destroy(pffrg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &s : c->slots) {
        (void)hipFree(s.Sigma);
        (void)hipFree(s.bare);
        (void)hipFree(s.gamma);
        (void)hipFree(s.gat32);
        (void)hipFree(s.gat64);
    }
    (void)hipFree(c->chi_mem);
    (void)hipFree(c->scan_mem);
    (void)hipFree(c->d_mesh);
    (void)hipFree(c->d_exch);
    (void)hipFree(c->d_ov_off);
    (void)hipFree(c->d_mult);
    (void)hipFree(c->d_ov_rows);
    (void)hipFree(c->d_corrs);
    (void)hipFree(c->d_count);
    (void)hipFree(c->d_max);
    (void)hipFree(c->d_el_half);
    (void)hipFree(c->d_el_full);
    (void)hipFree(c->d_el_red);
    for (int m = 0; m < 3; ++m) (void)hipFree(c->d_units[m]);
    for (int k = 0; k < 9; ++k) {
        (void)hipFree(c->d_ord[k][0]);
        (void)hipFree(c->d_ord[k][1]);
        (void)hipFree(c->d_cost[k]);
    }
    (void)hipFree(c->d_part);
    (void)hipFree(c->d_logpts);
    (void)hipFree(c->d_dbg_nodes);
    (void)hipFree(c->d_dbg_res);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *pffrg_last_error(const pffrg_ctx *c) { return c ? c->err.c_str() : "null context"; }

int32_t pffrg_n_components(const pffrg_ctx *c) { return c ? c->sys.nc : -1; }

int64_t pffrg_gamma_length(const pffrg_ctx *c) { return c ? c->sys.nc * c->sys.complen : -1; }

int pffrg_set_mesh(pffrg_ctx *c, const double *sigma, const double *Os, const double *ns, const double *Ot,
                   const double *nt) {
    if (!c || !sigma || !Os || !ns || !Ot || !nt) return fail(c, PFFRG_ERR_ARG, "null mesh pointer");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<double> h;
    int rc = mesh_table(c, sigma, Os, ns, Ot, nt, h);
    if (rc) return rc;
    return apply_mesh(c, h);
}

int pffrg_set_lattice(pffrg_ctx *c, const int64_t *exchange, const int64_t *mult, const int64_t *ov_off,
                      const int64_t *ov_rows, int32_t base) {
    if (!c || !exchange || !mult || !ov_off || !ov_rows) return fail(c, PFFRG_ERR_ARG, "null lattice pointer");
    if (base != 0 && base != 1) return fail(c, PFFRG_ERR_ARG, "index_base must be 0 or 1");
    HIPCHK(c, hipSetDevice(c->device));
    const int ns = c->sys.ns;
    std::vector<int> ex(ns), off(ns + 1);
    std::vector<double> mu(ns);
    for (int i = 0; i < ns; ++i) {
        ex[i] = (int)(exchange[i] - base);
        if (ex[i] < 0 || ex[i] >= ns) return fail(c, PFFRG_ERR_ARG, "exchange index out of range");
        // orbit sizes (Reduced_lattice.mult, Int64); the self-energy kernel weights sites by them in FP64
        if (mult[i] < 1 || mult[i] > (int64_t(1) << 40)) return fail(c, PFFRG_ERR_ARG, "site multiplicity out of range");
        mu[i] = (double)mult[i];
    }
    if (ov_off[0] != 0) return fail(c, PFFRG_ERR_ARG, "ov_off[0] must be 0");
    for (int i = 0; i <= ns; ++i) off[i] = (int)ov_off[i];
    const int R = off[ns];
    for (int i = 0; i < ns; ++i)
        if (off[i + 1] < off[i]) return fail(c, PFFRG_ERR_ARG, "ov_off must be non-decreasing");
    if ((size_t)pass_lds(c->sys, R, (int)sizeof(DescT<double>), false).total > 48 * 1024)
        return fail(c, PFFRG_ERR_LIMIT, "overlap table exceeds the LDS budget");
    std::vector<uint32_t> rows(R > 0 ? R : 1);
    for (int r = 0; r < R; ++r) {
        const long j1 = ov_rows[3 * r] - base, j2 = ov_rows[3 * r + 1] - base, m = ov_rows[3 * r + 2];
        if (j1 < 0 || j1 >= ns || j2 < 0 || j2 >= ns || m < 0 || m > 65535)
            return fail(c, PFFRG_ERR_ARG, "overlap row out of range");
        rows[r] = (uint32_t)j1 | ((uint32_t)j2 << 8) | ((uint32_t)m << 16);
    }
    // new tables are built aside and replace the old ones only when every allocation and copy
    // succeeded, so a failed call keeps the previous lattice usable
    int *n_exch = nullptr, *n_off = nullptr;
    uint32_t *n_rows = nullptr;
    double *n_mult = nullptr;
    bool ok = hipMalloc(&n_exch, ns * sizeof(int)) == hipSuccess &&
              hipMalloc(&n_off, (ns + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&n_rows, rows.size() * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&n_mult, ns * sizeof(double)) == hipSuccess &&
              hipMemcpy(n_exch, ex.data(), ns * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n_off, off.data(), (ns + 1) * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n_rows, rows.data(), rows.size() * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(n_mult, mu.data(), ns * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        const hipError_t e = hipGetLastError();
        (void)hipFree(n_exch);
        (void)hipFree(n_off);
        (void)hipFree(n_rows);
        (void)hipFree(n_mult);
        return fail(c, PFFRG_ERR_HIP, std::string("set_lattice: device tables not created: ") + hipGetErrorString(e));
    }
    // the old tables may still be read by queued work
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->d_exch);
    (void)hipFree(c->d_ov_off);
    (void)hipFree(c->d_ov_rows);
    (void)hipFree(c->d_mult);
    c->d_exch = n_exch;
    c->d_ov_off = n_off;
    c->d_ov_rows = n_rows;
    c->d_mult = n_mult;
    c->n_ov = R;
    c->have_lattice = true;
    return 0;
}

int pffrg_upload_action(pffrg_ctx *c, int32_t s, const double *Sigma, const double *bare, const double *gamma) {
    int rc = check_slot(c, s);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Slot &sl = c->slots[s];
    if (Sigma) HIPCHK(c, hipMemcpy(sl.Sigma, Sigma, c->sys.nsig * sizeof(double), hipMemcpyHostToDevice));
    if (bare) HIPCHK(c, hipMemcpy(sl.bare, bare, c->sys.nc * c->sys.ns * sizeof(double), hipMemcpyHostToDevice));
    if (gamma) HIPCHK(c, hipMemcpy(sl.gamma, gamma, c->sys.nc * c->sys.complen * sizeof(double), hipMemcpyHostToDevice));
    invalidate(c, s);
    return 0;
}

int pffrg_download_action(pffrg_ctx *c, int32_t s, double *Sigma, double *bare, double *gamma) {
    int rc = check_slot(c, s);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Slot &sl = c->slots[s];
    if (Sigma) HIPCHK(c, hipMemcpy(Sigma, sl.Sigma, c->sys.nsig * sizeof(double), hipMemcpyDeviceToHost));
    if (bare) HIPCHK(c, hipMemcpy(bare, sl.bare, c->sys.nc * c->sys.ns * sizeof(double), hipMemcpyDeviceToHost));
    if (gamma) HIPCHK(c, hipMemcpy(gamma, sl.gamma, c->sys.nc * c->sys.complen * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int pffrg_set_precision(pffrg_ctx *c, int32_t mode) {
    if (!c) return PFFRG_ERR_ARG;
    if (mode != PFFRG_FP64 && mode != PFFRG_FP32_MIXED && mode != PFFRG_FP32_ELEMENT)
        return fail(c, PFFRG_ERR_ARG, "unknown precision mode");
    c->precision = mode;
    return 0;
}

int pffrg_compute_dsigma(pffrg_ctx *c, double L, int32_t a, int32_t da, const double *tol) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a)) || (rc = check_slot(c, da))) return rc;
    if (!tol || !(L > 0.0)) return fail(c, PFFRG_ERR_ARG, "invalid Lambda or tolerance");
    HIPCHK(c, hipSetDevice(c->device));
    rc = launch_sigma<SK_DSIGMA>(c, L, a, a, c->slots[da].Sigma, nullptr, nullptr, tol);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_compute_dsigma_corr(pffrg_ctx *c, double L, int32_t a, int32_t da, int32_t daS, const double *tol) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a)) || (rc = check_slot(c, da)) || (rc = check_slot(c, daS))) return rc;
    if (!tol || !(L > 0.0)) return fail(c, PFFRG_ERR_ARG, "invalid Lambda or tolerance");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = launch_sigma<SK_CORR1>(c, L, a, daS, c->slots[daS].Sigma, nullptr, nullptr, tol))) return rc;
    if ((rc = launch_sigma<SK_CORR2>(c, L, a, a, c->slots[da].Sigma, c->slots[daS].Sigma, c->slots[daS].Sigma, tol)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int dgamma_1l(pffrg_ctx *c, double L, int a, int da, int eval, const double *tol) {
    int rc;
    if ((rc = launch_corrs(c, L, a, da, true, tol))) return rc;
    if ((rc = launch_pass<P_KAT>(c, L, a, da, da, eval, tol, 0))) return rc;
    return do_symmetrize(c, da);
}

static int dgamma_2l(pffrg_ctx *c, double L, int a, int da, int dal, int eval, const double *tol) {
    int rc;
    if ((rc = dgamma_1l(c, L, a, da, eval, tol))) return rc;
    if ((rc = launch_corrs(c, L, a, a, false, tol))) return rc;
    if ((rc = launch_pass<P_LEFT>(c, L, a, da, dal, eval, tol, 1))) return rc;
    return do_symmetrize_add(c, dal, da);
}

int pffrg_compute_dgamma_1l(pffrg_ctx *c, double L, int32_t a, int32_t da, int32_t eval, const double *tol) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a)) || (rc = check_slot(c, da))) return rc;
    if (!tol || !(L > 0.0) || a == da) return fail(c, PFFRG_ERR_ARG, "invalid arguments");
    if (eval < 1) return fail(c, PFFRG_ERR_ARG, "eval must be >= 1");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = dgamma_1l(c, L, a, da, eval, tol))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_set_shared_shape(pffrg_ctx *c, int32_t nw, int32_t kb, int32_t kmax) {
    int rc;
    if ((rc = ready(c))) return rc;
    if (c->precision != PFFRG_FP32_MIXED || !shared_fits(c))
        return fail(c, PFFRG_ERR_STATE, "the shared-node pass does not serve this context");
    if (!shared_shape_ok(c, nw, kb, kmax)) return fail(c, PFFRG_ERR_ARG, "infeasible shape for this problem");
    shared_set_shape(c, nw, kb, kmax);
    return 0;
}

int pffrg_get_shared_shape(pffrg_ctx *c, int32_t *shape) {
    if (!c || !shape) return PFFRG_ERR_ARG;
    shared_shape(c);
    shape[0] = c->sh_nw;
    shape[1] = c->sh_kb;
    shape[2] = c->sh_kmax;
    return 0;
}

// The autotuner: every feasible (waves per block, batch capacity, most nodes per batch) from a
// fixed candidate list, timed on this problem's own Katanin pass (compute_dΣ! once, then
// compute_dΓ_1l! of slot a into slot da at Lambda, one untimed and two timed runs each, the
// faster counted); the fastest is kept for every later shared-node pass of the context.
int pffrg_tune_shared(pffrg_ctx *c, double L, int32_t a, int32_t da, int32_t eval, const double *tol, int32_t *best,
                      double *ms, int32_t *n_cand) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a)) || (rc = check_slot(c, da))) return rc;
    if (!tol || !best || !n_cand || !(L > 0.0) || a == da || eval < 1) return fail(c, PFFRG_ERR_ARG, "invalid arguments");
    if (c->precision != PFFRG_FP32_MIXED || !shared_fits(c))
        return fail(c, PFFRG_ERR_STATE, "the shared-node pass does not serve this context");
    HIPCHK(c, hipSetDevice(c->device));
    static const int cand[][3] = {{8, 8, 64}, {8, 8, 32}, {8, 8, 16}, {8, 4, 64}, {8, 4, 32},
                                  {4, 8, 64}, {4, 8, 32}, {4, 4, 64}, {4, 4, 32}, {4, 4, 16}};
    const int nc = (int)(sizeof(cand) / sizeof(cand[0]));
    const int cap = *n_cand;
    if ((rc = launch_sigma<SK_DSIGMA>(c, L, a, a, c->slots[da].Sigma, nullptr, nullptr, tol))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const bool timing = c->timing;
    c->timing = true;
    int k_best = -1, n_run = 0;
    double t_best = 0.0;
    for (int k = 0; k < nc; ++k) {
        double t = -1.0;
        if (shared_shape_ok(c, cand[k][0], cand[k][1], cand[k][2])) {
            shared_set_shape(c, cand[k][0], cand[k][1], cand[k][2]);
            for (int r = 0; r < 3; ++r) {
                const double t0 = c->time_ms;
                if ((rc = dgamma_1l(c, L, a, da, eval, tol))) {
                    c->timing = timing;
                    return rc;
                }
                HIPCHK(c, hipStreamSynchronize(c->stream));
                const double dt = c->time_ms - t0;
                if (r > 0 && (t < 0.0 || dt < t)) t = dt;
            }
            if (k_best < 0 || t < t_best) {
                k_best = k;
                t_best = t;
            }
        }
        if (ms && n_run < cap) ms[n_run] = t;
        ++n_run;
    }
    c->timing = timing;
    *n_cand = n_run;
    shared_set_shape(c, cand[k_best][0], cand[k_best][1], cand[k_best][2]);
    best[0] = cand[k_best][0];
    best[1] = cand[k_best][1];
    best[2] = cand[k_best][2];
    return 0;
}

int pffrg_compute_dgamma_2l(pffrg_ctx *c, double L, int32_t a, int32_t da, int32_t dal, int32_t eval,
                            const double *tol) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a)) || (rc = check_slot(c, da)) || (rc = check_slot(c, dal))) return rc;
    if (!tol || !(L > 0.0) || a == da || a == dal || da == dal) return fail(c, PFFRG_ERR_ARG, "invalid arguments");
    if (eval < 1) return fail(c, PFFRG_ERR_ARG, "eval must be >= 1");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = dgamma_2l(c, L, a, da, dal, eval, tol))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_compute_dgamma_ml(pffrg_ctx *c, double L, int32_t loops, int32_t a, int32_t da, int32_t dal, int32_t dac,
                            int32_t datemp, int32_t daS, int32_t eval, const double *tol) {
    int rc;
    if ((rc = ready(c))) return rc;
    const int sl[6] = {a, da, dal, dac, datemp, daS};
    for (int k = 0; k < 6; ++k) {
        if ((rc = check_slot(c, sl[k]))) return rc;
        for (int j = 0; j < k; ++j)
            if (sl[j] == sl[k]) return fail(c, PFFRG_ERR_ARG, "multiloop slots must be distinct");
    }
    if (!tol || !(L > 0.0) || loops < 1) return fail(c, PFFRG_ERR_ARG, "invalid arguments");
    if (eval < 1) return fail(c, PFFRG_ERR_ARG, "eval must be >= 1");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = dgamma_2l(c, L, a, da, dal, eval, tol))) return rc;
    if ((rc = do_scale(c, datemp, 0.0, true))) return rc;
    if ((rc = do_scale(c, daS, 0.0, true))) return rc;
    if ((rc = do_symmetrize_add(c, dal, datemp))) return rc;
    for (int loop = 3; loop <= loops; ++loop) {
        if ((rc = launch_pass<P_CENTRAL>(c, L, a, dal, dac, eval, tol, 0))) return rc;
        if ((rc = launch_pass<P_LEFT>(c, L, a, datemp, dal, eval, tol, 1))) return rc;
        if ((rc = do_symmetrize(c, dac))) return rc;
        if ((rc = do_copy(c, datemp, dac, true))) return rc;
        if ((rc = do_symmetrize_add(c, dal, datemp))) return rc;
        if ((rc = do_axpy(c, dac, 1.0, daS, true))) return rc;
        if ((rc = do_axpy(c, datemp, 1.0, da, true))) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_compute_sigma(pffrg_ctx *c, double L, int32_t a1, int32_t a2, int32_t eval, const double *gtol,
                        const double *stol) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a1)) || (rc = check_slot(c, a2))) return rc;
    if (!gtol || !stol || !(L > 0.0) || a1 == a2) return fail(c, PFFRG_ERR_ARG, "invalid arguments");
    if (eval < 1) return fail(c, PFFRG_ERR_ARG, "eval must be >= 1");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = launch_corrs(c, L, a1, a1, false, gtol))) return rc;
    if ((rc = launch_pass<P_REDUCED>(c, L, a1, a1, a2, eval, gtol, 2))) return rc;
    if ((rc = launch_sigma<SK_SDE>(c, L, a1, a2, c->slots[a2].Sigma, nullptr, nullptr, stol))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_compute_gamma(pffrg_ctx *c, double L, int32_t a1, int32_t a2, int32_t eval, const double *tol) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a1)) || (rc = check_slot(c, a2))) return rc;
    if (!tol || !(L > 0.0) || a1 == a2) return fail(c, PFFRG_ERR_ARG, "invalid arguments");
    if (eval < 1) return fail(c, PFFRG_ERR_ARG, "eval must be >= 1");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = launch_pass<P_BSE>(c, L, a1, a1, a2, eval, tol, 0))) return rc;
    if ((rc = do_symmetrize(c, a2))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_symmetrize(pffrg_ctx *c, int32_t s) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, s))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = do_symmetrize(c, s))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_symmetrize_add_to(pffrg_ctx *c, int32_t l, int32_t a) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, l)) || (rc = check_slot(c, a))) return rc;
    if (l == a) return fail(c, PFFRG_ERR_ARG, "source and destination must differ");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = do_symmetrize_add(c, l, a))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_replace_with(pffrg_ctx *c, int32_t dst, int32_t src, int32_t vertex_only) {
    int rc;
    if ((rc = check_slot(c, dst)) || (rc = check_slot(c, src))) return rc;
    if (dst == src) return 0;
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = do_copy(c, dst, src, vertex_only != 0))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_mult_with(pffrg_ctx *c, int32_t s, double fac, int32_t vertex_only) {
    int rc;
    if ((rc = check_slot(c, s))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = do_scale(c, s, fac, vertex_only != 0))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_mult_with_add_to(pffrg_ctx *c, int32_t src, double fac, int32_t dst, int32_t vertex_only) {
    int rc;
    if ((rc = check_slot(c, src)) || (rc = check_slot(c, dst))) return rc;
    if (src == dst) return fail(c, PFFRG_ERR_ARG, "source and destination must differ");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = do_axpy(c, src, fac, dst, vertex_only != 0))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_abs_max(pffrg_ctx *c, int32_t s, double *out) {
    int rc;
    if ((rc = check_slot(c, s))) return rc;
    if (!out) return fail(c, PFFRG_ERR_ARG, "null output");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemsetAsync(c->d_max, 0, sizeof(unsigned long long), c->stream));
    const long n = 3L * c->sys.nc * c->sys.ns * c->sys.nO * c->sys.nn * c->sys.nn;
    hipLaunchKernelGGL(k_absmax_q3, dim3(grid_for(n)), dim3(256), 0, c->stream, c->slots[s].gamma, c->sys, c->d_max);
    HIPCHK(c, hipGetLastError());
    unsigned long long bits = 0;
    HIPCHK(c, hipMemcpyAsync(&bits, c->d_max, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *out = v;
    return 0;
}

int pffrg_get_sigma(pffrg_ctx *c, int32_t s, double *Sigma) {
    int rc;
    if ((rc = check_slot(c, s))) return rc;
    if (!Sigma) return fail(c, PFFRG_ERR_ARG, "null output");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(Sigma, c->slots[s].Sigma, c->sys.nsig * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int pffrg_set_sigma(pffrg_ctx *c, int32_t s, const double *Sigma) {
    int rc;
    if ((rc = check_slot(c, s))) return rc;
    if (!Sigma) return fail(c, PFFRG_ERR_ARG, "null input");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->slots[s].Sigma, Sigma, c->sys.nsig * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int pffrg_synchronize(pffrg_ctx *c) {
    if (!c) return PFFRG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pffrg_reset_stats(pffrg_ctx *c) {
    if (!c) return PFFRG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemset(c->d_count, 0, N_COUNTERS_ALLOC * sizeof(unsigned long long)));
    c->stat_elems = 0;
    c->stat_launches = 0;
    c->time_ms = 0.0;
    c->time_launches = 0;
    return 0;
}

int pffrg_get_stats(pffrg_ctx *c, int64_t *stats) {
    if (!c || !stats) return PFFRG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long n[2] = {0, 0};
    HIPCHK(c, hipMemcpy(n, c->d_count, sizeof(n), hipMemcpyDeviceToHost));
    stats[0] = (int64_t)n[0];
    stats[1] = c->stat_elems;
    stats[2] = c->stat_launches;
    stats[3] = (int64_t)n[1];
    return 0;
}

int pffrg_get_counters(pffrg_ctx *c, int64_t *out, int32_t n) {
    if (!c || !out || n < 1 || n > N_COUNTERS) return fail(c, PFFRG_ERR_ARG, "counters: 1..32 entries");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long v[N_COUNTERS];
    HIPCHK(c, hipMemcpy(v, c->d_count, sizeof(v), hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) out[k] = (int64_t)v[k];
    return 0;
}

#ifdef PFFRG_SHARED_PROFILE_UNITS
// measurement build only: the counters followed by the record count and the unit records
int pffrg_get_counters_ext(pffrg_ctx *c, int64_t *out, int32_t n) {
    if (!c || !out || n < 1 || n > N_COUNTERS_ALLOC) return PFFRG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->d_count, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost));
    return 0;
}
#endif

int pffrg_get_corrs(pffrg_ctx *c, double *out) {
    if (!c || !out) return PFFRG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->d_corrs, 2 * 3 * c->sys.nO * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int pffrg_probe_integrand(pffrg_ctx *c, int32_t pass, double L, int32_t a, int32_t da, int32_t ch, int32_t kernel,
                          int32_t w1, int32_t w2, int32_t w3, int32_t n, const double *v, double *out) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a)) || (rc = check_slot(c, da))) return rc;
    if (n < 1 || !v || !out || pass < 0 || pass > 4 || ch < 0 || ch > 2 || kernel < 1 || kernel > 4)
        return fail(c, PFFRG_ERR_ARG, "invalid probe arguments");
    if (w1 < 0 || w1 >= c->sys.nO || w2 >= c->sys.nn || w3 >= c->sys.nn) return fail(c, PFFRG_ERR_ARG, "index");
    HIPCHK(c, hipSetDevice(c->device));
    // the probe's buffers, released on every return
    struct Bufs {
        double *dv = nullptr, *dout = nullptr;
        ~Bufs() {
            (void)hipFree(dv);
            (void)hipFree(dout);
        }
    } B;
    double *&dv = B.dv, *&dout = B.dout;
    HIPCHK(c, hipMalloc(&dv, n * sizeof(double)));
    HIPCHK(c, hipMalloc(&dout, (size_t)n * c->sys.nc * c->sys.ns * sizeof(double)));
    HIPCHK(c, hipMemcpy(dv, v, n * sizeof(double), hipMemcpyHostToDevice));
    const uint32_t e = enc(ch, kernel, w1, w2, w3);
    const double tol[2] = {1e-8, 1e-4};
    const bool f32 = c->precision == PFFRG_FP32_MIXED;
    if ((rc = ensure_gather(c, a, f32)) || (rc = ensure_gather(c, da, f32))) return rc;
#define PROBE(PS)                                                                                                  \
    rc = f32 ? launch_pass_t<PS, float>(c, L, a, da, a, 1, tol, 0, dv, dout, n, e)                               \
             : launch_pass_t<PS, double>(c, L, a, da, a, 1, tol, 0, dv, dout, n, e);
    switch (pass) {
        case P_KAT: PROBE(P_KAT) break;
        case P_LEFT: PROBE(P_LEFT) break;
        case P_CENTRAL: PROBE(P_CENTRAL) break;
        case P_BSE: PROBE(P_BSE) break;
        default: PROBE(P_REDUCED) break;
    }
#undef PROBE
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, dout, (size_t)n * c->sys.nc * c->sys.ns * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int pffrg_debug_record(pffrg_ctx *c, int32_t *sub_nodes, double *sub_res, int64_t capacity) {
    if (!c || !sub_nodes || !sub_res || capacity < 1) return PFFRG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->d_dbg_nodes) return fail(c, PFFRG_ERR_STATE, "a recording is already armed");
    HIPCHK(c, hipMalloc(&c->d_dbg_nodes, capacity * sizeof(int)));
    HIPCHK(c, hipMalloc(&c->d_dbg_res, capacity * sizeof(double)));
    HIPCHK(c, hipMemset(c->d_dbg_nodes, 0, capacity * sizeof(int)));
    HIPCHK(c, hipMemset(c->d_dbg_res, 0, capacity * sizeof(double)));
    c->h_dbg_nodes = sub_nodes;
    c->h_dbg_res = sub_res;
    c->dbg_cap = capacity;
    return 0;
}

int pffrg_set_timing(pffrg_ctx *c, int32_t enable) {
    if (!c) return PFFRG_ERR_ARG;
    c->timing = enable != 0;
    return 0;
}

int pffrg_get_timing(pffrg_ctx *c, double *total_ms, int64_t *launches) {
    if (!c || !total_ms || !launches) return PFFRG_ERR_ARG;
    *total_ms = c->time_ms;
    *launches = c->time_launches;
    return 0;
}

int pffrg_gather_ceiling(pffrg_ctx *c, int32_t slot, int32_t precision, int32_t waves_per_simd, int64_t footprint,
                         int32_t n_iter, int32_t n_rep, double *gbps, double *ms_per_launch) {
    if (int e = check_slot(c, slot)) return e;
    if (!gbps || !ms_per_launch) return fail(c, PFFRG_ERR_ARG, "null output");
    if (precision != PFFRG_FP64 && precision != PFFRG_FP32_MIXED) return fail(c, PFFRG_ERR_ARG, "unknown precision");
    if (waves_per_simd < 0 || waves_per_simd > 8 || footprint < 0 || n_iter < 1 || n_iter > (1 << 20) || n_rep < 1 ||
        n_rep > 1000)
        return fail(c, PFFRG_ERR_ARG,
                    "gather ceiling: waves_per_simd in 0..8, footprint >= 0, n_iter in 1..2^20, n_rep in 1..1000");
    HIPCHK(c, hipSetDevice(c->device));
    const bool f32 = precision == PFFRG_FP32_MIXED;
    if (int e = ensure_gather(c, slot, f32)) return e;
    int cus = 0, lds_cu = 0;
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    HIPCHK(c, hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device));
    // two rounds of resident waves; a nonzero cap reserves LDS so that at most 4 * cap one-wave
    // blocks fit on a CU, i.e. cap waves per SIMD
    const int wps = waves_per_simd > 0 ? waves_per_simd : 8;
    const int n_blocks = 2 * cus * 4 * wps;
    const size_t lds = waves_per_simd > 0 ? (size_t)(lds_cu / (4 * wps)) & ~(size_t)255 : 0;
    float *out = nullptr;
    HIPCHK(c, hipMalloc(&out, (size_t)n_blocks * 64 * sizeof(float)));
    const void *g = f32 ? (const void *)c->slots[slot].gat32 : (const void *)c->slots[slot].gat64;
    // ν′ planes the lookups may touch: all, or as many as fit the footprint (corners included)
    const Sys &y = c->sys;
    const double plane = 3.0 * y.ns * y.nO * y.nn * 2.0 * (f32 ? 4.0 : 8.0);  // one ν′ plane, 3 channels
    int planes = y.nn - 1;
    if (footprint > 0) planes = std::max(1, std::min(planes, (int)(footprint / plane) - 1));
    hipError_t e = launch_gather_probe(g, f32, y, planes, n_blocks, lds, n_iter, 1u, out, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    for (int r = 0; r < n_rep && e == hipSuccess; ++r)
        e = launch_gather_probe(g, f32, y, planes, n_blocks, lds, n_iter, 2u + (uint32_t)r, out, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    (void)hipFree(out);
    if (e != hipSuccess) return fail(c, PFFRG_ERR_HIP, std::string("gather ceiling: ") + hipGetErrorString(e));
    const double bytes = (double)n_blocks * c->sys.G * c->sys.ns * (double)n_iter * 48.0 * (f32 ? 4.0 : 8.0);
    *ms_per_launch = ms / n_rep;
    *gbps = bytes / (*ms_per_launch * 1e-3) / 1e9;
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------------------- batched problems
namespace {
int bfail(pffrg_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}

// LDS of a one-wave vertex-pass block that still lets 12 blocks share a CU, i.e. keeps the pass
// at its register-bound 3 waves per SIMD (measured on MI355X: 12216 bytes give 12 blocks per
// CU, 12896 bytes 11).
constexpr size_t LDS_FULL_OCCUPANCY = 12800;

// Launch layout of a batch for one precision: the problems whose blocks fit 12 per CU first,
// then the others, each class in creation order (larger problems first).  The two classes run
// as two launches, each with the LDS of its own largest problem, so one problem that needs
// more LDS does not lower the occupancy of all the others.
template <typename TV>
int batch_layout(pffrg_batch *b) {
    const int n = (int)b->ctx.size();
    std::vector<size_t> cur(n);
    for (int k = 0; k < n; ++k) cur[k] = pass_lds_bytes<P_KAT, TV>(b->ctx[k]);
    if (cur == b->layout_lds) return 0;
    std::vector<int> idx(n);
    for (int k = 0; k < n; ++k) idx[k] = k;
    std::stable_partition(idx.begin(), idx.end(), [&](int k) { return cur[k] <= LDS_FULL_OCCUPANCY; });
    std::vector<pffrg_ctx *> ctx(n);
    std::vector<int> order(n);
    std::vector<uint32_t> codes;
    int n_full = 0;
    for (int k = 0; k < n; ++k) {
        ctx[k] = b->ctx[idx[k]];
        order[k] = b->order[idx[k]];
        for (uint32_t e : element_list(ctx[k]->sys.nO, ctx[k]->sys.nn, 0)) codes.push_back(e | ((uint32_t)k << 26));
        if (cur[idx[k]] <= LDS_FULL_OCCUPANCY) n_full = (int)codes.size();
    }
    if ((int)codes.size() != b->n_codes) return bfail(b, PFFRG_ERR_STATE, "batch element count changed");
    if (hipMemcpy(b->d_codes, codes.data(), codes.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(b, PFFRG_ERR_HIP, "element list upload failed");
    b->ctx = ctx;
    b->order = order;
    b->n_codes_full = n_full;
    b->layout_lds.resize(n);
    for (int k = 0; k < n; ++k) b->layout_lds[k] = cur[idx[k]];
    return 0;
}

// compute_dΣ! of every problem of a batch as one launch over all their σ points (each point as
// k_sigma computes it), on the batch stream, after the contexts' streams have drained
int batch_dsigma(pffrg_batch *b, const double *Lam, const int32_t *a, const int32_t *da, const double *stol) {
    const int n = (int)b->ctx.size();
    std::vector<SigmaArgs> tab(n);
    for (int k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        const int q = b->order[k];
        tab[k] = sigma_args(c, Lam[q], a[q], a[q], c->slots[da[q]].Sigma, nullptr, nullptr, stol);
    }
    if (!b->d_sig_pts) {
        std::vector<uint32_t> pts;
        for (int k = 0; k < n; ++k)
            for (int si = 1; si < b->ctx[k]->sys.nsig; ++si) pts.push_back(((uint32_t)k << 16) | (uint32_t)si);
        if (hipMalloc(&b->d_sig_pts, std::max<size_t>(1, pts.size()) * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc(&b->d_sig_tab, n * sizeof(SigmaArgs)) != hipSuccess)
            return bfail(b, PFFRG_ERR_HIP, "hipMalloc failed");
        if (!pts.empty() &&
            hipMemcpy(b->d_sig_pts, pts.data(), pts.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
            return bfail(b, PFFRG_ERR_HIP, "σ point list upload failed");
        b->n_sig_pts = (int)pts.size();
    }
    for (pffrg_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
    if (hipMemcpy(b->d_sig_tab, tab.data(), n * sizeof(SigmaArgs), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(b, PFFRG_ERR_HIP, "σ argument table upload failed");
    if (b->n_sig_pts > 0) {
        hipLaunchKernelGGL((k_sigma_batch<SK_DSIGMA>), dim3(b->n_sig_pts), dim3(SIGMA_T), 0, b->stream,
                           reinterpret_cast<const SigmaArgs *>(b->d_sig_tab), b->d_sig_pts);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) return bfail(b, PFFRG_ERR_HIP, std::string("batched dΣ: ") + hipGetErrorString(e));
    }
    return 0;
}

template <typename TV>
int batch_rhs_1l_t(pffrg_batch *b, const double *Lam, const int32_t *a, const int32_t *da, const int32_t *eval,
                   const double *stol, const double *gtol) {
    if (int rc = batch_layout<TV>(b)) return rc;
    if (int rc = batch_dsigma(b, Lam, a, da, stol)) return rc;
    const int n = (int)b->ctx.size();
    const bool f32 = std::is_same<TV, float>::value;
    std::vector<PassArgs<TV>> tab(n);
    size_t lds_full = 0, lds_rest = 0;
    for (int k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        const int q = b->order[k];
        int rc;
        if ((rc = launch_corrs(c, Lam[q], a[q], da[q], true, gtol))) return bfail(b, rc, c->err);
        if ((rc = ensure_gather(c, a[q], f32))) return bfail(b, rc, c->err);
        if ((rc = pass_args<P_KAT, TV>(c, Lam[q], a[q], da[q], da[q], eval[q], gtol, 0, tab[k])))
            return bfail(b, rc, c->err);
        const size_t l = pass_lds_bytes<P_KAT, TV>(c);
        if (l <= LDS_FULL_OCCUPANCY) lds_full = std::max(lds_full, l);
        else lds_rest = std::max(lds_rest, l);
    }
    for (pffrg_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
    const size_t bytes = n * sizeof(PassArgs<TV>);
    if (bytes > b->tab_bytes) {
        if (b->d_tab) (void)hipFree(b->d_tab);
        b->d_tab = nullptr;
        if (hipMalloc(&b->d_tab, bytes) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "hipMalloc failed");
        b->tab_bytes = bytes;
    }
    if (hipMemcpy(b->d_tab, tab.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        return bfail(b, PFFRG_ERR_HIP, "argument table upload failed");
    const PassArgs<TV> *dtab = reinterpret_cast<const PassArgs<TV> *>(b->d_tab);
    hipError_t e = hipSuccess;
    if (b->n_codes_full > 0)
        e = launch_vertex_pass_batch<P_KAT, TV>(dtab, b->d_codes, b->n_codes_full, lds_full, b->stream);
    if (e == hipSuccess && b->n_codes > b->n_codes_full)
        e = launch_vertex_pass_batch<P_KAT, TV>(dtab, b->d_codes + b->n_codes_full, b->n_codes - b->n_codes_full,
                                                lds_rest, b->stream);
    if (e != hipSuccess) return bfail(b, PFFRG_ERR_HIP, std::string("batched vertex pass: ") + hipGetErrorString(e));
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess)
        return bfail(b, PFFRG_ERR_HIP, std::string("batched vertex pass: ") + hipGetErrorString(e));
    for (int k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        const int q = b->order[k];
        c->stat_elems += c->n_half;
        c->stat_launches += 1;
        invalidate(c, da[q]);
        int rc;
        if ((rc = do_symmetrize(c, da[q]))) return bfail(b, rc, c->err);
    }
    for (pffrg_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
    return 0;
}

// The FP32-mixed batch: dΣ and the tail factors per problem, then the shared-node Katanin pass of
// every problem as one launch (the units of all problems, interleaved so that each part of the
// grid mixes problems), then each problem's partial sums and symmetrisation on its own stream.
int batch_rhs_1l_shared(pffrg_batch *b, const double *Lam, const int32_t *a, const int32_t *da, const int32_t *eval,
                        const double *stol, const double *gtol) {
    if (int rc = batch_dsigma(b, Lam, a, da, stol)) return rc;
    const int n = (int)b->ctx.size();
    std::vector<SharedArgs> tab(n);
    std::vector<int> spts(n);
    for (int k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        const int q = b->order[k];
        int rc;
        if ((rc = launch_corrs(c, Lam[q], a[q], da[q], true, gtol))) return bfail(b, rc, c->err);
        if ((rc = ensure_gather(c, a[q], true))) return bfail(b, rc, c->err);
        if ((rc = shared_args<P_KAT>(c, Lam[q], a[q], da[q], da[q], eval[q], gtol, 0, tab[k]))) return bfail(b, rc, c->err);
        spts[k] = tab[k].spt;
    }
    // The problems are grouped by sites per wave, which sizes the kernel's per-lane state: one
    // launch per group (each launch holds every problem of its group, its units interleaved
    // round-robin so that each part of the grid mixes problems), so that a problem with three
    // sites per wave does not give every other problem its larger state.  Each problem keeps its
    // own launch shape, LDS layout and batch size inside its group's launch.
    if (!b->d_sh_units || b->sh_spts != spts) {
        if (b->d_sh_units) (void)hipFree(b->d_sh_units);
        if (b->d_sh_units2) (void)hipFree(b->d_sh_units2);
        if (b->d_sh_cost) (void)hipFree(b->d_sh_cost);
        b->d_sh_units = nullptr;
        b->d_sh_units2 = nullptr;
        b->d_sh_cost = nullptr;
        std::vector<uint32_t> all;
        b->sh_groups.clear();
        for (int g = 1; g <= SH_SPT; ++g) {
            std::vector<int> mem;
            for (int k = 0; k < n; ++k)
                if (spts[k] == g) mem.push_back(k);
            if (mem.empty()) continue;
            std::vector<std::vector<uint32_t>> per;
            size_t total = 0;
            for (int k : mem) {
                per.push_back(shared_units(b->ctx[k], 0));
                total += per.back().size();
            }
            const size_t first = all.size();
            std::vector<size_t> pos(mem.size(), 0);
            while (all.size() - first < total)
                for (size_t j = 0; j < mem.size(); ++j)
                    if (pos[j] < per[j].size()) all.push_back(per[j][pos[j]++] | ((uint32_t)mem[j] << 26));
            if (total) b->sh_groups.push_back({g, (int)first, (int)total});
        }
        const size_t ubytes = std::max<size_t>(1, all.size()) * sizeof(uint32_t);
        if (hipMalloc(&b->d_sh_units, ubytes) != hipSuccess || hipMalloc(&b->d_sh_units2, ubytes) != hipSuccess ||
            hipMalloc(&b->d_sh_cost, ubytes) != hipSuccess ||
            (!b->d_sh_tab && hipMalloc(&b->d_sh_tab, n * sizeof(SharedArgs)) != hipSuccess))
            return bfail(b, PFFRG_ERR_HIP, "hipMalloc failed");
        std::vector<uint32_t> prior(all.size());
        for (size_t k = 0; k < all.size(); ++k) prior[k] = shared_unit_prior(all[k]);
        if (!all.empty() &&
            (hipMemcpy(b->d_sh_units, all.data(), all.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(b->d_sh_cost, prior.data(), all.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess))
            return bfail(b, PFFRG_ERR_HIP, "unit list upload failed");
        b->n_sh_units = (int)all.size();
        b->sh_spts = spts;
    }
    for (pffrg_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
    if (hipMemcpy(b->d_sh_tab, tab.data(), n * sizeof(SharedArgs), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(b, PFFRG_ERR_HIP, "argument table upload failed");
    hipError_t e = hipSuccess;
    std::vector<hipStream_t> used;
    for (const auto &grp : b->sh_groups) {
        // the group's launch: the threads and LDS of the member that needs most; the groups go to
        // different streams (the batch's, then a member's own) so that they can overlap
        size_t lds = 0;
        int nthreads = 64, first = -1;
        for (int k = 0; k < n; ++k)
            if (spts[k] == grp[0]) {
                nthreads = std::max(nthreads, shared_threads(tab[k]));
                lds = std::max(lds, shared_lds_bytes<P_KAT>(b->ctx[k], tab[k]));
                if (first < 0) first = k;
            }
        const hipStream_t st = used.empty() ? b->stream : b->ctx[first]->stream;
        used.push_back(st);
        // the group's units of the last launch, longest first, into the other list
        e = launch_order_units(b->d_sh_units + grp[1], b->d_sh_cost + grp[1], b->d_sh_units2 + grp[1], grp[2], st);
        if (e != hipSuccess) break;
        e = launch_shared_pass_batch<P_KAT>(reinterpret_cast<const SharedArgs *>(b->d_sh_tab), b->d_sh_units2 + grp[1],
                                            b->d_sh_cost + grp[1], grp[2], nthreads, grp[0], lds, st);
        if (e != hipSuccess) break;
    }
    // the next launch orders the lists just run (every group was reordered when no launch failed)
    if (e == hipSuccess) std::swap(b->d_sh_units, b->d_sh_units2);
    for (hipStream_t st : used)
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bfail(b, PFFRG_ERR_HIP, std::string("batched shared-node pass: ") + hipGetErrorString(e));
    for (int k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        const int q = b->order[k];
        if ((e = launch_shared_finish<P_KAT>(tab[k], c->stream)) != hipSuccess)
            return bfail(b, PFFRG_ERR_HIP, std::string("partial sums: ") + hipGetErrorString(e));
        c->stat_elems += c->n_half;
        c->stat_launches += 1;
        invalidate(c, da[q]);
        int rc;
        if ((rc = do_symmetrize(c, da[q]))) return bfail(b, rc, c->err);
    }
    for (pffrg_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
    return 0;
}

// ---- χ of one or several problems (compute_χ!, correlation_su2.jl:58-86)
struct ChiJob {
    pffrg_ctx *c;
    double L;
    int slot;
    const double *chi_mesh;
    int nchi;
    double *chi;
    int32_t *nreg;
    int comp;  // U(1)-DM correlation component (χxx, χzz, χxy)
};

#define CHICHK(expr)                                                      \
    do {                                                                  \
        hipError_t e_ = (expr);                                           \
        if (e_ != hipSuccess) {                                           \
            msg = std::string(#expr) + ": " + hipGetErrorString(e_);      \
            return PFFRG_ERR_HIP;                                         \
        }                                                                 \
    } while (0)

// every (problem, χ frequency, site, box) task of the jobs through one set of cubature rounds
int chi_run(void *&mem, size_t &bytes, hipStream_t stream, const ChiJob *jobs, int njobs, const double *tol,
            std::string &msg) {
    auto r256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    std::vector<ChiProb> probs(njobs);
    std::vector<int> task_prob, prop_prob;
    long tasks = 0;
    int nfreq = 0;
    for (int k = 0; k < njobs; ++k) {
        const pffrg_ctx *c = jobs[k].c;
        const long tk = (long)jobs[k].nchi * c->sys.ns * 9;
        probs[k].task0 = (int)tasks;
        probs[k].prop0 = nfreq;
        task_prob.insert(task_prob.end(), tk, k);
        prop_prob.insert(prop_prob.end(), jobs[k].nchi, k);
        tasks += tk;
        nfreq += jobs[k].nchi;
    }
    // every task first runs in a slot of CHI_CAP0 regions; the few that need more run again in
    // slots of the full budget, CHI_POOL_SLOTS at a time (the result does not depend on the slot size)
    const long CHI_CAP0 = 16384, CHI_POOL_SLOTS = 64;
    const long max_slots = std::max(tasks, CHI_POOL_SLOTS);
    const size_t pool = std::max((size_t)tasks * CHI_CAP0, (size_t)CHI_POOL_SLOTS * CHI_MAX_REGIONS);
    // split-table entries: at most one per slot plus one per CHI_RPB regions
    const size_t tab_cap = std::max((size_t)tasks * (CHI_CAP0 / CHI_RPB + 1),
                                    (size_t)CHI_POOL_SLOTS * (CHI_MAX_REGIONS / CHI_RPB + 1));
    // per call: problems, box/nreg/status/task_prob per task, prop/prop_prob/χ mesh per frequency;
    // per slot: task, active list, rn, rk, done; the split table; the region pool
    const size_t fixed = r256(njobs * sizeof(ChiProb)) + r256(tasks * 8) + 3 * r256(tasks * 4) +
                         2 * r256((size_t)nfreq * 8) + r256((size_t)nfreq * 4) + 3 * r256(max_slots * 4) +
                         2 * r256(max_slots * 8) + r256(tab_cap * 8);
    const size_t need = fixed + 6 * r256(pool * 8) + 2 * r256(pool * 4);
    if (need > bytes) {
        CHICHK(hipStreamSynchronize(stream));
        CHICHK(hipFree(mem));
        mem = nullptr;
        bytes = 0;
        CHICHK(hipMalloc(&mem, need));
        bytes = need;
    }
    char *p = (char *)mem;
    auto take = [&](size_t n) {
        char *q = p;
        p += r256(n);
        return (void *)q;
    };
    ChiArgs Q;
    ChiProb *d_probs = (ChiProb *)take(njobs * sizeof(ChiProb));
    Q.probs = d_probs;
    Q.box = (double *)take(tasks * 8);
    Q.nreg = (int *)take(tasks * 4);
    Q.status = (int *)take(tasks * 4);
    int *d_task_prob = (int *)take(tasks * 4);
    Q.task_prob = d_task_prob;
    Q.prop = (double *)take((size_t)nfreq * 8);
    double *d_chi = (double *)take((size_t)nfreq * 8);
    int *d_prop_prob = (int *)take((size_t)nfreq * 4);
    Q.prop_prob = d_prop_prob;
    Q.nfreq = nfreq;
    int *d_tasks = (int *)take(max_slots * 4);
    int *d_active = (int *)take(max_slots * 4);
    Q.rdone = (int *)take(max_slots * 4);
    Q.rn = (long *)take(max_slots * 8);
    Q.rk = (long *)take(max_slots * 8);
    int *d_tab = (int *)take(tab_cap * 8);
    Q.tasks = d_tasks;
    Q.active = d_active;
    Q.split_tab = d_tab;
    Q.rc0 = (double *)take(pool * 8);
    Q.rc1 = (double *)take(pool * 8);
    Q.rh0 = (double *)take(pool * 8);
    Q.rh1 = (double *)take(pool * 8);
    Q.rval = (double *)take(pool * 8);
    Q.rerr = (double *)take(pool * 8);
    Q.rsplit = (int *)take(pool * 4);
    Q.rsel = (int *)take(pool * 4);
    if ((size_t)(p - (char *)mem) > bytes) {
        msg = "chi scratch layout overflow";
        return PFFRG_ERR_LIMIT;
    }
    Q.atol = tol[0];
    Q.rtol = tol[1];
    std::vector<double> hchi(nfreq);
    for (int k = 0; k < njobs; ++k) {
        const pffrg_ctx *c = jobs[k].c;
        ChiProb &P = probs[k];
        P.y = c->sys;
        P.mesh = c->d_mesh;
        const Slot &S = c->slots[jobs[k].slot];
        P.A = ActV<double>{S.gamma, S.bare, S.Sigma};
        P.exch = c->d_exch;
        P.chi = d_chi + P.prop0;
        P.L = jobs[k].L;
        P.nchi = jobs[k].nchi;
        P.comp = jobs[k].comp;
        for (int w = 0; w < jobs[k].nchi; ++w) hchi[P.prop0 + w] = jobs[k].chi_mesh[w];
    }
    CHICHK(hipMemcpyAsync(d_probs, probs.data(), njobs * sizeof(ChiProb), hipMemcpyHostToDevice, stream));
    CHICHK(hipMemcpyAsync(d_task_prob, task_prob.data(), tasks * sizeof(int), hipMemcpyHostToDevice, stream));
    CHICHK(hipMemcpyAsync(d_prop_prob, prop_prob.data(), nfreq * sizeof(int), hipMemcpyHostToDevice, stream));
    CHICHK(hipMemcpyAsync(d_chi, hchi.data(), nfreq * sizeof(double), hipMemcpyHostToDevice, stream));
    CHICHK(launch_chi_prop(Q, stream));
    // the rounds of one pass over a list of tasks with slots of cap regions
    std::vector<int> act, next, tab, done;
    std::vector<long> rk;
    auto run_pass = [&](const int *list, int nslots, long cap) -> int {
        Q.cap = cap;
        Q.nslots = nslots;
        CHICHK(hipMemcpyAsync(d_tasks, list, nslots * sizeof(int), hipMemcpyHostToDevice, stream));
        CHICHK(launch_chi_init(Q, stream));
        act.resize(nslots);
        for (int k = 0; k < nslots; ++k) act[k] = k;
        rk.assign(nslots, 0);
        done.assign(nslots, 0);
        long rounds = 0;
        while (!act.empty()) {
            // every round adds a region to each unfinished slot, so cap + 1 rounds end any pass
            if (++rounds > cap + 1) {
                msg = "chi cubature made no progress";
                return PFFRG_ERR_STATE;
            }
            CHICHK(hipMemcpyAsync(d_active, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice, stream));
            CHICHK(launch_chi_round(Q, (int)act.size(), stream));
            CHICHK(hipMemcpyAsync(rk.data(), Q.rk, nslots * sizeof(long), hipMemcpyDeviceToHost, stream));
            CHICHK(hipMemcpyAsync(done.data(), Q.rdone, nslots * sizeof(int), hipMemcpyDeviceToHost, stream));
            CHICHK(hipStreamSynchronize(stream));
            next.clear();
            tab.clear();
            for (int sl : act) {
                if (done[sl]) continue;
                next.push_back(sl);
                for (long f = 0; f < rk[sl]; f += CHI_RPB) {
                    tab.push_back(sl);
                    tab.push_back((int)f);
                }
            }
            if (tab.size() / 2 > tab_cap) {
                msg = "chi split table overflow";
                return PFFRG_ERR_LIMIT;
            }
            if (!tab.empty()) {
                CHICHK(hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, stream));
                CHICHK(launch_chi_split(Q, (int)(tab.size() / 2), stream));
            }
            act.swap(next);
        }
        return 0;
    };
    int rc;
    std::vector<int> list(tasks);
    for (long t = 0; t < tasks; ++t) list[t] = (int)t;
    if ((rc = run_pass(list.data(), (int)tasks, CHI_CAP0))) return rc;
    std::vector<double> box(tasks), prop(nfreq);
    std::vector<int> nr(tasks), st(tasks);
    CHICHK(hipMemcpy(st.data(), Q.status, tasks * sizeof(int), hipMemcpyDeviceToHost));
    // tasks whose regions outgrew the first slot size: again with the full budget
    std::vector<int> again;
    for (long t = 0; t < tasks; ++t)
        if (st[t] != 0) again.push_back((int)t);
    for (size_t b0 = 0; b0 < again.size(); b0 += CHI_POOL_SLOTS) {
        const int nb = (int)std::min((size_t)CHI_POOL_SLOTS, again.size() - b0);
        if ((rc = run_pass(again.data() + b0, nb, CHI_MAX_REGIONS))) return rc;
    }
    CHICHK(hipMemcpyAsync(box.data(), Q.box, tasks * sizeof(double), hipMemcpyDeviceToHost, stream));
    CHICHK(hipMemcpyAsync(nr.data(), Q.nreg, tasks * sizeof(int), hipMemcpyDeviceToHost, stream));
    CHICHK(hipMemcpyAsync(st.data(), Q.status, tasks * sizeof(int), hipMemcpyDeviceToHost, stream));
    CHICHK(hipMemcpyAsync(prop.data(), Q.prop, nfreq * sizeof(double), hipMemcpyDeviceToHost, stream));
    CHICHK(hipStreamSynchronize(stream));
    for (long t = 0; t < tasks; ++t)
        if (st[t] != 0) {
            msg = "chi cubature slot overflow";
            return PFFRG_ERR_LIMIT;
        }
    for (int k = 0; k < njobs; ++k) {
        const int ns = jobs[k].c->sys.ns, t0 = probs[k].task0, f0 = probs[k].prop0;
        for (int w = 0; w < jobs[k].nchi; ++w)
            for (int i = 0; i < ns; ++i) {
                const long tb = t0 + ((long)w * ns + i) * 9;
                double I = 0.0;
                for (int b = 0; b < 9; ++b) I += box[tb + b];
                if (i == 0 && (jobs[k].c->sys.nc == 2 || jobs[k].comp < 2)) I += prop[f0 + w];
                jobs[k].chi[i + (long)ns * w] = I;
                if (jobs[k].nreg)
                    for (int b = 0; b < 9; ++b) jobs[k].nreg[((long)w * ns + i) * 9 + b] = nr[tb + b];
            }
    }
    return 0;
}

#undef CHICHK

}  // namespace

extern "C" {

int pffrg_batch_create(pffrg_ctx *const *ctxs, int32_t n, pffrg_batch **out) {
    if (!ctxs || !out || n < 1 || n > 64) return PFFRG_ERR_ARG;
    *out = nullptr;
    pffrg_batch *b = new pffrg_batch();
    b->order.resize(n);
    for (int k = 0; k < n; ++k) {
        if (!ctxs[k] || ready(ctxs[k])) {
            delete b;
            return PFFRG_ERR_STATE;
        }
        if (ctxs[k]->sys.nc != 2) {  // the batched launch serves the SU(2) action
            delete b;
            return PFFRG_ERR_ARG;
        }
        b->order[k] = k;
    }
    b->device = ctxs[0]->device;
    for (int k = 1; k < n; ++k)
        if (ctxs[k]->device != b->device) {
            delete b;
            return PFFRG_ERR_ARG;
        }
    // larger problems first, so the most expensive elements start early
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int x, int y) { return ctxs[x]->sys.ns * ctxs[x]->n_half > ctxs[y]->sys.ns * ctxs[y]->n_half; });
    std::vector<uint32_t> codes;
    for (int k = 0; k < n; ++k) {
        pffrg_ctx *c = ctxs[b->order[k]];
        b->ctx.push_back(c);
        for (uint32_t e : element_list(c->sys.nO, c->sys.nn, 0)) codes.push_back(e | ((uint32_t)k << 26));
    }
    b->n_codes = (int)codes.size();
    bool ok = hipSetDevice(b->device) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&b->d_codes, codes.size() * sizeof(uint32_t)) == hipSuccess;
    ok = ok && hipMemcpy(b->d_codes, codes.data(), codes.size() * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        pffrg_batch_destroy(b);
        return PFFRG_ERR_HIP;
    }
    *out = b;
    return 0;
}

void pffrg_batch_destroy(pffrg_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->d_codes) (void)hipFree(b->d_codes);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->d_sh_units) (void)hipFree(b->d_sh_units);
    if (b->d_sh_units2) (void)hipFree(b->d_sh_units2);
    if (b->d_sh_cost) (void)hipFree(b->d_sh_cost);
    if (b->d_sh_tab) (void)hipFree(b->d_sh_tab);
    if (b->d_sig_pts) (void)hipFree(b->d_sig_pts);
    if (b->d_sig_tab) (void)hipFree(b->d_sig_tab);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->chi_mem) (void)hipFree(b->chi_mem);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *pffrg_batch_last_error(const pffrg_batch *b) { return b ? b->err.c_str() : "null batch"; }

int pffrg_batch_rhs_1l(pffrg_batch *b, const double *Lambda, const int32_t *a, const int32_t *da, const int32_t *eval,
                       const double *sigma_tol, const double *gamma_tol) {
    if (!b || !Lambda || !a || !da || !eval || !sigma_tol || !gamma_tol) return PFFRG_ERR_ARG;
    const int n = (int)b->ctx.size();
    const int prec = b->ctx[0]->precision;
    for (int k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        const int q = b->order[k];
        if (c->precision != prec) return bfail(b, PFFRG_ERR_STATE, "all problems of a batch need the same precision");
        if (check_slot(c, a[q]) || check_slot(c, da[q]) || a[q] == da[q])
            return bfail(b, PFFRG_ERR_ARG, "invalid slot of problem " + std::to_string(q));
        if (!(Lambda[q] > 0.0) || eval[q] < 1) return bfail(b, PFFRG_ERR_ARG, "invalid Lambda or eval of problem " + std::to_string(q));
    }
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "hipSetDevice failed");
    bool all_fit = true;
    for (pffrg_ctx *c : b->ctx) all_fit = all_fit && shared_fits(c);
    if (prec == PFFRG_FP32_MIXED && all_fit) return batch_rhs_1l_shared(b, Lambda, a, da, eval, sigma_tol, gamma_tol);
    if (prec == PFFRG_FP32_MIXED) return batch_rhs_1l_t<float>(b, Lambda, a, da, eval, sigma_tol, gamma_tol);
    return prec == PFFRG_FP32_ELEMENT ? batch_rhs_1l_t<float>(b, Lambda, a, da, eval, sigma_tol, gamma_tol)
                                      : batch_rhs_1l_t<double>(b, Lambda, a, da, eval, sigma_tol, gamma_tol);
}

// The BS3 stage algebra of every problem of a batch (Action.jl:181-359): each problem's kernels go
// to its own stream and all streams are synchronised once at the end, where the per-context
// calls synchronise after every problem.  Every slot is checked before anything is queued.
int pffrg_batch_replace_with(pffrg_batch *b, int32_t dst, int32_t src, int32_t vertex_only) {
    if (!b) return PFFRG_ERR_ARG;
    for (pffrg_ctx *c : b->ctx)
        if (check_slot(c, dst) || check_slot(c, src)) return bfail(b, PFFRG_ERR_ARG, "invalid slot");
    if (dst == src) return 0;
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "hipSetDevice failed");
    for (pffrg_ctx *c : b->ctx)
        if (int rc = do_copy(c, dst, src, vertex_only != 0)) return bfail(b, rc, c->err);
    for (pffrg_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
    return 0;
}

int pffrg_batch_mult_with_add_to(pffrg_batch *b, int32_t src, const double *fac, int32_t dst, int32_t vertex_only) {
    if (!b) return PFFRG_ERR_ARG;
    if (!fac) return bfail(b, PFFRG_ERR_ARG, "null factors");
    if (src == dst) return bfail(b, PFFRG_ERR_ARG, "source and destination must differ");
    for (pffrg_ctx *c : b->ctx)
        if (check_slot(c, dst) || check_slot(c, src)) return bfail(b, PFFRG_ERR_ARG, "invalid slot");
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "hipSetDevice failed");
    for (size_t k = 0; k < b->ctx.size(); ++k)
        if (int rc = do_axpy(b->ctx[k], src, fac[b->order[k]], dst, vertex_only != 0)) return bfail(b, rc, b->ctx[k]->err);
    for (pffrg_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
    return 0;
}

int pffrg_batch_abs_max(pffrg_batch *b, int32_t s, double *out) {
    if (!b) return PFFRG_ERR_ARG;
    if (!out) return bfail(b, PFFRG_ERR_ARG, "null output");
    for (pffrg_ctx *c : b->ctx)
        if (check_slot(c, s)) return bfail(b, PFFRG_ERR_ARG, "invalid slot");
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "hipSetDevice failed");
    const size_t n = b->ctx.size();
    std::vector<unsigned long long> bits(n, 0ull);
    for (size_t k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        const long m = 3L * c->sys.nc * c->sys.ns * c->sys.nO * c->sys.nn * c->sys.nn;
        if (hipMemsetAsync(c->d_max, 0, sizeof(unsigned long long), c->stream) != hipSuccess)
            return bfail(b, PFFRG_ERR_HIP, "hipMemsetAsync failed");
        hipLaunchKernelGGL(k_absmax_q3, dim3(grid_for(m)), dim3(256), 0, c->stream, c->slots[s].gamma, c->sys, c->d_max);
        if (hipGetLastError() != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "k_absmax_q3 launch failed");
    }
    for (size_t k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        if (hipMemcpyAsync(&bits[k], c->d_max, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            return bfail(b, PFFRG_ERR_HIP, "read-back failed");
    }
    for (pffrg_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
    for (size_t k = 0; k < n; ++k) std::memcpy(&out[b->order[k]], &bits[k], sizeof(double));
    return 0;
}

// compute_χ! of every problem of a batch in one set of cubature rounds
int pffrg_batch_compute_chi(pffrg_batch *b, const double *Lambda, const int32_t *a, const double *const *chi_mesh,
                            const int32_t *n_chi, const double *tol, double *const *chi) {
    if (!b) return PFFRG_ERR_ARG;
    if (!Lambda || !a || !chi_mesh || !n_chi || !tol || !chi) return bfail(b, PFFRG_ERR_ARG, "null argument");
    const int n = (int)b->ctx.size();
    std::vector<ChiJob> jobs;
    for (int k = 0; k < n; ++k) {
        pffrg_ctx *c = b->ctx[k];
        const int q = b->order[k];
        if (n_chi[q] == 0) continue;  // this problem sits out
        int rc;
        if ((rc = ready(c)) || (rc = check_slot(c, a[q]))) return bfail(b, rc, c->err);
        if (!chi_mesh[q] || !chi[q] || n_chi[q] < 1 || !(Lambda[q] > 0.0))
            return bfail(b, PFFRG_ERR_ARG, "invalid chi arguments");
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "stream synchronize failed");
        jobs.push_back(ChiJob{c, Lambda[q], a[q], chi_mesh[q], n_chi[q], chi[q], nullptr, 0});
    }
    if (jobs.empty()) return 0;
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, PFFRG_ERR_HIP, "hipSetDevice failed");
    std::string msg;
    const int rc = chi_run(b->chi_mem, b->chi_bytes, b->stream, jobs.data(), (int)jobs.size(), tol, msg);
    return rc ? bfail(b, rc, msg) : 0;
}

// compute_χ! (correlation_su2.jl:58-86)
int pffrg_compute_chi(pffrg_ctx *c, double L, int32_t a, const double *chi_mesh, int32_t nchi, const double *tol,
                      double *chi, int32_t *nreg) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a))) return rc;
    if (!chi_mesh || !chi || !tol || nchi < 1 || !(L > 0.0)) return fail(c, PFFRG_ERR_ARG, "invalid chi arguments");
    HIPCHK(c, hipSetDevice(c->device));
    std::string msg;
    if (c->sys.nc == 6) {
        // U(1)-DM: χxx, χzz, χxy, each its own cubature (correlation_u1_dm.jl:138-190)
        const long blk = (long)c->sys.ns * nchi;
        ChiJob jobs[3];
        for (int k = 0; k < 3; ++k)
            jobs[k] = ChiJob{c, L, a, chi_mesh, nchi, chi + k * blk, nreg ? nreg + 9 * k * blk : nullptr, k};
        rc = chi_run(c->chi_mem, c->chi_bytes, c->stream, jobs, 3, tol, msg);
        return rc ? fail(c, rc, msg) : 0;
    }
    ChiJob job{c, L, a, chi_mesh, nchi, chi, nreg, 0};
    rc = chi_run(c->chi_mem, c->chi_bytes, c->stream, &job, 1, tol, msg);
    return rc ? fail(c, rc, msg) : 0;
}

// scan_channel's argmax and cuts of slot a into the scratch at mem (SCAN_CUTS_BYTES), then to the host
static int scan_cuts_into(pffrg_ctx *c, int a, char *mem, int32_t *idx, double *cuts) {
    ScanArgs A;
    A.y = c->sys;
    A.gamma = c->slots[a].gamma;
    A.pval = (double *)mem;
    A.pidx = (long long *)(mem + SCAN_OFF_PIDX);
    A.idx = (int *)(mem + SCAN_OFF_IDX);
    A.cuts = (double *)(mem + SCAN_OFF_CUTS);
    HIPCHK(c, launch_scan_cuts(A, c->stream));
    const int ncut = c->sys.nO + 3 * c->sys.nn, rows = 3 * c->sys.nc;
    HIPCHK(c, hipMemcpyAsync(idx, A.idx, 4 * rows * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cuts, A.cuts, rows * (size_t)ncut * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// scan_channel's argmax and cuts (Action.jl:449-480)
int pffrg_scan_cuts(pffrg_ctx *c, int32_t a, int32_t *idx, double *cuts) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a))) return rc;
    if (!idx || !cuts) return fail(c, PFFRG_ERR_ARG, "null output pointer");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = ensure_scan_mem(c))) return rc;
    return scan_cuts_into(c, a, c->scan_mem, idx, cuts);
}

// pffrg_scan_cuts on this device's static scratch: no device memory is allocated or freed
int pffrg_scan_cuts_noalloc(pffrg_ctx *c, int32_t a, int32_t *idx, double *cuts) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, a))) return rc;
    if (!idx || !cuts) return fail(c, PFFRG_ERR_ARG, "null output pointer");
    const size_t need = SCAN_OFF_CUTS + 3 * (size_t)c->sys.nc * (c->sys.nO + 3 * c->sys.nn) * sizeof(double);
    if (need > SCAN_CUTS_BYTES)
        return fail(c, PFFRG_ERR_LIMIT, "scan_cuts_noalloc: cuts of " + std::to_string(need) +
                                            " bytes exceed the static scratch of " + std::to_string(SCAN_CUTS_BYTES));
    if (!c->scan_scratch) return fail(c, PFFRG_ERR_STATE, "scan_cuts_noalloc: no static scratch on this device");
    HIPCHK(c, hipSetDevice(c->device));
    return scan_cuts_into(c, a, c->scan_scratch, idx, cuts);
}

// resample_from_to (Action.jl:531-556): slot src onto new meshes into slot dst
int pffrg_resample(pffrg_ctx *c, int32_t src, int32_t dst, const double *sigma, const double *Os, const double *ns,
                   const double *Ot, const double *nt) {
    int rc;
    if ((rc = ready(c)) || (rc = check_slot(c, src)) || (rc = check_slot(c, dst))) return rc;
    if (src == dst) return fail(c, PFFRG_ERR_ARG, "resample needs two different slots");
    if (!sigma || !Os || !ns || !Ot || !nt) return fail(c, PFFRG_ERR_ARG, "null mesh pointer");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<double> h;
    if ((rc = mesh_table(c, sigma, Os, ns, Ot, nt, h))) return rc;
    if ((rc = ensure_scan_mem(c))) return rc;
    double *mnew = (double *)(c->scan_mem + SCAN_OFF_MESH);
    HIPCHK(c, hipMemcpyAsync(mnew, h.data(), 5 * MAXN * sizeof(double), hipMemcpyHostToDevice, c->stream));
    ResampleArgs A;
    A.y = c->sys;
    A.mold = c->d_mesh;
    A.mnew = mnew;
    A.src_gamma = c->slots[src].gamma;
    A.src_sigma = c->slots[src].Sigma;
    A.dst_gamma = c->slots[dst].gamma;
    A.dst_sigma = c->slots[dst].Sigma;
    HIPCHK(c, launch_resample(A, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->slots[dst].bare, c->slots[src].bare, c->sys.nc * c->sys.ns * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
    invalidate(c, dst);
    // the host table is released only after the copy from it has completed (apply_mesh waits)
    return apply_mesh(c, h);
}

}  // extern "C"
