// pffrg_pass.h -- the vertex-pass kernel (one launch per pass over every output element).
// Instantiated per pass type in pass_*.hip so the translation units compile in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pffrg_device.h"

namespace pffrg {

template <typename TV>
struct PassArgs {
    Sys y;
    const double *mesh;  // sigma | Os | ns | Ot | nt (each padded to MAXN)
    const int *exch;
    const int *ov_off;
    const uint32_t *ov_rows;  // j1 | j2 << 8 | mult << 16
    ActV<TV> A;               // propagator action and "full" vertex
    ActV<TV> D;               // kat: dΣ; left: left vertex; central: right vertex
    double *out;              // output vertex (FP64)
    const double *corrs;      // [nO][3][2]
    const uint32_t *elems;
    int n_elems;
    // breakpoints of the logarithmic domains [2ref, val], (eval+1) per (mesh family, Ω index),
    // computed on the host with libm pow; entry 0 < 0 marks an empty domain
    const double *logpts;
    int per_xcd;              // elements handled by each XCD slice of the grid
    double L;
    int eval;
    double atol, rtol;
    unsigned long long *node_count;  // optional
    // integrand probe: evaluate element dbg_elem at dbg_v[blockIdx] (no quadrature)
    const double *dbg_v;
    double *dbg_out;  // [n][2][ns]
    uint32_t dbg_elem;
    // per (element, sub-interval) node counts and spin results of site 0 (optional)
    int *dbg_sub_nodes;  // device buffers of dbg_cap entries
    double *dbg_sub_res;
    long dbg_cap;
};

struct Lds {
    const double *Os, *ns, *Ot, *nt, *sig, *sigA, *sigD;
    const int *exch, *ov_off;
    const uint32_t *ov;
    double *scr;  // 64 x 4 doubles of lane exchange scratch
};

// Integrand of one element at loop frequency v for site i: returns Γs, Γd (before the
// dv weight).  Restates compute_{s,t,u}_{kat,left,central}! (flow_lib/flow_su2/*.jl),
// compute_{s,t,u}_BSE! (parquet_lib/parquet_su2/BSE_su2_*.jl) and compute_s_reduced!
// (SDE_su2.jl:2-43).  Must be executed by every lane of the wavefront (the t channel
// exchanges the site vectors of a group through LDS).
template <int PASS, typename TV>
__device__ __forceinline__ void node_eval(const PassArgs<TV> &P, const Lds &S, int ch, double W, double vw,
                                          double vwp, double v, int i, int slot, int gbase, TV &Gs, TV &Gd,
                                          int &nc) {
    const Sys &y = P.y;
    const Meshes M{S.sig, S.Os, S.ns, S.Ot, S.nt};
    double x1, x2;
    bubble_args(ch, W, v, x1, x2);
    double p;
    if (PASS == P_KAT) p = bubble_kat(P.L, x1, x2, S.sig, y.nsig, S.sigA, S.sigD);
    else p = bubble_plain(P.L, x1, x2, S.sig, y.nsig, S.sigA);
    const TV tp = (TV)p;
    TV v1s, v1d, v2s, v2d;
    nc = 0;
    int nc0 = 0;
    if (ch == 0) {
        const double s = W, vs = vw, vsp = vwp;
        const Buf bs2 = buf_s(M, y, s, v, vsp);
        const Buf bt2 = (PASS == P_CENTRAL) ? buf_empty() : buf_t(M, y, -v - vsp, 0.5 * (s + v - vsp), 0.5 * (s - v + vsp));
        const Buf bu2 = (PASS == P_CENTRAL) ? buf_empty() : buf_u(M, y, v - vsp, 0.5 * (s + v + vsp), 0.5 * (s - v - vsp));
        if (PASS == P_REDUCED) {
            v1s = (TV)P.A.bare[i];
            v1d = (TV)P.A.bare[y.ns + i];
            nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        } else {
            const Buf bs1 = (PASS == P_LEFT || PASS == P_BSE) ? buf_empty() : buf_s(M, y, s, vs, -v);
            const Buf bt1 = buf_t(M, y, v - vs, 0.5 * (s + v + vs), 0.5 * (s - v - vs));
            const Buf bu1 = buf_u(M, y, v + vs, 0.5 * (s - v + vs), 0.5 * (s + v - vs));
            if (PASS == P_KAT) {
                nc += gamma_site<TV, true, true, true>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
                nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
            } else if (PASS == P_LEFT) {
                nc += gamma_site<TV, false, true, true>(P.D, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
                nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
            } else if (PASS == P_CENTRAL) {
                nc += gamma_site<TV, true, true, true>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
                nc += gamma_site<TV, true, false, false>(P.D, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
            } else {  // BSE
                nc += gamma_site<TV, false, true, true>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
                nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
            }
        }
        Gs = -tp * ((TV)-2.0 * v1s * v2s + v1s * v2d + v1d * v2s);
        Gd = -tp * ((TV)3.0 * v1s * v2s + v1d * v2d);
    } else if (ch == 2) {
        const double u = W, vu = vw, vup = vwp;
        const Buf bs1 = buf_s(M, y, v + vu, 0.5 * (u - v + vu), 0.5 * (-u - v + vu));
        const Buf bt1 = buf_t(M, y, v - vu, 0.5 * (u + v + vu), 0.5 * (-u + v + vu));
        const Buf bu1 = (PASS == P_LEFT || PASS == P_BSE) ? buf_empty() : buf_u(M, y, u, vu, v);
        const Buf bs2 = (PASS == P_CENTRAL) ? buf_empty() : buf_s(M, y, v + vup, 0.5 * (u + v - vup), 0.5 * (-u + v - vup));
        const Buf bt2 = (PASS == P_CENTRAL) ? buf_empty() : buf_t(M, y, -v + vup, 0.5 * (u + v + vup), 0.5 * (-u + v + vup));
        const Buf bu2 = buf_u(M, y, u, v, vup);
        if (PASS == P_KAT) {
            nc += gamma_site<TV, true, true, true>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
            nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        } else if (PASS == P_LEFT) {
            nc += gamma_site<TV, true, true, false>(P.D, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
            nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        } else if (PASS == P_CENTRAL) {
            nc += gamma_site<TV, true, true, true>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
            nc += gamma_site<TV, false, false, true>(P.D, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        } else {
            nc += gamma_site<TV, true, true, false>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
            nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        }
        Gs = -tp * ((TV)2.0 * v1s * v2s + v1s * v2d + v1d * v2s);
        Gd = -tp * ((TV)3.0 * v1s * v2s + v1d * v2d);
    } else {
        const double t = W, vt = vw, vtp = vwp;
        const Buf bs1 = buf_s(M, y, v + vt, 0.5 * (-t - v + vt), 0.5 * (-t + v - vt));
        const Buf bt1 = (PASS == P_LEFT || PASS == P_BSE) ? buf_empty() : buf_t(M, y, t, vt, v);
        const Buf bu1 = buf_u(M, y, -v + vt, 0.5 * (-t + v + vt), 0.5 * (t + v + vt));
        const Buf bs2 = (PASS == P_CENTRAL) ? buf_empty() : buf_s(M, y, v + vtp, 0.5 * (-t + v - vtp), 0.5 * (-t - v + vtp));
        const Buf bt2 = buf_t(M, y, t, v, vtp);
        const Buf bu2 = (PASS == P_CENTRAL) ? buf_empty() : buf_u(M, y, v - vtp, 0.5 * (-t + v + vtp), 0.5 * (t + v + vtp));
        const Buf bs3 = buf_s(M, y, v + vt, 0.5 * (t + v - vt), 0.5 * (-t + v - vt));
        const Buf bt3 = buf_t(M, y, -v + vt, 0.5 * (t + v + vt), 0.5 * (-t + v + vt));
        const Buf bu3 = (PASS == P_LEFT || PASS == P_BSE) ? buf_empty() : buf_u(M, y, t, v, vt);
        const Buf bs4 = (PASS == P_CENTRAL) ? buf_empty() : buf_s(M, y, v + vtp, 0.5 * (t - v + vtp), 0.5 * (-t - v + vtp));
        const Buf bt4 = (PASS == P_CENTRAL) ? buf_empty() : buf_t(M, y, v - vtp, 0.5 * (t + v + vtp), 0.5 * (-t + v + vtp));
        const Buf bu4 = buf_u(M, y, t, vtp, v);
        TV v3s, v3d, v4s, v4d;
        if (PASS == P_KAT) {
            nc0 += gamma_site<TV, true, true, true>(P.A, bs3, bt3, bu3, 0, S.exch, y, v3s, v3d);
            nc0 += gamma_site<TV, true, true, true>(P.A, bs4, bt4, bu4, 0, S.exch, y, v4s, v4d);
            nc += gamma_site<TV, true, true, true>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
            nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        } else if (PASS == P_LEFT) {
            nc0 += gamma_site<TV, true, true, false>(P.D, bs3, bt3, bu3, 0, S.exch, y, v3s, v3d);
            nc0 += gamma_site<TV, true, true, true>(P.A, bs4, bt4, bu4, 0, S.exch, y, v4s, v4d);
            nc += gamma_site<TV, true, false, true>(P.D, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
            nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        } else if (PASS == P_CENTRAL) {
            nc0 += gamma_site<TV, true, true, true>(P.A, bs3, bt3, bu3, 0, S.exch, y, v3s, v3d);
            nc0 += gamma_site<TV, false, false, true>(P.D, bs4, bt4, bu4, 0, S.exch, y, v4s, v4d);
            nc += gamma_site<TV, true, true, true>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
            nc += gamma_site<TV, false, true, false>(P.D, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        } else {
            nc0 += gamma_site<TV, true, true, false>(P.A, bs3, bt3, bu3, 0, S.exch, y, v3s, v3d);
            nc0 += gamma_site<TV, true, true, true>(P.A, bs4, bt4, bu4, 0, S.exch, y, v4s, v4d);
            nc += gamma_site<TV, true, false, true>(P.A, bs1, bt1, bu1, i, S.exch, y, v1s, v1d);
            nc += gamma_site<TV, true, true, true>(P.A, bs2, bt2, bu2, i, S.exch, y, v2s, v2d);
        }
        TV gs = -tp * ((TV)-1.0 * v1s * v4s + v1s * v4d - (TV)1.0 * v3s * v2s + v3d * v2s);
        TV gd = -tp * ((TV)3.0 * v1d * v4s + v1d * v4d + (TV)3.0 * v3s * v2d + v3d * v2d);
        // overlap gather Σ_(j1,j2,m) -p (-2) m (2S) v1[j1] v2[j2]  (flow_su2_t.jl:59-74)
        double *sc = S.scr;
        sc[4 * slot + 0] = (double)v1s;
        sc[4 * slot + 1] = (double)v1d;
        sc[4 * slot + 2] = (double)v2s;
        sc[4 * slot + 3] = (double)v2d;
        __syncthreads();
        const TV c = -tp * (TV)(-2.0);
        const TV S2 = (TV)y.S2;
        const int r0 = S.ov_off[i], r1 = S.ov_off[i + 1];
        for (int r = r0; r < r1; ++r) {
            const uint32_t row = S.ov[r];
            const int j1 = row & 255, j2 = (row >> 8) & 255;
            const TV m = (TV)(double)(row >> 16);
            const double *e1 = sc + 4 * (gbase + j1);
            const double *e2 = sc + 4 * (gbase + j2);
            gs += c * m * S2 * (TV)e1[0] * (TV)e2[2];
            gd += c * m * S2 * (TV)e1[1] * (TV)e2[3];
        }
        __syncthreads();
        Gs = gs;
        Gd = gd;
        if (i == 0) nc += nc0;  // site-0 (local) vertices are gathered once per node
    }
}

// element code: ch 2 bits | kernel 3 bits | w1 7 bits | w2 7 bits | w3 7 bits
__device__ __forceinline__ void decode(uint32_t e, int &ch, int &kernel, int &w1, int &w2, int &w3) {
    ch = e & 3;
    kernel = (e >> 2) & 7;
    w1 = (e >> 5) & 127;
    w2 = (e >> 12) & 127;
    w3 = (e >> 19) & 127;
}

// node position a + c*h rounded as the reference computes it (product, then sum)
// (the pragma keeps hipcc from fusing it into an FMA; __dmul_rn/__dadd_rn are fused on gfx950)
__device__ __forceinline__ double node_at(double a, double c, double h) {
#pragma clang fp contract(off)
    return a + c * h;
}

// sub-interval j of the four loop-frequency domains (flow_s.jl:31-34 with
// quadrature.jl:84-139): [2ref,val] mirrored, [-2ref,0], [0,2ref], [2ref,val]
__device__ __forceinline__ bool subinterval(int j, int eval, double ref, const double *logpts, double &a,
                                            double &b, double &sgn) {
    const int d = j / eval, k = j % eval;
    if (d == 0 || d == 3) {
        // the last breakpoint lands on the mesh edge Ω_max+ν_max where the integrand jumps, so
        // the points come from the host's pow (the oracle's) rather than the device's
        if (!(logpts[0] > 0.0)) return false;
        a = logpts[k];
        b = logpts[k + 1];
        sgn = (d == 0) ? -1.0 : 1.0;
        return true;
    }
    const double lo = (d == 1) ? -2.0 * ref : 0.0 * ref;
    const double hi = (d == 1) ? 0.0 * ref : 2.0 * ref;
    if (!(hi > lo)) return false;
    const double h = (hi - lo) / eval;
    // separately rounded like the reference (no FMA): the last piece of [-2ref, 0] must end
    // where the reference's does, since the integrand jumps at v = 0 (sign fold of ν')
    a = node_at(lo, (double)k, h);
    b = node_at(lo, (double)(k + 1), h);
    sgn = 1.0;
    return true;
}

template <int PASS, typename TV>
__global__ __launch_bounds__(64) void k_vertex_pass(PassArgs<TV> P) {
    extern __shared__ uint32_t s_dyn[];  // overlap rows then ov_off
    __shared__ double s_mesh[5 * MAXN];
    __shared__ double s_sigA[MAXN], s_sigD[MAXN];
    __shared__ double s_scr[64 * 4];
    __shared__ int s_exch[64];
    __shared__ unsigned long long s_cnt, s_bytes;

    const Sys &y = P.y;
    const int bid = blockIdx.x;
    const int e_idx = (bid & 7) * P.per_xcd + (bid >> 3);  // XCD-contiguous element ranges
    if (!P.dbg_v && e_idx >= P.n_elems) return;
    const int lane = threadIdx.x;

    // ---- stage tables in LDS
    for (int k = lane; k < 5 * MAXN; k += 64) s_mesh[k] = P.mesh[k];
    for (int k = lane; k < y.nsig; k += 64) {
        s_sigA[k] = P.A.Sigma[k];
        s_sigD[k] = (PASS == P_KAT) ? P.D.Sigma[k] : 0.0;
    }
    if (lane < y.ns) s_exch[lane] = P.exch[lane];
    const int n_ov = P.ov_off[y.ns];
    for (int k = lane; k < n_ov; k += 64) s_dyn[k] = P.ov_rows[k];
    for (int k = lane; k <= y.ns; k += 64) s_dyn[n_ov + k] = (uint32_t)P.ov_off[k];
    if (lane == 0) { s_cnt = 0; s_bytes = 0; }
    __syncthreads();

    Lds S;
    S.sig = s_mesh;
    S.Os = s_mesh + MAXN;
    S.ns = s_mesh + 2 * MAXN;
    S.Ot = s_mesh + 3 * MAXN;
    S.nt = s_mesh + 4 * MAXN;
    S.sigA = s_sigA;
    S.sigD = s_sigD;
    S.exch = s_exch;
    S.ov = s_dyn;
    S.ov_off = reinterpret_cast<const int *>(s_dyn + n_ov);
    S.scr = s_scr;

    int ch, kernel, w1, w2, w3;
    decode(P.dbg_v ? P.dbg_elem : P.elems[e_idx], ch, kernel, w1, w2, w3);
    const double INF = __builtin_huge_val();
    double W, vw, vwp;
    if (PASS == P_REDUCED) {
        W = S.Os[w1];
        vw = 0.0;
        vwp = (kernel == 1) ? INF : S.ns[w3];
    } else {
        const double *Om = (ch == 1) ? S.Ot : S.Os;
        const double *nu = (ch == 1) ? S.nt : S.ns;
        W = Om[w1];
        vw = (kernel == 2 || kernel == 4) ? nu[w2] : INF;
        vwp = (kernel == 3 || kernel == 4) ? nu[w3] : INF;
    }
    const double ref = P.L + 0.5 * W;
    const double val = (ch == 1) ? y.valT : y.valS;
    const double *logpts = P.logpts + ((ch == 1 ? y.nO : 0) + w1) * (P.eval + 1);

    const int GS = y.GS, G = y.G;
    const int g = lane / GS;
    const bool real_lane = g < G;
    const int i = lane % GS;
    const int gbase = real_lane ? g * GS : 0;
    const int nsub = 4 * P.eval;
    const int rounds = (nsub + G - 1) / G;

    if (P.dbg_v) {
        TV Gs, Gd;
        int ncd;
        node_eval<PASS, TV>(P, S, ch, W, vw, vwp, P.dbg_v[bid], i, lane, gbase, Gs, Gd, ncd);
        if (g == 0) {
            P.dbg_out[(2 * bid + 0) * y.ns + i] = (double)Gs;
            P.dbg_out[(2 * bid + 1) * y.ns + i] = (double)Gd;
        }
        return;
    }
    double acc0 = 0.0, acc1 = 0.0;
    unsigned long long nodes = 0, corner_sum = 0;
    int nc = 0;

    for (int r = 0; r < rounds; ++r) {
        const int j = r * G + g;
        double a = 0.0, b = 0.0, sgn = 1.0;
        bool act = real_lane && j < nsub && subinterval(j, P.eval, ref, logpts, a, b, sgn);
        if (__ballot(act) == 0) continue;  // wave-uniform skip of empty rounds
        double b10 = 0.0, b11 = 0.0, b20 = 0.0, b21 = 0.0;
        const unsigned long long nodes_before = nodes;
        // initial Simpson S1 on (a, m, b) and S2 (simps!, quadrature.jl:57-68)
        for (int k = 0; k < 6; ++k) {
            double x, dx;
            if (k == 0) { x = a; dx = 1.0 * (b - a) / 6.0; }
            else if (k == 1) { x = 0.5 * (a + b); dx = 2.0 * (b - a) / 3.0; }
            else if (k == 2) { x = b; dx = 1.0 * (b - a) / 6.0; }
            else {
                const double h = 0.5 * (b - a) / 2;
                if (k == 3) { x = node_at(a, 2.0 * 1 - 1.0, h); dx = +4.0 * h / 3.0; }
                else if (k == 4) { x = node_at(a, 2.0 * 2 - 1.0, h); dx = +4.0 * h / 3.0; }
                else { x = node_at(a, 4.0 * 1 - 2.0, h); dx = -2.0 * h / 3.0; }
            }
            if (k == 3) { b20 = 0.5 * b10; b21 = 0.5 * b11; }
            TV Gs, Gd;
            node_eval<PASS, TV>(P, S, ch, W, vw, vwp, act ? sgn * x : 0.0, i, lane, gbase, Gs, Gd, nc);
            if (act) {
                if (k < 3) { b10 += dx * (double)Gs; b11 += dx * (double)Gd; }
                else { b20 += dx * (double)Gs; b21 += dx * (double)Gd; }
                ++nodes;
                corner_sum += nc;
            }
        }
        bool done = !act;
        double res0 = 0.0, res1 = 0.0;
        long n = 4;
        while (true) {
            // Richardson extrapolation and error estimate (quadrature.jl:72-88)
            double r10 = b10 * (-1.0 / 15.0), r11 = b11 * (-1.0 / 15.0);
            r10 += 16.0 / 15.0 * b20;
            r11 += 16.0 / 15.0 * b21;
            s_scr[4 * lane + 0] = r10;
            s_scr[4 * lane + 1] = r11;
            s_scr[4 * lane + 2] = b20;
            s_scr[4 * lane + 3] = b21;
            __syncthreads();
            double q1 = 0.0, q2 = 0.0, q3 = 0.0;
            if (real_lane) {
                for (int jj = 0; jj < y.ns; ++jj) {
                    const double *e = s_scr + 4 * (gbase + jj);
                    const double x0 = e[0], x1 = e[1], z0 = e[2], z1 = e[3];
                    q1 += x0 * x0; q1 += x1 * x1;
                    q2 += z0 * z0; q2 += z1 * z1;
                    const double d0 = x0 - z0, d1 = x1 - z1;
                    q3 += d0 * d0; q3 += d1 * d1;
                }
            }
            __syncthreads();
            const double n1 = sqrt(q1), n2 = sqrt(q2), adiff = sqrt(q3);
            const double rdiff = adiff / fmax(n1, n2);
            if (!done) {
                const double d0 = r10 - b20, d1 = r11 - b21;
                if (adiff < P.atol || rdiff < P.rtol || n > (1L << 20)) {
                    res0 = d0 + b20;
                    res1 = d1 + b21;
                    done = true;
                } else {
                    b10 = b20; b11 = b21;
                    b20 = 0.5 * b10; b21 = 0.5 * b11;
                }
            }
            if (__ballot(!done) == 0) break;
            // residual! with n new and n/2 re-weighted nodes (quadrature.jl:2-24)
            const double h = 0.5 * (b - a) / n;
            const long nk = n + n / 2;
            for (long k = 0; k < nk; ++k) {
                double x, dx;
                if (k < n) { x = node_at(a, 2.0 * (k + 1) - 1.0, h); dx = +4.0 * h / 3.0; }
                else { x = node_at(a, 4.0 * (k - n + 1) - 2.0, h); dx = -2.0 * h / 3.0; }
                TV Gs, Gd;
                node_eval<PASS, TV>(P, S, ch, W, vw, vwp, done ? 0.0 : sgn * x, i, lane, gbase, Gs, Gd, nc);
                if (!done) {
                    b20 += dx * (double)Gs;
                    b21 += dx * (double)Gd;
                    ++nodes;
                    corner_sum += nc;
                }
            }
            n *= 2;
        }
        if (P.dbg_sub_nodes && act && i == 0 && (long)e_idx * nsub + j < P.dbg_cap) {
            P.dbg_sub_nodes[(long)e_idx * nsub + j] = (int)(nodes - nodes_before);
            P.dbg_sub_res[(long)e_idx * nsub + j] = res0;
        }
        // ordered accumulation of this round's sub-interval results
        s_scr[2 * lane + 0] = res0;
        s_scr[2 * lane + 1] = res1;
        __syncthreads();
        for (int gg = 0; gg < G; ++gg) {
            const int jj = r * G + gg;
            double ta, tb, ts;
            if (jj < nsub && subinterval(jj, P.eval, ref, logpts, ta, tb, ts)) {
                acc0 += s_scr[2 * (gg * GS + i) + 0];
                acc1 += s_scr[2 * (gg * GS + i) + 1];
            }
        }
        __syncthreads();
    }
    // boundary corrections (flow_s.jl:37-38)
    {
        const double c0 = P.corrs[0 + 2 * (ch + 3 * w1)], c1 = P.corrs[1 + 2 * (ch + 3 * w1)];
        TV Gs, Gd;
        int nc1;
        node_eval<PASS, TV>(P, S, ch, W, vw, vwp, -val, i, lane, gbase, Gs, Gd, nc1);
        acc0 += c0 * (double)Gs;
        acc1 += c0 * (double)Gd;
        int nc2;
        node_eval<PASS, TV>(P, S, ch, W, vw, vwp, val, i, lane, gbase, Gs, Gd, nc2);
        acc0 += c1 * (double)Gs;
        acc1 += c1 * (double)Gd;
        if (g == 0) {
            nodes += 2;
            corner_sum += nc1 + nc2;
        }
    }
    // write the site vector (flow_s.jl:41-51)
    if (g == 0) {
        long base;
        if (kernel == 1) base = (long)y.ns * w1;
        else if (kernel == 2) base = y.q21 + (long)y.ns * (w1 + (long)y.nO * w2);
        else if (kernel == 3) base = y.q22 + (long)y.ns * (w1 + (long)y.nO * w3);
        else base = y.q3 + (long)y.ns * (w1 + (long)y.nO * (w2 + (long)y.nn * w3));
        double *o = P.out + ch * y.chlen + base + i;
        o[0] = acc0;
        o[y.complen] = acc1;
    }
    if (P.node_count) {
        // nodes: one count per group; gathered bytes: every lane's corners x 2 components
        if (real_lane && i == 0 && nodes) atomicAdd(&s_cnt, nodes);
        if (real_lane && corner_sum) atomicAdd(&s_bytes, corner_sum * 2ull * sizeof(TV));
        __syncthreads();
        if (lane == 0) {
            atomicAdd(P.node_count, s_cnt);
            atomicAdd(P.node_count + 1, s_bytes);
        }
    }
}


template <int PASS, typename TV>
hipError_t launch_vertex_pass(const PassArgs<TV> &P, int grid, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL((k_vertex_pass<PASS, TV>), dim3(grid), dim3(64), lds, stream, P);
    return hipGetLastError();
}

// explicit instantiations live in pass_*.hip; other translation units only link to them
#define PFFRG_PASS_INSTANCES(KW, PASS)                                                                      \
    KW template hipError_t launch_vertex_pass<PASS, double>(const PassArgs<double> &, int, size_t, hipStream_t); \
    KW template hipError_t launch_vertex_pass<PASS, float>(const PassArgs<float> &, int, size_t, hipStream_t);

PFFRG_PASS_INSTANCES(extern, P_KAT)
PFFRG_PASS_INSTANCES(extern, P_LEFT)
PFFRG_PASS_INSTANCES(extern, P_CENTRAL)
PFFRG_PASS_INSTANCES(extern, P_BSE)
PFFRG_PASS_INSTANCES(extern, P_REDUCED)

}  // namespace pffrg
