// chi.hip -- real-space correlations χ_ij(ω) (compute_χ!, correlation_su2.jl:58-86).
//
// Every (χ frequency, site, box) task runs the adaptive 2-D cubature of integrate_χ_boxes
// (correlation_lib.jl:2-30) on the kernel of compute_χ_kernel! (correlation_su2.jl:14-55) with
// the vectorised schedule of Cubature's hcubature_v: every round splits at once the
// largest-error regions that must be refined for the error of the others to meet the
// tolerance.  A round is two launches over all unfinished tasks: k_chi_round (one 256-thread
// block per task: totals, convergence, selection) and k_chi_split (one thread per Genz-Malik
// point of each half of each selected region, over as many blocks as the round has splits), so
// that a task with many regions spreads over the whole GPU.  Regions live in
// global memory; the totals are block sums in a fixed order, the selection takes whole
// quarter-octave error bins (fixed-point bin sums in LDS), and new regions are appended in
// index order, all as in the oracle's restatement.  This file is compiled without FMA
// contraction so that its arithmetic, and with it every adaptive decision, is the oracle's.
#include "pffrg_chi.h"

namespace pffrg {
namespace {

constexpr double GM_L2 = 0.35856858280031806;  // sqrt(9/70)
constexpr double GM_L4 = 0.9486832980505138;   // sqrt(9/10)
constexpr double GM_L5 = 0.6882472016116853;   // sqrt(9/19)

struct Reg {
    double c0, c1, h0, h1, val, err;
    int split;
};

__device__ __forceinline__ void gm_point(const Reg &R, int k, double &x, double &y) {
    // f0 | ±λ2 x, ±λ2 y | ±λ4 x, ±λ4 y | (±λ4, ±λ4) | (±λ5, ±λ5)
    const int sx = (k == 1 || k == 5 || k == 9 || k == 11 || k == 13 || k == 15) ? -1
                   : (k == 2 || k == 6 || k == 10 || k == 12 || k == 14 || k == 16) ? 1 : 0;
    const int sy = (k == 3 || k == 7 || k == 9 || k == 10 || k == 13 || k == 14) ? -1
                   : (k == 4 || k == 8 || k == 11 || k == 12 || k == 15 || k == 16) ? 1 : 0;
    const double l = (k < 5) ? GM_L2 : (k < 13) ? GM_L4 : GM_L5;
    x = R.c0 + (double)sx * (l * R.h0);
    y = R.c1 + (double)sy * (l * R.h1);
}

__device__ __forceinline__ void gm_combine(Reg &R, const double *f) {
    const double W1 = -3816.0 / 19683.0, W2 = 980.0 / 6561.0, W3 = 1020.0 / 19683.0, W4 = 200.0 / 19683.0,
                 W5 = 6859.0 / 19683.0 / 4.0;
    const double E1 = -971.0 / 729.0, E2 = 245.0 / 486.0, E3 = 65.0 / 1458.0, E4 = 25.0 / 729.0;
    const double f0 = f[0];
    const double s2 = ((f[1] + f[2]) + f[3]) + f[4];
    const double s3 = ((f[5] + f[6]) + f[7]) + f[8];
    const double s4 = ((f[9] + f[10]) + f[11]) + f[12];
    const double s5 = ((f[13] + f[14]) + f[15]) + f[16];
    const double vol = (2.0 * R.h0) * (2.0 * R.h1);
    const double r7 = vol * ((((W1 * f0 + W2 * s2) + W3 * s3) + W4 * s4) + W5 * s5);
    const double r5 = vol * (((E1 * f0 + E2 * s2) + E3 * s3) + E4 * s4);
    R.val = r7;
    R.err = fabs(r5 - r7);
    const double d0 = fabs(((f[1] + f[2]) - 2.0 * f0) - (1.0 / 7.0) * ((f[5] + f[6]) - 2.0 * f0));
    const double d1 = fabs(((f[3] + f[4]) - 2.0 * f0) - (1.0 / 7.0) * ((f[7] + f[8]) - 2.0 * f0));
    R.split = (d1 > d0 || (d1 == d0 && R.h1 > R.h0)) ? 1 : 0;
}

// compute_χ_kernel! at (x, y) of the unit square: SU(2) (correlation_su2.jl:14-55) or the
// U(1)-DM kernel of component Q.comp (compute_χ_kernel_{xx,zz,xy}!, correlation_u1_dm.jl:16-134)
__device__ double chi_point(const ChiProb &Q, const Meshes<> &M, const double *sA, const int *exch, int site,
                            double w, double xx, double yy) {
    const Sys &y = Q.y;
    const double L = Q.L;
    const double v = (2.0 * xx - 1.0) / ((1.0 - xx) * xx);
    const double vp = (2.0 * yy - 1.0) / ((1.0 - yy) * yy);
    const double dv = (2.0 * xx * xx - 2.0 * xx + 1) / ((1.0 - xx) * (1.0 - xx) * xx * xx);
    const double dvp = (2.0 * yy * yy - 2.0 * yy + 1) / ((1.0 - yy) * (1.0 - yy) * yy * yy);
    const Buf<> bs1 = buf_s(M, y, v + vp, 0.5 * (v - w - vp), 0.5 * (-v - w + vp));
    const Buf<> bt1 = buf_t(M, y, w, v, vp);
    const Buf<> bu1 = buf_u(M, y, v - vp, 0.5 * (v - w + vp), 0.5 * (v + w + vp));
    const double S2 = y.S2;
    const double tp2 = (2.0 * PI_D) * (2.0 * PI_D);
    double val;
    if (y.nc == 6) {
        double g6[6];
        gamma_site_u1<true, true, true>(Q.A, bs1, bt1, bu1, site, exch, y, g6);
        val = -g6[Q.comp == 2 ? U1_DM : Q.comp] / tp2;
        if (site == 0) {
            const Buf<> bs2 = buf_s(M, y, v + vp, 0.5 * (v - w - vp), 0.5 * (v + w - vp));
            const Buf<> bt2 = buf_t(M, y, -v + vp, 0.5 * (v - w + vp), 0.5 * (v + w + vp));
            const Buf<> bu2 = buf_u(M, y, -w, v, vp);
            gamma_site_u1<true, true, true>(Q.A, bs2, bt2, bu2, site, exch, y, g6);
            if (Q.comp == 0) val -= (g6[U1_ZZ] - g6[U1_DD]) / (2.0 * tp2);
            else if (Q.comp == 1) val -= (2.0 * g6[U1_XX] - g6[U1_ZZ] - g6[U1_DD]) / (2.0 * tp2);
            else val -= (g6[U1_ZD] - g6[U1_DZ]) / (2.0 * tp2);
        }
    } else {
        double vs, vd;
        gamma_site<double, true, true, true>(Q.A, bs1, bt1, bu1, site, exch, y, vs, vd);
        val = -(S2 * S2) * vs / tp2;
        if (site == 0) {
            const Buf<> bs2 = buf_s(M, y, v + vp, 0.5 * (v - w - vp), 0.5 * (v + w - vp));
            const Buf<> bt2 = buf_t(M, y, -v + vp, 0.5 * (v - w + vp), 0.5 * (v + w + vp));
            const Buf<> bu2 = buf_u(M, y, -w, v, vp);
            double s2, d2;
            gamma_site<double, true, true, true>(Q.A, bs2, bt2, bu2, site, exch, y, s2, d2);
            val -= S2 * (s2 - d2) / (2.0 * tp2);
        }
    }
    val *= G_full<double>(L, v - 0.5 * w, M.sig, y.nsig, sA) * G_full<double>(L, v + 0.5 * w, M.sig, y.nsig, sA) * dv;
    val *= G_full<double>(L, vp - 0.5 * w, M.sig, y.nsig, sA) * G_full<double>(L, vp + 0.5 * w, M.sig, y.nsig, sA) *
           dvp;
    return val;
}

__device__ __forceinline__ void chi_box(double width, int k, double *lo, double *hi) {
    const double widthp = (width - 2.0 + sqrt(width * width + 4.0)) / (2.0 * width);
    const double size = fabs(widthp - 0.5);
    const double a = 0.5 - size, b = 0.5 + size;
    const int kx = k % 3, ky = k / 3;
    lo[0] = (kx == 0) ? 0.0 : (kx == 1) ? a : b;
    hi[0] = (kx == 0) ? a : (kx == 1) ? b : 1.0;
    lo[1] = (ky == 0) ? b : (ky == 1) ? a : 0.0;
    hi[1] = (ky == 0) ? 1.0 : (ky == 1) ? b : a;
}

__device__ __forceinline__ void store_reg(const ChiArgs &Q, long base, long j, const Reg &R) {
    Q.rc0[base + j] = R.c0;
    Q.rc1[base + j] = R.c1;
    Q.rh0[base + j] = R.h0;
    Q.rh1[base + j] = R.h1;
    Q.rval[base + j] = R.val;
    Q.rerr[base + j] = R.err;
    Q.rsplit[base + j] = R.split;
}

// selection bins: quarter octaves below the largest error (bin 0 holds the largest, the last
// bin everything from 2^-63 of it down), summed in 44-bit fixed point relative to the largest
constexpr int CHI_BINS = 256;

__device__ __forceinline__ unsigned long long dbits(double e) { return (unsigned long long)__double_as_longlong(e); }

__device__ __forceinline__ int err_exp(double e) { return (int)((dbits(e) >> 52) & 0x7FF); }

__device__ __forceinline__ int err_bin(double e, int E) {
    const int d = (4 * E + 3) - (int)((dbits(e) >> 50) & 0x1FFF);
    return d < 0 ? 0 : (d > CHI_BINS - 1 ? CHI_BINS - 1 : d);
}

__device__ __forceinline__ unsigned long long err_fixed(double e, int E) {
    return (unsigned long long)ldexp(e, 1066 - (E > 1 ? E : 1));
}

// the region totals Σ val and Σ err in one pass: thread t adds j = t, t + CHI_T, ..., then a
// tree over the threads (p[t] += p[t + s] for s = CHI_T/2, ..., 1)
__device__ void block_sum2(long n, const double *val, const double *err, double *s_a, double *s_b, double &sv,
                           double &se) {
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (long j = t; j < n; j += CHI_T) {
        a += val[j];
        b += err[j];
    }
    s_a[t] = a;
    s_b[t] = b;
    __syncthreads();
    for (int s = CHI_T / 2; s >= 1; s >>= 1) {
        if (t < s) {
            s_a[t] += s_a[t + s];
            s_b[t] += s_b[t + s];
        }
        __syncthreads();
    }
    sv = s_a[0];
    se = s_b[0];
    __syncthreads();
}

// exclusive prefix of v over the block's threads (and the total in *total)
__device__ long block_scan(long v, long *s_scan, long *total) {
    const int t = threadIdx.x;
    s_scan[t] = v;
    __syncthreads();
    for (int s = 1; s < CHI_T; s <<= 1) {
        const long o = t >= s ? s_scan[t - s] : 0;
        __syncthreads();
        s_scan[t] += o;
        __syncthreads();
    }
    const long incl = s_scan[t];
    *total = s_scan[CHI_T - 1];
    __syncthreads();
    return incl - v;
}

// the LDS copies every χ kernel reads: meshes, Σ of the action, the exchange table
struct ChiLds {
    double mesh[5 * MAXN];
    double sA[MAXN];
    int exch[64];
};

__device__ __forceinline__ void load_lds(const ChiProb &Q, ChiLds &S) {
    for (int k = threadIdx.x; k < 5 * MAXN; k += blockDim.x) S.mesh[k] = Q.mesh[k];
    for (int k = threadIdx.x; k < Q.y.nsig; k += blockDim.x) S.sA[k] = Q.A.Sigma[k];
    if ((int)threadIdx.x < Q.y.ns) S.exch[threadIdx.x] = Q.exch[threadIdx.x];
    __syncthreads();
}

struct TaskGeo {
    int site;
    double W, lo[2], hi[2];
};

// geometry of the problem's local task t = (w * ns + site) * 9 + box
__device__ __forceinline__ TaskGeo task_geo(const ChiProb &Q, int t) {
    TaskGeo g;
    const int box = t % 9, q = t / 9;
    const int w = q / Q.y.ns;
    g.site = q % Q.y.ns;
    g.W = Q.chi[w];
    chi_box(4.0 * (Q.L + 0.5 * g.W), box, g.lo, g.hi);
    return g;
}

// the first region (the whole box) of every slot: one thread per rule point
__global__ __launch_bounds__(64) void k_chi_init(ChiArgs Q) {
    __shared__ ChiLds S;
    __shared__ double s_f[17];
    const int slot = blockIdx.x;
    const int t = Q.tasks[slot];
    const ChiProb &P = Q.probs[Q.task_prob[t]];
    load_lds(P, S);
    const Meshes<> M{S.mesh, S.mesh + MAXN, S.mesh + 2 * MAXN, S.mesh + 3 * MAXN, S.mesh + 4 * MAXN};
    const TaskGeo g = task_geo(P, t - P.task0);
    Reg r0;
    r0.c0 = 0.5 * (g.lo[0] + g.hi[0]);
    r0.c1 = 0.5 * (g.lo[1] + g.hi[1]);
    r0.h0 = 0.5 * (g.hi[0] - g.lo[0]);
    r0.h1 = 0.5 * (g.hi[1] - g.lo[1]);
    if (threadIdx.x < 17) {
        double xx, yy;
        gm_point(r0, threadIdx.x, xx, yy);
        s_f[threadIdx.x] = chi_point(P, M, S.sA, S.exch, g.site, g.W, xx, yy);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        gm_combine(r0, s_f);
        store_reg(Q, (long)slot * Q.cap, 0, r0);
        Q.rn[slot] = 1;
        Q.rk[slot] = 0;
        Q.rdone[slot] = 0;
    }
}

// one round of every active slot (Q.active[blockIdx.x]): append the previous round's halves,
// totals and convergence, else the selection of the regions to split next
__global__ __launch_bounds__(CHI_T) void k_chi_round(ChiArgs Q) {
    __shared__ double s_red[CHI_T], s_red2[CHI_T];
    __shared__ long s_scan[CHI_T];
    __shared__ unsigned long long s_hist[CHI_BINS];
    __shared__ int s_cnt[CHI_BINS];
    __shared__ int s_ctl[2];
    const int tid = threadIdx.x;
    const int slot = Q.active[blockIdx.x];
    const int t = Q.tasks[slot];
    const long base = (long)slot * Q.cap;
    const long n = Q.rn[slot] + Q.rk[slot];
    double tv, te;
    block_sum2(n, Q.rval + base, Q.rerr + base, s_red, s_red2, tv, te);
    if (te <= Q.atol || te <= fabs(tv) * Q.rtol || n >= CHI_MAX_REGIONS) {
        if (tid == 0) {
            Q.box[t] = tv;
            Q.nreg[t] = (int)n;
            Q.status[t] = 0;
            Q.rn[slot] = n;
            Q.rk[slot] = 0;
            Q.rdone[slot] = 1;
        }
        return;
    }
    // the fewest leading bins whose regions carry enough error for the rest to converge
    if (tid == 0) s_ctl[0] = 0;
    for (int b = tid; b < CHI_BINS; b += CHI_T) {
        s_hist[b] = 0ull;
        s_cnt[b] = 0;
    }
    __syncthreads();
    int em = 0;
    for (long j = tid; j < n; j += CHI_T) em = max(em, err_exp(Q.rerr[base + j]));
    atomicMax(&s_ctl[0], em);
    __syncthreads();
    const int E = s_ctl[0];
    for (long j = tid; j < n; j += CHI_T) {
        const double e = Q.rerr[base + j];
        const int b = err_bin(e, E);
        atomicAdd(&s_hist[b], err_fixed(e, E));
        atomicAdd(&s_cnt[b], 1);
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long Qs = 0ull, Cs = 0ull;
        for (int b = 0; b < CHI_BINS; ++b) Qs += s_hist[b];
        // from the first occupied bin on, so that a round splits at least one region
        int b0 = 0;
        while (b0 < CHI_BINS - 1 && s_cnt[b0] == 0) ++b0;
        for (int b = 0; b < b0; ++b) Cs += s_hist[b];
        int bstar = CHI_BINS - 1;
        for (int b = b0; b < CHI_BINS; ++b) {
            Cs += s_hist[b];
            const double rest = ldexp((double)(Qs - Cs), (E > 1 ? E : 1) - 1066);
            if (rest <= Q.atol || rest <= fabs(tv) * Q.rtol) {
                bstar = b;
                break;
            }
        }
        s_ctl[1] = bstar;
    }
    __syncthreads();
    const int bstar = s_ctl[1];
    // selected regions in index order: thread t scans the contiguous range [j0, j1)
    const long per = (n + CHI_T - 1) / CHI_T;
    const long j0 = (long)tid * per < n ? (long)tid * per : n;
    const long j1 = j0 + per < n ? j0 + per : n;
    long cnt = 0;
    for (long j = j0; j < j1; ++j) cnt += err_bin(Q.rerr[base + j], E) <= bstar ? 1 : 0;
    long total;
    long pos = block_scan(cnt, s_scan, &total);
    const long room = CHI_MAX_REGIONS - n;
    const long k = total < room ? total : room;
    if (n + k > Q.cap) {
        // this slot is too small: the host runs the task again with more room
        if (tid == 0) {
            Q.status[t] = 1;
            Q.rn[slot] = n;
            Q.rk[slot] = 0;
            Q.rdone[slot] = 1;
        }
        return;
    }
    for (long j = j0; j < j1 && pos < k; ++j)
        if (err_bin(Q.rerr[base + j], E) <= bstar) Q.rsel[base + pos++] = (int)j;
    if (tid == 0) {
        Q.rn[slot] = n;
        Q.rk[slot] = k;
    }
}

// the splits of a round: block b takes CHI_RPB selected regions of slot Q.split_tab[2b] from
// Q.split_tab[2b+1] on; one thread per rule point of each half (CHI_RPB x 34 threads), then one
// thread per half combines its 17 values; the lower half replaces the region, the upper half is
// appended at n + r
__global__ __launch_bounds__(CHI_SPLIT_T) void k_chi_split(ChiArgs Q) {
    __shared__ ChiLds S;
    __shared__ double s_f[CHI_RPB * 34];
    __shared__ Reg s_par[CHI_RPB];
    __shared__ long s_j[CHI_RPB];
    const int slot = Q.split_tab[2 * blockIdx.x];
    const int tsk = Q.tasks[slot];
    const ChiProb &P = Q.probs[Q.task_prob[tsk]];
    load_lds(P, S);
    const Meshes<> M{S.mesh, S.mesh + MAXN, S.mesh + 2 * MAXN, S.mesh + 3 * MAXN, S.mesh + 4 * MAXN};
    const long first = Q.split_tab[2 * blockIdx.x + 1];
    const long k = Q.rk[slot];
    const long base = (long)slot * Q.cap;
    const int tid = threadIdx.x;
    if (tid < CHI_RPB && first + tid < k) {
        const long j = Q.rsel[base + first + tid];
        Reg R;
        R.c0 = Q.rc0[base + j];
        R.c1 = Q.rc1[base + j];
        R.h0 = Q.rh0[base + j];
        R.h1 = Q.rh1[base + j];
        R.split = Q.rsplit[base + j];
        s_par[tid] = R;
        s_j[tid] = j;
    }
    __syncthreads();
    const TaskGeo g = task_geo(P, tsk - P.task0);
    auto half = [&](int q, int h) {
        const Reg &R = s_par[q];
        Reg X = R;
        if (R.split == 0) {
            X.h0 = 0.5 * R.h0;
            X.c0 = h ? R.c0 + X.h0 : R.c0 - X.h0;
        } else {
            X.h1 = 0.5 * R.h1;
            X.c1 = h ? R.c1 + X.h1 : R.c1 - X.h1;
        }
        return X;
    };
    {
        const int q = tid / 34, rem = tid % 34;
        if (q < CHI_RPB && first + q < k) {
            const Reg X = half(q, rem / 17);
            double xx, yy;
            gm_point(X, rem % 17, xx, yy);
            s_f[tid] = chi_point(P, M, S.sA, S.exch, g.site, g.W, xx, yy);
        }
    }
    __syncthreads();
    if (tid < 2 * CHI_RPB) {
        const int q = tid / 2, h = tid % 2;
        if (first + q < k) {
            Reg X = half(q, h);
            gm_combine(X, s_f + q * 34 + h * 17);
            store_reg(Q, base, h ? Q.rn[slot] + first + q : s_j[q], X);
        }
    }
}

// propagator term of site 0 (correlation_su2.jl:78-82): QuadGK over (-Inf, -2ref, 0, 2ref, Inf);
// block b handles χ frequency b of all problems' frequencies (Q.prop_prob names the problem)
__global__ __launch_bounds__(64) void k_chi_prop(ChiArgs Q) {
    __shared__ double s_mesh[5 * MAXN];
    __shared__ double s_sA[MAXN];
    __shared__ Seg segs[GK_MAX_SEGS];
    __shared__ double fv[64];
    __shared__ int ctl[4];
    __shared__ double ctld[2];
    const ChiProb &P = Q.probs[Q.prop_prob[blockIdx.x]];
    const Sys &y = P.y;
    const int lane = threadIdx.x;
    for (int k = lane; k < 5 * MAXN; k += 64) s_mesh[k] = P.mesh[k];
    for (int k = lane; k < y.nsig; k += 64) s_sA[k] = P.A.Sigma[k];
    __syncthreads();
    const double W = P.chi[blockIdx.x - P.prop0];
    const double L = P.L;
    const double ref = L + 0.5 * W;
    auto f = [&](double v) -> double {
        if (y.nc == 6)  // correlation_u1_dm.jl:178-183: no spin-length factor
            return G_full<double>(L, v - 0.5 * W, s_mesh, y.nsig, s_sA) *
                   G_full<double>(L, v + 0.5 * W, s_mesh, y.nsig, s_sA) / (4.0 * PI_D);
        return y.S2 * G_full<double>(L, v - 0.5 * W, s_mesh, y.nsig, s_sA) *
               G_full<double>(L, v + 0.5 * W, s_mesh, y.nsig, s_sA) / (4.0 * PI_D);
    };
    double bp[5];
    bp[0] = -1.0;
    const double inner[3] = {-2.0 * ref, 0.0, 2.0 * ref};
    for (int k = 0; k < 3; ++k) bp[k + 1] = 2.0 * inner[k] / (1.0 + hypot(1.0, 2.0 * inner[k]));
    bp[4] = 1.0;
    const double I = gk_block(f, 0, 0.0, bp, 5, Q.atol, Q.rtol, segs, fv, ctl, ctld);
    if (lane == 0) Q.prop[blockIdx.x] = I;
}

}  // namespace

hipError_t launch_chi_init(const ChiArgs &Q, hipStream_t stream) {
    hipLaunchKernelGGL(k_chi_init, dim3(Q.nslots), dim3(64), 0, stream, Q);
    return hipGetLastError();
}

hipError_t launch_chi_round(const ChiArgs &Q, int n_active, hipStream_t stream) {
    hipLaunchKernelGGL(k_chi_round, dim3(n_active), dim3(CHI_T), 0, stream, Q);
    return hipGetLastError();
}

hipError_t launch_chi_split(const ChiArgs &Q, int n_blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_chi_split, dim3(n_blocks), dim3(CHI_SPLIT_T), 0, stream, Q);
    return hipGetLastError();
}

hipError_t launch_chi_prop(const ChiArgs &Q, hipStream_t stream) {
    hipLaunchKernelGGL(k_chi_prop, dim3(Q.nfreq), dim3(64), 0, stream, Q);
    return hipGetLastError();
}

}  // namespace pffrg
