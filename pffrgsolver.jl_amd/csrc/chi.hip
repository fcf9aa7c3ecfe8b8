// chi.hip -- real-space correlations χ_ij(ω) (compute_χ!, correlation_su2.jl:58-86).
//
// One 256-thread block per (χ frequency, site, box) task runs the adaptive 2-D cubature of
// integrate_χ_boxes (correlation_lib.jl:2-30) on the kernel of compute_χ_kernel!
// (correlation_su2.jl:14-55) with the vectorised schedule of Cubature's hcubature_v: every
// round splits at once the largest-error regions that must be refined for the error of the
// others to meet the tolerance, each thread evaluating the 2 x 17 Genz-Malik points of the
// halves of its regions.  Regions live in global memory; the totals are block sums in a fixed
// order, the selection threshold is found by bisection over the top 16 bits of the error, and
// new regions are appended in index order, all as in the oracle's restatement.  This file is
// compiled without FMA contraction so that its arithmetic, and with it every adaptive
// decision, is the oracle's.
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

// compute_χ_kernel! at (x, y) of the unit square
__device__ double chi_point(const ChiArgs &Q, const Meshes<> &M, const double *sA, const int *exch, int site,
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
    double vs, vd;
    gamma_site<double, true, true, true>(Q.A, bs1, bt1, bu1, site, exch, y, vs, vd);
    double val = -(S2 * S2) * vs / tp2;
    if (site == 0) {
        const Buf<> bs2 = buf_s(M, y, v + vp, 0.5 * (v - w - vp), 0.5 * (v + w - vp));
        const Buf<> bt2 = buf_t(M, y, -v + vp, 0.5 * (v - w + vp), 0.5 * (v + w + vp));
        const Buf<> bu2 = buf_u(M, y, -w, v, vp);
        double s2, d2;
        gamma_site<double, true, true, true>(Q.A, bs2, bt2, bu2, site, exch, y, s2, d2);
        val -= S2 * (s2 - d2) / (2.0 * tp2);
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

// the selection key: the top 16 bits of a non-negative double
__device__ __forceinline__ unsigned err_key(double e) {
    return (unsigned)((unsigned long long)__double_as_longlong(e) >> 48);
}

// Σ_j x(j) over j < n: thread t adds j = t, t + CHI_T, ..., then a tree over the threads
template <class G>
__device__ double block_sum(long n, const G &x, double *s_red) {
    const int t = threadIdx.x;
    double a = 0.0;
    for (long j = t; j < n; j += CHI_T) a += x(j);
    s_red[t] = a;
    __syncthreads();
    for (int s = CHI_T / 2; s >= 1; s >>= 1) {
        if (t < s) s_red[t] += s_red[t + s];
        __syncthreads();
    }
    const double r = s_red[0];
    __syncthreads();
    return r;
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

__device__ __forceinline__ void eval_region(const ChiArgs &Q, const Meshes<> &M, const double *sA, const int *exch,
                                            int site, double W, Reg &R) {
    double f[17];
    for (int k = 0; k < 17; ++k) {
        double xx, yy;
        gm_point(R, k, xx, yy);
        f[k] = chi_point(Q, M, sA, exch, site, W, xx, yy);
    }
    gm_combine(R, f);
}

// one block per task slot; Q.tasks[slot] names the (frequency, site, box) task
__global__ __launch_bounds__(CHI_T) void k_chi(ChiArgs Q) {
    __shared__ double s_mesh[5 * MAXN];
    __shared__ double s_sA[MAXN];
    __shared__ int s_exch[64];
    __shared__ double s_red[CHI_T];
    __shared__ long s_scan[CHI_T];
    const Sys &y = Q.y;
    const int tid = threadIdx.x;
    const int slot = blockIdx.x;
    const int t = Q.tasks[slot];
    const int box = t % 9, q = t / 9;
    const int w = q / y.ns, site = q % y.ns;
    for (int k = tid; k < 5 * MAXN; k += CHI_T) s_mesh[k] = Q.mesh[k];
    for (int k = tid; k < y.nsig; k += CHI_T) s_sA[k] = Q.A.Sigma[k];
    if (tid < y.ns) s_exch[tid] = Q.exch[tid];
    __syncthreads();
    const Meshes<> M{s_mesh, s_mesh + MAXN, s_mesh + 2 * MAXN, s_mesh + 3 * MAXN, s_mesh + 4 * MAXN};
    const double W = Q.chi[w];
    const double ref = Q.L + 0.5 * W;
    double lo[2], hi[2];
    chi_box(4.0 * ref, box, lo, hi);
    const long base = (long)slot * Q.cap;
    if (tid == 0) {
        Reg r0;
        r0.c0 = 0.5 * (lo[0] + hi[0]);
        r0.c1 = 0.5 * (lo[1] + hi[1]);
        r0.h0 = 0.5 * (hi[0] - lo[0]);
        r0.h1 = 0.5 * (hi[1] - lo[1]);
        eval_region(Q, M, s_sA, s_exch, site, W, r0);
        store_reg(Q, base, 0, r0);
    }
    __syncthreads();
    long n = 1;
    double tv;
    int status = 0;
    while (true) {
        tv = block_sum(n, [&](long j) { return Q.rval[base + j]; }, s_red);
        const double te = block_sum(n, [&](long j) { return Q.rerr[base + j]; }, s_red);
        if (te <= Q.atol || te <= fabs(tv) * Q.rtol || n >= CHI_MAX_REGIONS) break;
        // the largest key K whose regions {err_key >= K} carry enough error (K = 0 takes all)
        unsigned klo = 0, khi = 65536;
        while (khi - klo > 1) {
            const unsigned mid = (klo + khi) / 2;
            const double S = block_sum(
                n, [&](long j) { const double e = Q.rerr[base + j]; return err_key(e) >= mid ? e : 0.0; }, s_red);
            const double rest = te - S;
            if (rest <= Q.atol || rest <= fabs(tv) * Q.rtol)
                klo = mid;
            else
                khi = mid;
        }
        // selected regions in index order: thread t scans the contiguous range [j0, j1)
        const long per = (n + CHI_T - 1) / CHI_T;
        const long j0 = (long)tid * per < n ? (long)tid * per : n;
        const long j1 = j0 + per < n ? j0 + per : n;
        long cnt = 0;
        for (long j = j0; j < j1; ++j) cnt += err_key(Q.rerr[base + j]) >= klo ? 1 : 0;
        long total;
        long pos = block_scan(cnt, s_scan, &total);
        const long room = CHI_MAX_REGIONS - n;
        const long k = total < room ? total : room;
        if (n + k > Q.cap) {
            status = 1;  // this slot is too small: the host runs the task again with more room
            break;
        }
        for (long j = j0; j < j1 && pos < k; ++j)
            if (err_key(Q.rerr[base + j]) >= klo) Q.rsel[base + pos++] = (int)j;
        __syncthreads();
        for (long r = tid; r < k; r += CHI_T) {
            const long j = Q.rsel[base + r];
            Reg R;
            R.c0 = Q.rc0[base + j];
            R.c1 = Q.rc1[base + j];
            R.h0 = Q.rh0[base + j];
            R.h1 = Q.rh1[base + j];
            const int d = Q.rsplit[base + j];
            Reg a = R, b = R;
            if (d == 0) {
                a.h0 = 0.5 * R.h0;
                b.h0 = a.h0;
                a.c0 = R.c0 - a.h0;
                b.c0 = R.c0 + a.h0;
            } else {
                a.h1 = 0.5 * R.h1;
                b.h1 = a.h1;
                a.c1 = R.c1 - a.h1;
                b.c1 = R.c1 + a.h1;
            }
            eval_region(Q, M, s_sA, s_exch, site, W, a);
            eval_region(Q, M, s_sA, s_exch, site, W, b);
            store_reg(Q, base, j, a);
            store_reg(Q, base, n + r, b);
        }
        __syncthreads();
        n += k;
    }
    if (tid == 0) {
        Q.box[t] = tv;
        Q.nreg[t] = (int)n;
        Q.status[t] = status;
    }
}

// propagator term of site 0 (correlation_su2.jl:78-82): QuadGK over (-Inf, -2ref, 0, 2ref, Inf)
__global__ __launch_bounds__(64) void k_chi_prop(ChiArgs Q) {
    __shared__ double s_mesh[5 * MAXN];
    __shared__ double s_sA[MAXN];
    __shared__ Seg segs[GK_MAX_SEGS];
    __shared__ double fv[64];
    __shared__ int ctl[4];
    __shared__ double ctld[2];
    const Sys &y = Q.y;
    const int lane = threadIdx.x;
    for (int k = lane; k < 5 * MAXN; k += 64) s_mesh[k] = Q.mesh[k];
    for (int k = lane; k < y.nsig; k += 64) s_sA[k] = Q.A.Sigma[k];
    __syncthreads();
    const double W = Q.chi[blockIdx.x];
    const double L = Q.L;
    const double ref = L + 0.5 * W;
    auto f = [&](double v) -> double {
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

hipError_t launch_chi(const ChiArgs &Q, hipStream_t stream) {
    hipLaunchKernelGGL(k_chi, dim3(Q.nslots), dim3(CHI_T), 0, stream, Q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !Q.with_prop) return e;
    hipLaunchKernelGGL(k_chi_prop, dim3(Q.nchi), dim3(64), 0, stream, Q);
    return hipGetLastError();
}

}  // namespace pffrg
