// rescan.hip -- adaptive frequency meshes between flow steps (resample_from_to,
// Action/Action.jl:483-557).
//
// k_scan_partial / k_scan_final find, for every (component, channel), the first maximum of
// |q3| in the reference's column-major order and gather the four cuts through it that
// scan_channel (Action.jl:449-480) scans.  pffrg_scan is that scan heuristic (Action.jl:
// 369-446) on the host.  k_resample interpolates a whole action onto new meshes, one output
// entry per thread (channel.jl:416-470, vertex.jl:148-160; Σ by get_Σ, Action.jl:8-26).
// The file is compiled without FMA contraction so that every interpolated value is the one
// the reference's sequence of products and sums gives.
#include <cmath>
#include <mutex>
#include <vector>

#include "pffrg.h"
#include "pffrg_rescan.h"

namespace pffrg {
namespace {

// position p beats q: larger |value|, ties to the smaller column-major index (Julia's argmax)
__device__ __forceinline__ bool beats(double pv, long long pi, double qv, long long qi) {
    return pv > qv || (pv == qv && pi < qi);
}

__device__ void block_best(double &v, long long &i, double *sv, long long *si) {
    const int t = threadIdx.x;
    sv[t] = v;
    si[t] = i;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s && beats(sv[t + s], si[t + s], sv[t], si[t])) {
            sv[t] = sv[t + s];
            si[t] = si[t + s];
        }
        __syncthreads();
    }
    v = sv[0];
    i = si[0];
    __syncthreads();
}

// grid (SCAN_BLOCKS, 3 nc), 256 threads: y = 3*comp + channel
__global__ __launch_bounds__(256) void k_scan_partial(ScanArgs A) {
    __shared__ double sv[256];
    __shared__ long long si[256];
    const int k = blockIdx.y;
    const Sys &y = A.y;
    const double *q3 = A.gamma + (long)(k / 3) * y.complen + (long)(k % 3) * y.chlen + y.q3;
    const long long n3 = (long long)y.ns * y.nO * y.nn * y.nn;
    double bv = -1.0;  // every |q3| beats it, so the first position seen is taken
    long long bi = 0;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n3;
         j += (long long)gridDim.x * blockDim.x) {
        const double v = fabs(q3[j]);
        if (v > bv) {  // positions of one thread increase, so ">" keeps the first maximum
            bv = v;
            bi = j;
        }
    }
    block_best(bv, bi, sv, si);
    if (threadIdx.x == 0) {
        A.pval[k * SCAN_BLOCKS + blockIdx.x] = bv;
        A.pidx[k * SCAN_BLOCKS + blockIdx.x] = bi;
    }
}

// 3 nc blocks of 256 threads: the maximum over the partials, then the cuts through it
__global__ __launch_bounds__(256) void k_scan_final(ScanArgs A) {
    __shared__ double sv[256];
    __shared__ long long si[256];
    const int k = blockIdx.x;
    const Sys &y = A.y;
    double bv = -2.0;
    long long bi = 0;
    for (int b = threadIdx.x; b < SCAN_BLOCKS; b += blockDim.x)
        if (beats(A.pval[k * SCAN_BLOCKS + b], A.pidx[k * SCAN_BLOCKS + b], bv, bi)) {
            bv = A.pval[k * SCAN_BLOCKS + b];
            bi = A.pidx[k * SCAN_BLOCKS + b];
        }
    block_best(bv, bi, sv, si);
    const int ns = y.ns, nO = y.nO, nn = y.nn;
    const int i = (int)(bi % ns);
    const int w = (int)((bi / ns) % nO);
    const int v = (int)((bi / ((long long)ns * nO)) % nn);
    const int vp = (int)(bi / ((long long)ns * nO * nn));
    if (threadIdx.x == 0) {
        A.idx[4 * k + 0] = i;
        A.idx[4 * k + 1] = w;
        A.idx[4 * k + 2] = v;
        A.idx[4 * k + 3] = vp;
    }
    const double *q3 = A.gamma + (long)(k / 3) * y.complen + (long)(k % 3) * y.chlen + y.q3;
    auto at = [&](int ww, int a, int b) { return q3[i + (long)ns * (ww + (long)nO * (a + (long)nn * b))]; };
    double *out = A.cuts + (long)k * (nO + 3 * nn);
    for (int t = threadIdx.x; t < nO + 3 * nn; t += blockDim.x) {
        if (t < nO) {
            out[t] = at(t, v, vp);
        } else {
            const int x = (t - nO) % nn, which = (t - nO) / nn;
            const int e = nn - 1;
            out[t] = which == 0 ? at(w, x, x) - at(w, e, e) : which == 1 ? at(w, x, vp) - at(w, e, vp)
                                                                       : at(w, v, x) - at(w, v, e);
        }
    }
}

// get_indices/get_param (param.jl:26-86) by the reference's linear scan; inv[j] = 1/(m[j+1]-m[j])
struct P1 {
    int lo, hi;
    double wl, wh;
};

__device__ __forceinline__ P1 param_of(double val, const double *m, const double *inv, int n) {
    P1 p;
    if (val >= m[n - 1]) {
        p.lo = p.hi = n - 1;
        p.wl = 1.0;
        p.wh = 0.0;
        return p;
    }
    int j = 0;
    while (val > m[j]) ++j;
    p.lo = p.hi = j;
    if (val < m[j]) p.lo = j - 1;
    if (p.lo < p.hi) {
        const double d = inv[p.lo];
        p.wl = d * (m[p.hi] - val);
        p.wh = d * (val - m[p.lo]);
    } else {
        p.wl = 1.0;
        p.wh = 0.0;
    }
    return p;
}

// one thread per output entry: the nc * complen vertex entries, then the n_sigma Σ entries
__global__ __launch_bounds__(256) void k_resample(ResampleArgs A) {
    const Sys &y = A.y;
    const long nv = (long)y.nc * y.complen;
    const long total = nv + y.nsig;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        if (t >= nv) {
            // get_Σ(σ_new[j], m_old, a_old) (Action.jl:8-26)
            const int j = (int)(t - nv);
            const double *so = A.mold;
            const double w = A.mnew[j];
            const double *S = A.src_sigma;
            double val;
            if (fabs(w) <= so[y.nsig - 1]) {
                const P1 p = param_of(fabs(w), so, A.mold + 5 * MAXN, y.nsig);
                const double sg = w > 0.0 ? 1.0 : (w < 0.0 ? -1.0 : w);
                val = sg * (p.wl * S[p.lo] + p.wh * S[p.hi]);
            } else {
                val = so[y.nsig - 1] * S[y.nsig - 1] / w;
            }
            A.dst_sigma[j] = val;
            continue;
        }
        const long comp = t / y.complen;
        const long r = t - comp * y.complen;
        const int ch = (int)(r / y.chlen);
        const long e = r - (long)ch * y.chlen;
        // s and u channels live on (Ωs, νs), t on (Ωt, νt)
        const int kO = ch == 1 ? 3 : 1, kn = ch == 1 ? 4 : 2;
        const double *Oo = A.mold + kO * MAXN, *no = A.mold + kn * MAXN;
        const double *iO = A.mold + (5 + kO) * MAXN, *in = A.mold + (5 + kn) * MAXN;
        const double *On = A.mnew + kO * MAXN, *nnw = A.mnew + kn * MAXN;
        const double *src = A.src_gamma + comp * y.complen + (long)ch * y.chlen;
        const int ns = y.ns, nO = y.nO, nn = y.nn;
        double val;
        if (e < y.q21) {
            // get_q1 (channel.jl:37-47)
            const int site = (int)(e % ns), w = (int)(e / ns);
            const P1 p1 = param_of(On[w], Oo, iO, nO);
            val = p1.wl * src[site + (long)ns * p1.lo];
            val += p1.wh * src[site + (long)ns * p1.hi];
        } else if (e < y.q3) {
            // get_q2_1 / get_q2_2 (channel.jl:94-107, 159-172)
            const long off = e < y.q22 ? y.q21 : y.q22;
            const long f = e - off;
            const int site = (int)(f % ns), w = (int)((f / ns) % nO), v = (int)(f / ((long)ns * nO));
            const P1 p1 = param_of(On[w], Oo, iO, nO);
            const P1 p2 = param_of(nnw[v], no, in, nn);
            const double *q = src + off;
            auto at = [&](int a, int b) { return q[site + (long)ns * (a + (long)nO * b)]; };
            val = p1.wl * p2.wl * at(p1.lo, p2.lo);
            val += p1.wh * p2.wl * at(p1.hi, p2.lo);
            val += p1.wl * p2.wh * at(p1.lo, p2.hi);
            val += p1.wh * p2.wh * at(p1.hi, p2.hi);
        } else {
            // get_q3 (channel.jl:224-242)
            const long f = e - y.q3;
            const int site = (int)(f % ns), w = (int)((f / ns) % nO), v = (int)((f / ((long)ns * nO)) % nn),
                      vp = (int)(f / ((long)ns * nO * nn));
            const P1 p1 = param_of(On[w], Oo, iO, nO);
            const P1 p2 = param_of(nnw[v], no, in, nn);
            const P1 p3 = param_of(nnw[vp], no, in, nn);
            const double *q = src + y.q3;
            auto at = [&](int a, int b, int c) { return q[site + (long)ns * (a + (long)nO * (b + (long)nn * c))]; };
            val = p1.wl * p2.wl * p3.wl * at(p1.lo, p2.lo, p3.lo);
            val += p1.wh * p2.wl * p3.wl * at(p1.hi, p2.lo, p3.lo);
            val += p1.wl * p2.wh * p3.wl * at(p1.lo, p2.hi, p3.lo);
            val += p1.wh * p2.wh * p3.wl * at(p1.hi, p2.hi, p3.lo);
            val += p1.wl * p2.wl * p3.wh * at(p1.lo, p2.lo, p3.hi);
            val += p1.wh * p2.wl * p3.wh * at(p1.hi, p2.lo, p3.hi);
            val += p1.wl * p2.wh * p3.wh * at(p1.lo, p2.hi, p3.hi);
            val += p1.wh * p2.wh * p3.wh * at(p1.hi, p2.hi, p3.hi);
        }
        A.dst_gamma[t] = val;
    }
}

}  // namespace

hipError_t launch_scan_cuts(const ScanArgs &A, hipStream_t stream) {
    hipLaunchKernelGGL(k_scan_partial, dim3(SCAN_BLOCKS, 3 * A.y.nc), dim3(256), 0, stream, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_final, dim3(3 * A.y.nc), dim3(256), 0, stream, A);
    return hipGetLastError();
}

// Scratch of the allocation-free scan_cuts entry (pffrg_scan_cuts_noalloc): the argmax partials,
// the maximum's position and the cuts of one call, SCAN_CUTS_BYTES (54.5 KB).  Each GPU holds its
// own copy for the life of the process, shared by every context on that GPU: calls that overlap
// on one GPU race.  Every byte a call reads is written by the same call first, so nothing needs
// clearing between calls.  Host code must not pass `scratch` itself to a kernel (on the host the
// name stands for a host-side shadow, not the device copy): it uses the address that
// scan_scratch_address looks up on each device.
static __device__ __align__(256) unsigned char scratch[SCAN_CUTS_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

hipError_t scan_scratch_address(int device, void **ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_resample(const ResampleArgs &A, hipStream_t stream) {
    const long total = (long)A.y.nc * A.y.complen + A.y.nsig;
    long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_resample, dim3((unsigned)blocks), dim3(256), 0, stream, A);
    return hipGetLastError();
}

}  // namespace pffrg

namespace {

// get_mesh (mesh.jl:86-121) into out[0..num]; false where the reference asserts
bool lin_log_mesh(double linear, double upper, int num, double p, double *out) {
    if (!(linear > 0.0) || !(upper > 0.0) || !(linear < upper)) return false;
    const int num_lin = (int)std::ceil(p * num);
    const int num_log = num - num_lin;
    if (num_log <= 0 || num_lin <= 0) return false;
    const double h = linear / num_lin;
    const double xi = std::pow(upper / linear, 1.0 / num_log);
    out[0] = 0.0;
    for (int i = 1; i <= num_lin; ++i) out[i] = i * h;
    for (int i = 1; i <= num_log; ++i) out[num_lin + i] = std::pow(xi, (double)i) * linear;
    return true;
}

// Julia's sign: ±1, and zero (of either sign) or NaN returned as is
double jl_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }

double jl_max(double a, double b) {
    if (std::isnan(a) || std::isnan(b)) return NAN;
    return a > b ? a : (b > a ? b : (std::signbit(a) ? b : a));
}

double jl_min(double a, double b) {
    if (std::isnan(a) || std::isnan(b)) return NAN;
    return a < b ? a : (b < a ? b : (std::signbit(a) ? a : b));
}

// linear-scan get_param (param.jl:26-86) and the interpolated value of y at val
double interp_at(double val, const double *x, const double *y, int n) {
    if (val >= x[n - 1]) return 1.0 * y[n - 1] + 0.0 * y[n - 1];
    int j = 0;
    while (val > x[j]) ++j;
    int lo = j;
    const int hi = j;
    if (val < x[j]) lo = j - 1;
    double wl = 1.0, wh = 0.0;
    if (lo < hi) {
        const double d = 1.0 / (x[hi] - x[lo]);
        wl = d * (x[hi] - val);
        wh = d * (val - x[lo]);
    }
    return wl * y[lo] + wh * y[hi];
}

constexpr long SCAN_MAX_ITER = 1000000;  // the origin heuristic's 2 % steps; a finite cut needs far fewer

}  // namespace

extern "C" int pffrg_scan(const double *x, const double *y, int32_t n, const double *p, double *delta) {
    if (!x || !y || !p || !delta || n < 3) return PFFRG_ERR_ARG;
    const double p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3], p4 = p[4];
    const int num_lin = (int)std::ceil(p0 * (n - 1));

    // a) first finite-frequency peak: δa = where the cut has decayed to 75 % of it
    double da = INFINITY;
    int idx = 0;
    double slope = (y[1] - y[0]) / (x[1] - x[0]);
    for (int i = 1; i < n - 1; ++i) {
        const double s = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        if (jl_sign(s) != jl_sign(slope)) {
            idx = i;
            break;
        }
        slope = s;
    }
    if (idx > 0) {
        for (int i = idx + 1; i < n; ++i)
            if (std::fabs(y[i] / y[idx]) <= 0.75) {
                // at most half of the linear points below the peak
                da = jl_max(x[i], num_lin * x[idx] / std::ceil(0.5 * num_lin));
                break;
            }
    }

    // b) finite value at the origin: keep its relative step Δ to the first point in (p1, p2)
    double db = num_lin * x[1];
    double ymax = 0.0;
    bool nan = false;
    for (int i = 0; i < n; ++i) {
        if (std::isnan(y[i])) nan = true;
        ymax = std::fabs(y[i]) > ymax ? std::fabs(y[i]) : ymax;
    }
    if (nan) ymax = NAN;
    if (std::fabs(y[0]) / ymax > 1e-2) {
        double D = std::fabs(y[1] - y[0]) / jl_max(std::fabs(y[1]), std::fabs(y[0]));
        std::vector<double> xp(n);
        long it = 0;
        while (!(p1 < D && D < p2)) {
            if (D >= p2)
                db *= 0.98;
            else if (D <= p1)
                db *= 1.02;
            if (db >= x[n - 1]) break;
            if (++it > SCAN_MAX_ITER) return PFFRG_ERR_LIMIT;
            if (!lin_log_mesh(db, x[n - 1], n - 1, p0, xp.data())) return PFFRG_ERR_ARG;
            const double y0 = interp_at(xp[0], x, y, n), y1 = interp_at(xp[1], x, y, n);
            D = std::fabs(y1 - y0) / jl_max(std::fabs(y1), std::fabs(y0));
        }
    }

    // c) merge and bound
    *delta = jl_min(jl_max(num_lin * p3, jl_min(da, db)), p4);
    return 0;
}
