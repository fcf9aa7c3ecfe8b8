// pffrg_device.h -- device-side building blocks of the gfx950 flow right-hand side:
// mesh lookup, frequency folding, vertex interpolation, propagators and the
// Gauss-Kronrod block integrator.  Included by pffrg.hip only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pffrg_math.h"

namespace pffrg {

constexpr int MAXN = 64;      // mesh points per axis held in LDS
constexpr int WAVE = 64;      // gfx950 wavefront
constexpr int N_COUNTERS = 32;  // device measurement counters of a context (nodes, bytes, phase cycles)
#ifdef PFFRG_SHARED_PROFILE_UNITS
// measurement build: the counters, a record count and four words per shared-pass unit run
constexpr int N_COUNTERS_ALLOC = N_COUNTERS + 1 + 4 * 16384;
#else
constexpr int N_COUNTERS_ALLOC = N_COUNTERS;
#endif

// Bracket hints of the five meshes (σ, Ωs, νs, Ωt, νt), stored after the meshes and their
// inverse spacings in the device mesh table.  A positive double's bits shifted right by 49 are
// its exponent and three leading mantissa bits: a cell of 1/8 octave with exact binary bounds.
// Cell c of mesh k holds every value whose key minus hint_c0[k] is c, clamped to
// [0, hint_nc[k]); tab[k][c] is the first point index >= the cell's lower bound (clamped to
// n-1), with bit 7 set when the cell's answers span more than three indices.  96 cells cover
// 12 octaves at full resolution; the table and its header take 500 bytes of LDS.
constexpr int HINT_CELLS = 96;
constexpr int HINT_WORDS = (20 + 5 * HINT_CELLS) / 4;  // int16 c0[5], nc[5], then uint8 tab[5][HINT_CELLS]
constexpr int HINT_DOUBLES = (HINT_WORDS + 1) / 2;
constexpr int MESH_TABLE = 10 * MAXN + HINT_DOUBLES;  // doubles of the device mesh table

__host__ __device__ inline int hint_key(double v) {
    long long b;
    __builtin_memcpy(&b, &v, sizeof b);
    return (int)(b >> 49);
}
constexpr double PI_D = 3.14159265358979323846;

// A value equal on every lane of the wave, moved to scalar registers.  Values read from LDS
// count as per-lane to the compiler, which then keeps them in vector registers for the whole
// element; only for wave-uniform values (the result is lane 0's).
__device__ __forceinline__ double wave_uniform(double v) {
    unsigned long long b;
    __builtin_memcpy(&b, &v, sizeof b);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    b = ((unsigned long long)hi << 32) | lo;
    __builtin_memcpy(&v, &b, sizeof v);
    return v;
}

enum PassType { P_KAT = 0, P_LEFT = 1, P_CENTRAL = 2, P_BSE = 3, P_REDUCED = 4 };

// Channel/array offsets of the packed vertex (site fastest, column-major).
struct Sys {
    int ns, nO, nn, nsig;
    int G, GS;            // site groups per wavefront and lanes per group
    double S2;            // 2S
    long q21, q22, q3;    // offsets of q2_1, q2_2, q3 inside a channel block
    long chlen, complen;  // channel block length, component block length
    double valS, valT;    // Ω_max + ν_max of the s/u and t meshes
    int nc;               // vertex components: 2 (SU(2): spin, density) or 6 (U(1)-DM: xx, zz, DM, dd, zd, dz)
};

// U(1)-DM component indices (action_u1_dm.jl:1-11)
enum { U1_XX = 0, U1_ZZ = 1, U1_DM = 2, U1_DD = 3, U1_ZD = 4, U1_DZ = 5 };

// Mesh views; in the pass kernels they point into LDS.  ST is the scalar type of the
// per-node work (double for the FP64 path, float for the FP32-mixed path).
template <typename ST = double>
struct Meshes {
    const ST *sig, *Os, *ns, *Ot, *nt;
};

template <typename ST = double>
struct Param {
    int lo, hi;
    ST wl, wh;
};

// get_indices/get_param semantics (Frequency/param.jl:26-86) with a bisection over the
// LDS-resident mesh instead of the reference's linear scan: upper = first index with
// mesh[idx] >= val, lower = upper unless val < mesh[upper]; clamp to the last point.
template <typename ST>
__device__ __forceinline__ Param<ST> find_param(ST val, const ST *m, int n) {
    Param<ST> p;
    if (val >= m[n - 1]) {
        p.lo = p.hi = n - 1;
        p.wl = (ST)1;
        p.wh = (ST)0;
        return p;
    }
    int l = -1, h = n - 1;
    while (h - l > 1) {
        int mid = (l + h) >> 1;
        if (m[mid] >= val) h = mid;
        else l = mid;
    }
    p.hi = h;
    p.lo = (val < m[h]) ? h - 1 : h;
    if (p.lo < p.hi) {
        const ST d = (ST)1 / (m[p.hi] - m[p.lo]);
        p.wl = d * (m[p.hi] - val);
        p.wh = d * (val - m[p.lo]);
    } else {
        p.wl = (ST)1;
        p.wh = (ST)0;
    }
    return p;
}

// Access buffer of Frequency/buffer.jl:16-26: kernel 0 = nothing, 1 = q1, 2 = q2_1, 3 = q2_2, 4 = q3.
template <typename ST = double>
struct Buf {
    int kernel;
    bool exch, map, smu, snu;
    bool smunu;  // read by the U(1)-DM flags only
    Param<ST> p1, p2, p3;
};

template <typename ST = double>
__device__ __forceinline__ Buf<ST> buf_empty() {
    Buf<ST> b;
    b.kernel = 0;
    b.exch = b.map = b.smu = b.snu = b.smunu = false;
    b.p1 = b.p2 = b.p3 = Param<ST>{0, 0, (ST)0, (ST)0};
    return b;
}

// get_buffer, Frequency/buffer.jl:51-86
template <typename ST>
__device__ __forceinline__ Buf<ST> make_buf(ST w, ST v, ST vp, const ST *Om, const ST *nu, int nO, int nn, bool ex,
                                            bool mp, bool smu, bool snu) {
    Buf<ST> b = buf_empty<ST>();
    if (Om[nO - 1] < fabs(w)) return b;
    b.exch = ex;
    b.map = mp;
    b.smu = smu;
    b.snu = snu;
    const bool av = nu[nn - 1] < fabs(v);
    const bool avp = nu[nn - 1] < fabs(vp);
    b.p1 = find_param(w, Om, nO);
    if (av) {
        if (avp) {
            b.kernel = 1;
        } else {
            b.kernel = 3;
            b.p3 = find_param(vp, nu, nn);
        }
    } else {
        b.p2 = find_param(v, nu, nn);
        if (avp) {
            b.kernel = 2;
        } else {
            b.kernel = 4;
            b.p3 = find_param(vp, nu, nn);
        }
    }
    return b;
}

// get_buffer_s (buffer.jl:88-124)
template <typename ST>
__device__ __forceinline__ Buf<ST> buf_s(const Meshes<ST> &M, const Sys &y, ST w, ST v, ST vp) {
    bool ex = false, mp = false, smu = false, snu = false;
    if (w < (ST)0) { w = -w; ex = !ex; }
    if (v < (ST)0) { v = -v; ex = !ex; mp = !mp; smu = !smu; }
    if (vp < (ST)0) { vp = -vp; mp = !mp; snu = !snu; }
    return make_buf(w, v, vp, M.Os, M.ns, y.nO, y.nn, ex, mp, smu, snu);
}

// get_buffer_t (buffer.jl:127-160)
template <typename ST>
__device__ __forceinline__ Buf<ST> buf_t(const Meshes<ST> &M, const Sys &y, ST w, ST v, ST vp) {
    bool smu = false, snu = false, smunu = false;
    if (w < (ST)0) { w = -w; smunu = !smunu; }
    if (v < (ST)0) { v = -v; smu = !smu; }
    if (vp < (ST)0) { vp = -vp; snu = !snu; }
    Buf<ST> b = make_buf(w, v, vp, M.Ot, M.nt, y.nO, y.nn, false, false, smu, snu);
    b.smunu = b.kernel ? smunu : false;
    return b;
}

// get_buffer_u (buffer.jl:163-199)
template <typename ST>
__device__ __forceinline__ Buf<ST> buf_u(const Meshes<ST> &M, const Sys &y, ST w, ST v, ST vp) {
    bool ex = false, mp = false, snu = false, smunu = false;
    if (w < (ST)0) { w = -w; ex = !ex; smunu = !smunu; }
    if (v < (ST)0) { v = -v; ex = !ex; mp = !mp; snu = !snu; }
    if (vp < (ST)0) { vp = -vp; mp = !mp; snu = !snu; }
    Buf<ST> b = make_buf(w, v, vp, M.Os, M.ns, y.nO, y.nn, ex, mp, false, snu);
    b.smunu = b.kernel ? smunu : false;
    return b;
}

// Interpolate one channel at one site for both SU(2) components (channel.jl:42-305):
// c0/c1 point at the channel block of component 0/1.  Corner order and weight
// products follow the reference so FP64 results agree to rounding.
template <typename TV>
__device__ __forceinline__ void interp2(const TV *__restrict__ c0, const TV *__restrict__ c1, const Buf<> &b, int site,
                                        const Sys &y, TV &r0, TV &r1) {
    const long ns = y.ns, nO = y.nO, nn = y.nn;
    if (b.kernel == 1) {
        const TV wl = (TV)b.p1.wl, wh = (TV)b.p1.wh;
        const long o0 = site + ns * b.p1.lo, o1 = site + ns * b.p1.hi;
        r0 = wl * c0[o0];
        r0 += wh * c0[o1];
        r1 = wl * c1[o0];
        r1 += wh * c1[o1];
    } else if (b.kernel == 4) {
        const TV *q0 = c0 + y.q3, *q1 = c1 + y.q3;
        const double a1 = b.p1.wl, b1 = b.p1.wh, a2 = b.p2.wl, b2 = b.p2.wh, a3 = b.p3.wl, b3 = b.p3.wh;
        const TV w000 = (TV)(a1 * a2 * a3), w100 = (TV)(b1 * a2 * a3), w010 = (TV)(a1 * b2 * a3),
                 w110 = (TV)(b1 * b2 * a3), w001 = (TV)(a1 * a2 * b3), w101 = (TV)(b1 * a2 * b3),
                 w011 = (TV)(a1 * b2 * b3), w111 = (TV)(b1 * b2 * b3);
        const long l1 = b.p1.lo, h1 = b.p1.hi;
        const long l2 = nO * b.p2.lo, h2 = nO * b.p2.hi;
        const long l3 = nO * nn * b.p3.lo, h3 = nO * nn * b.p3.hi;
        const long o000 = site + ns * (l1 + l2 + l3), o100 = site + ns * (h1 + l2 + l3);
        const long o010 = site + ns * (l1 + h2 + l3), o110 = site + ns * (h1 + h2 + l3);
        const long o001 = site + ns * (l1 + l2 + h3), o101 = site + ns * (h1 + l2 + h3);
        const long o011 = site + ns * (l1 + h2 + h3), o111 = site + ns * (h1 + h2 + h3);
        TV v0 = w000 * q0[o000];
        TV v1 = w000 * q1[o000];
        v0 += w100 * q0[o100]; v1 += w100 * q1[o100];
        v0 += w010 * q0[o010]; v1 += w010 * q1[o010];
        v0 += w110 * q0[o110]; v1 += w110 * q1[o110];
        v0 += w001 * q0[o001]; v1 += w001 * q1[o001];
        v0 += w101 * q0[o101]; v1 += w101 * q1[o101];
        v0 += w011 * q0[o011]; v1 += w011 * q1[o011];
        v0 += w111 * q0[o111]; v1 += w111 * q1[o111];
        r0 = v0;
        r1 = v1;
    } else if (b.kernel == 2 || b.kernel == 3) {
        const long off = (b.kernel == 2) ? y.q21 : y.q22;
        const Param<> &pv = (b.kernel == 2) ? b.p2 : b.p3;
        const TV *q0 = c0 + off, *q1 = c1 + off;
        const double a1 = b.p1.wl, b1 = b.p1.wh, a2 = pv.wl, b2 = pv.wh;
        const TV w00 = (TV)(a1 * a2), w10 = (TV)(b1 * a2), w01 = (TV)(a1 * b2), w11 = (TV)(b1 * b2);
        const long o00 = site + ns * (b.p1.lo + nO * pv.lo), o10 = site + ns * (b.p1.hi + nO * pv.lo);
        const long o01 = site + ns * (b.p1.lo + nO * pv.hi), o11 = site + ns * (b.p1.hi + nO * pv.hi);
        TV v0 = w00 * q0[o00];
        TV v1 = w00 * q1[o00];
        v0 += w10 * q0[o10]; v1 += w10 * q1[o10];
        v0 += w01 * q0[o01]; v1 += w01 * q1[o01];
        v0 += w11 * q0[o11]; v1 += w11 * q1[o11];
        r0 = v0;
        r1 = v1;
    } else {
        r0 = (TV)0;
        r1 = (TV)0;
    }
}

// apply_flags_su2 (action_su2.jl:73-92): density sign from the two spin flags
__device__ __forceinline__ double dens_sign(const Buf<> &b) {
    double s = 1.0;
    if (b.smu) s *= -1.0;
    if (b.snu) s *= -1.0;
    return s;
}

// apply_flags_u1_dm (action_u1_dm.jl:79-118) from the buffer's flags (bit 0 exchange, 1 sgn_μ,
// 2 sgn_ν, 3 sgn_μν): the sign of the channel's contribution to component comp and the component
// src it is read from (the site exchange flips DM and swaps zd and dz)
__device__ __forceinline__ double flag_u1(int flags, int comp, int &src) {
    double sgn = 1.0;
    src = comp;
    if ((comp == U1_ZD || comp == U1_DZ) && (flags & 8)) sgn *= -1.0;
    if ((comp == U1_DD || comp == U1_DZ) && (flags & 2)) sgn *= -1.0;
    if ((comp == U1_DD || comp == U1_ZD) && (flags & 4)) sgn *= -1.0;
    if (flags & 1) {
        if (comp == U1_DM) sgn *= -1.0;
        else if (comp == U1_ZD) src = U1_DZ;
        else if (comp == U1_DZ) src = U1_ZD;
    }
    return sgn;
}

__device__ __forceinline__ int buf_flags(const Buf<> &b) {
    return (b.exch ? 1 : 0) | (b.smu ? 2 : 0) | (b.snu ? 4 : 0) | (b.smunu ? 8 : 0);
}

// One action as seen by a kernel: vertex channels (TV), bare and Σ (FP64).
template <typename TV>
struct ActV {
    const TV *gamma;
    const double *bare;
    const double *Sigma;
};

// get_Γ_comp / get_Γ_comp_avx! (Action/Action.jl:29-167) for one site and both
// components, with channel masks MS/MT/MU (the reference's ch_s/ch_t/ch_u keywords).
__device__ __forceinline__ int corners(int kernel) { return kernel == 4 ? 8 : (kernel == 1 ? 2 : (kernel ? 4 : 0)); }

// Returns the number of interpolation corners gathered per component (for the
// algorithmic-bytes counter).
template <typename TV, bool MS, bool MT, bool MU>
__device__ __forceinline__ int gamma_site(const ActV<TV> &A, const Buf<> &bs, const Buf<> &bt, const Buf<> &bu, int i,
                                          const int *exch, const Sys &y, TV &vs, TV &vd) {
    vs = (TV)A.bare[i];
    vd = (TV)A.bare[y.ns + i];
    if (MS && bs.kernel) {
        const int s = bs.exch ? exch[i] : i;
        const long cb = (bs.map ? 2 : 0) * y.chlen;
        TV r0, r1;
        interp2(A.gamma + cb, A.gamma + y.complen + cb, bs, s, y, r0, r1);
        vs += r0;
        vd += (TV)dens_sign(bs) * r1;
    }
    if (MT && bt.kernel) {
        const int s = bt.exch ? exch[i] : i;
        const long cb = y.chlen;
        TV r0, r1;
        interp2(A.gamma + cb, A.gamma + y.complen + cb, bt, s, y, r0, r1);
        vs += r0;
        vd += (TV)dens_sign(bt) * r1;
    }
    if (MU && bu.kernel) {
        const int s = bu.exch ? exch[i] : i;
        const long cb = (bu.map ? 0 : 2) * y.chlen;
        TV r0, r1;
        interp2(A.gamma + cb, A.gamma + y.complen + cb, bu, s, y, r0, r1);
        vs += r0;
        vd += (TV)dens_sign(bu) * r1;
    }
    return (MS ? corners(bs.kernel) : 0) + (MT ? corners(bt.kernel) : 0) + (MU ? corners(bu.kernel) : 0);
}

// One channel of one component at one site (channel.jl:42-305) on the plain FP64 layout.
__device__ __forceinline__ double interp1(const double *__restrict__ c, const Buf<> &b, int site, const Sys &y) {
    double r0, r1;
    interp2<double>(c, c, b, site, y, r0, r1);
    return r0;
}

// get_Γ_comp for the six U(1)-DM components at one site (Action.jl:29-107 with
// apply_flags_u1_dm), channel masks MS/MT/MU; v[comp].
template <bool MS, bool MT, bool MU>
__device__ __forceinline__ void gamma_site_u1(const ActV<double> &A, const Buf<> &bs, const Buf<> &bt, const Buf<> &bu,
                                              int i, const int *exch, const Sys &y, double *v) {
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = A.bare[k * y.ns + i];
    const Buf<> *bb[3] = {&bs, &bt, &bu};
    const bool on[3] = {MS, MT, MU};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const Buf<> &b = *bb[c];
        if (!on[c] || !b.kernel) continue;
        const int s = b.exch ? exch[i] : i;
        const int chn = (c == 1) ? 1 : ((c == 0) == b.map ? 2 : 0);
        const int fl = buf_flags(b);
        double r[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) r[k] = interp1(A.gamma + k * y.complen + chn * y.chlen, b, s, y);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            int src;
            const double sg = flag_u1(fl, k, src);
            v[k] += sg * r[src];
        }
    }
}

// ---------------------------------------------------------------- propagators (Flow/bubbles.jl)
// get_Σ (Action.jl:8-26)
template <typename ST>
__device__ __forceinline__ ST sigma_at(ST w, const ST *sm, int n, const ST *Sig) {
    if (fabs(w) <= sm[n - 1]) {
        const Param<ST> p = find_param<ST>(fabs(w), sm, n);
        const ST sg = (w > (ST)0) ? (ST)1 : ((w < (ST)0) ? (ST)-1 : w);
        return sg * (p.wl * Sig[p.lo] + p.wh * Sig[p.hi]);
    }
    return sm[n - 1] * Sig[n - 1] / w;
}

// exp and expm1: the written-out double versions shared with the oracle, the device library's
// in single precision
template <typename ST>
__device__ __forceinline__ ST exp_st(ST x) {
    if constexpr (std::is_same<ST, double>::value) return pf_exp(x);
    else return expf(x);
}

template <typename ST>
__device__ __forceinline__ ST expm1_st(ST x) {
    if constexpr (std::is_same<ST, double>::value) return pf_expm1(x);
    else return expm1f(x);
}

template <typename ST>
__device__ __forceinline__ ST G_bare(ST L, ST w) {
    return (fabs(w) > (ST)1e-8) ? -expm1_st<ST>(-(w * w) / (L * L)) / w : (ST)0;
}

template <typename ST>
__device__ __forceinline__ ST G_full(ST L, ST w, const ST *sm, int n, const ST *Sig) {
    if (!(fabs(w) > (ST)1e-8)) return (ST)0;
    const ST s = sigma_at<ST>(w, sm, n, Sig);
    const ST G0 = G_bare<ST>(L, w);
    const ST denom = (w * s >= (ST)0) ? (ST)1 / G0 + s : (ST)1 / G0;
    return (ST)1 / denom;
}

__device__ __forceinline__ double S_single(double L, double w, const double *sm, int n, const double *Sig) {
    if (!(fabs(w) > 1e-8)) return 0.0;
    const double G = G_full<double>(L, w, sm, n, Sig);
    const double G0 = G_bare<double>(L, w);
    const double r = G / G0;
    return r * r * pf_exp(-(w * w) / (L * L)) * 2.0 * w / (L * L * L);
}

// symmetrised Katanin bubble K(x1,x2) + K(x2,x1), K = (S + G²dΣ) G' / 2π (bubbles.jl:64-80)
__device__ __forceinline__ double bubble_kat(double L, double x1, double x2, const double *sm, int n,
                                             const double *Sig, const double *dSig) {
    const double G1 = G_full<double>(L, x1, sm, n, Sig), G2 = G_full<double>(L, x2, sm, n, Sig);
    const double S1 = S_single(L, x1, sm, n, Sig), S2 = S_single(L, x2, sm, n, Sig);
    const double d1 = sigma_at<double>(x1, sm, n, dSig), d2 = sigma_at<double>(x2, sm, n, dSig);
    return (S1 + G1 * G1 * d1) * G2 / (2.0 * PI_D) + (S2 + G2 * G2 * d2) * G1 / (2.0 * PI_D);
}

// -G1 G2 / 2π (bubbles.jl:83-96, with the sign of the callers)
__device__ __forceinline__ double bubble_plain(double L, double x1, double x2, const double *sm, int n,
                                               const double *Sig) {
    const double G1 = G_full<double>(L, x1, sm, n, Sig), G2 = G_full<double>(L, x2, sm, n, Sig);
    return -(G1 * G2 / (2.0 * PI_D));
}

// fermionic arguments of the bubble per channel (flow_su2_{s,t,u}.jl:18)
template <typename ST>
__device__ __forceinline__ void bubble_args(int ch, ST W, ST v, ST &x1, ST &x2) {
    const ST h = (ST)0.5 * W;
    if (ch == 0) { x1 = v + h; x2 = h - v; }
    else if (ch == 1) { x1 = v + h; x2 = v - h; }
    else { x1 = v - h; x2 = v + h; }
}

// ---------------------------------------------------------------- Gauss-Kronrod (QuadGK.jl restated)
__device__ static const double GKX[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                                         0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                                         0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                                         0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
__device__ static const double GKWK[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                          0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                          0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                          0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
__device__ static const double GKWG[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                          0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

constexpr int GK_MAX_SEGS = 512;

struct Seg {
    double a, b, I, E;
};

// node k (0..14) of the 15-point rule on [a, b]; order matches the oracle's evaluation order
__device__ __forceinline__ double gk_node(double a, double b, int k) {
    const double s = 0.5 * (b - a);
    if (k < 12) {
        const int i = k >> 2, r = k & 3;
        const double x = (r < 2) ? GKX[2 * i + 1] : GKX[2 * i];
        return (r & 1) ? a + (1.0 + x) * s : a + (1.0 - x) * s;
    }
    if (k == 12) return a + s;
    if (k == 13) return a + (1.0 - GKX[6]) * s;
    return a + (1.0 + GKX[6]) * s;
}

__device__ __forceinline__ Seg gk_combine(double a, double b, const double *F) {
    const double s = 0.5 * (b - a);
    double Ig = 0.0, Ik = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double fg = F[4 * i] + F[4 * i + 1];
        const double fk = F[4 * i + 2] + F[4 * i + 3];
        Ig += fg * GKWG[i];
        Ik += fg * GKWK[2 * i + 1] + fk * GKWK[2 * i];
    }
    const double f0 = F[12];
    Ig += f0 * GKWG[3];
    Ik += f0 * GKWK[7] + (F[13] + F[14]) * GKWK[6];
    const double Iks = Ik * s, Igs = Ig * s;
    return Seg{a, b, Iks, fabs(Iks - Igs)};
}

// Variable maps of quadgk for infinite endpoints: mode 0 = (-Inf, Inf) with t/(1-t²),
// mode 1 = (s0, Inf) with s0 + t/(1-t), mode 2 = (-Inf, s0) with s0 - t/(1-t).
template <class F>
__device__ __forceinline__ double gk_mapped(const F &f, int mode, double s0, double t) {
    if (mode == 0) {
        const double t2 = t * t;
        const double den = 1.0 / (1.0 - t2);
        return f(t * den) * (1.0 + t2) * den * den;
    }
    const double den = 1.0 / (1.0 - t);
    return (mode == 1 ? f(s0 + t * den) : f(s0 - t * den)) * den * den;
}

// Adaptive Gauss-Kronrod over one interval family, evaluated by one 64-lane workgroup:
// lanes evaluate the 15 (initial: up to 60, bisection: 30) rule nodes in parallel,
// lane 0 keeps the segment list.  The segment with the largest error estimate is
// bisected until E <= atol or E <= rtol |I| (QuadGK's criterion) or the cap is hit;
// the total is re-summed in segment order.  Must be called by all 64 lanes.
template <class F>
__device__ double gk_block(const F &f, int mode, double s0, const double *bp, int nbp, double atol, double rtol,
                           Seg *segs, double *fv, int *ctl, double *ctld) {
    const int lane = threadIdx.x;
    const int nseg0 = nbp - 1;
    // initial segments
    if (lane < 15 * nseg0) {
        const int q = lane / 15, k = lane % 15;
        fv[lane] = gk_mapped(f, mode, s0, gk_node(bp[q], bp[q + 1], k));
    }
    __syncthreads();
    if (lane == 0) {
        double I = 0.0, E = 0.0;
        for (int q = 0; q < nseg0; ++q) {
            segs[q] = gk_combine(bp[q], bp[q + 1], fv + 15 * q);
            if (q == 0) { I = segs[0].I; E = segs[0].E; }
            else { I += segs[q].I; E += segs[q].E; }
        }
        ctl[0] = nseg0;
        ctld[0] = I;
        ctld[1] = E;
        ctl[1] = (E <= atol || E <= rtol * fabs(I)) ? 0 : 1;
    }
    __syncthreads();
    while (true) {
        if (lane == 0) {
            const double I = ctld[0], E = ctld[1];
            int go = ctl[1] && (E > atol && E > rtol * fabs(I) && ctl[0] < GK_MAX_SEGS);
            if (go) {
                int k = 0;
                for (int j = 1; j < ctl[0]; ++j)
                    if (segs[j].E > segs[k].E) k = j;
                ctl[2] = k;
            }
            ctl[1] = go;
        }
        __syncthreads();
        if (!ctl[1]) break;
        const int k = ctl[2];
        const double a = segs[k].a, b = segs[k].b;
        const double mid = 0.5 * (a + b);
        if (lane < 30) {
            const int q = lane / 15, kk = lane % 15;
            fv[lane] = gk_mapped(f, mode, s0, q == 0 ? gk_node(a, mid, kk) : gk_node(mid, b, kk));
        }
        __syncthreads();
        if (lane == 0) {
            const Seg s = segs[k];
            const Seg s1 = gk_combine(a, mid, fv), s2 = gk_combine(mid, b, fv + 15);
            ctld[0] = (ctld[0] - s.I) + s1.I + s2.I;
            ctld[1] = (ctld[1] - s.E) + s1.E + s2.E;
            segs[k] = s1;
            segs[ctl[0]] = s2;
            ctl[0] += 1;
        }
        __syncthreads();
    }
    double I = 0.0;
    if (lane == 0) {
        I = segs[0].I;
        for (int j = 1; j < ctl[0]; ++j) I += segs[j].I;
    }
    return I;  // valid on lane 0
}

// gk_block with the rule nodes evaluated by the whole block: ev(n, ts, fv) is called by every
// thread and writes fv[q], the mapped integrand at rule node ts[q] (q < n; ts and fv in LDS, room
// for 60 nodes).  The segment list, the bisection choice and every sum are gk_block's (thread 0).
template <typename EV>
__device__ double gk_block_coop(const EV &ev, const double *bp, int nbp, double atol, double rtol, Seg *segs,
                                double *fv, double *ts, int *ctl, double *ctld) {
    const int lane = threadIdx.x;
    const int nseg0 = nbp - 1;
    if (lane < 15 * nseg0) {
        const int q = lane / 15, k = lane % 15;
        ts[lane] = gk_node(bp[q], bp[q + 1], k);
    }
    __syncthreads();
    ev(15 * nseg0, ts, fv);
    __syncthreads();
    if (lane == 0) {
        double I = 0.0, E = 0.0;
        for (int q = 0; q < nseg0; ++q) {
            segs[q] = gk_combine(bp[q], bp[q + 1], fv + 15 * q);
            if (q == 0) { I = segs[0].I; E = segs[0].E; }
            else { I += segs[q].I; E += segs[q].E; }
        }
        ctl[0] = nseg0;
        ctld[0] = I;
        ctld[1] = E;
        ctl[1] = (E <= atol || E <= rtol * fabs(I)) ? 0 : 1;
    }
    __syncthreads();
    __shared__ double s_be[32];
    __shared__ int s_bk[32];
    const int nwv = (blockDim.x + 63) / 64;
    while (true) {
        // the segment with the largest error, the first one on ties (gk_block's serial choice):
        // each thread scans its stride, then the waves and the block reduce the (error, index) pairs
        const int nseg = ctl[0];
        double eb = 0.0;
        int kb = -1;
        for (int j = lane; j < nseg; j += blockDim.x)
            if (kb < 0 || segs[j].E > eb) {
                eb = segs[j].E;
                kb = j;
            }
        for (int off = 32; off > 0; off >>= 1) {
            const double eo = __shfl_xor(eb, off);
            const int ko = __shfl_xor(kb, off);
            if (ko >= 0 && (kb < 0 || eo > eb || (eo == eb && ko < kb))) {
                eb = eo;
                kb = ko;
            }
        }
        if ((lane & 63) == 0) {
            s_be[lane >> 6] = eb;
            s_bk[lane >> 6] = kb;
        }
        __syncthreads();
        if (lane == 0) {
            const double I = ctld[0], E = ctld[1];
            int go = ctl[1] && (E > atol && E > rtol * fabs(I) && ctl[0] < GK_MAX_SEGS);
            if (go) {
                int k = s_bk[0];
                double ek = s_be[0];
                for (int q = 1; q < nwv; ++q)
                    if (s_bk[q] >= 0 && (k < 0 || s_be[q] > ek || (s_be[q] == ek && s_bk[q] < k))) {
                        ek = s_be[q];
                        k = s_bk[q];
                    }
                ctl[2] = k;
            }
            ctl[1] = go;
        }
        __syncthreads();
        if (!ctl[1]) break;
        const int k = ctl[2];
        const double a = segs[k].a, b = segs[k].b;
        const double mid = 0.5 * (a + b);
        if (lane < 30) {
            const int q = lane / 15, kk = lane % 15;
            ts[lane] = q == 0 ? gk_node(a, mid, kk) : gk_node(mid, b, kk);
        }
        __syncthreads();
        ev(30, ts, fv);
        __syncthreads();
        if (lane == 0) {
            const Seg s = segs[k];
            const Seg s1 = gk_combine(a, mid, fv), s2 = gk_combine(mid, b, fv + 15);
            ctld[0] = (ctld[0] - s.I) + s1.I + s2.I;
            ctld[1] = (ctld[1] - s.E) + s1.E + s2.E;
            segs[k] = s1;
            segs[ctl[0]] = s2;
            ctl[0] += 1;
        }
        __syncthreads();
    }
    double I = 0.0;
    if (lane == 0) {
        I = segs[0].I;
        for (int j = 1; j < ctl[0]; ++j) I += segs[j].I;
    }
    return I;  // valid on thread 0
}

}  // namespace pffrg
