// pffrg_oracle.cpp -- FP64 CPU restatement of the PFFRGSolver.jl flow right-hand side for the
// SU(2) action (2 vertex components: spin, density) and the U(1)-symmetric Dzyaloshinskii-Moriya
// action (6 components: xx, zz, DM, dd, zd, dz; action_u1_dm.jl:1-11).
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU baseline
// ("port") for the HIP library.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it.  It is never part of the product path.
//
// Every routine follows one reference function; the citation is given as
// file:line relative to the reference repository root.  Arithmetic is kept in
// the reference's operation order (left-associative products, sequential
// accumulation) so the GPU FP64 path can be compared to near machine precision.
//
// Data layout (same as the reference's column-major Julia arrays, site fastest):
//   action.gamma = [comp 0..nc-1][channel s,t,u][ q1 | q2_1 | q2_2 | q3 ]
//     q1  [site, w]          -> site + ns*w
//     q2_x[site, w, v]       -> site + ns*(w + nO*v)
//     q3  [site, w, v, vp]   -> site + ns*(w + nO*(v + nn*vp))
//   action.bare  = [comp][site];  action.Sigma = [n_sigma]
// Lattice tables are 0-based here (the reference's are 1-based).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "pffrg_math.h"  // exp/expm1 with the device's exact bits (pffrgsolver.jl_amd/csrc)
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct {
    int32_t n_sites, n_sigma, n_Omega, n_nu;
    double S;
    const double *sigma, *Os, *ns, *Ot, *nt;  // meshes, lengths n_sigma / n_Omega / n_nu
    const int64_t *exchange;                   // [n_sites], 0-based
    const double *mult;                        // [n_sites]
    const int64_t *ov_off;                     // [n_sites+1] row offsets into ov_rows
    const int64_t *ov_rows;                    // [3*rows] (j1, j2 0-based, multiplicity)
    int32_t n_comp;                            // vertex components: 2 (SU(2)) or 6 (U(1)-DM)
} orc_system;

typedef struct {
    double *Sigma;  // [n_sigma]
    double *bare;   // [n_comp*n_sites]
    double *gamma;  // packed channels, see header
} orc_action;

}  // extern "C"

namespace {

const double PI = 3.14159265358979323846;

// ---------------------------------------------------------------- Param (param.jl)
struct Param {
    int lo, hi;
    double wl, wh;
};

// get_indices + get_param, Frequency/param.jl:26-86 (linear search kept as in the reference)
inline Param get_param(double val, const double *list, int n) {
    Param p{0, 0, 0.0, 0.0};
    if (val >= list[n - 1]) {
        p.lo = p.hi = n - 1;
    } else {
        int idx = 0;
        while (val > list[idx]) ++idx;
        p.lo = p.hi = idx;
        if (val < list[idx]) p.lo -= 1;
    }
    if (p.lo < p.hi) {
        double d = 1.0 / (list[p.hi] - list[p.lo]);
        p.wl = d * (list[p.hi] - val);
        p.wh = d * (val - list[p.lo]);
    } else {
        p.wl = 1.0;
    }
    return p;
}

// ---------------------------------------------------------------- Buffer (buffer.jl)
struct Buffer {
    bool exch, map, sgn_mu, sgn_nu;
    int kernel;
    Param p1, p2, p3;
    bool sgn_munu;  // read by the U(1)-DM flags only
};

inline Buffer buffer_empty() {
    return Buffer{false, false, false, false, 0, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, false};
}

// get_buffer, Frequency/buffer.jl:51-86
inline Buffer get_buffer(double w, double v, double vp, const double *Om, const double *nu, int nO, int nn,
                         bool exch, bool map, bool smu, bool snu, bool smunu = false) {
    Buffer b = buffer_empty();
    if (Om[nO - 1] < std::fabs(w)) return b;
    b.exch = exch; b.map = map; b.sgn_mu = smu; b.sgn_nu = snu; b.sgn_munu = smunu;
    if (nu[nn - 1] < std::fabs(v)) {
        if (nu[nn - 1] < std::fabs(vp)) {
            b.kernel = 1; b.p1 = get_param(w, Om, nO);
        } else {
            b.kernel = 3; b.p1 = get_param(w, Om, nO); b.p3 = get_param(vp, nu, nn);
        }
    } else {
        if (nu[nn - 1] < std::fabs(vp)) {
            b.kernel = 2; b.p1 = get_param(w, Om, nO); b.p2 = get_param(v, nu, nn);
        } else {
            b.kernel = 4; b.p1 = get_param(w, Om, nO); b.p2 = get_param(v, nu, nn); b.p3 = get_param(vp, nu, nn);
        }
    }
    return b;
}

// get_buffer_s, Frequency/buffer.jl:88-124
inline Buffer buffer_s(const orc_system &y, double w, double v, double vp) {
    bool ex = false, mp = false, smu = false, snu = false;
    if (w < 0.0) { w *= -1.0; ex = !ex; }
    if (v < 0.0) { v *= -1.0; ex = !ex; mp = !mp; smu = !smu; }
    if (vp < 0.0) { vp *= -1.0; mp = !mp; snu = !snu; }
    return get_buffer(w, v, vp, y.Os, y.ns, y.n_Omega, y.n_nu, ex, mp, smu, snu);
}

// get_buffer_t, Frequency/buffer.jl:127-160 (sgn_munu is read by the U(1)-DM flags only)
inline Buffer buffer_t(const orc_system &y, double w, double v, double vp) {
    bool smu = false, snu = false, smunu = false;
    if (w < 0.0) { w *= -1.0; smunu = !smunu; }
    if (v < 0.0) { v *= -1.0; smu = !smu; }
    if (vp < 0.0) { vp *= -1.0; snu = !snu; }
    return get_buffer(w, v, vp, y.Ot, y.nt, y.n_Omega, y.n_nu, false, false, smu, snu, smunu);
}

// get_buffer_u, Frequency/buffer.jl:163-199
inline Buffer buffer_u(const orc_system &y, double w, double v, double vp) {
    bool ex = false, mp = false, snu = false, smunu = false;
    if (w < 0.0) { w *= -1.0; ex = !ex; smunu = !smunu; }
    if (v < 0.0) { v *= -1.0; ex = !ex; mp = !mp; snu = !snu; }
    if (vp < 0.0) { vp *= -1.0; mp = !mp; snu = !snu; }
    return get_buffer(w, v, vp, y.Os, y.ns, y.n_Omega, y.n_nu, ex, mp, false, snu, smunu);
}

// ---------------------------------------------------------------- Channel access (channel.jl)
struct Dims {
    int ns, nO, nn, nc;
    long q1, q21, q22, q3, ch, comp;
};

inline Dims dims_of(const orc_system &y) {
    Dims d;
    d.ns = y.n_sites; d.nO = y.n_Omega; d.nn = y.n_nu; d.nc = y.n_comp;
    d.q1 = 0;
    d.q21 = (long)d.ns * d.nO;
    d.q22 = d.q21 + (long)d.ns * d.nO * d.nn;
    d.q3 = d.q22 + (long)d.ns * d.nO * d.nn;
    d.ch = d.q3 + (long)d.ns * d.nO * d.nn * d.nn;
    d.comp = 3 * d.ch;
    return d;
}

inline const double *chan(const Dims &d, const double *gamma, int comp, int ch) { return gamma + comp * d.comp + ch * d.ch; }
inline double *chan(const Dims &d, double *gamma, int comp, int ch) { return gamma + comp * d.comp + ch * d.ch; }

// get_q1 / get_q2_1 / get_q2_2 / get_q3 and get_channel, Action/channel.jl:42-327
inline double get_channel(const Dims &d, const double *c, const Buffer &b, int site) {
    const int ns = d.ns, nO = d.nO, nn = d.nn;
    const Param &p1 = b.p1, &p2 = b.p2, &p3 = b.p3;
    if (b.kernel == 1) {
        const double *q = c + d.q1;
        double val = p1.wl * q[site + ns * p1.lo];
        val += p1.wh * q[site + ns * p1.hi];
        return val;
    } else if (b.kernel == 2 || b.kernel == 3) {
        const double *q = c + (b.kernel == 2 ? d.q21 : d.q22);
        const Param &pv = (b.kernel == 2) ? p2 : p3;
        auto at = [&](int w, int v) { return q[site + (long)ns * (w + (long)nO * v)]; };
        double val = p1.wl * pv.wl * at(p1.lo, pv.lo);
        val += p1.wh * pv.wl * at(p1.hi, pv.lo);
        val += p1.wl * pv.wh * at(p1.lo, pv.hi);
        val += p1.wh * pv.wh * at(p1.hi, pv.hi);
        return val;
    } else if (b.kernel == 4) {
        const double *q = c + d.q3;
        auto at = [&](int w, int v, int vp) { return q[site + (long)ns * (w + (long)nO * (v + (long)nn * vp))]; };
        double val = p1.wl * p2.wl * p3.wl * at(p1.lo, p2.lo, p3.lo);
        val += p1.wh * p2.wl * p3.wl * at(p1.hi, p2.lo, p3.lo);
        val += p1.wl * p2.wh * p3.wl * at(p1.lo, p2.hi, p3.lo);
        val += p1.wh * p2.wh * p3.wl * at(p1.hi, p2.hi, p3.lo);
        val += p1.wl * p2.wl * p3.wh * at(p1.lo, p2.lo, p3.hi);
        val += p1.wh * p2.wl * p3.wh * at(p1.hi, p2.lo, p3.hi);
        val += p1.wl * p2.wh * p3.wh * at(p1.lo, p2.hi, p3.hi);
        val += p1.wh * p2.wh * p3.wh * at(p1.hi, p2.hi, p3.hi);
        return val;
    }
    return 0.0;
}

// apply_flags_su2 (Action/action_lib/action_su2.jl:73-92) and apply_flags_u1_dm
// (action_u1_dm.jl:79-118): the sign of a channel's contribution to component comp and the
// component src it is read from (U(1)-DM: the site exchange swaps zd and dz and flips DM)
inline double flag_apply(int nc, const Buffer &b, int comp, int &src) {
    double sgn = 1.0;
    src = comp;
    if (nc == 2) {
        if (comp == 1 && b.sgn_mu) sgn *= -1.0;
        if (comp == 1 && b.sgn_nu) sgn *= -1.0;
        return sgn;
    }
    if ((comp == 4 || comp == 5) && b.sgn_munu) sgn *= -1.0;
    if ((comp == 3 || comp == 5) && b.sgn_mu) sgn *= -1.0;
    if ((comp == 3 || comp == 4) && b.sgn_nu) sgn *= -1.0;
    if (b.exch) {
        if (comp == 2) sgn *= -1.0;
        else if (comp == 4) src = 5;
        else if (comp == 5) src = 4;
    }
    return sgn;
}

struct Mask {
    bool s, t, u;
};
const Mask ALL{true, true, true};

// get_Γ_comp, Action/Action.jl:29-107 (scalar, one site)
inline double gamma_comp(const orc_system &y, const Dims &d, const orc_action &a, int comp, int site, const Buffer &bs,
                         const Buffer &bt, const Buffer &bu, Mask m) {
    double val = a.bare[comp * d.ns + site];
    int src;
    if (m.s) {
        int s = bs.exch ? (int)y.exchange[site] : site;
        double sgn = flag_apply(d.nc, bs, comp, src);
        val += sgn * get_channel(d, chan(d, a.gamma, src, bs.map ? 2 : 0), bs, s);
    }
    if (m.t) {
        int s = bt.exch ? (int)y.exchange[site] : site;
        double sgn = flag_apply(d.nc, bt, comp, src);
        val += sgn * get_channel(d, chan(d, a.gamma, src, 1), bt, s);
    }
    if (m.u) {
        int s = bu.exch ? (int)y.exchange[site] : site;
        double sgn = flag_apply(d.nc, bu, comp, src);
        val += sgn * get_channel(d, chan(d, a.gamma, src, bu.map ? 0 : 2), bu, s);
    }
    return val;
}

// get_Γ_comp_avx! + get_Γ_avx!, Action/Action.jl:110-167, action_su2.jl:115-134 and
// action_u1_dm.jl:141-167.  temp[comp*ns + i]; the bare value is never permuted, the channels are.
inline void gamma_vec(const orc_system &y, const Dims &d, const orc_action &a, const Buffer &bs, const Buffer &bt,
                      const Buffer &bu, Mask m, double *temp) {
    for (int comp = 0; comp < d.nc; ++comp) {
        double *t = temp + comp * d.ns;
        int src;
        for (int i = 0; i < d.ns; ++i) t[i] = a.bare[comp * d.ns + i];
        if (m.s && bs.kernel) {
            double sgn = flag_apply(d.nc, bs, comp, src);
            const double *c = chan(d, a.gamma, src, bs.map ? 2 : 0);
            for (int i = 0; i < d.ns; ++i) t[i] += sgn * get_channel(d, c, bs, bs.exch ? (int)y.exchange[i] : i);
        }
        if (m.t && bt.kernel) {
            double sgn = flag_apply(d.nc, bt, comp, src);
            const double *c = chan(d, a.gamma, src, 1);
            for (int i = 0; i < d.ns; ++i) t[i] += sgn * get_channel(d, c, bt, bt.exch ? (int)y.exchange[i] : i);
        }
        if (m.u && bu.kernel) {
            double sgn = flag_apply(d.nc, bu, comp, src);
            const double *c = chan(d, a.gamma, src, bu.map ? 0 : 2);
            for (int i = 0; i < d.ns; ++i) t[i] += sgn * get_channel(d, c, bu, bu.exch ? (int)y.exchange[i] : i);
        }
    }
}

// ---------------------------------------------------------------- self energy + propagators
// get_Σ, Action/Action.jl:8-26
inline double get_sigma(const orc_system &y, const double *Sig, double w) {
    const int n = y.n_sigma;
    if (std::fabs(w) <= y.sigma[n - 1]) {
        Param p = get_param(std::fabs(w), y.sigma, n);
        double sg = (w > 0.0) ? 1.0 : ((w < 0.0) ? -1.0 : w);  // Julia sign(): sign(±0.0) = ±0.0
        return sg * (p.wl * Sig[p.lo] + p.wh * Sig[p.hi]);
    }
    return y.sigma[n - 1] * Sig[n - 1] / w;
}

// get_G_bare, Flow/bubbles.jl:2-14
inline double G_bare(double L, double w) {
    double val = 0.0;
    if (std::fabs(w) > 1e-8) val = -pffrg::pf_expm1(-(w * w) / (L * L)) / w;
    return val;
}

// get_G, Flow/bubbles.jl:17-42
inline double G_full(const orc_system &y, double L, double w, const double *Sig) {
    double val = 0.0;
    if (std::fabs(w) > 1e-8) {
        double s = get_sigma(y, Sig, w);
        double G0 = G_bare(L, w);
        double denom;
        if (w * s >= 0.0) denom = 1.0 / G0 + s;
        else denom = 1.0 / G0;
        val = 1.0 / denom;
    }
    return val;
}

// get_S, Flow/bubbles.jl:45-61
inline double S_single(const orc_system &y, double L, double w, const double *Sig) {
    double val = 0.0;
    if (std::fabs(w) > 1e-8) {
        double G = G_full(y, L, w, Sig);
        double G0 = G_bare(L, w);
        double r = G / G0;
        val = r * r * pffrg::pf_exp(-(w * w) / (L * L)) * 2.0 * w / (L * L * L);
    }
    return val;
}

// get_propagator_kat, Flow/bubbles.jl:64-80
inline double prop_kat(const orc_system &y, double L, double w1, double w2, const double *Sig, const double *dSig) {
    double dS = get_sigma(y, dSig, w1);
    double S1 = S_single(y, L, w1, Sig);
    double G1 = G_full(y, L, w1, Sig);
    double G2 = G_full(y, L, w2, Sig);
    return (S1 + G1 * G1 * dS) * G2 / (2.0 * PI);
}

// get_propagator, Flow/bubbles.jl:83-96
inline double prop(const orc_system &y, double L, double w1, double w2, const double *Sig) {
    double G1 = G_full(y, L, w1, Sig);
    double G2 = G_full(y, L, w2, Sig);
    return G1 * G2 / (2.0 * PI);
}

// ---------------------------------------------------------------- adaptive Gauss-Kronrod (QuadGK.jl 2.x)
// QuadGK.jl is an un-vendored dependency ("QuadGK" compat "2.4", Project.toml).  Its
// published algorithm is restated: order-7 Gauss / 15-point Kronrod rule, infinite
// endpoints mapped by x = t/(1-t^2) (both infinite) or x = s0 -/+ t/(1-t) (one
// infinite), initial segments at the given breakpoints, then repeated bisection of
// the segment with the largest error estimate until E <= atol or E <= rtol*|I|.
// Segments are kept in insertion order (the popped segment is replaced by its left
// half, the right half appended) and re-summed in that order at the end; the same
// policy is used by the HIP library.  A segment cap bounds memory on both sides.
const double GK_X[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                        0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                        0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                        0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
const double GK_WK[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                         0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                         0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                         0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
const double GK_WG[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                         0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
const int GK_MAX_SEGS = 512;

struct Seg {
    double a, b, I, E;
};

template <class F>
Seg gk_rule(const F &f, double a, double b) {
    // QuadGK evalrule for odd order (n = 7): Gauss nodes are the even Kronrod nodes
    double s = 0.5 * (b - a);
    double Ig = 0.0, Ik = 0.0;
    for (int i = 0; i < 3; ++i) {
        double xg = GK_X[2 * i + 1], xk = GK_X[2 * i];
        double fg = f(a + (1.0 - xg) * s) + f(a + (1.0 + xg) * s);
        double fk = f(a + (1.0 - xk) * s) + f(a + (1.0 + xk) * s);
        Ig += fg * GK_WG[i];
        Ik += fg * GK_WK[2 * i + 1] + fk * GK_WK[2 * i];
    }
    double f0 = f(a + s);
    Ig += f0 * GK_WG[3];
    Ik += f0 * GK_WK[7] + (f(a + (1.0 - GK_X[6]) * s) + f(a + (1.0 + GK_X[6]) * s)) * GK_WK[6];
    double Iks = Ik * s, Igs = Ig * s;
    return Seg{a, b, Iks, std::fabs(Iks - Igs)};
}

template <class F>
double gk_adapt(const F &f, const double *bp, int nbp, double atol, double rtol) {
    std::vector<Seg> segs;
    segs.reserve(GK_MAX_SEGS);
    double I = 0.0, E = 0.0;
    for (int i = 0; i + 1 < nbp; ++i) segs.push_back(gk_rule(f, bp[i], bp[i + 1]));
    I = segs[0].I; E = segs[0].E;
    for (size_t i = 1; i < segs.size(); ++i) { I += segs[i].I; E += segs[i].E; }
    if (E <= atol || E <= rtol * std::fabs(I)) return I;
    while (E > atol && E > rtol * std::fabs(I) && (int)segs.size() < GK_MAX_SEGS) {
        int k = 0;
        for (int i = 1; i < (int)segs.size(); ++i)
            if (segs[i].E > segs[k].E) k = i;
        Seg s = segs[k];
        double mid = 0.5 * (s.a + s.b);
        Seg s1 = gk_rule(f, s.a, mid), s2 = gk_rule(f, mid, s.b);
        I = (I - s.I) + s1.I + s2.I;
        E = (E - s.E) + s1.E + s2.E;
        segs[k] = s1;
        segs.push_back(s2);
    }
    I = segs[0].I;
    for (size_t i = 1; i < segs.size(); ++i) I += segs[i].I;
    return I;
}

// quadgk over (-Inf, x1, ..., Inf): both endpoints infinite
template <class F>
double quadgk_inf_inf(const F &f, const double *inner, int ninner, double atol, double rtol) {
    std::vector<double> bp;
    bp.push_back(-1.0);
    for (int i = 0; i < ninner; ++i) {
        double x = inner[i];
        bp.push_back(2.0 * x / (1.0 + std::hypot(1.0, 2.0 * x)));
    }
    bp.push_back(1.0);
    auto g = [&](double t) {
        double t2 = t * t;
        double den = 1.0 / (1.0 - t2);
        return f(t * den) * (1.0 + t2) * den * den;
    };
    return gk_adapt(g, bp.data(), (int)bp.size(), atol, rtol);
}

// quadgk over (s0, Inf) (upper infinite): x = s0 + t/(1-t), t in [0,1]
template <class F>
double quadgk_to_inf(const F &f, double s0, double atol, double rtol) {
    double bp[2] = {0.0, 1.0};
    auto g = [&](double t) {
        double den = 1.0 / (1.0 - t);
        return f(s0 + t * den) * den * den;
    };
    return gk_adapt(g, bp, 2, atol, rtol);
}

// quadgk over (-Inf, s0) (lower infinite): x = s0 - t/(1-t), t in [0,1]
template <class F>
double quadgk_from_inf(const F &f, double s0, double atol, double rtol) {
    double bp[2] = {0.0, 1.0};
    auto g = [&](double t) {
        double den = 1.0 / (1.0 - t);
        return f(s0 - t * den) * den * den;
    };
    return gk_adapt(g, bp, 2, atol, rtol);
}

// ---------------------------------------------------------------- adaptive Simpson (quadrature.jl)
// buffers are (2 x ns) matrices stored comp-fastest: b[comp + 2*site]
inline double norm2(const double *b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += b[i] * b[i];
    return std::sqrt(s);
}

// simps!, Flow/quadrature.jl:27-80 (residual! at :2-24 inlined)
template <class F>
void simps(const F &f, double *b1, double *b2, int n, double a, double b, double atol, double rtol, long *nodes) {
    for (int i = 0; i < n; ++i) b1[i] = 0.0;
    double m = 0.5 * (a + b);
    f(b1, a, 1.0 * (b - a) / 6.0);
    f(b1, m, 2.0 * (b - a) / 3.0);
    f(b1, b, 1.0 * (b - a) / 6.0);
    *nodes += 3;
    for (int i = 0; i < n; ++i) b2[i] = 0.5 * b1[i];
    auto residual = [&](long nn) {
        double h = 0.5 * (b - a) / nn;
        for (long i = 1; i <= nn; ++i) f(b2, a + (2.0 * i - 1.0) * h, +4.0 * h / 3.0);
        for (long i = 1; i <= nn / 2; ++i) f(b2, a + (4.0 * i - 2.0) * h, -2.0 * h / 3.0);
        *nodes += nn + nn / 2;
    };
    residual(2);
    long nn = 4;
    while (true) {
        for (int i = 0; i < n; ++i) b1[i] *= -1.0 / 15.0;
        for (int i = 0; i < n; ++i) b1[i] += 16.0 / 15.0 * b2[i];
        double n1 = norm2(b1, n), n2 = norm2(b2, n);
        for (int i = 0; i < n; ++i) b1[i] -= b2[i];
        double adiff = norm2(b1, n);
        double rdiff = adiff / std::max(n1, n2);
        if (adiff < atol || rdiff < rtol || nn > (1L << 20)) {
            for (int i = 0; i < n; ++i) b1[i] += b2[i];
            break;
        }
        for (int i = 0; i < n; ++i) b1[i] = b2[i];
        for (int i = 0; i < n; ++i) b2[i] = 0.5 * b1[i];
        residual(nn);
        nn *= 2;
    }
}

// integrate_lin!, Flow/quadrature.jl:84-108
thread_local long *g_rec_nodes = nullptr;  // optional per-sub-interval recording (diagnostics)
thread_local double *g_rec_res = nullptr;

template <class F>
void integrate_lin(const F &f, double *t1, double *t2, double *t3, int n, double a, double b, int eval, double atol,
                   double rtol, long *nodes) {
    if (b > a) {
        double h = (b - a) / eval;
        for (int i = 1; i <= eval; ++i) {
            long before = *nodes;
            simps(f, t2, t3, n, a + (i - 1) * h, a + i * h, atol, rtol, nodes);
            if (g_rec_nodes) { *g_rec_nodes++ = *nodes - before; *g_rec_res++ = t2[0]; }
            for (int k = 0; k < n; ++k) t1[k] += t2[k];
        }
    } else if (g_rec_nodes) {
        for (int i = 0; i < eval; ++i) { *g_rec_nodes++ = 0; *g_rec_res++ = 0.0; }
    }
}

// integrate_log!, Flow/quadrature.jl:113-139 (sgn maps [a,b] to [-b,-a])
template <class F>
void integrate_log(const F &f, double *t1, double *t2, double *t3, int n, double a, double b, int eval, double atol,
                   double rtol, double sgn, long *nodes) {
    if (b > a) {
        double xi = std::pow(b / a, 1.0 / eval);
        auto g = [&](double *buf, double x, double dx) { f(buf, sgn * x, dx); };
        for (int i = 1; i <= eval; ++i) {
            long before = *nodes;
            simps(g, t2, t3, n, std::pow(xi, i - 1) * a, std::pow(xi, i) * a, atol, rtol, nodes);
            if (g_rec_nodes) { *g_rec_nodes++ = *nodes - before; *g_rec_res++ = t2[0]; }
            for (int k = 0; k < n; ++k) t1[k] += t2[k];
        }
    } else if (g_rec_nodes) {
        for (int i = 0; i < eval; ++i) { *g_rec_nodes++ = 0; *g_rec_res++ = 0.0; }
    }
}

// ---------------------------------------------------------------- pass types
enum PassType { P_KAT = 0, P_LEFT = 1, P_CENTRAL = 2, P_BSE = 3, P_REDUCED = 4 };

struct PassCtx {
    const orc_system *y;
    Dims d;
    double L;
    int pass;
    const orc_action *a;   // propagator action and "full" vertex
    const orc_action *da;  // kat: dΣ source; left: left vertex; central: right vertex (da_l)
    const double *corrs;   // [side][ch][nO], side-fastest as in corrs[2,3,nΩ]
};

inline double corr_at(const PassCtx &c, int side, int ch, int w1) { return c.corrs[side + 2 * (ch + 3 * w1)]; }

// per-channel bosonic pair (w1, w2) for the bubble: s (flow_su2_s.jl:18), t (flow_su2_t.jl:18), u (flow_su2_u.jl:18)
inline void bubble_args(int ch, double W, double v, double &x1, double &x2) {
    if (ch == 0) { x1 = v + 0.5 * W; x2 = 0.5 * W - v; }
    else if (ch == 1) { x1 = v + 0.5 * W; x2 = v - 0.5 * W; }
    else { x1 = v - 0.5 * W; x2 = v + 0.5 * W; }
}

inline double bubble(const PassCtx &c, int ch, double W, double v) {
    double x1, x2;
    bubble_args(ch, W, v, x1, x2);
    const orc_system &y = *c.y;
    if (c.pass == P_KAT) return prop_kat(y, c.L, x1, x2, c.a->Sigma, c.da->Sigma) + prop_kat(y, c.L, x2, x1, c.a->Sigma, c.da->Sigma);
    return -prop(y, c.L, x1, x2, c.a->Sigma);
}

// Component products of a left vertex a and a right vertex b at one site, the brackets that
// the integrands multiply by -p.  SU(2): flow_su2_s.jl:38-39 and flow_su2_u.jl:38-39; U(1)-DM
// (components xx, zz, DM, dd, zd, dz): flow_u1_dm_s.jl:52-57 and flow_u1_dm_u.jl:52-57.
enum { XX = 0, ZZ = 1, DM = 2, DD = 3, ZD = 4, DZ = 5 };

inline void prod_s(int nc, const double *a, const double *b, double *g) {
    if (nc == 2) {
        g[0] = -2.0 * a[0] * b[0] + a[0] * b[1] + a[1] * b[0];
        g[1] = 3.0 * a[0] * b[0] + a[1] * b[1];
        return;
    }
    g[XX] = a[DM] * b[DZ] - a[DM] * b[ZD] + a[XX] * b[DD] + a[DD] * b[XX] - a[XX] * b[ZZ] - a[DZ] * b[DM] + a[ZD] * b[DM] - a[ZZ] * b[XX];
    g[ZZ] = a[ZZ] * b[DD] + a[DD] * b[ZZ] - a[DZ] * b[ZD] - 2.0 * a[XX] * b[XX] - 2.0 * a[DM] * b[DM] - a[ZD] * b[DZ];
    g[DM] = a[DD] * b[DM] + a[DM] * b[DD] - a[ZD] * b[XX] - a[XX] * b[DZ] - a[ZZ] * b[DM] + a[XX] * b[ZD] - a[DM] * b[ZZ] + a[DZ] * b[XX];
    g[DD] = -a[DZ] * b[DZ] + 2.0 * a[XX] * b[XX] + a[DD] * b[DD] + 2.0 * a[DM] * b[DM] - a[ZD] * b[ZD] + a[ZZ] * b[ZZ];
    g[ZD] = a[ZD] * b[DD] + a[DD] * b[ZD] - 2.0 * a[DM] * b[XX] + a[DZ] * b[ZZ] + a[ZZ] * b[DZ] + 2.0 * a[XX] * b[DM];
    g[DZ] = a[ZD] * b[ZZ] + a[DD] * b[DZ] + a[ZZ] * b[ZD] + 2.0 * a[DM] * b[XX] + a[DZ] * b[DD] - 2.0 * a[XX] * b[DM];
}

inline void prod_u(int nc, const double *a, const double *b, double *g) {
    if (nc == 2) {
        g[0] = 2.0 * a[0] * b[0] + a[0] * b[1] + a[1] * b[0];
        g[1] = 3.0 * a[0] * b[0] + a[1] * b[1];
        return;
    }
    g[XX] = -a[DM] * b[DZ] - a[DM] * b[ZD] + a[XX] * b[DD] + a[DD] * b[XX] + a[XX] * b[ZZ] + a[DZ] * b[DM] + a[ZD] * b[DM] + a[ZZ] * b[XX];
    g[ZZ] = a[ZZ] * b[DD] + a[DD] * b[ZZ] - a[DZ] * b[ZD] + 2.0 * a[XX] * b[XX] + 2.0 * a[DM] * b[DM] - a[ZD] * b[DZ];
    g[DM] = a[DD] * b[DM] + a[DM] * b[DD] - a[ZD] * b[XX] + a[XX] * b[DZ] + a[ZZ] * b[DM] + a[XX] * b[ZD] + a[DM] * b[ZZ] - a[DZ] * b[XX];
    g[DD] = -a[DZ] * b[DZ] + 2.0 * a[XX] * b[XX] + a[DD] * b[DD] + 2.0 * a[DM] * b[DM] - a[ZD] * b[ZD] + a[ZZ] * b[ZZ];
    g[ZD] = a[ZD] * b[DD] + a[DD] * b[ZD] - 2.0 * a[DM] * b[XX] + a[DZ] * b[ZZ] + a[ZZ] * b[DZ] + 2.0 * a[XX] * b[DM];
    g[DZ] = a[ZD] * b[ZZ] + a[DD] * b[DZ] + a[ZZ] * b[ZD] - 2.0 * a[DM] * b[XX] + a[DZ] * b[DD] + 2.0 * a[XX] * b[DM];
}

// t channel, local part: non-local left vertex a (site i) with the local right vertex c4, and
// the local left vertex c3 with the non-local right vertex b (flow_su2_t.jl:57-58,
// flow_u1_dm_t.jl:56-67)
inline void prod_t_local(int nc, const double *a, const double *c4, const double *c3, const double *b, double *g) {
    if (nc == 2) {
        g[0] = -1.0 * a[0] * c4[0] + a[0] * c4[1] - 1.0 * c3[0] * b[0] + c3[1] * b[0];
        g[1] = 3.0 * a[1] * c4[0] + a[1] * c4[1] + 3.0 * c3[0] * b[1] + c3[1] * b[1];
        return;
    }
    g[XX] = a[DM] * c4[DZ] - a[DM] * c4[ZD] + a[XX] * c4[DD] - a[XX] * c4[ZZ] + c3[DD] * b[XX] + c3[DZ] * b[DM] -
            c3[ZD] * b[DM] - c3[ZZ] * b[XX];
    g[ZZ] = a[ZZ] * c4[DD] + 2.0 * a[ZD] * c4[DM] - a[ZD] * c4[ZD] + a[ZZ] * c4[ZZ] - 2.0 * a[ZZ] * c4[XX] -
            a[ZD] * c4[DZ] - 2.0 * c3[DM] * b[DZ] - c3[DZ] * b[DZ] + c3[DD] * b[ZZ] - 2.0 * c3[XX] * b[ZZ] +
            c3[ZZ] * b[ZZ] - c3[ZD] * b[DZ];
    g[DM] = a[DM] * c4[DD] - a[XX] * c4[DZ] + a[XX] * c4[ZD] - a[DM] * c4[ZZ] - c3[DZ] * b[XX] + c3[DD] * b[DM] +
            c3[ZD] * b[XX] - c3[ZZ] * b[DM];
    g[DD] = -a[DZ] * c4[DZ] + 2.0 * a[DD] * c4[XX] + a[DD] * c4[ZZ] - a[DZ] * c4[ZD] - 2.0 * a[DZ] * c4[DM] +
            a[DD] * c4[DD] + 2.0 * c3[DM] * b[ZD] + 2.0 * c3[XX] * b[DD] + c3[ZZ] * b[DD] - c3[DZ] * b[ZD] +
            c3[DD] * b[DD] - c3[ZD] * b[ZD];
    g[ZD] = a[ZD] * c4[ZZ] + a[ZD] * c4[DD] + 2.0 * a[ZD] * c4[XX] + a[ZZ] * c4[ZD] + 2.0 * a[ZZ] * c4[DM] +
            a[ZZ] * c4[DZ] + c3[ZD] * b[DD] + 2.0 * c3[DM] * b[DD] + c3[DD] * b[ZD] + c3[ZZ] * b[ZD] -
            2.0 * c3[XX] * b[ZD] + c3[DZ] * b[DD];
    g[DZ] = -2.0 * a[DD] * c4[DM] + a[DD] * c4[ZD] + a[DD] * c4[DZ] + a[DZ] * c4[ZZ] + a[DZ] * c4[DD] -
            2.0 * a[DZ] * c4[XX] + c3[ZD] * b[ZZ] + 2.0 * c3[XX] * b[DZ] + c3[DD] * b[DZ] + c3[DZ] * b[ZZ] -
            2.0 * c3[DM] * b[ZZ] + c3[ZZ] * b[DZ];
}

// t channel, one overlap row (j1, j2, m): G += -p (-2) m [(2S)] (a[j1] . b[j2])
// (flow_su2_t.jl:70-71, flow_u1_dm_t.jl:88-93; the U(1)-DM sum carries no 2S)
inline void add_t_row(int nc, double p, double m, double S2, const double *a, const double *b, double *G) {
    if (nc == 2) {
        G[0] += -p * (-2.0) * m * S2 * a[0] * b[0];
        G[1] += -p * (-2.0) * m * S2 * a[1] * b[1];
        return;
    }
    G[XX] += -p * (-2.0) * m * (a[XX] * b[XX] - a[DM] * b[DM]);
    G[ZZ] += -p * (-2.0) * m * (a[ZZ] * b[ZZ] - a[ZD] * b[DZ]);
    G[DM] += -p * (-2.0) * m * (a[DM] * b[XX] + a[XX] * b[DM]);
    G[DD] += -p * (-2.0) * m * (a[DD] * b[DD] - a[DZ] * b[ZD]);
    G[ZD] += -p * (-2.0) * m * (a[ZD] * b[DD] + a[ZZ] * b[ZD]);
    G[DZ] += -p * (-2.0) * m * (a[DD] * b[DZ] + a[DZ] * b[ZZ]);
}

// Integrands: compute_{s,t,u}_{kat,left,central}! (Flow/flow_lib/flow_su2/flow_su2_{s,t,u}.jl,
// flow_lib/flow_u1_dm/flow_u1_dm_{s,t,u}.jl), compute_{s,t,u}_BSE! (Flow/parquet_lib/
// parquet_su2/BSE_su2_{s,t,u}.jl, parquet_u1_dm/BSE_u1_dm_{s,t,u}.jl) and compute_s_reduced!
// (SDE_su2.jl:2-43, SDE_u1_dm.jl:2-43).  The buffers and channel masks are the same for both
// actions; the component products differ.  buff[nc*i + comp].
struct Integrand {
    const PassCtx *c;
    int ch;
    double W, vw, vwp;  // element frequencies (Ω, ν, ν')
    mutable std::vector<double> t1, t2;  // vertex caches [comp*ns + i]

    void operator()(double *buff, double v, double dv) const {
        const PassCtx &C = *c;
        const orc_system &y = *C.y;
        const Dims &d = C.d;
        const int ns = d.ns, nc = d.nc;
        const double p = bubble(C, ch, W, v);
        const orc_action &A = *C.a;
        const orc_action &D = *C.da;
        double a[6], b[6], g[6];
        auto site = [&](const std::vector<double> &t, int i, double *x) {
            for (int k = 0; k < nc; ++k) x[k] = t[k * ns + i];
        };
        if (ch == 0 || ch == 2) {
            if (ch == 0) {
                const double s = W, vs = vw, vsp = vwp;
                Buffer bs2 = buffer_s(y, s, v, vsp);
                Buffer bt2 = buffer_t(y, -v - vsp, 0.5 * (s + v - vsp), 0.5 * (s - v + vsp));
                Buffer bu2 = buffer_u(y, v - vsp, 0.5 * (s + v + vsp), 0.5 * (s - v - vsp));
                if (C.pass == P_REDUCED) {
                    // the left vertex is the bare coupling; U(1)-DM keeps xx, zz, dd (SDE_u1_dm.jl:36-42)
                    gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data());
                    for (int i = 0; i < ns; ++i) {
                        for (int k = 0; k < nc; ++k) a[k] = A.bare[k * ns + i];
                        site(t2, i, b);
                        prod_s(nc, a, b, g);
                        for (int k = 0; k < nc; ++k)
                            if (nc == 2 || k == XX || k == ZZ || k == DD) buff[nc * i + k] += dv * (-p * g[k]);
                    }
                    return;
                }
                Buffer bs1 = buffer_s(y, s, vs, -v);
                Buffer bt1 = buffer_t(y, v - vs, 0.5 * (s + v + vs), 0.5 * (s - v - vs));
                Buffer bu1 = buffer_u(y, v + vs, 0.5 * (s - v + vs), 0.5 * (s + v - vs));
                switch (C.pass) {
                    case P_KAT: gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data()); break;
                    case P_LEFT: gamma_vec(y, d, D, bs1, bt1, bu1, Mask{false, true, true}, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data()); break;
                    case P_CENTRAL: gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, D, bs2, bt2, bu2, Mask{true, false, false}, t2.data()); break;
                    case P_BSE: gamma_vec(y, d, A, bs1, bt1, bu1, Mask{false, true, true}, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data()); break;
                }
            } else {
                const double u = W, vu = vw, vup = vwp;
                Buffer bs1 = buffer_s(y, v + vu, 0.5 * (u - v + vu), 0.5 * (-u - v + vu));
                Buffer bt1 = buffer_t(y, v - vu, 0.5 * (u + v + vu), 0.5 * (-u + v + vu));
                Buffer bu1 = buffer_u(y, u, vu, v);
                Buffer bs2 = buffer_s(y, v + vup, 0.5 * (u + v - vup), 0.5 * (-u + v - vup));
                Buffer bt2 = buffer_t(y, -v + vup, 0.5 * (u + v + vup), 0.5 * (-u + v + vup));
                Buffer bu2 = buffer_u(y, u, v, vup);
                switch (C.pass) {
                    case P_KAT: gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data()); break;
                    case P_LEFT: gamma_vec(y, d, D, bs1, bt1, bu1, Mask{true, true, false}, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data()); break;
                    case P_CENTRAL: gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, D, bs2, bt2, bu2, Mask{false, false, true}, t2.data()); break;
                    case P_BSE: gamma_vec(y, d, A, bs1, bt1, bu1, Mask{true, true, false}, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data()); break;
                }
            }
            for (int i = 0; i < ns; ++i) {
                site(t1, i, a);
                site(t2, i, b);
                if (ch == 0) prod_s(nc, a, b, g);
                else prod_u(nc, a, b, g);
                for (int k = 0; k < nc; ++k) buff[nc * i + k] += dv * (-p * g[k]);
            }
        } else {
            const double t = W, vt = vw, vtp = vwp;
            Buffer bs1 = buffer_s(y, v + vt, 0.5 * (-t - v + vt), 0.5 * (-t + v - vt));
            Buffer bt1 = buffer_t(y, t, vt, v);
            Buffer bu1 = buffer_u(y, -v + vt, 0.5 * (-t + v + vt), 0.5 * (t + v + vt));
            Buffer bs2 = buffer_s(y, v + vtp, 0.5 * (-t + v - vtp), 0.5 * (-t - v + vtp));
            Buffer bt2 = buffer_t(y, t, v, vtp);
            Buffer bu2 = buffer_u(y, v - vtp, 0.5 * (-t + v + vtp), 0.5 * (t + v + vtp));
            Buffer bs3 = buffer_s(y, v + vt, 0.5 * (t + v - vt), 0.5 * (-t + v - vt));
            Buffer bt3 = buffer_t(y, -v + vt, 0.5 * (t + v + vt), 0.5 * (-t + v + vt));
            Buffer bu3 = buffer_u(y, t, v, vt);
            Buffer bs4 = buffer_s(y, v + vtp, 0.5 * (t - v + vtp), 0.5 * (-t - v + vtp));
            Buffer bt4 = buffer_t(y, v - vtp, 0.5 * (t + v + vtp), 0.5 * (-t + v + vtp));
            Buffer bu4 = buffer_u(y, t, vtp, v);
            // local (site-0) vertices v3, v4 and the non-local caches of each pass
            double c3[6], c4[6];
            const orc_action *V3 = &A, *V4 = &A;
            Mask m3 = ALL, m4 = ALL;
            switch (C.pass) {
                case P_KAT:
                    gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data());
                    break;
                case P_LEFT:
                    V3 = &D; m3 = Mask{true, true, false};
                    gamma_vec(y, d, D, bs1, bt1, bu1, Mask{true, false, true}, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data());
                    break;
                case P_CENTRAL:
                    V4 = &D; m4 = Mask{false, false, true};
                    gamma_vec(y, d, A, bs1, bt1, bu1, ALL, t1.data()); gamma_vec(y, d, D, bs2, bt2, bu2, Mask{false, true, false}, t2.data());
                    break;
                default:  // P_BSE
                    m3 = Mask{true, true, false};
                    gamma_vec(y, d, A, bs1, bt1, bu1, Mask{true, false, true}, t1.data()); gamma_vec(y, d, A, bs2, bt2, bu2, ALL, t2.data());
                    break;
            }
            for (int k = 0; k < nc; ++k) {
                c3[k] = gamma_comp(y, d, *V3, k, 0, bs3, bt3, bu3, m3);
                c4[k] = gamma_comp(y, d, *V4, k, 0, bs4, bt4, bu4, m4);
            }
            const double S2 = 2.0 * y.S;
            double G[6], w1v[6], w2v[6];
            for (int i = 0; i < ns; ++i) {
                site(t1, i, a);
                site(t2, i, b);
                prod_t_local(nc, a, c4, c3, b, g);
                for (int k = 0; k < nc; ++k) G[k] = -p * g[k];
                for (long r = y.ov_off[i]; r < y.ov_off[i + 1]; ++r) {
                    const int j1 = (int)y.ov_rows[3 * r], j2 = (int)y.ov_rows[3 * r + 1];
                    const double m = (double)y.ov_rows[3 * r + 2];
                    site(t1, j1, w1v);
                    site(t2, j2, w2v);
                    add_t_row(nc, p, m, S2, w1v, w2v, G);
                }
                for (int k = 0; k < nc; ++k) buff[nc * i + k] += dv * G[k];
            }
        }
    }
};

// get_kernel_args, Frequency/mesh.jl:30-79 (u uses the s meshes); indices 0-based, -1 = "Inf"
inline void kernel_args(const orc_system &y, int ch, int kernel, int w1, int w2, int w3, double &W, double &v, double &vp) {
    const double *Om = (ch == 1) ? y.Ot : y.Os;
    const double *nu = (ch == 1) ? y.nt : y.ns;
    W = Om[w1];
    v = (kernel == 2 || kernel == 4) ? nu[w2] : INFINITY;
    vp = (kernel == 3 || kernel == 4) ? nu[w3] : INFINITY;
}

// compute_channel_{s,t,u}_{kat,left,central}! (Flow/flow_{s,t,u}.jl), compute_channel_*_BSE!
// (Flow/BSE_*.jl) and compute_channel_s_reduced! (Flow/SDE.jl:2-56): one output element.
long compute_element(const PassCtx &C, int ch, int kernel, int w1, int w2, int w3, orc_action &out, double *t1, double *t2,
                     double *t3, int eval, double atol, double rtol) {
    const orc_system &y = *C.y;
    const Dims &d = C.d;
    const int n = d.nc * d.ns;
    for (int i = 0; i < n; ++i) t1[i] = 0.0;
    double W, vw, vwp;
    if (C.pass == P_REDUCED) {
        // SDE.jl:13-19: kernel 1 -> (Ωs[w1], Inf); else (Ωs[w1], νs[w3])
        W = y.Os[w1]; vw = 0.0; vwp = (kernel == 1) ? INFINITY : y.ns[w3];
    } else {
        kernel_args(y, ch, kernel, w1, w2, w3, W, vw, vwp);
    }
    Integrand f{&C, ch, W, vw, vwp, std::vector<double>(n), std::vector<double>(n)};
    const double ref = C.L + 0.5 * W;
    const double val = (ch == 1) ? (y.Ot[d.nO - 1] + y.nt[d.nn - 1]) : (y.Os[d.nO - 1] + y.ns[d.nn - 1]);
    long nodes = 0;
    integrate_log(f, t1, t2, t3, n, 2.0 * ref, 1.0 * val, eval, atol, rtol, -1.0, &nodes);
    integrate_lin(f, t1, t2, t3, n, -2.0 * ref, 0.0 * ref, eval, atol, rtol, &nodes);
    integrate_lin(f, t1, t2, t3, n, 0.0 * ref, 2.0 * ref, eval, atol, rtol, &nodes);
    integrate_log(f, t1, t2, t3, n, 2.0 * ref, 1.0 * val, eval, atol, rtol, 1.0, &nodes);
    f(t1, -val, corr_at(C, 0, ch, w1));
    f(t1, val, corr_at(C, 1, ch, w1));
    nodes += 2;
    // parse result (flow_s.jl:41-51)
    for (int comp = 0; comp < d.nc; ++comp) {
        double *c = chan(d, out.gamma, comp, ch);
        long base;
        if (kernel == 1) base = d.q1 + (long)d.ns * w1;
        else if (kernel == 2) base = d.q21 + (long)d.ns * (w1 + (long)d.nO * w2);
        else if (kernel == 3) base = d.q22 + (long)d.ns * (w1 + (long)d.nO * w3);
        else base = d.q3 + (long)d.ns * (w1 + (long)d.nO * (w2 + (long)d.nn * w3));
        for (int i = 0; i < d.ns; ++i) c[base + i] = t1[comp + d.nc * i];
    }
    return nodes;
}

struct Elem {
    int ch, kernel, w1, w2, w3;
};

// element enumeration of compute_Γ!/compute_dΓ_*! (Flow/compute.jl:2-286): half passes
// have w2 >= w3 for q3, full (left) passes all (w2, w3)
std::vector<Elem> elements(const Dims &d, bool full) {
    std::vector<Elem> e;
    for (int w1 = 0; w1 < d.nO; ++w1) {
        for (int ch = 0; ch < 3; ++ch) e.push_back({ch, 1, w1, -1, -1});
        for (int w3 = 0; w3 < d.nn; ++w3) {
            for (int ch = 0; ch < 3; ++ch) e.push_back({ch, 2, w1, w3, -1});
            for (int ch = 0; ch < 3; ++ch) e.push_back({ch, 3, w1, -1, w3});
            for (int w2 = full ? 0 : w3; w2 < d.nn; ++w2)
                for (int ch = 0; ch < 3; ++ch) e.push_back({ch, 4, w1, w2, w3});
        }
    }
    return e;
}

long run_pass(const PassCtx &C, const std::vector<Elem> &E, orc_action &out, int eval, double atol, double rtol) {
    long total = 0;
    const int n = C.d.nc * C.d.ns;
#pragma omp parallel reduction(+ : total)
    {
        std::vector<double> t1(n), t2(n), t3(n);
#pragma omp for schedule(dynamic, 1)
        for (long k = 0; k < (long)E.size(); ++k) {
            const Elem &e = E[k];
            total += compute_element(C, e.ch, e.kernel, e.w1, e.w2, e.w3, out, t1.data(), t2.data(), t3.data(), eval, atol, rtol);
        }
    }
    return total;
}

// The ν <-> ν' exchange of component comp in channel ch: the right part reads component
// sym_src with sign sym_sign (action_su2.jl:141-250; action_u1_dm.jl:170-339: DM changes sign
// and zd, dz swap with a sign in the s and t channels, zd and dz change sign in the u channel)
inline int sym_src(int nc, int ch, int comp) {
    if (nc == 6 && ch != 2 && comp == ZD) return DZ;
    if (nc == 6 && ch != 2 && comp == DZ) return ZD;
    return comp;
}
inline double sym_sign(int nc, int ch, int comp) {
    if (nc != 6) return 1.0;
    if (ch != 2) return (comp == DM || comp == ZD || comp == DZ) ? -1.0 : 1.0;
    return (comp == ZD || comp == DZ) ? -1.0 : 1.0;
}

// symmetrize!, Action/action_lib/action_su2.jl:141-173 and action_u1_dm.jl:170-215
void symmetrize(const orc_system &y, const Dims &d, orc_action &a) {
    // the upper triangle (v < v') is written from lower-triangle entries only, also where the
    // U(1)-DM s and t channels read the other component of the pair (zd, dz)
    for (int comp = 0; comp < d.nc; ++comp)
        for (int ch = 0; ch < 3; ++ch) {
            double *q3 = chan(d, a.gamma, comp, ch) + d.q3;
            const double sg = sym_sign(d.nc, ch, comp);
            const double *src = chan(d, a.gamma, sym_src(d.nc, ch, comp), ch) + d.q3;
            for (int v = 0; v < d.nn; ++v)
                for (int vp = v + 1; vp < d.nn; ++vp)
                    for (int w = 0; w < d.nO; ++w)
                        for (int i = 0; i < d.ns; ++i) {
                            int s = (ch == 2) ? i : (int)y.exchange[i];
                            const double x = src[s + (long)d.ns * (w + (long)d.nO * (vp + (long)d.nn * v))];
                            q3[i + (long)d.ns * (w + (long)d.nO * (v + (long)d.nn * vp))] = (d.nc == 2) ? x : sg * x;
                        }
        }
}

// symmetrize_add_to!, Action/action_lib/action_su2.jl:176-250 and action_u1_dm.jl:218-339
void symmetrize_add_to(const orc_system &y, const Dims &d, const orc_action &al, orc_action &a) {
    for (int comp = 0; comp < d.nc; ++comp)
        for (int ch = 0; ch < 3; ++ch) {
            const double *l = chan(d, al.gamma, comp, ch);
            const double *r = chan(d, al.gamma, sym_src(d.nc, ch, comp), ch);
            const double sg = sym_sign(d.nc, ch, comp);
            double *o = chan(d, a.gamma, comp, ch);
            auto ex = [&](int i) { return (ch == 2) ? i : (int)y.exchange[i]; };
            auto pair = [&](double x, double xr) { return (d.nc == 2) ? x + xr : (sg > 0.0 ? x + xr : x - xr); };
            for (int w = 0; w < d.nO; ++w)
                for (int i = 0; i < d.ns; ++i)
                    o[d.q1 + i + (long)d.ns * w] += pair(l[d.q1 + i + (long)d.ns * w], r[d.q1 + ex(i) + (long)d.ns * w]);
            for (int v = 0; v < d.nn; ++v)
                for (int w = 0; w < d.nO; ++w)
                    for (int i = 0; i < d.ns; ++i) {
                        long k = i + (long)d.ns * (w + (long)d.nO * v), kx = ex(i) + (long)d.ns * (w + (long)d.nO * v);
                        double a21 = pair(l[d.q21 + k], r[d.q22 + kx]);
                        double a22 = pair(l[d.q22 + k], r[d.q21 + kx]);
                        o[d.q21 + k] += a21;
                        o[d.q22 + k] += a22;
                    }
            for (int vp = 0; vp < d.nn; ++vp)
                for (int v = 0; v < d.nn; ++v)
                    for (int w = 0; w < d.nO; ++w)
                        for (int i = 0; i < d.ns; ++i) {
                            long k = i + (long)d.ns * (w + (long)d.nO * (v + (long)d.nn * vp));
                            long kx = ex(i) + (long)d.ns * (w + (long)d.nO * (vp + (long)d.nn * v));
                            o[d.q3 + k] += pair(l[d.q3 + k], r[d.q3 + kx]);
                        }
        }
}

long gamma_size(const Dims &d) { return d.nc * d.comp; }

// compute_corrs_kat! / compute_corrs!, Flow/bubbles.jl:99-220
void compute_corrs(const orc_system &y, double L, const orc_action &a, const orc_action *da, bool kat, double *corrs,
                   double atol, double rtol) {
    const int nO = y.n_Omega, nn = y.n_nu;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nO; ++i) {
        for (int ch = 0; ch < 3; ++ch) {
            const double W = (ch == 1) ? y.Ot[i] : y.Os[i];
            const double val = (ch == 1) ? (y.Ot[nO - 1] + y.nt[nn - 1]) : (y.Os[nO - 1] + y.ns[nn - 1]);
            auto f = [&](double v) {
                double x1, x2;
                bubble_args(ch, W, v, x1, x2);
                if (kat) return prop_kat(y, L, x1, x2, a.Sigma, da->Sigma) + prop_kat(y, L, x2, x1, a.Sigma, da->Sigma);
                return -prop(y, L, x1, x2, a.Sigma);
            };
            double bm = f(-val), bp = f(val);
            corrs[0 + 2 * (ch + 3 * i)] = 0.0;
            corrs[1 + 2 * (ch + 3 * i)] = 0.0;
            if (std::fabs(bm) > 1e-8) corrs[0 + 2 * (ch + 3 * i)] = quadgk_from_inf(f, -val, atol, rtol) / bm;
            if (std::fabs(bp) > 1e-8) corrs[1 + 2 * (ch + 3 * i)] = quadgk_to_inf(f, val, atol, rtol) / bp;
        }
    }
}

// vertex buffers shared by the self-energy kernels, Flow/flow_lib/flow_su2/flow_su2_self.jl:11-19

// ---------------------------------------------------------------- adaptive 2-D h-cubature
// Restatement of the h-adaptive cubature of Cubature.jl 1.5 (a wrapper of S. G. Johnson's C
// library "cubature", not part of the reference sources) as called by integrate_χ_boxes
// (correlation_lib.jl:2-30): hcubature_v with maxevals = 10^7.  Each region is integrated by
// the degree-7 Genz-Malik rule with its embedded degree-5 rule as the error estimate (A. C.
// Genz, A. A. Malik, J. Comput. Appl. Math. 6 (1980) 295-302) and bisected along its axis of
// largest fourth difference.  hcubature_v runs the library's vectorised schedule (I. Gladwell's,
// as described by J. M. Bull and T. L. Freeman, 1994): every round splits at once the
// largest-error regions that must be refined for the error of the others to meet
// max(atol, rtol |I|), until the total error meets it or the evaluation budget is spent.
// Fixed here, and shared with the GPU kernel so both produce the same regions and sums:
//   - the selection works on quarter-octave bins of the error below the largest one, with
//     the bin sums in fixed point, so a round splits every region of the bins it needs;
//   - split regions keep their index, the upper halves are appended in index order;
//   - totals are summed in the order of a 256-thread block (thread t adds indices t, t+256, ...,
//     then p[t] += p[t+s] for s = 128, 64, ..., 1).
struct Region {
    double c[2], h[2], val, err;
    int split;
};

const double GM_L2 = 0.35856858280031806;  // sqrt(9/70)
const double GM_L4 = 0.9486832980505138;   // sqrt(9/10)
const double GM_L5 = 0.6882472016116853;   // sqrt(9/19)
const double GM_W1 = -3816.0 / 19683.0, GM_W2 = 980.0 / 6561.0, GM_W3 = 1020.0 / 19683.0, GM_W4 = 200.0 / 19683.0,
             GM_W5 = 6859.0 / 19683.0 / 4.0;
const double GM_E1 = -971.0 / 729.0, GM_E2 = 245.0 / 486.0, GM_E3 = 65.0 / 1458.0, GM_E4 = 25.0 / 729.0;

// the 17 rule points of a region in the fixed order f0 | ±λ2 x, ±λ2 y | ±λ4 x, ±λ4 y |
// (±λ4, ±λ4) | (±λ5, ±λ5)
inline void gm_point(const Region &R, int k, double &x, double &y) {
    static const signed char sx[17] = {0, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
    static const signed char sy[17] = {0, 0, 0, -1, 1, 0, 0, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1};
    const double l = (k < 5) ? GM_L2 : (k < 13) ? GM_L4 : GM_L5;
    x = R.c[0] + sx[k] * (l * R.h[0]);
    y = R.c[1] + sy[k] * (l * R.h[1]);
}

// rule values from the 17 function values (same arithmetic on the GPU)
inline void gm_combine(Region &R, const double *f) {
    const double f0 = f[0];
    const double s2 = ((f[1] + f[2]) + f[3]) + f[4];
    const double s3 = ((f[5] + f[6]) + f[7]) + f[8];
    const double s4 = ((f[9] + f[10]) + f[11]) + f[12];
    const double s5 = ((f[13] + f[14]) + f[15]) + f[16];
    const double vol = (2.0 * R.h[0]) * (2.0 * R.h[1]);
    const double r7 = vol * ((((GM_W1 * f0 + GM_W2 * s2) + GM_W3 * s3) + GM_W4 * s4) + GM_W5 * s5);
    const double r5 = vol * (((GM_E1 * f0 + GM_E2 * s2) + GM_E3 * s3) + GM_E4 * s4);
    R.val = r7;
    R.err = std::fabs(r5 - r7);
    const double d0 = std::fabs(((f[1] + f[2]) - 2.0 * f0) - (1.0 / 7.0) * ((f[5] + f[6]) - 2.0 * f0));
    const double d1 = std::fabs(((f[3] + f[4]) - 2.0 * f0) - (1.0 / 7.0) * ((f[7] + f[8]) - 2.0 * f0));
    R.split = (d1 > d0 || (d1 == d0 && R.h[1] > R.h[0])) ? 1 : 0;
}

inline bool cub_converged(double val, double err, double atol, double rtol) {
    return err <= atol || err <= std::fabs(val) * rtol;
}

const long CUB_MAX_EVALS = 10000000;                          // maxevals of integrate_χ_boxes
const long CUB_MAX_REGIONS = 1 + (CUB_MAX_EVALS - 17) / 34;   // regions that budget allows
const int CUB_T = 256;

// selection bins of a round: errors are binned by quarter octaves below the largest one (bin 0
// holds the largest, bin 255 everything from 2^-63 of it down) and summed per bin in 44-bit fixed
// point relative to the largest (integer sums, so the order of accumulation does not matter)
const int CUB_BINS = 256;

inline uint64_t dbits(double e) {
    uint64_t b;
    std::memcpy(&b, &e, 8);
    return b;
}

inline int err_exp(double e) { return (int)((dbits(e) >> 52) & 0x7FF); }

inline int err_bin(double e, int E) {
    const int d = (4 * E + 3) - (int)((dbits(e) >> 50) & 0x1FFF);
    return d < 0 ? 0 : (d > CUB_BINS - 1 ? CUB_BINS - 1 : d);
}

inline uint64_t err_fixed(double e, int E) { return (uint64_t)std::ldexp(e, 1066 - (E > 1 ? E : 1)); }

// Σ_j x[j] over j < n in the GPU block's order
template <class G>
double block_sum(long n, const G &x) {
    double p[CUB_T];
    for (int t = 0; t < CUB_T; ++t) {
        double a = 0.0;
        for (long j = t; j < n; j += CUB_T) a += x(j);
        p[t] = a;
    }
    for (int s = CUB_T / 2; s >= 1; s >>= 1)
        for (int t = 0; t < s; ++t) p[t] += p[t + s];
    return p[0];
}

// returns the integral over [lo, hi]; *nreg receives the final region count
template <class F>
double hcubature2(const F &f, const double *lo, const double *hi, double atol, double rtol, int *nreg = nullptr) {
    std::vector<Region> R;
    auto eval = [&](Region &r) {
        double fv[17];
        for (int k = 0; k < 17; ++k) {
            double x, y;
            gm_point(r, k, x, y);
            fv[k] = f(x, y);
        }
        gm_combine(r, fv);
    };
    Region r0;
    r0.c[0] = 0.5 * (lo[0] + hi[0]);
    r0.c[1] = 0.5 * (lo[1] + hi[1]);
    r0.h[0] = 0.5 * (hi[0] - lo[0]);
    r0.h[1] = 0.5 * (hi[1] - lo[1]);
    eval(r0);
    R.push_back(r0);
    double tv = r0.val;
    std::vector<long> sel;
    while (true) {
        const long n = (long)R.size();
        tv = block_sum(n, [&](long j) { return R[j].val; });
        const double te = block_sum(n, [&](long j) { return R[j].err; });
        if (cub_converged(tv, te, atol, rtol) || n >= CUB_MAX_REGIONS) break;
        // the fewest leading bins whose regions carry enough error for the rest to converge
        int E = 0;
        for (long j = 0; j < n; ++j) E = std::max(E, err_exp(R[j].err));
        uint64_t H[CUB_BINS] = {0};
        long N[CUB_BINS] = {0};
        for (long j = 0; j < n; ++j) {
            H[err_bin(R[j].err, E)] += err_fixed(R[j].err, E);
            N[err_bin(R[j].err, E)] += 1;
        }
        uint64_t Q = 0;
        for (int b = 0; b < CUB_BINS; ++b) Q += H[b];
        // from the first occupied bin on, so that a round splits at least one region
        int b0 = 0;
        while (b0 < CUB_BINS - 1 && N[b0] == 0) ++b0;
        int bstar = CUB_BINS - 1;
        uint64_t Cs = 0;
        for (int b = 0; b < b0; ++b) Cs += H[b];
        for (int b = b0; b < CUB_BINS; ++b) {
            Cs += H[b];
            const double rest = std::ldexp((double)(Q - Cs), (E > 1 ? E : 1) - 1066);
            if (cub_converged(tv, rest, atol, rtol)) {
                bstar = b;
                break;
            }
        }
        sel.clear();
        const long room = CUB_MAX_REGIONS - n;
        for (long j = 0; j < n && (long)sel.size() < room; ++j)
            if (err_bin(R[j].err, E) <= bstar) sel.push_back(j);
        R.resize(n + sel.size());
        for (size_t q = 0; q < sel.size(); ++q) {
            const Region r = R[sel[q]];
            Region a = r, b = r;
            const int d = r.split;
            a.h[d] = 0.5 * r.h[d];
            b.h[d] = a.h[d];
            a.c[d] = r.c[d] - a.h[d];
            b.c[d] = r.c[d] + a.h[d];
            eval(a);
            eval(b);
            R[sel[q]] = a;
            R[n + q] = b;
        }
    }
    if (nreg) *nreg = (int)R.size();
    return tv;
}

// ---------------------------------------------------------------- correlations
// compute_χ_kernel! (correlation_su2.jl:14-55) at one point (x, y) of the unit square, and the
// U(1)-DM kernels compute_χ_kernel_{xx,zz,xy}! (correlation_u1_dm.jl:16-134) for comp 0, 1, 2
inline double chi_kernel(const orc_system &y, const Dims &d, double L, int site, double w, const orc_action &a,
                         double xx, double yy, int comp = 0) {
    const double v = (2.0 * xx - 1.0) / ((1.0 - xx) * xx);
    const double vp = (2.0 * yy - 1.0) / ((1.0 - yy) * yy);
    const double dv = (2.0 * xx * xx - 2.0 * xx + 1) / ((1.0 - xx) * (1.0 - xx) * xx * xx);
    const double dvp = (2.0 * yy * yy - 2.0 * yy + 1) / ((1.0 - yy) * (1.0 - yy) * yy * yy);
    const Buffer bs1 = buffer_s(y, v + vp, 0.5 * (v - w - vp), 0.5 * (-v - w + vp));
    const Buffer bt1 = buffer_t(y, w, v, vp);
    const Buffer bu1 = buffer_u(y, v - vp, 0.5 * (v - w + vp), 0.5 * (v + w + vp));
    const double tp2 = (2.0 * PI) * (2.0 * PI);
    double val;
    if (d.nc == 2) {
        const double S2 = 2.0 * y.S;
        val = -(S2 * S2) * gamma_comp(y, d, a, 0, site, bs1, bt1, bu1, ALL) / tp2;
        if (site == 0) {
            const Buffer bs2 = buffer_s(y, v + vp, 0.5 * (v - w - vp), 0.5 * (v + w - vp));
            const Buffer bt2 = buffer_t(y, -v + vp, 0.5 * (v - w + vp), 0.5 * (v + w + vp));
            const Buffer bu2 = buffer_u(y, -w, v, vp);
            const double vs = gamma_comp(y, d, a, 0, site, bs2, bt2, bu2, ALL);
            const double vd = gamma_comp(y, d, a, 1, site, bs2, bt2, bu2, ALL);
            val -= S2 * (vs - vd) / (2.0 * tp2);
        }
    } else {
        val = -gamma_comp(y, d, a, comp == 2 ? DM : comp, site, bs1, bt1, bu1, ALL) / tp2;
        if (site == 0) {
            const Buffer bs2 = buffer_s(y, v + vp, 0.5 * (v - w - vp), 0.5 * (v + w - vp));
            const Buffer bt2 = buffer_t(y, -v + vp, 0.5 * (v - w + vp), 0.5 * (v + w + vp));
            const Buffer bu2 = buffer_u(y, -w, v, vp);
            if (comp == 0) {
                const double vzz = gamma_comp(y, d, a, ZZ, site, bs2, bt2, bu2, ALL);
                const double vdd = gamma_comp(y, d, a, DD, site, bs2, bt2, bu2, ALL);
                val -= (vzz - vdd) / (2.0 * tp2);
            } else if (comp == 1) {
                const double vxx = gamma_comp(y, d, a, XX, site, bs2, bt2, bu2, ALL);
                const double vzz = gamma_comp(y, d, a, ZZ, site, bs2, bt2, bu2, ALL);
                const double vdd = gamma_comp(y, d, a, DD, site, bs2, bt2, bu2, ALL);
                val -= (2.0 * vxx - vzz - vdd) / (2.0 * tp2);
            } else {
                const double vzd = gamma_comp(y, d, a, ZD, site, bs2, bt2, bu2, ALL);
                const double vdz = gamma_comp(y, d, a, DZ, site, bs2, bt2, bu2, ALL);
                val -= (vzd - vdz) / (2.0 * tp2);
            }
        }
    }
    val *= G_full(y, L, v - 0.5 * w, a.Sigma) * G_full(y, L, v + 0.5 * w, a.Sigma) * dv;
    val *= G_full(y, L, vp - 0.5 * w, a.Sigma) * G_full(y, L, vp + 0.5 * w, a.Sigma) * dvp;
    return val;
}

// the nine boxes of integrate_χ_boxes (correlation_lib.jl:2-30), box k = [lo, hi]
inline void chi_box(double width, int k, double *lo, double *hi) {
    const double widthp = (width - 2.0 + std::sqrt(width * width + 4.0)) / (2.0 * width);
    const double size = std::fabs(widthp - 0.5);
    const double a = 0.5 - size, b = 0.5 + size;
    const double xl[3] = {0.0, a, b}, xh[3] = {a, b, 1.0};
    const double yl[3] = {b, a, 0.0}, yh[3] = {1.0, b, a};
    lo[0] = xl[k % 3];
    hi[0] = xh[k % 3];
    lo[1] = yl[k / 3];
    hi[1] = yh[k / 3];
}

struct SigmaBufs {
    Buffer b1s, b1t, b1u, b2s, b2t, b2u;
};
inline SigmaBufs sigma_bufs(const orc_system &y, double w, double v) {
    SigmaBufs b;
    b.b1s = buffer_s(y, v + w, 0.5 * (-v + w), 0.5 * (v - w));
    b.b1t = buffer_t(y, 0.0, w, v);
    b.b1u = buffer_u(y, -v + w, 0.5 * (v + w), 0.5 * (v + w));
    b.b2s = buffer_s(y, v + w, 0.5 * (-v + w), 0.5 * (-v + w));
    b.b2t = buffer_t(y, v - w, 0.5 * (v + w), 0.5 * (v + w));
    b.b2u = buffer_u(y, 0.0, w, v);
    return b;
}

// loop value of the self-energy kernels from the local (b2) and non-local (b1) buffers:
// SU(2) 3Γs + Γd - Σ_j 2 m_j (2S) Γd_j (flow_su2_self.jl:22-27); U(1)-DM 2Γxx + Γzz + Γdd -
// Σ_j 2 m_j Γdd_j (flow_u1_dm_self.jl:22-28)
inline double sigma_loop2(const orc_system &y, const Dims &d, const orc_action &a, const Buffer &b2s, const Buffer &b2t,
                          const Buffer &b2u, const Buffer &b1s, const Buffer &b1t, const Buffer &b1u, Mask ml, Mask mn) {
    if (d.nc == 2) {
        double val = 3.0 * gamma_comp(y, d, a, 0, 0, b2s, b2t, b2u, ml) + gamma_comp(y, d, a, 1, 0, b2s, b2t, b2u, ml);
        for (int j = 0; j < d.ns; ++j) val -= 2.0 * y.mult[j] * (2.0 * y.S) * gamma_comp(y, d, a, 1, j, b1s, b1t, b1u, mn);
        return val;
    }
    double val = 2.0 * gamma_comp(y, d, a, XX, 0, b2s, b2t, b2u, ml) + gamma_comp(y, d, a, ZZ, 0, b2s, b2t, b2u, ml) +
                 gamma_comp(y, d, a, DD, 0, b2s, b2t, b2u, ml);
    for (int j = 0; j < d.ns; ++j) val -= 2.0 * y.mult[j] * gamma_comp(y, d, a, DD, j, b1s, b1t, b1u, mn);
    return val;
}
inline double sigma_loop(const orc_system &y, const Dims &d, const orc_action &a, const SigmaBufs &b, Mask ml, Mask mn) {
    return sigma_loop2(y, d, a, b.b2s, b.b2t, b.b2u, b.b1s, b.b1t, b.b1u, ml, mn);
}

const double SIGMA_INNER_BP_SCALE[3] = {-2.0, 0.0, 2.0};

}  // namespace

// ======================================================================== C API
extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// get_param (param.jl:65-86): returns lo, hi (0-based) and weights
void orc_get_param(double val, const double *list, int n, int32_t *lo, int32_t *hi, double *wl, double *wh) {
    Param p = get_param(val, list, n);
    *lo = p.lo; *hi = p.hi; *wl = p.wl; *wh = p.wh;
}

// get_buffer_{s,t,u} (buffer.jl:88-199): out_i = [exch, map, sgn_mu, sgn_nu, kernel, lo1, hi1, lo2, hi2, lo3, hi3, sgn_munu]
void orc_get_buffer(const orc_system *y, int ch, double w, double v, double vp, int32_t *out_i, double *out_w) {
    Buffer b = (ch == 0) ? buffer_s(*y, w, v, vp) : (ch == 1) ? buffer_t(*y, w, v, vp) : buffer_u(*y, w, v, vp);
    out_i[0] = b.exch; out_i[1] = b.map; out_i[2] = b.sgn_mu; out_i[3] = b.sgn_nu; out_i[4] = b.kernel;
    out_i[5] = b.p1.lo; out_i[6] = b.p1.hi; out_i[7] = b.p2.lo; out_i[8] = b.p2.hi; out_i[9] = b.p3.lo; out_i[10] = b.p3.hi;
    out_i[11] = b.sgn_munu;
    out_w[0] = b.p1.wl; out_w[1] = b.p1.wh; out_w[2] = b.p2.wl; out_w[3] = b.p2.wh; out_w[4] = b.p3.wl; out_w[5] = b.p3.wh;
}

// vertex component on all sites for a buffer triple built from (ch, w, v, vp) frequencies of the
// three channels: get_Γ_avx! with every channel enabled. freqs = [ws, vs, vps, wt, vt, vpt, wu, vu, vpu]
void orc_get_gamma_avx(const orc_system *y, const orc_action *a, const double *freqs, int32_t mask, double *out) {
    Dims d = dims_of(*y);
    Buffer bs = buffer_s(*y, freqs[0], freqs[1], freqs[2]);
    Buffer bt = buffer_t(*y, freqs[3], freqs[4], freqs[5]);
    Buffer bu = buffer_u(*y, freqs[6], freqs[7], freqs[8]);
    gamma_vec(*y, d, *a, bs, bt, bu, Mask{(mask & 1) != 0, (mask & 2) != 0, (mask & 4) != 0}, out);
}

// vertex value of one channel (vertex.jl get_vertex) at one site, no flags applied
double orc_get_vertex(const orc_system *y, const orc_action *a, int comp, int ch, double w, double v, double vp, int site) {
    Dims d = dims_of(*y);
    Buffer b = (ch == 0) ? buffer_s(*y, w, v, vp) : (ch == 1) ? buffer_t(*y, w, v, vp) : buffer_u(*y, w, v, vp);
    return get_channel(d, chan(d, a->gamma, comp, ch), b, site);
}

double orc_get_sigma(const orc_system *y, const double *Sig, double w) { return get_sigma(*y, Sig, w); }
double orc_G_bare(double L, double w) { return G_bare(L, w); }
double orc_G(const orc_system *y, double L, double w, const double *Sig) { return G_full(*y, L, w, Sig); }
double orc_S(const orc_system *y, double L, double w, const double *Sig) { return S_single(*y, L, w, Sig); }
double orc_prop_kat(const orc_system *y, double L, double w1, double w2, const double *Sig, const double *dSig) {
    return prop_kat(*y, L, w1, w2, Sig, dSig);
}
double orc_prop(const orc_system *y, double L, double w1, double w2, const double *Sig) { return prop(*y, L, w1, w2, Sig); }

// Flow/test.jl test kernel: b[i] += dx * x^i exp(-x^2) sin(2πx), i = 1..n, integrated by
// integrate_lin! (mode 0) or integrate_log! (mode 1) with the given eval and tolerances.
void orc_test_integrate(int mode, int n, double a, double b, int eval, double atol, double rtol, double *out) {
    std::vector<double> t2(n), t3(n);
    for (int i = 0; i < n; ++i) out[i] = 0.0;
    auto f = [&](double *buf, double x, double dx) {
        for (int i = 0; i < n; ++i) buf[i] += dx * std::pow(x, i + 1) * std::exp(-x * x) * std::sin(2.0 * PI * x);
    };
    long nodes = 0;
    if (mode == 0) integrate_lin(f, out, t2.data(), t3.data(), n, a, b, eval, atol, rtol, &nodes);
    else integrate_log(f, out, t2.data(), t3.data(), n, a, b, eval, atol, rtol, 1.0, &nodes);
}

// quadgk restatement on closed-form test integrands: kind 0 = exp(-x^2) over (-Inf, bp.., Inf);
// kind 1 = 1/(1+x^2)^2 over (s0, Inf); kind 2 = exp(x) over (-Inf, s0)
double orc_test_quadgk(int kind, double s0, double atol, double rtol) {
    if (kind == 0) {
        double inner[3] = {-2.0 * s0, 0.0, 2.0 * s0};
        return quadgk_inf_inf([](double x) { return std::exp(-x * x); }, inner, 3, atol, rtol);
    } else if (kind == 1) {
        return quadgk_to_inf([](double x) { return 1.0 / ((1.0 + x * x) * (1.0 + x * x)); }, s0, atol, rtol);
    }
    return quadgk_from_inf([](double x) { return std::exp(x); }, s0, atol, rtol);
}

void orc_compute_corrs(const orc_system *y, double L, const orc_action *a, const orc_action *da, int kat, double *corrs,
                       const double *tol) {
    compute_corrs(*y, L, *a, da, kat != 0, corrs, tol[0], tol[1]);
}

// one output element of any pass (for fine-grained parity checks); returns the node count
long orc_compute_element(const orc_system *y, int pass, double L, const orc_action *a, const orc_action *da,
                         const double *corrs, int ch, int kernel, int w1, int w2, int w3, int eval, const double *tol,
                         orc_action *out) {
    PassCtx C{y, dims_of(*y), L, pass, a, da, corrs};
    const int n = y->n_comp * y->n_sites;
    std::vector<double> t1(n), t2(n), t3(n);
    return compute_element(C, ch, kernel, w1, w2, w3, *out, t1.data(), t2.data(), t3.data(), eval, tol[0], tol[1]);
}

// one element with per-sub-interval node counts and spin results of site 0 recorded (4*eval each)
long orc_compute_element_rec(const orc_system *y, int pass, double L, const orc_action *a, const orc_action *da,
                             const double *corrs, int ch, int kernel, int w1, int w2, int w3, int eval,
                             const double *tol, orc_action *out, long *rec_nodes, double *rec_res) {
    g_rec_nodes = rec_nodes;
    g_rec_res = rec_res;
    long n = orc_compute_element(y, pass, L, a, da, corrs, ch, kernel, w1, w2, w3, eval, tol, out);
    g_rec_nodes = nullptr;
    g_rec_res = nullptr;
    return n;
}

// integrand of one element at loop frequency v (dv = 1): out[comp*ns + i]
void orc_integrand(const orc_system *y, int pass, double L, const orc_action *a, const orc_action *da, int ch,
                   int kernel, int w1, int w2, int w3, double v, double *out) {
    PassCtx C{y, dims_of(*y), L, pass, a, da, nullptr};
    double W, vw, vwp;
    if (pass == P_REDUCED) {
        W = y->Os[w1]; vw = 0.0; vwp = (kernel == 1) ? INFINITY : y->ns[w3];
    } else {
        kernel_args(*y, ch, kernel, w1, w2, w3, W, vw, vwp);
    }
    const int nc = y->n_comp, n = nc * y->n_sites;
    Integrand f{&C, ch, W, vw, vwp, std::vector<double>(n), std::vector<double>(n)};
    std::vector<double> b(n, 0.0);
    f(b.data(), v, 1.0);
    for (int i = 0; i < y->n_sites; ++i)
        for (int k = 0; k < nc; ++k) out[k * y->n_sites + i] = b[nc * i + k];
}

// a whole pass without symmetrisation: pass in {KAT, LEFT, CENTRAL, BSE, REDUCED}; full = all (w2,w3)
long orc_run_pass(const orc_system *y, int pass, double L, const orc_action *a, const orc_action *da, const double *corrs,
                  int full, int eval, const double *tol, orc_action *out) {
    Dims d = dims_of(*y);
    PassCtx C{y, d, L, pass, a, da, corrs};
    std::vector<Elem> E;
    if (pass == P_REDUCED) {
        for (int w1 = 0; w1 < d.nO; ++w1) {
            E.push_back({0, 1, w1, -1, -1});
            for (int w3 = 0; w3 < d.nn; ++w3) E.push_back({0, 3, w1, -1, w3});
        }
    } else {
        E = elements(d, full != 0);
    }
    return run_pass(C, E, *out, eval, tol[0], tol[1]);
}

void orc_symmetrize(const orc_system *y, orc_action *a) { symmetrize(*y, dims_of(*y), *a); }
void orc_symmetrize_add_to(const orc_system *y, const orc_action *al, orc_action *a) {
    symmetrize_add_to(*y, dims_of(*y), *al, *a);
}

// compute_dΣ!, Flow/flow_self.jl:2-20 with compute_dΣ_kernel (flow_su2_self.jl:2-33)
void orc_compute_dsigma(const orc_system *y, double L, const orc_action *a, orc_action *da, const double *tol) {
    Dims d = dims_of(*y);
    double inner[3] = {-2.0 * L, 0.0, 2.0 * L};
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 1; i < y->n_sigma; ++i) {
        const double w = y->sigma[i];
        auto f = [&](double v) {
            SigmaBufs b = sigma_bufs(*y, w, v);
            double val = sigma_loop(*y, d, *a, b, ALL, ALL);
            val *= S_single(*y, L, v, a->Sigma) / (2.0 * PI);
            return val;
        };
        da->Sigma[i] = quadgk_inf_inf(f, inner, 3, tol[0], tol[1]);
    }
}

// compute_dΣ_corr!, Flow/flow_self.jl:23-50 with the kernels at flow_su2_self.jl:36-101
void orc_compute_dsigma_corr(const orc_system *y, double L, const orc_action *a, orc_action *da, orc_action *daS,
                             const double *tol) {
    Dims d = dims_of(*y);
    double inner[3] = {-2.0 * L, 0.0, 2.0 * L};
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 1; i < y->n_sigma; ++i) {
        const double w = y->sigma[i];
        auto f = [&](double v) {
            SigmaBufs b = sigma_bufs(*y, w, v);
            double val = sigma_loop(*y, d, *daS, b, Mask{true, true, false}, Mask{true, false, true});
            val *= -G_full(*y, L, v, a->Sigma) / (2.0 * PI);
            return val;
        };
        daS->Sigma[i] = quadgk_inf_inf(f, inner, 3, tol[0], tol[1]);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 1; i < y->n_sigma; ++i) {
        const double w = y->sigma[i];
        auto f = [&](double v) {
            SigmaBufs b = sigma_bufs(*y, w, v);
            double val = sigma_loop(*y, d, *a, b, ALL, ALL);
            double G = G_full(*y, L, v, a->Sigma);
            val *= G * G * get_sigma(*y, daS->Sigma, v) / (2.0 * PI);
            return val;
        };
        da->Sigma[i] += daS->Sigma[i];
        da->Sigma[i] += quadgk_inf_inf(f, inner, 3, tol[0], tol[1]);
    }
}

// compute_dΓ_1l!, Flow/compute.jl:57-112; returns node count
long orc_compute_dgamma_1l(const orc_system *y, double L, const orc_action *a, orc_action *da, int eval,
                           const double *tol, double *corrs) {
    compute_corrs(*y, L, *a, da, true, corrs, tol[0], tol[1]);
    long n = orc_run_pass(y, P_KAT, L, a, da, corrs, 0, eval, tol, da);
    symmetrize(*y, dims_of(*y), *da);
    return n;
}

// compute_dΓ_2l!, Flow/compute.jl:115-174
long orc_compute_dgamma_2l(const orc_system *y, double L, const orc_action *a, orc_action *da, orc_action *dal, int eval,
                           const double *tol, double *corrs) {
    long n = orc_compute_dgamma_1l(y, L, a, da, eval, tol, corrs);
    compute_corrs(*y, L, *a, nullptr, false, corrs, tol[0], tol[1]);
    n += orc_run_pass(y, P_LEFT, L, a, da, corrs, 1, eval, tol, dal);
    symmetrize_add_to(*y, dims_of(*y), *dal, *da);
    return n;
}

// compute_dΓ_ml!, Flow/compute.jl:177-286
long orc_compute_dgamma_ml(const orc_system *y, double L, int loops, const orc_action *a, orc_action *da,
                           orc_action *dal, orc_action *dac, orc_action *datemp, orc_action *daS, int eval,
                           const double *tol, double *corrs) {
    Dims d = dims_of(*y);
    const long G = gamma_size(d);
    long n = orc_compute_dgamma_2l(y, L, a, da, dal, eval, tol, corrs);
    for (long k = 0; k < G; ++k) { datemp->gamma[k] *= 0.0; daS->gamma[k] *= 0.0; }
    symmetrize_add_to(*y, d, *dal, *datemp);
    for (int loop = 3; loop <= loops; ++loop) {
        n += orc_run_pass(y, P_CENTRAL, L, a, dal, corrs, 0, eval, tol, dac);
        n += orc_run_pass(y, P_LEFT, L, a, datemp, corrs, 1, eval, tol, dal);
        symmetrize(*y, d, *dac);
        for (long k = 0; k < G; ++k) datemp->gamma[k] = dac->gamma[k];
        symmetrize_add_to(*y, d, *dal, *datemp);
        for (long k = 0; k < G; ++k) daS->gamma[k] += 1.0 * dac->gamma[k];
        for (long k = 0; k < G; ++k) da->gamma[k] += 1.0 * datemp->gamma[k];
    }
    return n;
}

// compute_Σ!, Flow/SDE.jl:58-96 with compute_Σ_kernel (SDE_su2.jl:45-77, SDE_u1_dm.jl:45-91)
long orc_compute_sigma_sde(const orc_system *y, double L, const orc_action *a1, orc_action *a2, int eval,
                           const double *gtol, const double *stol, double *corrs) {
    Dims d = dims_of(*y);
    compute_corrs(*y, L, *a1, nullptr, false, corrs, gtol[0], gtol[1]);
    long n = orc_run_pass(y, P_REDUCED, L, a1, a1, corrs, 0, eval, gtol, a2);
    double inner[3] = {-2.0 * L, 0.0, 2.0 * L};
    const Mask SONLY{true, false, false};
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 1; i < y->n_sigma; ++i) {
        const double w = y->sigma[i];
        auto f = [&](double v) {
            Buffer b1s = buffer_s(*y, v + w, INFINITY, 0.5 * (v - w));
            Buffer b2s = buffer_s(*y, v + w, INFINITY, 0.5 * (-v + w));
            Buffer e = buffer_empty();
            double val = sigma_loop2(*y, d, *a2, b2s, e, e, b1s, e, e, SONLY, SONLY);
            val *= -G_full(*y, L, v, a1->Sigma) / (2.0 * PI);
            return val;
        };
        a2->Sigma[i] = quadgk_inf_inf(f, inner, 3, stol[0], stol[1]);
    }
    return n;
}

// compute_Γ!, Flow/compute.jl:2-54 (uses the corrs left by the preceding compute_Σ!)
long orc_compute_gamma_bse(const orc_system *y, double L, const orc_action *a1, orc_action *a2, int eval,
                           const double *tol, const double *corrs) {
    long n = orc_run_pass(y, P_BSE, L, a1, a1, corrs, 0, eval, tol, a2);
    symmetrize(*y, dims_of(*y), *a2);
    return n;
}

// compute_χ! (correlation_su2.jl:58-86): chi[i + ns*w] for the irreducible sites and the χ mesh
void orc_compute_chi(const orc_system *y, double L, const orc_action *a, const double *chi_mesh, int nchi,
                     const double *tol, double *chi, int32_t *nreg) {
    // U(1)-DM (compute_χ!, correlation_u1_dm.jl:138-190): χxx, χzz, χxy, each its own cubature;
    // chi[c*ns*nchi + i + ns*w], nreg[9*(c*ns*nchi + t) + k]
    Dims d = dims_of(*y);
    const int ns = d.ns;
    const int ncc = (d.nc == 2) ? 1 : 3;
#pragma omp parallel for schedule(dynamic, 1)
    for (int tc = 0; tc < ncc * nchi * ns; ++tc) {
        const int c = tc / (nchi * ns), t = tc % (nchi * ns);
        const int w = t / ns, i = t % ns;
        const double W = chi_mesh[w];
        const double ref = L + 0.5 * W;
        double I = 0.0;
        for (int k = 0; k < 9; ++k) {
            double lo[2], hi[2];
            chi_box(4.0 * ref, k, lo, hi);
            int nr = 0;
            I += hcubature2([&](double xx, double yy) { return chi_kernel(*y, d, L, i, W, *a, xx, yy, c); }, lo, hi,
                            tol[0], tol[1], &nr);
            if (nreg) nreg[9 * (long)tc + k] = nr;
        }
        if (i == 0 && c < 2) {
            const double inner[3] = {-2.0 * ref, 0.0, 2.0 * ref};
            const double S2 = (d.nc == 2) ? 2.0 * y->S : 1.0;
            auto f = [&](double v) {
                if (d.nc == 2)
                    return S2 * G_full(*y, L, v - 0.5 * W, a->Sigma) * G_full(*y, L, v + 0.5 * W, a->Sigma) / (4.0 * PI);
                return G_full(*y, L, v - 0.5 * W, a->Sigma) * G_full(*y, L, v + 0.5 * W, a->Sigma) / (4.0 * PI);
            };
            I += quadgk_inf_inf(f, inner, 3, tol[0], tol[1]);
        }
        chi[(long)c * ns * nchi + i + ns * w] = I;
    }
}

// 2-D cubature of a test integrand over [lo, hi]: kind 0 = x^p y^q, kind 1 = exp(-(x^2+y^2))
double orc_test_hcubature(int kind, int p, int q, const double *lo, const double *hi, double atol, double rtol,
                          int32_t *nreg) {
    int nr = 0;
    double I;
    if (kind == 0)
        I = hcubature2([&](double x, double y) { return std::pow(x, p) * std::pow(y, q); }, lo, hi, atol, rtol, &nr);
    else
        I = hcubature2([&](double x, double y) { return std::exp(-(x * x + y * y)); }, lo, hi, atol, rtol, &nr);
    if (nreg) *nreg = nr;
    return I;
}


// ======================================================= mesh rescanning (Action.jl:369-557)
// get_mesh, Frequency/mesh.jl:86-121 (returns false where the reference asserts)
static bool orc_mesh(double linear, double upper, int num, double p, std::vector<double> &m) {
    if (!(linear > 0.0) || !(upper > 0.0) || !(linear < upper)) return false;
    const int num_lin = (int)std::ceil(p * num);
    const int num_log = num - num_lin;
    if (num_log <= 0) return false;
    const double h = linear / num_lin;
    const double xi = std::pow(upper / linear, 1.0 / num_log);
    m.assign(num + 1, 0.0);
    for (int i = 1; i <= num_lin; ++i) m[i] = i * h;
    for (int i = 1; i <= num_log; ++i) m[num_lin + i] = std::pow(xi, (double)i) * linear;
    return true;
}

static double jsign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }
static double jmax(double a, double b) {
    if (std::isnan(a) || std::isnan(b)) return NAN;
    if (a == b) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}
static double jmin(double a, double b) {
    if (std::isnan(a) || std::isnan(b)) return NAN;
    if (a == b) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}

// scan, Action/Action.jl:369-446.  Returns 0, or -1 where the reference asserts.
int orc_scan(const double *x, const double *y, int n, const double *prm, double *delta) {
    const double p0 = prm[0], p1 = prm[1], p2 = prm[2], p3 = prm[3], p4 = prm[4];
    const int num_lin = (int)std::ceil(p0 * (n - 1));
    // a) (1-based idx of the reference = i + 1 here)
    double da = INFINITY;
    int idx = 0;
    double slope = (y[1] - y[0]) / (x[1] - x[0]);
    for (int i = 1; i <= n - 2; ++i) {
        if (jsign((y[i + 1] - y[i]) / (x[i + 1] - x[i])) != jsign(slope)) {
            idx = i;
            break;
        } else {
            slope = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        }
    }
    if (idx > 0) {
        for (int i = idx + 1; i <= n - 1; ++i) {
            if (std::fabs(y[i] / y[idx]) <= 0.75) {
                da = jmax(x[i], num_lin * x[idx] / std::ceil(0.5 * num_lin));
                break;
            }
        }
    }
    // b)
    double db = num_lin * x[1];
    double ymax = std::fabs(y[0]);
    for (int i = 1; i < n; ++i) ymax = jmax(ymax, std::fabs(y[i]));
    if (std::fabs(y[0]) / ymax > 1e-2) {
        double D = std::fabs(y[1] - y[0]) / jmax(std::fabs(y[1]), std::fabs(y[0]));
        std::vector<double> xp, yp(n);
        while ((p1 < D && D < p2) == false) {
            if (D >= p2) {
                db *= 0.98;
            } else if (D <= p1) {
                db *= 1.02;
            }
            if (db >= x[n - 1]) break;
            if (!orc_mesh(db, x[n - 1], n - 1, p0, xp)) return -1;
            for (int i = 0; i < n; ++i) {
                Param p = get_param(xp[i], x, n);
                yp[i] = p.wl * y[p.lo] + p.wh * y[p.hi];
            }
            D = std::fabs(yp[1] - yp[0]) / jmax(std::fabs(yp[1]), std::fabs(yp[0]));
        }
    }
    // c)
    *delta = jmin(jmax(num_lin * p3, jmin(da, db)), p4);
    return 0;
}

// scan_channel's argmax and cuts, Action/Action.jl:449-476: k = 3*comp + channel, k < 3*n_comp
void orc_scan_cuts(const orc_system *y, const orc_action *a, int32_t *idx, double *cuts) {
    const Dims d = dims_of(*y);
    const int ns = d.ns, nO = d.nO, nn = d.nn;
    const long n3 = (long)ns * nO * nn * nn;
    for (int k = 0; k < 3 * d.nc; ++k) {
        const double *q3 = chan(d, a->gamma, k / 3, k % 3) + d.q3;
        long best = 0;
        for (long j = 1; j < n3; ++j)
            if (std::fabs(q3[j]) > std::fabs(q3[best])) best = j;
        const int i = (int)(best % ns), w = (int)((best / ns) % nO), v = (int)((best / ((long)ns * nO)) % nn),
                  vp = (int)(best / ((long)ns * nO * nn));
        idx[4 * k] = i; idx[4 * k + 1] = w; idx[4 * k + 2] = v; idx[4 * k + 3] = vp;
        auto at = [&](int ww, int x1, int x2) { return q3[i + (long)ns * (ww + (long)nO * (x1 + (long)nn * x2))]; };
        double *o = cuts + (long)k * (nO + 3 * nn);
        for (int x = 0; x < nO; ++x) o[x] = at(x, v, vp);
        for (int x = 0; x < nn; ++x) o[nO + x] = at(w, x, x) - at(w, nn - 1, nn - 1);
        for (int x = 0; x < nn; ++x) o[nO + nn + x] = at(w, x, vp) - at(w, nn - 1, vp);
        for (int x = 0; x < nn; ++x) o[nO + 2 * nn + x] = at(w, v, x) - at(w, v, nn - 1);
    }
}

// resample_from_to!, Action/channel.jl:416-470, for one channel
static void resample_channel(const Dims &d, const double *Oo, const double *no, const double *c_old,
                             const double *On, const double *nnw, double *c_new) {
    const int ns = d.ns, nO = d.nO, nn = d.nn;
    for (int w = 0; w < nO; ++w) {
        Param p1 = get_param(On[w], Oo, nO);
        for (int s = 0; s < ns; ++s) {
            const double *q = c_old + d.q1;
            double val = p1.wl * q[s + ns * p1.lo];
            val += p1.wh * q[s + ns * p1.hi];
            c_new[d.q1 + s + ns * w] = val;
        }
    }
    for (int v = 0; v < nn; ++v)
        for (int w = 0; w < nO; ++w) {
            Param p1 = get_param(On[w], Oo, nO);
            Param p2 = get_param(nnw[v], no, nn);
            for (int s = 0; s < ns; ++s)
                for (int which = 0; which < 2; ++which) {
                    const long off = which == 0 ? d.q21 : d.q22;
                    const double *q = c_old + off;
                    auto at = [&](int a, int b) { return q[s + (long)ns * (a + (long)nO * b)]; };
                    double val = p1.wl * p2.wl * at(p1.lo, p2.lo);
                    val += p1.wh * p2.wl * at(p1.hi, p2.lo);
                    val += p1.wl * p2.wh * at(p1.lo, p2.hi);
                    val += p1.wh * p2.wh * at(p1.hi, p2.hi);
                    c_new[off + s + (long)ns * (w + (long)nO * v)] = val;
                }
        }
    for (int vp = 0; vp < nn; ++vp)
        for (int v = 0; v < nn; ++v)
            for (int w = 0; w < nO; ++w) {
                Param p1 = get_param(On[w], Oo, nO);
                Param p2 = get_param(nnw[v], no, nn);
                Param p3 = get_param(nnw[vp], no, nn);
                const double *q = c_old + d.q3;
                for (int s = 0; s < ns; ++s) {
                    const double *qs = q + s;
                    auto g = [&](int a, int b, int c) { return qs[(long)ns * (a + (long)nO * (b + (long)nn * c))]; };
                    double val = p1.wl * p2.wl * p3.wl * g(p1.lo, p2.lo, p3.lo);
                    val += p1.wh * p2.wl * p3.wl * g(p1.hi, p2.lo, p3.lo);
                    val += p1.wl * p2.wh * p3.wl * g(p1.lo, p2.hi, p3.lo);
                    val += p1.wh * p2.wh * p3.wl * g(p1.hi, p2.hi, p3.lo);
                    val += p1.wl * p2.wl * p3.wh * g(p1.lo, p2.lo, p3.hi);
                    val += p1.wh * p2.wl * p3.wh * g(p1.hi, p2.lo, p3.hi);
                    val += p1.wl * p2.wh * p3.wh * g(p1.lo, p2.hi, p3.hi);
                    val += p1.wh * p2.wh * p3.wh * g(p1.hi, p2.hi, p3.hi);
                    c_new[d.q3 + s + (long)ns * (w + (long)nO * (v + (long)nn * vp))] = val;
                }
            }
}

// resample_from_to, Action/Action.jl:543-553 (Σ by get_Σ) and vertex.jl:148-160 (channels):
// a_old on the meshes of y_old onto the new meshes into a_new; bare copied
void orc_resample(const orc_system *y_old, const double *sigma, const double *Os, const double *ns_,
                  const double *Ot, const double *nt, const orc_action *a_old, orc_action *a_new) {
    const Dims d = dims_of(*y_old);
    for (int w = 0; w < y_old->n_sigma; ++w) a_new->Sigma[w] = get_sigma(*y_old, a_old->Sigma, sigma[w]);
    for (int c = 0; c < d.nc; ++c) {
        resample_channel(d, y_old->Os, y_old->ns, chan(d, a_old->gamma, c, 0), Os, ns_, chan(d, a_new->gamma, c, 0));
        resample_channel(d, y_old->Ot, y_old->nt, chan(d, a_old->gamma, c, 1), Ot, nt, chan(d, a_new->gamma, c, 1));
        resample_channel(d, y_old->Os, y_old->ns, chan(d, a_old->gamma, c, 2), Os, ns_, chan(d, a_new->gamma, c, 2));
    }
    for (int k = 0; k < d.nc * d.ns; ++k) a_new->bare[k] = a_old->bare[k];
}

// the shared written-out exp / expm1 (pffrg_math.h), for their accuracy test
double orc_pf_expm1(double x) { return pffrg::pf_expm1(x); }
double orc_pf_exp(double x) { return pffrg::pf_exp(x); }

}  // extern "C"
