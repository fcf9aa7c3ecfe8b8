// pffrg_pass.h -- the vertex-pass kernel (one launch per pass over every output element).
// Instantiated per pass type in pass_*.hip so the translation units compile in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pffrg_device.h"

namespace pffrg {

#ifndef PFFRG_PASS_WAVES_PER_EU
#define PFFRG_PASS_WAVES_PER_EU 3
#endif

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

// Interpolation descriptor of one buffer (one channel of one vertex slot) at one node, built
// once per node by one lane and read by every site lane of the group.
template <typename TV>
struct DescT {
    int meta;            // kernel (bits 0-2) | exchange << 3 | map << 4 | negative density sign << 5 |
                         // upper corner one step up along Ω << 6, ν << 7, ν' << 8 (else the same point) |
                         // sgn_μ << 9 | sgn_ν << 10 | sgn_μν << 11 (the U(1)-DM flags)
    int off;             // offset of the first corner inside the channel block (site 0)
    int pad[sizeof(TV) == 4 ? 1 : 0];  // 44- or 72-byte stride (11 or 18 dwords): the descriptors of up to
                                       // 16 (FP32) or 8 (FP64) groups start in distinct LDS banks
    TV w[8];             // corner weights in the reference's corner order
};

// Offset steps to the upper corner along Ω, ν, ν' (0 where the bracket is a single point).
__device__ __forceinline__ void corner_steps(int meta, const Sys &y, uint32_t &c1, uint32_t &c2, uint32_t &c3) {
    const uint32_t s1 = (uint32_t)y.ns, s2 = s1 * (uint32_t)y.nO, s3 = s2 * (uint32_t)y.nn;
    c1 = (meta & 64) ? s1 : 0u;
    c2 = (meta & 128) ? s2 : 0u;
    c3 = (meta & 256) ? s3 : 0u;
}

// Scalar type of the per-node work (bubble legs, frequency transfer, brackets, weights).  FP64
// in both modes: an FP32 variant measured no faster, the pass being latency-bound, and FP64
// keeps the FP32-mixed mode's node placement and adaptive decisions those of the FP64 path.
template <typename TV>
struct PassScalar {
    using T = double;
};

template <typename TV>
struct Vec2;
template <>
struct Vec2<float> {
    using T = float2;
};
template <>
struct Vec2<double> {
    using T = double2;
};

constexpr int MAX_GROUPS = 16;  // site groups per wavefront
constexpr int MAX_DESC = 12;    // 4 vertex slots x 3 channels

// Meshes, inverse-spacing tables and the two Σ arrays in LDS, in the scalar type of the
// per-node work (FP64 path: the FP64 tables; FP32-mixed path: FP32 copies of them).
template <typename ST>
struct MeshV {
    const ST *sig, *Os, *ns, *Ot, *nt, *inv, *sigA, *sigD;
    int mp;
    const int16_t *hint;  // c0[5], nc[5], then uint8 tab[5][HINT_CELLS] (FP64 tables only; else null)
};

// Published state of one sub-interval owner (one per group), read by the groups that
// evaluate its pooled nodes.
struct Ctl {
    double a, b, h, sgn;
    int n, init, issued, pend;
};

struct Lds {
    const double *Os, *ns, *Ot, *nt, *sig, *sigA, *sigD;
    const int *exch, *ov_off;
    const uint32_t *ov;
    const double *inv;      // inverse mesh spacings: σ | Ωs | νs | Ωt | νt, mp each
    int mp;                 // stride of the mesh tables
    const int16_t *hint;    // bracket hints of the five meshes
    MeshV<float> mf;        // FP32 copies (FP32-mixed passes only)
    const uint32_t *code;   // [MAX_DESC] frequency-transfer codes of this element's vertex slots
    Ctl *ctl;      // [group]
    int *own;      // [group]: owner whose node the group evaluated this step, -1 = none
    double *scr;   // lane exchange scratch, [4][sr] doubles (one row per quantity); lane (g, i)
                   // at g*(GS+1)+i so that equal sites of different groups fall in different banks
    void *desc;    // DescT<TV>[group][MAX_DESC]
    double *bub;   // [group][2][2]: (G, K) of the two bubble legs
};

// Byte offsets of the dynamic LDS block of k_vertex_pass, sized to the problem: the five
// meshes and their inverse-spacing tables (stride mp doubles), Σ (and dΣ in the Katanin pass), the lane
// exchange rows, bubble legs, owner states, descriptors, overlap rows and row offsets, the
// site-exchange table and the bracket hints.  At 3 waves per SIMD a block may take 1/12 of the
// CU's LDS in 1280-byte granules (12800 bytes): the square L=6, N=40 FP32-mixed pass takes
// 11628 (the Katanin pass; the others keep no dΣ).
struct PassLds {
    int mp, sr, o_sig, o_scr, o_bub, o_ctl, o_desc, o_mf, o_ov, o_ovoff, o_exch, o_hint, total;
};

// sig_d: the pass reads dΣ as well as Σ (the Katanin pass); the others keep only Σ in LDS.
__host__ __device__ inline PassLds pass_lds(const Sys &y, int n_ov, int desc_size, bool f32, bool sig_d = true) {
    PassLds L;
    int mp = y.nsig;
    if (y.nO > mp) mp = y.nO;
    if (y.nn > mp) mp = y.nn;
    L.mp = (mp + 1) & ~1;
    int o = 10 * L.mp * 8;
    L.o_sig = o; o += (sig_d ? 2 : 1) * L.mp * 8;
    // exchange rows: lane (g, i) at g*(GS+1)+i, then one row for the idle lanes; 2 nc rows (the
    // Richardson check reads two values per component); the block also holds the 6 nc local-vertex
    // values per group of the t channel
    L.sr = (y.G * (y.GS + 1) + 2) & ~1;
    L.o_scr = o; o += (2 * y.nc * L.sr > 6 * y.nc * y.G ? 2 * y.nc * L.sr : 6 * y.nc * y.G) * 8;
    L.o_bub = o; o += y.G * 4 * 8;
    L.o_ctl = o; o += y.G * (int)sizeof(Ctl);
    L.o_desc = o; o += y.G * MAX_DESC * desc_size;
    L.o_mf = o; o += f32 ? 12 * L.mp * 4 : 0;
    L.o_ov = o; o += n_ov * 4;
    L.o_ovoff = o; o += (y.ns + 1) * 4;
    L.o_exch = o; o += y.ns * 4;
    L.o_hint = o; o += HINT_WORDS * 4;
    L.total = o;
    return L;
}

// Vertex-slot configuration of an integrand: bits 0-2 = enabled channels (s, t, u; the
// reference's ch_s/ch_t/ch_u keywords), bit 3 = the slot reads the second action D instead
// of A.  Slots 0/1 are the left/right vertices on every site, slots 2/3 the local (site 0)
// vertices of the t channel.  A slot with no channel is the bare vertex (reduced s bubble).
template <int PASS>
__device__ __forceinline__ int vertex_cfg(int ch, int slot) {
    constexpr int D = 8;
    if (ch == 1) {  // t: flow_su2_t.jl, BSE_su2_t.jl
        if (PASS == P_LEFT) return slot == 0 ? (5 | D) : slot == 1 ? 7 : slot == 2 ? (3 | D) : 7;
        if (PASS == P_CENTRAL) return slot == 0 ? 7 : slot == 1 ? (2 | D) : slot == 2 ? 7 : (4 | D);
        if (PASS == P_BSE) return slot == 0 ? 5 : slot == 1 ? 7 : slot == 2 ? 3 : 7;
        return 7;
    }
    if (ch == 0) {  // s: flow_su2_s.jl, BSE_su2_s.jl, SDE_su2.jl
        if (PASS == P_LEFT) return slot == 0 ? (6 | D) : 7;
        if (PASS == P_CENTRAL) return slot == 0 ? 7 : (1 | D);
        if (PASS == P_BSE) return slot == 0 ? 6 : 7;
        if (PASS == P_REDUCED) return slot == 0 ? 0 : 7;
        return 7;
    }
    // u: flow_su2_u.jl, BSE_su2_u.jl
    if (PASS == P_LEFT) return slot == 0 ? (3 | D) : 7;
    if (PASS == P_CENTRAL) return slot == 0 ? 7 : (4 | D);
    if (PASS == P_BSE) return slot == 0 ? 3 : 7;
    return 7;
}

// Frequency arguments (w, v, v') of the buffer of channel c for vertex slot `slot`, written
// with the reference's operation order (flow_su2_{s,t,u}.jl:21-39).
__device__ __forceinline__ void slot_freqs(int ch, int slot, int c, double W, double a1, double a2, double v,
                                           double &w, double &x, double &y) {
    if (ch == 0) {  // s = W, vs = a1, vsp = a2
        const double s = W;
        if (slot == 0) {
            if (c == 0) { w = s; x = a1; y = -v; }
            else if (c == 1) { w = v - a1; x = 0.5 * (s + v + a1); y = 0.5 * (s - v - a1); }
            else { w = v + a1; x = 0.5 * (s - v + a1); y = 0.5 * (s + v - a1); }
        } else {
            if (c == 0) { w = s; x = v; y = a2; }
            else if (c == 1) { w = -v - a2; x = 0.5 * (s + v - a2); y = 0.5 * (s - v + a2); }
            else { w = v - a2; x = 0.5 * (s + v + a2); y = 0.5 * (s - v - a2); }
        }
    } else if (ch == 2) {  // u = W, vu = a1, vup = a2
        const double u = W;
        if (slot == 0) {
            if (c == 0) { w = v + a1; x = 0.5 * (u - v + a1); y = 0.5 * (-u - v + a1); }
            else if (c == 1) { w = v - a1; x = 0.5 * (u + v + a1); y = 0.5 * (-u + v + a1); }
            else { w = u; x = a1; y = v; }
        } else {
            if (c == 0) { w = v + a2; x = 0.5 * (u + v - a2); y = 0.5 * (-u + v - a2); }
            else if (c == 1) { w = -v + a2; x = 0.5 * (u + v + a2); y = 0.5 * (-u + v + a2); }
            else { w = u; x = v; y = a2; }
        }
    } else {  // t = W, vt = a1, vtp = a2
        const double t = W;
        if (slot == 0) {
            if (c == 0) { w = v + a1; x = 0.5 * (-t - v + a1); y = 0.5 * (-t + v - a1); }
            else if (c == 1) { w = t; x = a1; y = v; }
            else { w = -v + a1; x = 0.5 * (-t + v + a1); y = 0.5 * (t + v + a1); }
        } else if (slot == 1) {
            if (c == 0) { w = v + a2; x = 0.5 * (-t + v - a2); y = 0.5 * (-t - v + a2); }
            else if (c == 1) { w = t; x = v; y = a2; }
            else { w = v - a2; x = 0.5 * (-t + v + a2); y = 0.5 * (t + v + a2); }
        } else if (slot == 2) {
            if (c == 0) { w = v + a1; x = 0.5 * (t + v - a1); y = 0.5 * (-t + v - a1); }
            else if (c == 1) { w = -v + a1; x = 0.5 * (t + v + a1); y = 0.5 * (-t + v + a1); }
            else { w = t; x = v; y = a1; }
        } else {
            if (c == 0) { w = v + a2; x = 0.5 * (t - v + a2); y = 0.5 * (-t - v + a2); }
            else if (c == 1) { w = v - a2; x = 0.5 * (t + v + a2); y = 0.5 * (-t + v + a2); }
            else { w = t; x = a2; y = v; }
        }
    }
}

template <typename ST>
__device__ __forceinline__ MeshV<ST> meshv(const Lds &S);
template <>
__device__ __forceinline__ MeshV<double> meshv<double>(const Lds &S) {
    return MeshV<double>{S.sig, S.Os, S.ns, S.Ot, S.nt, S.inv, S.sigA, S.sigD, S.mp, S.hint};
}
template <>
__device__ __forceinline__ MeshV<float> meshv<float>(const Lds &S) {
    return S.mf;
}

// Frequency-transfer code of (vertex slot k, channel c): bit 0 = channel enabled, bit 1 = the
// slot reads action D; bytes 1-3 encode the buffer arguments (w, v, v') of slot_freqs, each
// as ((sW·W + sv·v) + sa·A)·h with signs s ∈ {0, +, -}, A ∈ {a1, a2}, h ∈ {1, 1/2}
// (bits 0-1 sW, 2-3 sv, 4-5 sa, 6 A = a2, 7 h = 1/2).  The code is read off slot_freqs itself
// by probing it with unit arguments, so both forms are the same formulas; every formula there
// has this shape, and the IEEE sum of the signed terms in this order equals the original
// expression exactly.
template <int PASS>
__device__ uint32_t slot_code(int ch, int k, int c) {
    const int cfg = vertex_cfg<PASS>(ch, k);
    uint32_t code = (uint32_t)((cfg >> c) & 1) | ((cfg & 8) ? 2u : 0u);
    double o[4][3];
    slot_freqs(ch, k, c, 1.0, 0.0, 0.0, 0.0, o[0][0], o[0][1], o[0][2]);
    slot_freqs(ch, k, c, 0.0, 0.0, 0.0, 1.0, o[1][0], o[1][1], o[1][2]);
    slot_freqs(ch, k, c, 0.0, 1.0, 0.0, 0.0, o[2][0], o[2][1], o[2][2]);
    slot_freqs(ch, k, c, 0.0, 0.0, 1.0, 0.0, o[3][0], o[3][1], o[3][2]);
    for (int q = 0; q < 3; ++q) {
        uint32_t e = 0;
        bool half = false;
        const double tw = o[0][q], tv = o[1][q], t1 = o[2][q], t2 = o[3][q];
        if (tw != 0.0) { e |= (tw > 0.0) ? 1u : 2u; half = half || fabs(tw) == 0.5; }
        if (tv != 0.0) { e |= (tv > 0.0) ? 4u : 8u; half = half || fabs(tv) == 0.5; }
        if (t1 != 0.0) { e |= (t1 > 0.0) ? 16u : 32u; half = half || fabs(t1) == 0.5; }
        if (t2 != 0.0) { e |= ((t2 > 0.0) ? 16u : 32u) | 64u; half = half || fabs(t2) == 0.5; }
        if (half) e |= 128u;
        code |= e << (8 + 8 * q);
    }
    return code;
}

template <typename ST>
__device__ __forceinline__ ST coord(uint32_t e, ST W, ST v, ST a1, ST a2) {
    const ST A = (e & 64u) ? a2 : a1;
    const ST tW = (e & 1u) ? W : ((e & 2u) ? -W : (ST)0);
    const ST tv = (e & 4u) ? v : ((e & 8u) ? -v : (ST)0);
    const ST ta = (e & 16u) ? A : ((e & 32u) ? -A : (ST)0);
    const ST r = (tW + tv) + ta;
    return (e & 128u) ? (ST)0.5 * r : r;
}

// get_indices + get_param (param.jl:26-86) for val >= 0 on a mesh of n <= 64 points starting
// at 0: a fixed six-step search for the first point >= val, then the reference's bracket and
// weights, with the reciprocal spacing read from a table built by the same IEEE division.
template <typename ST>
__device__ __forceinline__ void bracket(ST val, const ST *m, const ST *inv, int n, int &lo, int &hi, ST &wl,
                                        ST &wh) {
    int base = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const int t = base + step;
        const int tc = (t <= n - 1) ? t : n - 1;
        if (t <= n - 1 && m[tc - 1] < val) base = t;
    }
    hi = base;
    const ST mh = m[base];
    lo = (val < mh) ? base - 1 : base;
    if (lo < hi) {
        const ST d = inv[lo];
        wl = d * (mh - val);
        wh = d * (val - m[lo]);
    } else {
        wl = (ST)1;
        wh = (ST)0;
    }
}

// bracket() with the mesh's hint table (k = mesh family σ, Ωs, νs, Ωt, νt): the value's cell
// gives the first candidate index A and the answer lies in [A, A+3], so three independent LDS
// reads replace the six dependent steps of the search.  The bracket, the weights and their
// operation order are bracket()'s; cells whose answers span more indices use bracket().
__device__ __forceinline__ void bracket_hint(double val, const double *m, const double *inv, int n,
                                             const int16_t *hint, int k, int &lo, int &hi, double &wl, double &wh) {
    const int nc = hint[5 + k];
    int cell = hint_key(val) - hint[k];
    cell = cell < 0 ? 0 : (cell > nc - 1 ? nc - 1 : cell);
    const int hb = reinterpret_cast<const uint8_t *>(hint + 10)[k * HINT_CELLS + cell];
    if (hb & 128) {
        bracket<double>(val, m, inv, n, lo, hi, wl, wh);
        return;
    }
    // points of the window below val: m[A], m[A+1], m[A+2]; the answer is A + their count
    const int n1 = n - 1;
    const int j2 = hb + 1 < n1 ? hb + 1 : n1, j3 = hb + 2 < n1 ? hb + 2 : n1;
    const int cnt = (m[hb] < val ? 1 : 0) + (m[j2] < val ? 1 : 0) + (m[j3] < val ? 1 : 0);
    const int base = (hb + cnt < n1) ? hb + cnt : n1;
    const double mh = m[base];
    hi = base;
    lo = (val < mh) ? base - 1 : base;
    const double ml = m[lo], dl = inv[lo];
    if (lo < hi) {
        wl = dl * (mh - val);
        wh = dl * (val - ml);
    } else {
        wl = 1.0;
        wh = 0.0;
    }
}

// bracket of mesh family k through its hint table where one exists (FP64 tables)
template <typename ST>
__device__ __forceinline__ void bracket_k(ST val, const MeshV<ST> &M, const ST *m, const ST *inv, int n, int k, int &lo,
                                          int &hi, ST &wl, ST &wh) {
    if constexpr (std::is_same<ST, double>::value) {
        if (M.hint) {
            bracket_hint(val, m, inv, n, M.hint, k, lo, hi, wl, wh);
            return;
        }
    }
    bracket<ST>(val, m, inv, n, lo, hi, wl, wh);
}

// Descriptor of one buffer in straight-line form: frequency transfer from the slot code, the
// sign folding of get_buffer_{s,t,u} (buffer.jl:88-199), the asymptotic-kernel choice of
// get_buffer (buffer.jl:51-86), three brackets, and the corner weights and offsets of the
// interpolation (channel.jl:42-305) in the reference's product order.  The unused dimensions of
// q1/q2 carry the neutral bracket (index 0, weights 1 and 0), which reproduces the reference's
// weights exactly and leaves the extra corners at weight zero.
// The descriptor's parts: meta, the offset of the first corner and the three 1-D brackets'
// weights (lower, upper) along Ω, ν and ν'.
template <typename ST>
struct DescParts {
    int meta, off;
    ST a1w, b1w, a2w, b2w, a3w, b3w;
};

template <typename ST>
__device__ __forceinline__ DescParts<ST> desc_parts(const Lds &S, const Sys &y, uint32_t code, int c, ST W, ST a1, ST a2,
                                                     ST v) {
    const MeshV<ST> M = meshv<ST>(S);
    ST w = coord<ST>(code >> 8 & 255u, W, v, a1, a2);
    ST x = coord<ST>(code >> 16 & 255u, W, v, a1, a2);
    ST z = coord<ST>(code >> 24, W, v, a1, a2);
    const bool nw = w < (ST)0, nx = x < (ST)0, nz = z < (ST)0;
    w = nw ? -w : w;
    x = nx ? -x : x;
    z = nz ? -z : z;
    const bool tch = (c == 1);
    const bool ex = !tch && (nw != nx), mp = !tch && (nx != nz), negd = (nx != nz);
    const int nO = y.nO, nn = y.nn, ns = y.ns;
    const ST *Om = tch ? M.Ot : M.Os;
    const ST *nu = tch ? M.nt : M.ns;
    const ST *iO = M.inv + (tch ? 3 : 1) * M.mp;
    const ST *inu = M.inv + (tch ? 4 : 2) * M.mp;
    const bool out = Om[nO - 1] < w;
    const bool av = nu[nn - 1] < x, avp = nu[nn - 1] < z;
    const int kernel = out ? 0 : (av ? (avp ? 1 : 3) : (avp ? 2 : 4));
    int l1, h1, l2, h2, l3, h3;
    ST a1w, b1w, a2w, b2w, a3w, b3w;
    bracket_k<ST>(w, M, Om, iO, nO, tch ? 3 : 1, l1, h1, a1w, b1w);
    bracket_k<ST>(x, M, nu, inu, nn, tch ? 4 : 2, l2, h2, a2w, b2w);
    bracket_k<ST>(z, M, nu, inu, nn, tch ? 4 : 2, l3, h3, a3w, b3w);
    if (kernel == 3) { l2 = l3; h2 = h3; a2w = a3w; b2w = b3w; }
    if (kernel != 4) { l3 = 0; h3 = 0; a3w = (ST)1; b3w = (ST)0; }
    if (kernel == 1) { l2 = 0; h2 = 0; a2w = (ST)1; b2w = (ST)0; }
    const int base = (kernel == 2) ? (int)y.q21 : (kernel == 3) ? (int)y.q22 : (kernel == 4) ? (int)y.q3 : 0;
    // the flags of get_buffer_{s,t,u}: s: sgn_μ = ν < 0, sgn_ν = ν' < 0; t: also sgn_μν = Ω < 0;
    // u: sgn_ν = (ν < 0) xor (ν' < 0), sgn_μν = Ω < 0 (buffer.jl:88-199)
    const bool smu = (c != 2) && nx, snu = (c == 2) ? (nx != nz) : nz, smunu = (c != 0) && nw;
    DescParts<ST> p;
    p.meta = ((code & 1u) && kernel) ? (kernel | (ex ? 8 : 0) | (mp ? 16 : 0) | (negd ? 32 : 0) |
                                        (h1 > l1 ? 64 : 0) | (h2 > l2 ? 128 : 0) | (h3 > l3 ? 256 : 0) |
                                        (smu ? 512 : 0) | (snu ? 1024 : 0) | (smunu ? 2048 : 0))
                                     : 0;
    p.off = base + ns * (l1 + nO * (l2 + nn * l3));
    p.a1w = a1w; p.b1w = b1w; p.a2w = a2w; p.b2w = b2w; p.a3w = a3w; p.b3w = b3w;
    return p;
}

template <typename TV, typename ST>
__device__ __forceinline__ void build_desc_u(const Lds &S, const Sys &y, uint32_t code, int c, ST W, ST a1, ST a2, ST v,
                                             DescT<TV> &d) {
    const DescParts<ST> p = desc_parts<ST>(S, y, code, c, W, a1, a2, v);
    d.meta = p.meta;
    d.off = p.off;
    d.w[0] = (TV)(p.a1w * p.a2w * p.a3w); d.w[1] = (TV)(p.b1w * p.a2w * p.a3w);
    d.w[2] = (TV)(p.a1w * p.b2w * p.a3w); d.w[3] = (TV)(p.b1w * p.b2w * p.a3w);
    d.w[4] = (TV)(p.a1w * p.a2w * p.b3w); d.w[5] = (TV)(p.b1w * p.a2w * p.b3w);
    d.w[6] = (TV)(p.a1w * p.b2w * p.b3w); d.w[7] = (TV)(p.b1w * p.b2w * p.b3w);
}

// Corners any group of the wavefront needs for one channel (wave-uniform): 8, 4, 2 or 0.
__device__ __forceinline__ int wave_corners(int meta) {
    const int k = meta & 7;
    if (__ballot(k == 4) != 0ull) return 8;
    if (__ballot(k == 2 || k == 3) != 0ull) return 4;
    if (__ballot(k == 1) != 0ull) return 2;
    return 0;
}

template <typename V2>
__device__ __forceinline__ void load_corners(const V2 *__restrict__ q, uint32_t o, uint32_t c1, uint32_t c2, uint32_t c3,
                                             int n, V2 *x) {
    x[0] = q[o];
    x[1] = q[o + c1];
    if (n >= 4) {
        x[2] = q[o + c2];
        x[3] = q[o + c1 + c2];
    }
    if (n >= 8) {
        x[4] = q[o + c3];
        x[5] = q[o + c1 + c3];
        x[6] = q[o + c2 + c3];
        x[7] = q[o + c1 + c2 + c3];
    }
}

// Interpolated value of both SU(2) components: the first corner initialises, the rest
// accumulate in the reference's corner order (channel.jl:42-305).  Corners beyond the
// buffer's own count carry weight zero.
template <typename TV, typename V2>
__device__ __forceinline__ void corner_sum(const DescT<TV> &d, int n, const V2 *x, TV &r0, TV &r1) {
    r0 = d.w[0] * x[0].x;
    r1 = d.w[0] * x[0].y;
    r0 += d.w[1] * x[1].x; r1 += d.w[1] * x[1].y;
    if (n >= 4) {
        r0 += d.w[2] * x[2].x; r1 += d.w[2] * x[2].y;
        r0 += d.w[3] * x[3].x; r1 += d.w[3] * x[3].y;
    }
    if (n >= 8) {
        r0 += d.w[4] * x[4].x; r1 += d.w[4] * x[4].y;
        r0 += d.w[5] * x[5].x; r1 += d.w[5] * x[5].y;
        r0 += d.w[6] * x[6].x; r1 += d.w[6] * x[6].y;
        r0 += d.w[7] * x[7].x; r1 += d.w[7] * x[7].y;
    }
}

// Vertex of one slot at one site: bare + Σ_channels sign·interpolation (get_Γ_avx!,
// action_su2.jl:95-134).  The gather copy stores (spin, density) pairs, so one load fetches
// both components of a corner.  In FP32 the loads of all three channels are issued before
// any is consumed; in FP64 one channel at a time to bound registers.
template <typename TV>
__device__ __forceinline__ void gather_slot(const DescT<TV> *ds, const TV *gamma, const int *exch, int site,
                                            const Sys &y, TV &vs, TV &vd, int &cn) {
    using V2 = typename Vec2<TV>::T;
    const V2 *q = reinterpret_cast<const V2 *>(gamma);
    int meta[3], n[3];
    uint32_t o[3], c1[3], c2[3], c3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const DescT<TV> &d = ds[c];
        meta[c] = d.meta;
        n[c] = wave_corners(meta[c]);
        const bool on = (meta[c] & 7) != 0;
        const int sx = (meta[c] & 8) ? exch[site] : site;
        const bool map = (meta[c] & 16) != 0;
        const int chn = (c == 1) ? 1 : ((c == 0) == map ? 2 : 0);
        o[c] = on ? (uint32_t)(chn * y.chlen + d.off + sx) : 0u;
        corner_steps(meta[c], y, c1[c], c2[c], c3[c]);
    }
    cn = 0;
    if (sizeof(TV) == 4) {
        V2 x[3][8];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (n[c]) load_corners(q, o[c], c1[c], c2[c], c3[c], n[c], x[c]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (meta[c] & 7) {
                TV r0, r1;
                corner_sum(ds[c], n[c], x[c], r0, r1);
                vs += r0;
                vd += (meta[c] & 32) ? -r1 : r1;
                cn += corners(meta[c] & 7);
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (!n[c]) continue;
            V2 x[8];
            load_corners(q, o[c], c1[c], c2[c], c3[c], n[c], x);
            if (meta[c] & 7) {
                TV r0, r1;
                corner_sum(ds[c], n[c], x, r0, r1);
                vs += r0;
                vd += (meta[c] & 32) ? -r1 : r1;
                cn += corners(meta[c] & 7);
            }
        }
    }
}

// One channel of one vertex slot at one site: the part of gather_slot for channel c, for lanes
// that each take a different (slot, channel).  The corner count is uniform over the lanes that
// take part; (r0, r1) is the interpolation with the corner weights, before sign and density
// flags; cn the corners it read (0 for a disabled channel).
template <typename TV>
__device__ __forceinline__ void gather_channel(const DescT<TV> &d, const TV *gamma, const int *exch, int site,
                                               const Sys &y, int c, TV &r0, TV &r1, int &cn) {
    using V2 = typename Vec2<TV>::T;
    const V2 *q = reinterpret_cast<const V2 *>(gamma);
    const int meta = d.meta;
    const int n = wave_corners(meta);
    const bool on = (meta & 7) != 0;
    const int sx = (meta & 8) ? exch[site] : site;
    const bool map = (meta & 16) != 0;
    const int chn = (c == 1) ? 1 : ((c == 0) == map ? 2 : 0);
    const uint32_t o = on ? (uint32_t)(chn * y.chlen + d.off + sx) : 0u;
    uint32_t c1, c2, c3;
    corner_steps(meta, y, c1, c2, c3);
    r0 = (TV)0;
    r1 = (TV)0;
    cn = 0;
    if (n) {
        V2 x[8];
        load_corners(q, o, c1, c2, c3, n, x);
        if (on) {
            corner_sum(d, n, x, r0, r1);
            cn = corners(meta & 7);
        }
    }
}

// U(1)-DM gathers.  The gather copy stores the six components of an entry next to each other,
// three pairs per corner; r[k] is channel c's interpolation of component k before the flags.
template <typename TV>
__device__ __forceinline__ void corners6(const DescT<TV> &d, const TV *gamma, const int *exch, int site, const Sys &y,
                                         int c, TV *r, int &cn) {
    using V2 = typename Vec2<TV>::T;
    const V2 *q = reinterpret_cast<const V2 *>(gamma);
    const int meta = d.meta;
    const int n = wave_corners(meta);
    const bool on = (meta & 7) != 0;
    const int sx = (meta & 8) ? exch[site] : site;
    const bool map = (meta & 16) != 0;
    const int chn = (c == 1) ? 1 : ((c == 0) == map ? 2 : 0);
    const uint32_t o = on ? 3u * (uint32_t)(chn * y.chlen + d.off + sx) : 0u;
    uint32_t c1, c2, c3;
    corner_steps(meta, y, c1, c2, c3);
    c1 *= 3u; c2 *= 3u; c3 *= 3u;
#pragma unroll
    for (int k = 0; k < 6; ++k) r[k] = (TV)0;
    cn = 0;
    if (!n) return;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        V2 x[8];
        load_corners(q, o + (uint32_t)j, c1, c2, c3, n, x);
        if (on) corner_sum(d, n, x, r[2 * j], r[2 * j + 1]);
    }
    if (on) cn = corners(meta & 7);
}

// add one channel's interpolations r to the six components v with apply_flags_u1_dm
// (action_u1_dm.jl:79-118); the flags are the descriptor's exchange and sign bits
template <typename TV>
__device__ __forceinline__ void add_flags6(int meta, const TV *r, TV *v) {
    const int fl = ((meta >> 3) & 1) | (((meta >> 9) & 7) << 1);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int src;
        const double sg = flag_u1(fl, k, src);
        v[k] += (sg < 0.0) ? -r[src] : r[src];
    }
}

// gather_slot for the six U(1)-DM components: bare + the enabled channels in s, t, u order
template <typename TV>
__device__ __forceinline__ void gather_slot6(const DescT<TV> *ds, const TV *gamma, const int *exch, int site,
                                             const Sys &y, TV *v, int &cn) {
    cn = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        TV r[6];
        int n;
        corners6<TV>(ds[c], gamma, exch, site, y, c, r, n);
        if (ds[c].meta & 7) {
            add_flags6<TV>(ds[c].meta, r, v);
            cn += n;
        }
    }
}

// One leg x of the bubble: G(x) and, for the Katanin bubble, K(x) = S(x) + G(x)² dΣ(x)
// (bubbles.jl:17-80; S reuses G as the reference's get_S recomputes the same value).  Σ and
// dΣ are interpolated with one shared bracket of |x| (get_Σ, Action.jl:8-26).
template <bool KAT, typename ST>
__device__ __forceinline__ void bubble_leg(ST L, ST x, const MeshV<ST> &M, int n, ST &G, ST &K) {
    const ST ax = fabs(x);
    ST s, ds = (ST)0;
    if (ax <= M.sig[n - 1]) {
        int lo, hi;
        ST wl, wh;
        bracket_k<ST>(ax, M, M.sig, M.inv, n, 0, lo, hi, wl, wh);
        const ST sg = (x > (ST)0) ? (ST)1 : ((x < (ST)0) ? (ST)-1 : x);
        s = sg * (wl * M.sigA[lo] + wh * M.sigA[hi]);
        if (KAT) ds = sg * (wl * M.sigD[lo] + wh * M.sigD[hi]);
    } else {
        s = M.sig[n - 1] * M.sigA[n - 1] / x;
        if (KAT) ds = M.sig[n - 1] * M.sigD[n - 1] / x;
    }
    G = (ST)0;
    ST G0 = (ST)0;
    if (ax > (ST)1e-8) {
        // get_G (bubbles.jl:17-36)
        G0 = G_bare<ST>(L, x);
        const ST denom = (x * s >= (ST)0) ? (ST)1 / G0 + s : (ST)1 / G0;
        G = (ST)1 / denom;
    }
    K = (ST)0;
    if (KAT) {
        ST Sx = (ST)0;
        if (ax > (ST)1e-8) {
            const ST r = G / G0;
            Sx = r * r * exp_st<ST>(-(x * x) / (L * L)) * (ST)2 * x / (L * L * L);
        }
        K = Sx + G * G * ds;
    }
}

// U(1)-DM component products (components xx, zz, DM, dd, zd, dz), the brackets multiplied by -p:
// s and u channels (flow_u1_dm_s.jl:52-57, flow_u1_dm_u.jl:52-57), the local part of the t
// channel (flow_u1_dm_t.jl:56-67) and one overlap row of its non-local sum (:88-93, no 2S).
template <typename TV>
__device__ __forceinline__ void prod_su6(bool u, const TV *a, const TV *b, TV *g) {
    const TV two = (TV)2.0;
    const TV xx1 = a[0], zz1 = a[1], DM1 = a[2], dd1 = a[3], zd1 = a[4], dz1 = a[5];
    const TV xx2 = b[0], zz2 = b[1], DM2 = b[2], dd2 = b[3], zd2 = b[4], dz2 = b[5];
    if (!u) {
        g[0] = DM1 * dz2 - DM1 * zd2 + xx1 * dd2 + dd1 * xx2 - xx1 * zz2 - dz1 * DM2 + zd1 * DM2 - zz1 * xx2;
        g[1] = zz1 * dd2 + dd1 * zz2 - dz1 * zd2 - two * xx1 * xx2 - two * DM1 * DM2 - zd1 * dz2;
        g[2] = dd1 * DM2 + DM1 * dd2 - zd1 * xx2 - xx1 * dz2 - zz1 * DM2 + xx1 * zd2 - DM1 * zz2 + dz1 * xx2;
        g[3] = -dz1 * dz2 + two * xx1 * xx2 + dd1 * dd2 + two * DM1 * DM2 - zd1 * zd2 + zz1 * zz2;
        g[4] = zd1 * dd2 + dd1 * zd2 - two * DM1 * xx2 + dz1 * zz2 + zz1 * dz2 + two * xx1 * DM2;
        g[5] = zd1 * zz2 + dd1 * dz2 + zz1 * zd2 + two * DM1 * xx2 + dz1 * dd2 - two * xx1 * DM2;
    } else {
        g[0] = -DM1 * dz2 - DM1 * zd2 + xx1 * dd2 + dd1 * xx2 + xx1 * zz2 + dz1 * DM2 + zd1 * DM2 + zz1 * xx2;
        g[1] = zz1 * dd2 + dd1 * zz2 - dz1 * zd2 + two * xx1 * xx2 + two * DM1 * DM2 - zd1 * dz2;
        g[2] = dd1 * DM2 + DM1 * dd2 - zd1 * xx2 + xx1 * dz2 + zz1 * DM2 + xx1 * zd2 + DM1 * zz2 - dz1 * xx2;
        g[3] = -dz1 * dz2 + two * xx1 * xx2 + dd1 * dd2 + two * DM1 * DM2 - zd1 * zd2 + zz1 * zz2;
        g[4] = zd1 * dd2 + dd1 * zd2 - two * DM1 * xx2 + dz1 * zz2 + zz1 * dz2 + two * xx1 * DM2;
        g[5] = zd1 * zz2 + dd1 * dz2 + zz1 * zd2 - two * DM1 * xx2 + dz1 * dd2 + two * xx1 * DM2;
    }
}

template <typename TV>
__device__ __forceinline__ void prod_t_local6(const TV *a, const TV *c4, const TV *c3, const TV *b, TV *g) {
    const TV two = (TV)2.0;
    const TV xx1 = a[0], zz1 = a[1], DM1 = a[2], dd1 = a[3], zd1 = a[4], dz1 = a[5];
    const TV xx2 = b[0], zz2 = b[1], DM2 = b[2], dd2 = b[3], zd2 = b[4], dz2 = b[5];
    const TV xx3 = c3[0], zz3 = c3[1], DM3 = c3[2], dd3 = c3[3], zd3 = c3[4], dz3 = c3[5];
    const TV xx4 = c4[0], zz4 = c4[1], DM4 = c4[2], dd4 = c4[3], zd4 = c4[4], dz4 = c4[5];
    g[0] = DM1 * dz4 - DM1 * zd4 + xx1 * dd4 - xx1 * zz4 + dd3 * xx2 + dz3 * DM2 - zd3 * DM2 - zz3 * xx2;
    g[1] = zz1 * dd4 + two * zd1 * DM4 - zd1 * zd4 + zz1 * zz4 - two * zz1 * xx4 - zd1 * dz4 - two * DM3 * dz2 -
           dz3 * dz2 + dd3 * zz2 - two * xx3 * zz2 + zz3 * zz2 - zd3 * dz2;
    g[2] = DM1 * dd4 - xx1 * dz4 + xx1 * zd4 - DM1 * zz4 - dz3 * xx2 + dd3 * DM2 + zd3 * xx2 - zz3 * DM2;
    g[3] = -dz1 * dz4 + two * dd1 * xx4 + dd1 * zz4 - dz1 * zd4 - two * dz1 * DM4 + dd1 * dd4 + two * DM3 * zd2 +
           two * xx3 * dd2 + zz3 * dd2 - dz3 * zd2 + dd3 * dd2 - zd3 * zd2;
    g[4] = zd1 * zz4 + zd1 * dd4 + two * zd1 * xx4 + zz1 * zd4 + two * zz1 * DM4 + zz1 * dz4 + zd3 * dd2 +
           two * DM3 * dd2 + dd3 * zd2 + zz3 * zd2 - two * xx3 * zd2 + dz3 * dd2;
    g[5] = -two * dd1 * DM4 + dd1 * zd4 + dd1 * dz4 + dz1 * zz4 + dz1 * dd4 - two * dz1 * xx4 + zd3 * zz2 +
           two * xx3 * dz2 + dd3 * dz2 + dz3 * zz2 - two * DM3 * zz2 + zz3 * dz2;
}

template <typename TV>
__device__ __forceinline__ void add_t_row6(TV cm, const TV *a, const TV *b, TV *G) {
    G[0] += cm * (a[0] * b[0] - a[2] * b[2]);
    G[1] += cm * (a[1] * b[1] - a[4] * b[5]);
    G[2] += cm * (a[2] * b[0] + a[0] * b[2]);
    G[3] += cm * (a[3] * b[3] - a[5] * b[4]);
    G[4] += cm * (a[4] * b[3] + a[1] * b[4]);
    G[5] += cm * (a[3] * b[5] + a[5] * b[1]);
}

// Integrand of one element at loop frequency v for site i: G[comp] (before the dv weight).
// Restates compute_{s,t,u}_{kat,left,central}! (flow_lib/flow_su2/*.jl, flow_u1_dm/*.jl),
// compute_{s,t,u}_BSE! (parquet_lib/parquet_su2/BSE_su2_*.jl, parquet_u1_dm/BSE_u1_dm_*.jl) and
// compute_s_reduced! (SDE_su2.jl:2-43, SDE_u1_dm.jl:2-43); NC = 2 (SU(2)) or 6 (U(1)-DM), whose
// buffers and channel masks are the same.  The per-node scalar work (two bubble legs, one buffer
// per vertex slot and channel) is split over the lanes of the group and published in LDS; every
// lane then gathers its own site.  Must be executed by every lane of the wavefront.
template <int PASS, typename TV, int NC>
__device__ __forceinline__ void node_eval(const PassArgs<TV> &P, const Lds &S, int ch, double W, double vw,
                                          double vwp, double v, int i, int slot, int g, int gbase, bool real_lane,
                                          TV *G, int &nc) {
    const Sys &y = P.y;
    using ST = typename PassScalar<TV>::T;
    const int nslot = (ch == 1) ? 4 : 2;
    const int ntask = 2 + 3 * nslot;
    const int gs = real_lane ? g : 0;
    DescT<TV> *dg = reinterpret_cast<DescT<TV> *>(S.desc) + gs * MAX_DESC;
    double *bg = S.bub + gs * 4;
    if (real_lane) {
#pragma clang loop unroll(disable)
        for (int t = i; t < ntask; t += y.GS) {
            if (t < 2) {
                ST x1, x2;
                bubble_args<ST>(ch, (ST)W, (ST)v, x1, x2);
                ST Gl, K;
                bubble_leg<PASS == P_KAT, ST>((ST)P.L, t == 0 ? x1 : x2, meshv<ST>(S), y.nsig, Gl, K);
                bg[2 * t] = (double)Gl;
                bg[2 * t + 1] = (double)K;
            } else {
                build_desc_u<TV, ST>(S, y, S.code[t - 2], (t - 2) % 3, (ST)W, (ST)vw, (ST)vwp, (ST)v, dg[t - 2]);
            }
        }
    }
    __syncthreads();
    const ST G1 = (ST)bg[0], K1 = (ST)bg[1], G2 = (ST)bg[2], K2 = (ST)bg[3];
    const ST twopi = (ST)2 * (ST)PI_D;
    const TV tp = (TV)((PASS == P_KAT) ? K1 * G2 / twopi + K2 * G1 / twopi : -(G1 * G2 / twopi));
    TV v1[NC], v2[NC], v3[NC], v4[NC];
    nc = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int cfg = vertex_cfg<PASS>(ch, k);
        const ActV<TV> &V = (cfg & 8) ? P.D : P.A;
        TV *vk = (k == 0) ? v1 : v2;
        int cn;
        if constexpr (NC == 2) {
            TV vs = (TV)V.bare[i], vd = (TV)V.bare[y.ns + i];
            gather_slot<TV>(dg + 3 * k, V.gamma, S.exch, i, y, vs, vd, cn);
            vk[0] = vs;
            vk[1] = vd;
        } else {
#pragma unroll
            for (int q = 0; q < NC; ++q) vk[q] = (TV)V.bare[q * y.ns + i];
            gather_slot6<TV>(dg + 3 * k, V.gamma, S.exch, i, y, vk, cn);
        }
        nc += cn;
    }
    if (ch == 1) {
        // local (site-0) vertices of slots 2 and 3, the same for every lane of a group: task
        // q < 6 (channel q % 3 of slot 2 + q / 3) goes to lane q mod GS of the group, which
        // publishes its interpolation; every lane then adds the three channels to bare[0] in
        // gather_slot's order
        double *loc = S.scr + 6 * NC * gs;
        if (real_lane) {
#pragma clang loop unroll(disable)
            for (int q = i; q < 6; q += y.GS) {
                const int k = 2 + q / 3, c = q - 3 * (q / 3);
                const ActV<TV> &V = (vertex_cfg<PASS>(ch, k) & 8) ? P.D : P.A;
                int cn;
                if constexpr (NC == 2) {
                    TV r0, r1;
                    gather_channel<TV>(dg[3 * k + c], V.gamma, S.exch, 0, y, c, r0, r1, cn);
                    loc[2 * q] = (double)r0;
                    loc[2 * q + 1] = (double)r1;
                } else {
                    TV r[6];
                    corners6<TV>(dg[3 * k + c], V.gamma, S.exch, 0, y, c, r, cn);
#pragma unroll
                    for (int j = 0; j < 6; ++j) loc[6 * q + j] = (double)r[j];
                }
                nc += cn;  // each local corner counted once per node
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 2; k < 4; ++k) {
            const ActV<TV> &V = (vertex_cfg<PASS>(ch, k) & 8) ? P.D : P.A;
            TV *vk = (k == 2) ? v3 : v4;
            if constexpr (NC == 2) {
                TV vs = (TV)V.bare[0], vd = (TV)V.bare[y.ns];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int meta = dg[3 * k + c].meta;
                    if (meta & 7) {
                        const TV r0 = (TV)loc[2 * (3 * (k - 2) + c)], r1 = (TV)loc[2 * (3 * (k - 2) + c) + 1];
                        vs += r0;
                        vd += (meta & 32) ? -r1 : r1;
                    }
                }
                vk[0] = vs;
                vk[1] = vd;
            } else {
#pragma unroll
                for (int q = 0; q < NC; ++q) vk[q] = (TV)V.bare[q * y.ns];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int meta = dg[3 * k + c].meta;
                    if (meta & 7) {
                        TV r[6];
#pragma unroll
                        for (int j = 0; j < 6; ++j) r[j] = (TV)loc[6 * (3 * (k - 2) + c) + j];
                        add_flags6<TV>(meta, r, vk);
                    }
                }
            }
        }
        __syncthreads();  // the scratch is rewritten by the overlap exchange below
    }
    if constexpr (NC == 2) {
        const TV v1s = v1[0], v1d = v1[1], v2s = v2[0], v2d = v2[1];
        if (ch == 0) {
            G[0] = -tp * ((TV)-2.0 * v1s * v2s + v1s * v2d + v1d * v2s);
            G[1] = -tp * ((TV)3.0 * v1s * v2s + v1d * v2d);
        } else if (ch == 2) {
            G[0] = -tp * ((TV)2.0 * v1s * v2s + v1s * v2d + v1d * v2s);
            G[1] = -tp * ((TV)3.0 * v1s * v2s + v1d * v2d);
        } else {
            const TV v3s = v3[0], v3d = v3[1], v4s = v4[0], v4d = v4[1];
            TV gsum = -tp * ((TV)-1.0 * v1s * v4s + v1s * v4d - (TV)1.0 * v3s * v2s + v3d * v2s);
            TV gdsum = -tp * ((TV)3.0 * v1d * v4s + v1d * v4d + (TV)3.0 * v3s * v2d + v3d * v2d);
            // overlap gather Σ_(j1,j2,m) -p (-2) m (2S) v1[j1] v2[j2]  (flow_su2_t.jl:59-74).  The
            // scratch holds (v1s, v1d, v2s, v2d) of each lane next to each other, so one LDS read
            // fetches both components of a site; rows go four at a time, their reads issued
            // together, and are added in row order.
            using V2 = typename Vec2<TV>::T;
            V2 *sx = reinterpret_cast<V2 *>(S.scr);
            sx[2 * slot] = V2{v1s, v1d};
            sx[2 * slot + 1] = V2{v2s, v2d};
            __syncthreads();
            const TV cc = -tp * (TV)(-2.0);
            const TV S2 = (TV)y.S2;
            const int r0 = S.ov_off[i], r1 = S.ov_off[i + 1];
            const V2 *xg = sx + 2 * gbase;
            int r = r0;
            for (; r + 4 <= r1; r += 4) {
                uint32_t row[4];
                V2 e1[4], e2[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) row[k] = S.ov[r + k];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    e1[k] = xg[2 * (int)(row[k] & 255)];
                    e2[k] = xg[2 * (int)((row[k] >> 8) & 255) + 1];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const TV m = (TV)(double)(row[k] >> 16);
                    gsum += cc * m * S2 * e1[k].x * e2[k].x;
                    gdsum += cc * m * S2 * e1[k].y * e2[k].y;
                }
            }
            for (; r < r1; ++r) {
                const uint32_t row = S.ov[r];
                const V2 e1 = xg[2 * (int)(row & 255)], e2 = xg[2 * (int)((row >> 8) & 255) + 1];
                const TV m = (TV)(double)(row >> 16);
                gsum += cc * m * S2 * e1.x * e2.x;
                gdsum += cc * m * S2 * e1.y * e2.y;
            }
            G[0] = gsum;
            G[1] = gdsum;
        }
    } else {
        TV gb[6];
        if (ch == 0 || ch == 2) {
            if (PASS == P_REDUCED) {
                // the left vertex is the bare coupling; xx, zz and dd are kept (SDE_u1_dm.jl:36-42)
                prod_su6<TV>(false, v1, v2, gb);
#pragma unroll
                for (int q = 0; q < 6; ++q) G[q] = (q == 0 || q == 1 || q == 3) ? -tp * gb[q] : (TV)0;
            } else {
                prod_su6<TV>(ch == 2, v1, v2, gb);
#pragma unroll
                for (int q = 0; q < 6; ++q) G[q] = -tp * gb[q];
            }
        } else {
            prod_t_local6<TV>(v1, v4, v3, v2, gb);
#pragma unroll
            for (int q = 0; q < 6; ++q) G[q] = -tp * gb[q];
            // overlap gather Σ_(j1,j2,m) -p (-2) m (a[j1] . b[j2])  (flow_u1_dm_t.jl:71-94): each
            // lane publishes its twelve components (v1 then v2) in the scratch
            TV *xs = reinterpret_cast<TV *>(S.scr);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                xs[12 * slot + q] = v1[q];
                xs[12 * slot + 6 + q] = v2[q];
            }
            __syncthreads();
            const TV cc = -tp * (TV)(-2.0);
            const TV *xg = xs + 12 * gbase;
            for (int r = S.ov_off[i]; r < S.ov_off[i + 1]; ++r) {
                const uint32_t row = S.ov[r];
                const TV m = (TV)(double)(row >> 16);
                add_t_row6<TV>(cc * m, xg + 12 * (int)(row & 255), xg + 12 * (int)((row >> 8) & 255) + 6, G);
            }
        }
    }
    // the descriptors and bubble legs of this node are read; the next node may overwrite them
    __syncthreads();
}

// element code: ch 2 bits | kernel 3 bits | w1 7 bits | w2 7 bits | w3 7 bits
__device__ __forceinline__ void decode(uint32_t e, int &ch, int &kernel, int &w1, int &w2, int &w3) {
    ch = e & 3;
    kernel = (e >> 2) & 7;
    w1 = (e >> 5) & 127;
    w2 = (e >> 12) & 127;
    w3 = (e >> 19) & 127;
}

// node position a + c*h rounded as the reference computes it (product, then sum); the
// pragma keeps hipcc from fusing it into an FMA (__dmul_rn/__dadd_rn are fused on gfx950)
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

// One output element of one pass (block = one wavefront).  e_idx is the element's position in
// its pass (and the probe index in probe mode), code its encoded (ch, kernel, w1, w2, w3); NC
// vertex components (2: SU(2), 6: U(1)-DM).
template <int PASS, typename TV, int NC = 2>
__device__ __forceinline__ void vertex_pass_body(const PassArgs<TV> &P, int e_idx, uint32_t code) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ uint32_t s_code[MAX_DESC];
    __shared__ int s_own[MAX_GROUPS];
    __shared__ unsigned long long s_cnt, s_bytes;

    const Sys &y = P.y;
    const int bid = e_idx;
    const int lane = threadIdx.x;
    const int n_ov = P.ov_off[y.ns];
    constexpr bool ST32 = sizeof(typename PassScalar<TV>::T) == 4;
    const PassLds Lay = pass_lds(y, n_ov, (int)sizeof(DescT<TV>), ST32, PASS == P_KAT);
    const int mp = Lay.mp;
    double *s_mesh = reinterpret_cast<double *>(s_dyn);
    double *s_sigA = reinterpret_cast<double *>(s_dyn + Lay.o_sig);
    double *s_sigD = (PASS == P_KAT) ? s_sigA + mp : s_sigA;  // read by the Katanin bubble only
    uint32_t *s_ov = reinterpret_cast<uint32_t *>(s_dyn + Lay.o_ov);
    int *s_ovoff = reinterpret_cast<int *>(s_dyn + Lay.o_ovoff);

    // ---- stage tables in LDS
    for (int k = lane; k < 10 * mp; k += 64) s_mesh[k] = P.mesh[(k / mp) * MAXN + (k % mp)];
    for (int k = lane; k < y.nsig; k += 64) {
        s_sigA[k] = P.A.Sigma[k];
        if (PASS == P_KAT) s_sigD[k] = P.D.Sigma[k];
    }
    if (ST32) {
        float *mf = reinterpret_cast<float *>(s_dyn + Lay.o_mf);
        for (int k = lane; k < 10 * mp; k += 64) mf[k] = (float)P.mesh[(k / mp) * MAXN + (k % mp)];
        for (int k = lane; k < y.nsig; k += 64) {
            mf[10 * mp + k] = (float)P.A.Sigma[k];
            mf[11 * mp + k] = (PASS == P_KAT) ? (float)P.D.Sigma[k] : 0.0f;
        }
    }
    int *s_exch = reinterpret_cast<int *>(s_dyn + Lay.o_exch);
    if (lane < y.ns) s_exch[lane] = P.exch[lane];
    for (int k = lane; k < n_ov; k += 64) s_ov[k] = P.ov_rows[k];
    for (int k = lane; k <= y.ns; k += 64) s_ovoff[k] = P.ov_off[k];
    int32_t *s_hint = reinterpret_cast<int32_t *>(s_dyn + Lay.o_hint);  // copied as words
    {
        const int32_t *gh = reinterpret_cast<const int32_t *>(P.mesh + 10 * MAXN);
        for (int k = lane; k < HINT_WORDS; k += 64) s_hint[k] = gh[k];
    }
    if (lane == 0) { s_cnt = 0; s_bytes = 0; }
    __syncthreads();

    Lds S;
    S.sig = s_mesh;
    S.Os = s_mesh + mp;
    S.ns = s_mesh + 2 * mp;
    S.Ot = s_mesh + 3 * mp;
    S.nt = s_mesh + 4 * mp;
    S.inv = s_mesh + 5 * mp;
    S.mp = mp;
    S.code = s_code;
    S.sigA = s_sigA;
    S.sigD = s_sigD;
    S.exch = s_exch;
    S.ov = s_ov;
    S.ov_off = s_ovoff;
    S.hint = reinterpret_cast<const int16_t *>(s_hint);
    S.scr = reinterpret_cast<double *>(s_dyn + Lay.o_scr);
    S.desc = reinterpret_cast<void *>(s_dyn + Lay.o_desc);
    S.bub = reinterpret_cast<double *>(s_dyn + Lay.o_bub);
    S.ctl = reinterpret_cast<Ctl *>(s_dyn + Lay.o_ctl);
    S.own = s_own;
    {
        const float *mf = reinterpret_cast<const float *>(s_dyn + Lay.o_mf);
        S.mf = MeshV<float>{mf, mf + mp, mf + 2 * mp, mf + 3 * mp, mf + 4 * mp, mf + 5 * mp, mf + 10 * mp,
                            mf + 11 * mp, mp, nullptr};
    }

    int ch, kernel, w1, w2, w3;
    decode(code, ch, kernel, w1, w2, w3);
    if (lane < MAX_DESC) s_code[lane] = slot_code<PASS>(ch, lane / 3, lane % 3);
    __syncthreads();
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
    W = wave_uniform(W);  // one element per wave
    vw = wave_uniform(vw);
    vwp = wave_uniform(vwp);
    const double ref = P.L + 0.5 * W;
    const double val = (ch == 1) ? y.valT : y.valS;
    const double *logpts = P.logpts + ((ch == 1 ? y.nO : 0) + w1) * (P.eval + 1);

    const int GS = y.GS, G = y.G;
    const int g = lane / GS;
    const bool real_lane = g < G;
    const int i = lane % GS;
    const int SR = Lay.sr;  // exchange row length
    const int gbase = real_lane ? g * (GS + 1) : 0;  // padded row of the group in the exchange scratch
    const int ri = real_lane ? gbase + i : SR - 1;
    const int nsub = 4 * P.eval;
    const int rounds = (nsub + G - 1) / G;
    const double c0 = P.corrs[0 + 2 * (ch + 3 * w1)], c1 = P.corrs[1 + 2 * (ch + 3 * w1)];
    Ctl *ctl = S.ctl;

    // Work layout.  A round hands sub-interval r*G+g to group g (its owner), whose lanes keep
    // the Simpson state of their site.  The quadrature nodes pending on all owners of the round
    // form one pool; every wave step gives each group the next pooled node (owners in order,
    // nodes in ascending order), so groups whose sub-interval converged early keep working on
    // the others.  Each owner adds the contributions of its nodes in node order, which keeps
    // the reference's summation order exactly.  A single node_eval call site serves the
    // probe, the pooled steps and the tail corrections.
    enum { M_ROUND, M_STEP, M_TAIL, M_PROBE, M_END };
    int mode = P.dbg_v ? M_PROBE : M_ROUND;
    int r = 0, j = 0, tk = 0;
    long n = 4;             // owner: current level's n (quadrature.jl:2-24)
    int lvl = 0, cnt = 0;   // owner: nodes in the current level, contributions received
    bool init = true, done = true, act = false;
    double a = 0.0, b = 0.0, sgn = 1.0, h = 0.0;
    double b1[NC], b2[NC], res[NC], acc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) b1[k] = b2[k] = res[k] = acc[k] = 0.0;
    unsigned long long nodes = 0, corner_sum = 0;
    int sub_nodes = 0;
    while (mode != M_END) {
        if (mode == M_ROUND) {
            if (r >= rounds) {
                mode = M_TAIL;
                continue;
            }
            j = r * G + g;
            act = real_lane && j < nsub && subinterval(j, P.eval, ref, logpts, a, b, sgn);
            if (__ballot(act) == 0) {  // wave-uniform skip of empty rounds
                ++r;
                continue;
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) b1[k] = b2[k] = res[k] = 0.0;
            done = !act;
            init = true;
            n = 4;
            lvl = act ? 6 : 0;
            cnt = 0;
            sub_nodes = 0;
            mode = M_STEP;
            continue;
        }
        // ---- one node per group
        double x = 0.0, dx = 0.0;
        bool use = false;
        int own = -1, kk = 0;
        if (mode == M_PROBE) {
            x = P.dbg_v[bid];
            use = true;
        } else if (mode == M_TAIL) {
            // boundary corrections (flow_s.jl:37-38), -val then +val; node tk+g for group g
            if (real_lane && tk + g < 2) {
                x = (tk + g == 0) ? -val : val;
                dx = (tk + g == 0) ? c0 : c1;
                use = true;
            }
        } else {
            // publish each owner's state, then claim the g-th pooled node
            if (real_lane && i == 0) {
                Ctl &C = ctl[g];
                C.a = a; C.b = b; C.h = h; C.sgn = sgn;
                C.n = (int)n; C.init = init ? 1 : 0;
                C.issued = cnt; C.pend = done ? 0 : lvl - cnt;
            }
            __syncthreads();
            // group g takes pooled node g: the first owner o with g < (pending nodes of owners
            // before o) + (pending of o).  The pending counts are the same on every lane, so the
            // loop runs on scalar registers and ends once all G nodes are handed out.
            {
                int cum = 0;
                for (int o = 0; o < G && cum < G; ++o) {
                    const int p = __builtin_amdgcn_readfirstlane(ctl[o].pend);
                    if (real_lane && own < 0 && g < cum + p) {
                        own = o;
                        kk = ctl[o].issued + (g - cum);
                    }
                    cum += p;
                }
            }
            if (own >= 0) {
                const Ctl &C = ctl[own];
                const double ca = C.a, cb = C.b, ch_ = C.h;
                if (C.init) {
                    // initial Simpson S1 on (a, m, b) and S2 (simps!, quadrature.jl:57-68)
                    if (kk == 0) { x = ca; dx = 1.0 * (cb - ca) / 6.0; }
                    else if (kk == 1) { x = 0.5 * (ca + cb); dx = 2.0 * (cb - ca) / 3.0; }
                    else if (kk == 2) { x = cb; dx = 1.0 * (cb - ca) / 6.0; }
                    else {
                        const double h2 = 0.5 * (cb - ca) / 2;
                        if (kk == 3) { x = node_at(ca, 2.0 * 1 - 1.0, h2); dx = +4.0 * h2 / 3.0; }
                        else if (kk == 4) { x = node_at(ca, 2.0 * 2 - 1.0, h2); dx = +4.0 * h2 / 3.0; }
                        else { x = node_at(ca, 4.0 * 1 - 2.0, h2); dx = -2.0 * h2 / 3.0; }
                    }
                } else {
                    // residual!: n new nodes, then n/2 re-weighted ones (quadrature.jl:2-24)
                    const int cn = C.n;
                    if (kk < cn) { x = node_at(ca, 2.0 * (kk + 1) - 1.0, ch_); dx = +4.0 * ch_ / 3.0; }
                    else { x = node_at(ca, 4.0 * (kk - cn + 1) - 2.0, ch_); dx = -2.0 * ch_ / 3.0; }
                }
                x *= C.sgn;
                use = true;
            }
        }
        TV Gv[NC];
        int nc;
        node_eval<PASS, TV, NC>(P, S, ch, W, vw, vwp, use ? x : 0.0, i, ri, g, gbase, real_lane, Gv, nc);
        if (mode == M_PROBE) {
            if (g == 0) {
#pragma unroll
                for (int k = 0; k < NC; ++k) P.dbg_out[(NC * bid + k) * y.ns + i] = (double)Gv[k];
            }
            return;
        }
        if (use && real_lane) {
            if (i == 0) ++nodes;
            corner_sum += nc;
        }
        // contributions dv * Γ of every group, in group (= pooled node) order
        double *sc = S.scr;
#pragma unroll
        for (int k = 0; k < NC; ++k) sc[k * SR + ri] = use ? dx * (double)Gv[k] : 0.0;
        if (real_lane && i == 0) S.own[g] = use ? own : -1;
        __syncthreads();
        if (mode == M_TAIL) {
            for (int e = 0; e < G && tk + e < 2; ++e) {
#pragma unroll
                for (int k = 0; k < NC; ++k) acc[k] += sc[k * SR + e * (GS + 1) + i];
            }
            __syncthreads();
            tk += G;
            if (tk >= 2) mode = M_END;
            continue;
        }
        bool need_check = false;
        if (real_lane && !done) {
            for (int e = 0; e < G; ++e) {
                if (__builtin_amdgcn_readfirstlane(S.own[e]) != g) continue;
                if (init && cnt < 3) {
#pragma unroll
                    for (int k = 0; k < NC; ++k) b1[k] += sc[k * SR + e * (GS + 1) + i];
                } else {
                    if (init && cnt == 3) {
#pragma unroll
                        for (int k = 0; k < NC; ++k) b2[k] = 0.5 * b1[k];
                    }
#pragma unroll
                    for (int k = 0; k < NC; ++k) b2[k] += sc[k * SR + e * (GS + 1) + i];
                }
                ++cnt;
                ++sub_nodes;
            }
            if (cnt == lvl) {
                if (!init) n *= 2;
                need_check = true;
            }
        }
        if (__ballot(need_check) != 0) {
            // Richardson extrapolation and error estimate (quadrature.jl:72-88); the norms run
            // over the owner's sites in site order, components fastest, as the reference's do
            double rr[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                rr[k] = b1[k] * (-1.0 / 15.0);
                rr[k] += 16.0 / 15.0 * b2[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                sc[k * SR + ri] = rr[k];
                sc[(NC + k) * SR + ri] = b2[k];
            }
            __syncthreads();
            if (need_check) {
                double q1 = 0.0, q2 = 0.0, q3 = 0.0;
                for (int jj = 0; jj < y.ns; ++jj) {
                    const double *e = sc + gbase + jj;
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        const double x0 = e[k * SR], z0 = e[(NC + k) * SR];
                        q1 += x0 * x0;
                        q2 += z0 * z0;
                        const double d0 = x0 - z0;
                        q3 += d0 * d0;
                    }
                }
                const double n1 = sqrt(q1), n2 = sqrt(q2), adiff = sqrt(q3);
                const double rdiff = adiff / fmax(n1, n2);
                if (adiff < P.atol || rdiff < P.rtol || n > (1L << 20)) {
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        const double d0 = rr[k] - b2[k];
                        res[k] = d0 + b2[k];
                    }
                    done = true;
                    lvl = 0;
                    if (P.dbg_sub_nodes && i == 0 && (long)e_idx * nsub + j < P.dbg_cap) {
                        P.dbg_sub_nodes[(long)e_idx * nsub + j] = sub_nodes;
                        P.dbg_sub_res[(long)e_idx * nsub + j] = res[0];
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        b1[k] = b2[k];
                        b2[k] = 0.5 * b1[k];
                    }
                    init = false;
                    h = 0.5 * (b - a) / n;
                    lvl = (int)(n + n / 2);
                    cnt = 0;
                }
            }
            __syncthreads();
        }
        if (__ballot(real_lane && !done) != 0) continue;
        // round complete: ordered accumulation of its sub-interval results
#pragma unroll
        for (int k = 0; k < NC; ++k) sc[k * SR + ri] = res[k];
        __syncthreads();
        for (int gg = 0; gg < G; ++gg) {
            const int jj = r * G + gg;
            double ta, tb, ts;
            if (jj < nsub && subinterval(jj, P.eval, ref, logpts, ta, tb, ts)) {
#pragma unroll
                for (int k = 0; k < NC; ++k) acc[k] += sc[k * SR + gg * (GS + 1) + i];
            }
        }
        __syncthreads();
        ++r;
        mode = M_ROUND;
    }
    // write the site vector (flow_s.jl:41-51)
    if (g == 0) {
        long base;
        if (kernel == 1) base = (long)y.ns * w1;
        else if (kernel == 2) base = y.q21 + (long)y.ns * (w1 + (long)y.nO * w2);
        else if (kernel == 3) base = y.q22 + (long)y.ns * (w1 + (long)y.nO * w3);
        else base = y.q3 + (long)y.ns * (w1 + (long)y.nO * (w2 + (long)y.nn * w3));
        double *o = P.out + ch * y.chlen + base + i;
#pragma unroll
        for (int k = 0; k < NC; ++k) o[k * y.complen] = acc[k];
    }
    if (P.node_count) {
        // nodes: one count per group; gathered bytes: every lane's corners x NC components
        if (real_lane && i == 0 && nodes) atomicAdd(&s_cnt, nodes);
        if (real_lane && corner_sum) atomicAdd(&s_bytes, corner_sum * (unsigned long long)NC * sizeof(TV));
        __syncthreads();
        if (lane == 0) {
            atomicAdd(P.node_count, s_cnt);
            atomicAdd(P.node_count + 1, s_bytes);
        }
    }
}


// One pass of one problem.  Consecutive blocks go to different XCDs, so each XCD receives an
// even sample of the element list (channels and Ω differ widely in cost).
template <int PASS, typename TV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PFFRG_PASS_WAVES_PER_EU))) void k_vertex_pass(
    PassArgs<TV> P) {
    const int e_idx = blockIdx.x;
    if (!P.dbg_v && e_idx >= P.n_elems) return;
    vertex_pass_body<PASS, TV>(P, e_idx, __builtin_amdgcn_readfirstlane(P.dbg_v ? P.dbg_elem : P.elems[e_idx]));
}

// The same pass of several independent problems in one launch: element code bits 26-31 select
// the problem's arguments in tab (batched coupling sweeps and truncations).
template <int PASS, typename TV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PFFRG_PASS_WAVES_PER_EU))) void k_vertex_pass_batch(
    const PassArgs<TV> *__restrict__ tab, const uint32_t *__restrict__ codes, int n) {
    const int bid = blockIdx.x;
    if (bid >= n) return;
    const uint32_t c = __builtin_amdgcn_readfirstlane(codes[bid]);
    vertex_pass_body<PASS, TV>(tab[c >> 26], bid, c & 0x03ffffffu);
}

// The U(1)-DM pass (six components): three times the accumulators and gathers of the SU(2)
// pass, so it takes the register file of two waves per SIMD.
template <int PASS, typename TV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_vertex_pass_u1(PassArgs<TV> P) {
    const int e_idx = blockIdx.x;
    if (!P.dbg_v && e_idx >= P.n_elems) return;
    vertex_pass_body<PASS, TV, 6>(P, e_idx, __builtin_amdgcn_readfirstlane(P.dbg_v ? P.dbg_elem : P.elems[e_idx]));
}

template <int PASS, typename TV>
hipError_t launch_vertex_pass(const PassArgs<TV> &P, int grid, size_t lds, hipStream_t stream) {
    if (P.y.nc == 6) hipLaunchKernelGGL((k_vertex_pass_u1<PASS, TV>), dim3(grid), dim3(64), lds, stream, P);
    else hipLaunchKernelGGL((k_vertex_pass<PASS, TV>), dim3(grid), dim3(64), lds, stream, P);
    return hipGetLastError();
}

template <int PASS, typename TV>
hipError_t launch_vertex_pass_batch(const PassArgs<TV> *tab, const uint32_t *codes, int n, size_t lds,
                                    hipStream_t stream) {
    hipLaunchKernelGGL((k_vertex_pass_batch<PASS, TV>), dim3(n), dim3(64), lds, stream, tab, codes, n);
    return hipGetLastError();
}

// explicit instantiations live in pass_*.hip; other translation units only link to them
#define PFFRG_PASS_INSTANCES(KW, PASS)                                                                      \
    KW template hipError_t launch_vertex_pass<PASS, double>(const PassArgs<double> &, int, size_t, hipStream_t); \
    KW template hipError_t launch_vertex_pass<PASS, float>(const PassArgs<float> &, int, size_t, hipStream_t);   \
    KW template hipError_t launch_vertex_pass_batch<PASS, double>(const PassArgs<double> *, const uint32_t *, int, \
                                                                  size_t, hipStream_t);                        \
    KW template hipError_t launch_vertex_pass_batch<PASS, float>(const PassArgs<float> *, const uint32_t *, int,   \
                                                                 size_t, hipStream_t);

PFFRG_PASS_INSTANCES(extern, P_KAT)
PFFRG_PASS_INSTANCES(extern, P_LEFT)
PFFRG_PASS_INSTANCES(extern, P_CENTRAL)
PFFRG_PASS_INSTANCES(extern, P_BSE)
PFFRG_PASS_INSTANCES(extern, P_REDUCED)

}  // namespace pffrg
