// pffrg_shared.h -- the shared-node vertex pass of the FP32-mixed mode.
//
// All output elements of one (channel, Ω) block integrate over the same loop-frequency
// sub-intervals (flow_s.jl:31-34) with the same Simpson nodes at every refinement level
// (quadrature.jl:2-139): an element only decides how many levels it uses.  At a node, the
// integrand of element (ν, ν') multiplies a vertex slot whose frequency arguments contain ν but
// not ν' (the left vertex, e.g. flow_su2_s.jl:21-28) with one that contains ν' but not ν (the
// right vertex).  This pass therefore integrates a tile of 16 x 16 elements together: per node
// it interpolates the 32 site vectors once, where the per-element pass interpolates 512.  The
// index ν = n_ν stands for the asymptotic argument (ν = ∞), so the q1, q2 and q3 elements of a
// block share the tiles.
//
// At one node and site the 16 x 16 integrands of a tile are a sum of products of a ν-indexed and
// a ν'-indexed factor: two terms (the spin and density components) in the s and u channels
// (flow_su2_s.jl:41-42), two local terms plus one per overlap row in the t channel
// (flow_su2_t.jl:56-74).  The pass forms them on the matrix cores: v_mfma_f32_16x16x4f32 with the
// tile's ν index as rows, ν' as columns and (node, term) as the inner dimension, so that a batch
// of nodes is summed inside the product.
//
// Every element keeps its own adaptive state: the Simpson sums of its last two levels, and the
// Richardson convergence test of quadrature.jl:54-65 on its own sites.  It stops at its own
// level, exactly when the reference would; a tile keeps refining while any of its elements is
// unconverged, and only the rows and columns of unconverged elements are interpolated.  The
// reweighted nodes of a level (the previous level's new nodes, quadrature.jl:16-21) are not
// evaluated again: their raw sum is kept from the previous level.  Sums over levels and
// sub-intervals, the Simpson sums and the convergence norms are FP64; the vertex copy, the
// interpolation, the integrand products and the sum over one batch of nodes are FP32.  Results
// differ from the per-element FP64 pass by FP32 rounding and summation order only.
//
// Lane l of wave w owns elements (ν row 4 (l/16) + r, ν' column l mod 16), r = 0..3, on the
// wave's sites (at most spt per wave, assigned so that the waves' overlap rows balance): the
// layout of the matrix cores' result.
#pragma once

#include <type_traits>

#include "pffrg_pass.h"

namespace pffrg {

constexpr int SH_MAXT = 512;  // threads per block, at most (8 waves)
constexpr int SH_SPT = 3;     // sites per wave, at most (the pass takes up to 8 * SH_SPT sites)
constexpr int SH_T = 16;      // tile edge (the matrix cores' 16 x 16 result)
constexpr int SH_NPART = 5;   // partial sums per element: the four domains, then the two tail nodes
constexpr int SH_NODES = 256; // nodes whose positions and bubbles are prepared together
constexpr int SH_KMAX = 64;   // nodes per batch, at most (levels with few active vectors), by default
// static LDS of k_shared_pass beside the dynamic block (small tables and flags), with some slack:
// the launch needs dynamic + static <= 160 KB
constexpr int SH_STATIC_LDS = 2048;

typedef float sh_v4f __attribute__((ext_vector_type(4)));

// quadrature nodes of a batch with all 32 row and column vectors active, by default: 8, or 4 at
// three sites per wave (whose per-lane sums take the LDS that more nodes would need); a batch
// holds this many nodes (SharedArgs::kb) times 32 site vectors, so levels with fewer active
// vectors take more nodes per batch, up to SharedArgs::kmax
__host__ __device__ constexpr int shared_batch_nodes(int spt) { return spt >= 3 ? 4 : 8; }

// Descriptor of one buffer lookup in this pass, with everything that does not depend on the site
// resolved once per (node, vector, channel) where it is built: meta as in DescT (bits 0-11) with
// the step to the upper Ω corner in bits 16-23, the offset of the first corner inside the vertex
// copy (channel block included, bit 31 set when the site is taken through the exchange table),
// the steps to the upper ν and ν' corners, and the eight corner weights, zero beyond the
// buffer's own corners and for a disabled or out-of-range channel.  48 bytes, three 16-byte
// LDS reads; a lookup then always reads eight corners (repeated ones at step zero).
struct __align__(16) DescG {
    int meta, base, c2, c3;
    float w[8];
};

__device__ __forceinline__ void desc_shared(const Lds &S, const Sys &y, uint32_t code, int c, double W, double a1,
                                            double a2, double v, DescG &d) {
    const DescParts<double> p = desc_parts<double>(S, y, code, c, W, a1, a2, v);
    const int mc = p.meta;
    const bool map = (mc & 16) != 0;
    const int chn = (c == 1) ? 1 : ((c == 0) == map ? 2 : 0);
    uint32_t c1, c2, c3;
    corner_steps(mc, y, c1, c2, c3);
    const int n = corners(mc & 7);
    d.meta = mc | (int)(c1 << 16);
    d.base = (mc & 7) ? (int)((uint32_t)(chn * y.chlen + p.off) | ((mc & 8) ? 0x80000000u : 0u)) : 0;
    d.c2 = (int)c2;
    d.c3 = (int)c3;
    d.w[0] = n ? (float)(p.a1w * p.a2w * p.a3w) : 0.f;
    d.w[1] = n ? (float)(p.b1w * p.a2w * p.a3w) : 0.f;
    d.w[2] = n >= 4 ? (float)(p.a1w * p.b2w * p.a3w) : 0.f;
    d.w[3] = n >= 4 ? (float)(p.b1w * p.b2w * p.a3w) : 0.f;
    d.w[4] = n >= 8 ? (float)(p.a1w * p.a2w * p.b3w) : 0.f;
    d.w[5] = n >= 8 ? (float)(p.b1w * p.a2w * p.b3w) : 0.f;
    d.w[6] = n >= 8 ? (float)(p.a1w * p.b2w * p.b3w) : 0.f;
    d.w[7] = n >= 8 ? (float)(p.b1w * p.b2w * p.b3w) : 0.f;
}

// offset of the first corner of descriptor d for site i (exchange partner sx)
__device__ __forceinline__ uint32_t desc_site_offset(const DescG &d, int i, int sx) {
    return ((uint32_t)d.base & 0x7fffffffu) + (uint32_t)(d.base < 0 ? sx : i);
}

// the eight corners of descriptor d at corner offset o (repeated corners where a step is zero)
__device__ __forceinline__ void load_corners_g(const float2 *__restrict__ q, const DescG &d, uint32_t o, float2 *x) {
    const uint32_t c1 = ((uint32_t)d.meta >> 16) & 255u, c2 = (uint32_t)d.c2, c3 = (uint32_t)d.c3;
    x[0] = q[o];
    x[1] = q[o + c1];
    x[2] = q[o + c2];
    x[3] = q[o + c1 + c2];
    x[4] = q[o + c3];
    x[5] = q[o + c1 + c3];
    x[6] = q[o + c2 + c3];
    x[7] = q[o + c1 + c2 + c3];
}

// Interpolated (spin, density) pair from n corners, added in the reference's order
// (channel.jl:42-305); corners beyond the buffer's own count carry weight zero.
__device__ __forceinline__ void corner_sum_g(const DescG &d, int n, const float2 *x, float &r0, float &r1) {
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

// row stride (in float2) of the site vectors in LDS: odd, so that rows of neighbouring tile
// positions start in different LDS banks
__host__ __device__ inline int shared_vrow(int ns) { return ns | 1; }

struct SharedArgs {
    PassArgs<float> P;       // tables, FP32 gather copies, Λ, eval, tolerances, output slot
    const uint32_t *units;   // work units (shared_unit codes)
    int n_units;
    int mode;                // 0 half pass (q3 with ν >= ν'), 1 full pass, 2 reduced s bubble
    int NX2, NX3;            // extents of the extended ν and ν' indices (n_ν + 1; reduced pass: NX2 = 1)
    int nw, spt;             // waves per block and sites per wave
    int kb, kmax;            // batch capacity in nodes of 32 vectors, most nodes per batch
    double *part;            // [SH_NPART][3][nO][NX2][NX3][2][ns] partial sums
    unsigned long long *count;  // optional: element-nodes, gathered bytes, phase cycles
    uint32_t *cost;          // optional: each unit's run time (clock ticks) at its place in units
};

// work unit: ch 2 bits | Ω index 7 | tile along ν 4 | tile along ν' 4 | domain 2
__host__ __device__ inline uint32_t shared_unit(int ch, int w1, int t2, int t3, int d) {
    return (uint32_t)ch | ((uint32_t)w1 << 2) | ((uint32_t)t2 << 9) | ((uint32_t)t3 << 13) | ((uint32_t)d << 17);
}

__host__ __device__ inline long shared_part_index(int d, int ch, int w1, int w2x, int w3x, int nO, int NX2, int NX3,
                                                  int ns) {
    return (((((long)d * 3 + ch) * nO + w1) * NX2 + w2x) * (long)NX3 + w3x) * 2L * ns;
}

// element (ν index w2x, ν' index w3x) of the extended grid belongs to the pass's element list
// (compute.jl; SDE.jl:156-166): index n_ν is the asymptotic argument
__host__ __device__ inline bool shared_in_pass(int mode, int w2x, int w3x, int nn) {
    if (mode == 1 || mode == 2) return true;
    return w2x == nn || w3x == nn || w2x >= w3x;
}

// dynamic LDS of k_shared_pass (byte offsets), sized to the problem.  The descriptors, site
// vectors and local vectors of a batch share their space with the cross-wave norm sums of a
// convergence test, which never overlaps a batch.
struct SharedLds {
    int mp, o_sig, o_mf, o_ov, o_ovoff, o_exch, o_hint, o_desc, o_V, o_Z, o_loc, o_red, o_nodes, o_b1, total;
};

__host__ __device__ inline SharedLds shared_lds(const Sys &y, int n_ov, int nw, bool sig_d, int spt, int kb) {
    SharedLds L;
    int mp = y.nsig;
    if (y.nO > mp) mp = y.nO;
    if (y.nn > mp) mp = y.nn;
    L.mp = (mp + 1) & ~1;
    int o = 10 * L.mp * 8;
    L.o_sig = o; o += (sig_d ? 2 : 1) * L.mp * 8;
    L.o_mf = o; o += 12 * L.mp * 4;
    L.o_ov = o; o += n_ov * 8;  // (row, multiplicity factor) pairs
    L.o_ovoff = o; o += ((y.ns + 2) & ~1) * 4;
    L.o_exch = o; o += ((y.ns + 1) & ~1) * 4;
    L.o_hint = o; o += ((HINT_WORDS + 1) & ~1) * 4;
    o = (o + 15) & ~15;
    const int base = o;
    const int K = kb;
    L.o_desc = o; o += K * 4 * SH_T * 3 * (int)sizeof(DescG);
    o = (o + 15) & ~15;
    L.o_V = o; o += K * 2 * SH_T * shared_vrow(y.ns) * 8;
    L.o_Z = o; o += shared_vrow(y.ns) * 8;  // a row of zeros: the site vector of an inactive position
    L.o_loc = o; o += K * 2 * SH_T * 8;
    L.o_red = base;
    const int red_end = base + nw * SH_T * SH_T * 3 * 8;
    if (red_end > o) o = red_end;
    o = (o + 15) & ~15;
    L.o_nodes = o; o += SH_NODES * (8 + 8 + 4 + 4);  // positions, weights, bubbles, reweight flags
    L.o_b1 = o; o += nw * 64 * spt * 2 * 4 * 4;    // each lane's previous-level Simpson sums (FP32)
    L.total = o;
    return L;
}

// Vertex of one slot at one site from its three channel descriptors (get_Γ_avx!,
// action_su2.jl:95-134): bare + Σ_channels sign·interpolation.
__device__ __forceinline__ void gather_site(const DescG *ds, const float *gamma, const int *exch, int site,
                                            float &vs, float &vd) {
    const float2 *q = reinterpret_cast<const float2 *>(gamma);
    DescG d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = ds[c];
    // all 24 corners in flight together (eight per channel, zero weights beyond a buffer's own)
    float2 x[3][8];
    const int sx = exch[site];
#pragma unroll
    for (int c = 0; c < 3; ++c) load_corners_g(q, d[c], desc_site_offset(d[c], site, sx), x[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float r0, r1;
        corner_sum_g(d[c], 8, x[c], r0, r1);
        vs += r0;
        vd += (d[c].meta & 32) ? -r1 : r1;
    }
}

// Phase timing of the shared-node pass (a measurement build with PFFRG_SHARED_PROFILE defined):
// thread 0 of every block adds the shader-clock cycles of each phase to count[2 + phase] (units of
// the s and u channels) or count[2 + PR_N + phase] (t channel).
#ifdef PFFRG_SHARED_PROFILE
#define SH_PROF_T0() const unsigned long long _t0 = (t == 0) ? clock64() : 0ull
#define SH_PROF_ADD(k) do { if (t == 0 && A.count) s_prof[prof_base + (k)] += clock64() - _t0; } while (0)
#else
#define SH_PROF_T0() do { } while (0)
#define SH_PROF_ADD(k) do { } while (0)
#endif
enum { PR_ACTIVE, PR_PREP, PR_DESC, PR_GATHER, PR_CONTRACT, PR_CHECK, PR_BATCHES, PR_TOTAL, PR_N };

// One work unit (code uc) of one problem.  Every thread of the block calls it.  SPT: the largest
// number of sites per wave the launch serves (the per-lane quadrature state is sized by it).
template <int PASS, int SPT>
__device__ __forceinline__ void shared_pass_body(const SharedArgs &A, uint32_t uc, uint32_t *cost_slot) {
    const unsigned long long rt_unit = wall_clock64();
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ uint32_t s_code[MAX_DESC];
    __shared__ double s_a1[SH_T], s_a2[SH_T];
    __shared__ int s_act2[SH_T], s_act3[SH_T], s_nact[2];
    __shared__ int s_rowf[SH_T], s_colf[SH_T], s_rmap[SH_T], s_cmap[SH_T], s_kb;
    __shared__ int s_site[8 * SH_SPT], s_nsw[8];
    __shared__ uint8_t s_conv[SH_T * SH_T];
    __shared__ unsigned long long s_cnt[4];
    __shared__ unsigned long long s_prof[2 * PR_N + 13];

    const PassArgs<float> &P = A.P;
    const Sys &y = P.y;
    const int t = threadIdx.x;
    const int NT = blockDim.x;
    const int NTp = 64 * A.nw;
    const int lane = t & 63;
    const int wv = t >> 6;  // this thread's wave
    const int nw = A.nw;
    const int n_ov = P.ov_off[y.ns];
    // the LDS layout and batch size are the problem's own (A.spt), whatever SPT a batched launch uses
    const SharedLds Lay = shared_lds(y, n_ov, nw, PASS == P_KAT, A.spt, A.kb);
    const int mp = Lay.mp;
    double *s_mesh = reinterpret_cast<double *>(s_dyn);
    double *s_sigA = reinterpret_cast<double *>(s_dyn + Lay.o_sig);
    double *s_sigD = (PASS == P_KAT) ? s_sigA + mp : s_sigA;
    uint2 *s_ov = reinterpret_cast<uint2 *>(s_dyn + Lay.o_ov);
    int *s_ovoff = reinterpret_cast<int *>(s_dyn + Lay.o_ovoff);
    int *s_exch = reinterpret_cast<int *>(s_dyn + Lay.o_exch);
    int32_t *s_hint = reinterpret_cast<int32_t *>(s_dyn + Lay.o_hint);
    float *mf = reinterpret_cast<float *>(s_dyn + Lay.o_mf);
    DescG *s_desc = reinterpret_cast<DescG *>(s_dyn + Lay.o_desc);
    float2 *s_V = reinterpret_cast<float2 *>(s_dyn + Lay.o_V);
    float2 *s_Z = reinterpret_cast<float2 *>(s_dyn + Lay.o_Z);
    float2 *s_loc = reinterpret_cast<float2 *>(s_dyn + Lay.o_loc);
    double *s_red = reinterpret_cast<double *>(s_dyn + Lay.o_red);
    double *s_xs = reinterpret_cast<double *>(s_dyn + Lay.o_nodes);  // node positions
    double *s_wq = s_xs + SH_NODES;                                   // node weights
    float *s_tp = reinterpret_cast<float *>(s_wq + SH_NODES);         // bubbles
    int *s_rw = reinterpret_cast<int *>(s_tp + SH_NODES);             // node sums kept for the next level
    const int VR = shared_vrow(y.ns);
    // this lane's previous-level Simpson sums, output o at b1[o * NTp] (consecutive lanes adjacent);
    // NTp = the problem's own threads (a batched launch may give the block more, which hold no sites)
    float *b1 = reinterpret_cast<float *>(s_dyn + Lay.o_b1) + t;

    // ---- tables in LDS (as the per-element pass stages them)
    for (int k = t; k < 10 * mp; k += NT) s_mesh[k] = P.mesh[(k / mp) * MAXN + (k % mp)];
    for (int k = t; k < y.nsig; k += NT) {
        s_sigA[k] = P.A.Sigma[k];
        if (PASS == P_KAT) s_sigD[k] = P.D.Sigma[k];
    }
    for (int k = t; k < 10 * mp; k += NT) mf[k] = (float)P.mesh[(k / mp) * MAXN + (k % mp)];
    for (int k = t; k < y.nsig; k += NT) {
        mf[10 * mp + k] = (float)P.A.Sigma[k];
        mf[11 * mp + k] = (PASS == P_KAT) ? (float)P.D.Sigma[k] : 0.0f;
    }
    for (int k = t; k < y.ns; k += NT) s_exch[k] = P.exch[k];
    {
        // overlap rows (j1 | j2 << 8 | m << 16) beside the FP32 factor -2 S² m of their term
        // (flow_su2_t.jl:59-74), each site's rows ordered by j1 (stable: by table position
        // within one j1) and the last row of every j1 run flagged in bit 31, so that the
        // contraction sums the ν' vectors of a run before one product with the ν vector of j1
        const float mfac = (float)(-2.0 * y.S2);
        for (int k = t; k < n_ov; k += NT) {
            const uint32_t r = P.ov_rows[k];
            int i = 0;
            while (P.ov_off[i + 1] <= k) ++i;
            const int j1 = (int)(r & 255u);
            int pos = P.ov_off[i];
            bool last = true;
            for (int q = P.ov_off[i]; q < P.ov_off[i + 1]; ++q) {
                const int jq = (int)(P.ov_rows[q] & 255u);
                pos += (jq < j1 || (jq == j1 && q < k)) ? 1 : 0;
                last = last && !(jq == j1 && q > k);
            }
            s_ov[pos] = uint2{r | (last ? 0x80000000u : 0u), __float_as_uint(mfac * (float)((r >> 16) & 0x7fffu))};
        }
    }
    for (int k = t; k <= y.ns; k += NT) s_ovoff[k] = P.ov_off[k];
    {
        const int32_t *gh = reinterpret_cast<const int32_t *>(P.mesh + 10 * MAXN);
        for (int k = t; k < HINT_WORDS; k += NT) s_hint[k] = gh[k];
    }
    if (t < 4) s_cnt[t] = 0;
    for (int k = t; k < VR; k += NT) s_Z[k] = float2{0.f, 0.f};
    if (t < 2 * PR_N + 13) s_prof[t] = 0;
#ifdef PFFRG_SHARED_PROFILE
    const unsigned long long t_begin = clock64(), rt_begin = wall_clock64();
#endif

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
    S.ov = nullptr;  // the contraction reads the rows from s_ov
    S.ov_off = s_ovoff;
    S.hint = reinterpret_cast<const int16_t *>(s_hint);
    S.mf = MeshV<float>{mf, mf + mp, mf + 2 * mp, mf + 3 * mp, mf + 4 * mp, mf + 5 * mp, mf + 10 * mp, mf + 11 * mp,
                        mp, nullptr};

    // ---- the unit
    const int ch = uc & 3, w1 = (uc >> 2) & 127, t2 = (uc >> 9) & 15, t3 = (uc >> 13) & 15, dom = (uc >> 17) & 3;
#ifdef PFFRG_SHARED_PROFILE
    const int prof_base = (ch == 1) ? PR_N : 0;
#endif
    const int nn = y.nn;
    if (t < MAX_DESC) s_code[t] = slot_code<PASS>(ch, t / 3, t % 3);
    if (t == 0) {
        // sites to waves, at most spt each, balancing the overlap rows that the t-channel
        // contraction of a site runs through: the site with the most rows first, each to the
        // least-loaded wave with room (lowest index on ties)
        int load[8], cntw[8];
        uint32_t used = 0u;
        for (int w = 0; w < nw; ++w) { load[w] = 0; cntw[w] = 0; }
        for (int n = 0; n < y.ns; ++n) {
            int best = 0, br = -1;
            for (int j = 0; j < y.ns; ++j) {
                const int rj = P.ov_off[j + 1] - P.ov_off[j];
                if (!((used >> j) & 1u) && rj > br) { br = rj; best = j; }
            }
            used |= 1u << best;
            int bw = -1;
            for (int w = 0; w < nw; ++w)
                if (cntw[w] < A.spt && (bw < 0 || load[w] < load[bw])) bw = w;
            s_site[bw * SH_SPT + cntw[bw]] = best;
            cntw[bw] += 1;
            load[bw] += br;
        }
        for (int w = 0; w < 8; ++w) s_nsw[w] = w < nw ? cntw[w] : 0;
    }
    const double INF = __builtin_huge_val();
    const double *Om = (PASS == P_REDUCED || ch != 1) ? S.Os : S.Ot;
    const double *nu = (PASS == P_REDUCED || ch != 1) ? S.ns : S.nt;
    __syncthreads();
    const double W = wave_uniform(Om[w1]);
    if (t < SH_T) {
        const int w2x = t2 * SH_T + t, w3x = t3 * SH_T + t;
        s_a1[t] = (PASS == P_REDUCED) ? 0.0 : (w2x >= nn ? INF : nu[w2x]);
        s_a2[t] = (w3x >= nn) ? INF : nu[w3x];
    }
    const double ref = P.L + 0.5 * W;
    const double val = (PASS != P_REDUCED && ch == 1) ? y.valT : y.valS;
    const double *logpts = P.logpts + ((PASS != P_REDUCED && ch == 1 ? y.nO : 0) + w1) * (P.eval + 1);
    const int nslot = (ch == 1) ? 4 : 2;

    // ---- this lane's elements (4 rows, one column of the tile) and sites
    const int e3 = lane & 15, r2 = 4 * (lane >> 4);
    const int w3x = t3 * SH_T + e3;
    bool valid[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int w2x = t2 * SH_T + r2 + r;
        valid[r] = w2x < A.NX2 && w3x < A.NX3 && shared_in_pass(A.mode, w2x, w3x, nn);
    }
    auto part_at = [&](int d, int r) {
        return A.part + shared_part_index(d, ch, w1, t2 * SH_T + r2 + r, w3x, y.nO, A.NX2, A.NX3, y.ns);
    };
    // sites of this wave (wave-uniform)
    const int nsw = __builtin_amdgcn_readfirstlane(s_nsw[wv]);
    int msite[SPT];
#pragma unroll
    for (int s = 0; s < SPT; ++s) msite[s] = __builtin_amdgcn_readfirstlane(s < nsw ? s_site[wv * SH_SPT + s] : 0);
#pragma unroll
    for (int r = 0; r < 4; ++r)
        if (valid[r]) {
            double *pp = part_at(dom, r);
#pragma unroll
            for (int s = 0; s < SPT; ++s) {
                if (s >= nsw) break;
                pp[msite[s]] = 0.0;
                pp[y.ns + msite[s]] = 0.0;
            }
        }

    // per-lane quadrature state: output o = (site slot s, component c, element r) at (s * 2 + c) * 4 + r
    constexpr int NO = SPT * 2 * 4;
    const int Kcap = A.kb * 32;  // site vectors of one batch (node x vector)
    double b2[NO];
    float Pp[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) { b2[o] = 0.0; Pp[o] = 0.f; }
    bool act[4] = {false, false, false, false};
    unsigned long long my_nodes = 0;

    // Positions, weights and bubbles of up to SH_NODES nodes, prepared together (one thread per
    // node): node q of the chunk is gen(q); rw marks the nodes whose raw sum the next level
    // reweights.  The bubble is the propagator product of the integrand (bubbles.jl:64-96; the
    // symmetrised Katanin sum in the Katanin pass, flow_su2_s.jl:18).
    auto prep_nodes = [&](int n, auto gen) {
        SH_PROF_T0();
        for (int q = t; q < n; q += NT) {
            double x, w;
            int rw;
            gen(q, x, w, rw);
            double x1, x2, G1, K1, G2, K2;
            bubble_args<double>(ch, W, x, x1, x2);
            bubble_leg<PASS == P_KAT, double>(P.L, x1, meshv<double>(S), y.nsig, G1, K1);
            bubble_leg<PASS == P_KAT, double>(P.L, x2, meshv<double>(S), y.nsig, G2, K2);
            const double twopi = 2.0 * PI_D;
            s_xs[q] = x;
            s_wq[q] = w;
            s_rw[q] = rw;
            s_tp[q] = (float)((PASS == P_KAT) ? K1 * G2 / twopi + K2 * G1 / twopi : -(G1 * G2 / twopi));
        }
        __syncthreads();
        SH_PROF_ADD(PR_PREP);
    };

    // Evaluate the integrands of the tile at the prepared nodes q0 .. q0+nb-1, which share one
    // weight w and reweight flag, and accumulate b2 += w Σ_q f_q (FP64) and, when flagged,
    // Pp += Σ_q f_q.  The active rows and columns are numbered compactly (rows first): vector
    // v of node k has its site vector at s_V[(k nAS + v) VR + i], its descriptors (slot pair pr)
    // at s_desc[((k npair + pr) nAS + v) 3 + c] and its t-channel local vertex at s_loc[k nAS + v].
    auto eval_batch = [&](int q0, int nb) {
#ifdef PFFRG_SHARED_PROFILE
        unsigned long long _tb = (t == 0) ? clock64() : 0ull;
        if (t == 0 && A.count) s_prof[prof_base + PR_BATCHES] += 1;
        if (wv == 0 && A.count) {
            // active elements of the tile in this batch: a histogram and their sum
            int na = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) na += __popcll(__ballot(act[r]));
            if (t == 0) {
                const int bk = na <= 4 ? 0 : (na <= 16 ? 1 : (na <= 64 ? 2 : (na <= 128 ? 3 : 4)));
                s_prof[2 * PR_N + bk] += 1;
                s_prof[2 * PR_N + 5] += (unsigned long long)na * nb;
                s_prof[2 * PR_N + 6] += (unsigned long long)(s_nact[0] + s_nact[1]) * nb;
            }
        }
#define SH_PROF_LAP(k) do { if (t == 0 && A.count) { const unsigned long long _n = clock64(); s_prof[prof_base + (k)] += _n - _tb; _tb = _n; } } while (0)
#else
#define SH_PROF_LAP(k) do { } while (0)
#endif
        const int nA2 = s_nact[0], nA3 = s_nact[1];
        const int nAS = nA2 + nA3;
        const int npair = nslot / 2;
        // descriptors (node, slot pair, vector, channel)
        {
            const int per_k = 3 * npair * nAS;
            const int ntask = nb * per_k;
            // the gathered-bytes counter: each descriptor of pair 0 is read on every site, of pair
            // 1 (the t channel's local vertices) on site 0
            unsigned long long cbd = 0;
#if defined(PFFRG_SHARED_PROFILE) && defined(PFFRG_SHARED_PROFILE_DESC)
            const unsigned long long _d0 = clock64();
            int _nd = 0;
#endif
            for (int tau = t; tau < ntask; tau += NT) {
#if defined(PFFRG_SHARED_PROFILE) && defined(PFFRG_SHARED_PROFILE_DESC)
                ++_nd;
#endif
                const int k = tau / per_k, r = tau - k * per_k;
                const int c = r % 3, q = r / 3;
                const int pr = q / nAS, v = q - pr * nAS;
                const bool row = v < nA2;
                const int s = 2 * pr + (row ? 0 : 1);
                const int ee = row ? s_act2[v] : s_act3[v - nA2];
                DescG &dg = s_desc[((k * npair + pr) * nAS + v) * 3 + c];
                desc_shared(S, y, s_code[3 * s + c], c, W, row ? s_a1[ee] : 0.0, row ? 0.0 : s_a2[ee], s_xs[q0 + k], dg);
                if (A.count) cbd += (unsigned long long)corners(dg.meta & 7) * (pr == 0 ? y.ns : 1);
            }
            if (A.count && cbd) atomicAdd(&s_cnt[1], cbd * 2ull * sizeof(float));
#if defined(PFFRG_SHARED_PROFILE) && defined(PFFRG_SHARED_PROFILE_DESC)
            // thread 0's own descriptor work (no barrier) and its task count; the most tasks a thread has
            if (t == 0 && A.count) {
                s_prof[2 * PR_N + 7 + (ch == 1 ? 2 : 0)] += clock64() - _d0;
                s_prof[2 * PR_N + 8 + (ch == 1 ? 2 : 0)] += (unsigned long long)_nd;
            }
#endif
        }
        __syncthreads();
#if defined(PFFRG_SHARED_PROFILE) && defined(PFFRG_SHARED_PROFILE_PAIRS)
        if (A.count) {
            // node pairs (k-1, k) of the batch whose descriptors of one (slot, vector) read the
            // same corners in every channel: [0] pairs, [1] same in all channels, [2] channels same
            const int per_k = npair * nAS;
            unsigned long long np_ = 0, ns_ = 0, nc_ = 0;
            for (int tau = t; tau < (nb - 1) * per_k; tau += NT) {
                const int k = 1 + tau / per_k, q = tau % per_k;
                const DescG *d1 = s_desc + (k * per_k + q) * 3;
                const DescG *d0 = s_desc + ((k - 1) * per_k + q) * 3;
                int same = 0;
                for (int c = 0; c < 3; ++c) same += (d1[c].meta == d0[c].meta && d1[c].base == d0[c].base);
                np_ += 1; ns_ += (same == 3); nc_ += same;
            }
            if (np_) { atomicAdd(&s_prof[2 * PR_N + 7], np_); atomicAdd(&s_prof[2 * PR_N + 8], ns_); atomicAdd(&s_prof[2 * PR_N + 9], nc_); }
        }
#endif
        SH_PROF_LAP(PR_DESC);
        // site vectors of slots 0 (rows) and 1 (columns) on every site, and of the t channel's
        // local slots 2 and 3 on site 0.  A thread takes one (vector, site) through a run of the
        // batch's nodes: consecutive nodes mostly interpolate in the same mesh cells (the same
        // corner offsets and flags in all three channels), and then the corners loaded for the
        // previous node are reused with the new node's weights.
        {
            const int ntask = nAS * y.ns;
            const bool rowD = (vertex_cfg<PASS>(ch, 0) & 8) != 0, colD = (vertex_cfg<PASS>(ch, 1) & 8) != 0;
            const float2 *qA = reinterpret_cast<const float2 *>(P.A.gamma);
            const float2 *qD = reinterpret_cast<const float2 *>(P.D.gamma);
            // the batch's nodes of one task are cut into R runs (R threads per task) when there
            // are fewer tasks than threads, so that sparse batches keep every thread busy
            int R = 1;
            while (R < nb && 2 * R * ntask <= NT) R *= 2;
            const int C = (nb + R - 1) / R;
            for (int tau = t; tau < ntask * R; tau += NT) {
                const int task = tau / R, run = tau - task * R;
                const int v = task / y.ns, i = task - v * y.ns;
                const int s = v < nA2 ? 0 : 1;
                const bool useD = s ? colD : rowD;
                const float2 *q = useD ? qD : qA;
                const double *bare = useD ? P.D.bare : P.A.bare;
                const float b0 = (float)bare[i], b1v = (float)bare[y.ns + i];
                const int sx0 = s_exch[i];
                float2 x[3][8];
                int pm[3] = {-1, -1, -1}, po[3] = {0, 0, 0};
                const int k1 = (run + 1) * C < nb ? (run + 1) * C : nb;
                for (int k = run * C; k < k1; ++k) {
#if defined(PFFRG_SHARED_PROFILE) && defined(PFFRG_SHARED_PROFILE_CELLS)
                    const unsigned long long _g0 = clock64();
                    bool _rl = false;
#endif
                    // the three descriptors are read together, then the corners of every channel
                    // whose cell changed are loaded (all channels' loads in flight together), then
                    // summed
                    // the three descriptors are read together; when any channel's cell changed,
                    // all 24 corners are loaded (in flight together), then summed with the node's
                    // weights in straight-line code
                    const DescG *ds = s_desc + (k * npair * nAS + v) * 3;
                    DescG d[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) d[c] = ds[c];
                    const bool rl = d[0].meta != pm[0] || d[0].base != po[0] || d[1].meta != pm[1] ||
                                    d[1].base != po[1] || d[2].meta != pm[2] || d[2].base != po[2];
                    if (rl) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            load_corners_g(q, d[c], desc_site_offset(d[c], i, sx0), x[c]);
                            pm[c] = d[c].meta;
                            po[c] = d[c].base;
                        }
#if defined(PFFRG_SHARED_PROFILE) && defined(PFFRG_SHARED_PROFILE_CELLS)
                        _rl = true;
#endif
                    }
                    float vs = b0, vd = b1v;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        float r0, r1;
                        corner_sum_g(d[c], 8, x[c], r0, r1);
                        vs += r0;
                        vd += (d[c].meta & 32) ? -r1 : r1;
                    }
                    if (s == 1) {
                        // every term's ν' factor carries the node's -tp
                        const float mt = -s_tp[q0 + k];
                        vs *= mt;
                        vd *= mt;
                    }
                    s_V[(k * nAS + v) * VR + i] = float2{vs, vd};
#if defined(PFFRG_SHARED_PROFILE) && defined(PFFRG_SHARED_PROFILE_CELLS)
                    if (A.count) {
                        // wave node steps, those where a lane reloaded corners, and their cycles
                        const unsigned long long am = __ballot(1), rb = __ballot(_rl);
                        if (lane == __ffsll((long long)am) - 1) {
                            atomicAdd(&s_prof[2 * PR_N + 7], 1ull);
                            atomicAdd(&s_prof[2 * PR_N + 8], rb ? 1ull : 0ull);
                            atomicAdd(&s_prof[2 * PR_N + (rb ? 9 : 10)], clock64() - _g0);
                        }
                    }
#endif
                }
            }
            if (ch == 1) {
                const bool locD2 = (vertex_cfg<PASS>(ch, 2) & 8) != 0, locD3 = (vertex_cfg<PASS>(ch, 3) & 8) != 0;
                const int ntl = nb * nAS;
                for (int tau = t; tau < ntl; tau += NT) {
                    const int k = tau / nAS, v = tau - k * nAS;
                    const int s = v < nA2 ? 2 : 3;
                    const ActV<float> &V = (s == 2 ? locD2 : locD3) ? P.D : P.A;
                    float vs = (float)V.bare[0], vd = (float)V.bare[y.ns];
                    gather_site(s_desc + ((k * 2 + 1) * nAS + v) * 3, V.gamma, s_exch, 0, vs, vd);
                    if (s == 3) {
                        const float mt = -s_tp[q0 + k];
                        vs *= mt;
                        vd *= mt;
                    }
                    s_loc[k * nAS + v] = float2{vs, vd};
                }
                // the t contraction sums over the node slots of whole groups of four: those beyond
                // the batch hold zeros
                const int ngr = 4 * ((nb + 3) / 4);
                if (nb < ngr) {
                    const int nz = (ngr - nb) * nAS;
                    for (int tau = t; tau < nz * y.ns; tau += NT) {
                        const int kv = tau / y.ns, i = tau - kv * y.ns;
                        s_V[(nb * nAS + kv) * VR + i] = float2{0.f, 0.f};
                    }
                    for (int tau = t; tau < nz; tau += NT) s_loc[nb * nAS + tau] = float2{0.f, 0.f};
                }
            }
        }
        __syncthreads();
        SH_PROF_LAP(PR_GATHER);
        // the tile's integrands on this wave's sites, summed over the batch on the matrix cores:
        // A rows = ν positions, B columns = ν' positions, inner index = (node, term).  A row or
        // column without an active element reads the zero row (node stride 0).
        {
            const double wq = s_wq[q0];
            const bool addp = s_rw[q0] != 0;
            const int pos = lane & 15, kq = lane >> 4;
            const int vr = s_rmap[pos], vc = s_cmap[pos];
            const float2 *Rb = vr >= 0 ? s_V + vr * VR : s_Z;
            const float2 *Cb = vc >= 0 ? s_V + (nA2 + vc) * VR : s_Z;
            const int Rs = vr >= 0 ? nAS * VR : 0, Cs = vc >= 0 ? nAS * VR : 0;
            // the t channel with at most 8 active rows and 8 active columns runs on 4 x 4 blocks
            // (v_mfma_f32_4x4x1f32, 16 blocks per product): block b = lane / 4 holds node
            // kb4 = b / nsub4 of the products' node group and 4 x 4 sub-tile st4 = b mod nsub4 of the
            // compact active rows and columns; lane 4b + e supplies row u4 = 4 p + e to A and column
            // v4 = 4 q + e to B, and receives D[b][j][e] (row 4 p + j, column 4 q + e) in register j
            const bool sparse4 = ch == 1 && nA2 <= 8 && nA3 <= 8;
            const int nq4 = (nA3 + 3) >> 2, nsub4 = ((nA2 + 3) >> 2) * nq4, npi4 = 16 / (nsub4 > 0 ? nsub4 : 1);
            const int kb4 = (lane >> 2) / (nsub4 > 0 ? nsub4 : 1), st4 = (lane >> 2) - kb4 * nsub4;
            const int u4 = 4 * (st4 / (nq4 > 0 ? nq4 : 1)) + (lane & 3), v4 = 4 * (st4 % (nq4 > 0 ? nq4 : 1)) + (lane & 3);
            const bool ua4 = u4 < nA2, va4 = v4 < nA3;
#pragma unroll
            for (int s = 0; s < SPT; ++s) {
                if (s >= nsw) break;
                const int i = msite[s];
                sh_v4f Ds = {0.f, 0.f, 0.f, 0.f}, Dd = {0.f, 0.f, 0.f, 0.f};
                if (sparse4) {
                    sh_v4f Es = {0.f, 0.f, 0.f, 0.f}, Ed = {0.f, 0.f, 0.f, 0.f};
                    const int r0 = __builtin_amdgcn_readfirstlane(s_ovoff[i]);
                    const int r1 = __builtin_amdgcn_readfirstlane(s_ovoff[i + 1]);
                    for (int k0 = 0; k0 < nb; k0 += npi4) {
                        const int k = k0 + kb4;
                        const bool kon = k < nb;
                        const float2 *R0 = (kon && ua4) ? s_V + (k * nAS + u4) * VR : s_Z;
                        const float2 *C0 = (kon && va4) ? s_V + (k * nAS + nA2 + v4) * VR : s_Z;
                        {
                            // the two local terms (flow_su2_t.jl:56-58)
                            const float2 v1i = R0[i], v2i = C0[i];
                            const float2 l3 = (kon && ua4) ? s_loc[k * nAS + u4] : float2{0.f, 0.f};
                            const float2 l4 = (kon && va4) ? s_loc[k * nAS + nA2 + v4] : float2{0.f, 0.f};
                            Es = __builtin_amdgcn_mfma_f32_4x4x1f32(v1i.x, -1.0f * l4.x + l4.y, Es, 0, 0, 0);
                            Ed = __builtin_amdgcn_mfma_f32_4x4x1f32(v1i.y, 3.0f * l4.x + l4.y, Ed, 0, 0, 0);
                            Es = __builtin_amdgcn_mfma_f32_4x4x1f32(-1.0f * l3.x + l3.y, v2i.x, Es, 0, 0, 0);
                            Ed = __builtin_amdgcn_mfma_f32_4x4x1f32(3.0f * l3.x + l3.y, v2i.y, Ed, 0, 0, 0);
                        }
                        // the overlap rows in j1 runs, as in the 16 x 16 form
                        float bx = 0.f, by = 0.f;
                        for (int cb = r0; cb < r1; cb += 64) {
                            const int nr = (r1 - cb) < 64 ? r1 - cb : 64;
                            const uint2 rl = (lane < nr) ? s_ov[cb + lane] : uint2{0u, 0u};
                            const int rj = (int)rl.x, rm = (int)rl.y;
                            auto step = [&](uint32_t jj, float mm, float2 f2) {
                                bx += mm * f2.x;
                                by += mm * f2.y;
                                if (jj >> 31) {
                                    const float2 e1 = R0[jj & 255];
                                    Es = __builtin_amdgcn_mfma_f32_4x4x1f32(e1.x, bx, Es, 0, 0, 0);
                                    Ed = __builtin_amdgcn_mfma_f32_4x4x1f32(e1.y, by, Ed, 0, 0, 0);
                                    bx = 0.f;
                                    by = 0.f;
                                }
                            };
                            int r = 0;
                            for (; r + 4 <= nr; r += 4) {
                                uint32_t jj[4];
                                float mm[4];
                                float2 f2[4];
#pragma unroll
                                for (int u = 0; u < 4; ++u) {
                                    jj[u] = (uint32_t)__builtin_amdgcn_readlane(rj, r + u);
                                    mm[u] = __int_as_float(__builtin_amdgcn_readlane(rm, r + u));
                                    f2[u] = C0[(jj[u] >> 8) & 255];
                                }
#pragma unroll
                                for (int u = 0; u < 4; ++u) step(jj[u], mm[u], f2[u]);
                            }
                            for (; r < nr; ++r) {
                                const uint32_t jj = (uint32_t)__builtin_amdgcn_readlane(rj, r);
                                const float mm = __int_as_float(__builtin_amdgcn_readlane(rm, r));
                                step(jj, mm, C0[(jj >> 8) & 255]);
                            }
                        }
                    }
                    // the blocks of one sub-tile (lanes 4 nsub4 apart) summed over their nodes
                    for (int off = 4 * nsub4; off < 64; off <<= 1) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            Es[r] += __shfl_xor(Es[r], off);
                            Ed[r] += __shfl_xor(Ed[r], off);
                        }
                    }
                    // each owner lane (tile row 4 (lane / 16) + r, column lane mod 16) reads its element
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int ur = s_rmap[4 * (lane >> 4) + r], vcol = s_cmap[lane & 15];
                        const bool on = ur >= 0 && vcol >= 0;
                        const int src = on ? 4 * ((ur >> 2) * nq4 + (vcol >> 2)) + (vcol & 3) : 0;
                        const int jr = ur & 3;
                        float gs = 0.f, gd = 0.f;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const float ts = __shfl(Es[j], src), td = __shfl(Ed[j], src);
                            if (j == jr) {
                                gs = ts;
                                gd = td;
                            }
                        }
                        Ds[r] = on ? gs : 0.f;
                        Dd[r] = on ? gd : 0.f;
                    }
                } else if (ch != 1) {
                    // s and u (flow_su2_s.jl:41-42, flow_su2_u.jl:41-42): Γs = v1s (-tp (∓2 v2s + v2d))
                    // + v1d (-tp v2s), Γd = v1s (-3 tp v2s) + v1d (-tp v2d), with -tp already in
                    // v2; two nodes per product
                    const float sgn2 = (ch == 0) ? -2.0f : 2.0f;
                    for (int st = 0; st < (nb + 1) / 2; ++st) {
                        const int kap = 4 * st + kq;
                        const int n = kap >> 1, tau = kap & 1;
                        float as = 0.f, bs = 0.f, bd = 0.f;
                        if (n < nb) {
                            const float2 v1 = Rb[n * Rs + i];
                            const float2 v2 = Cb[n * Cs + i];
                            as = tau ? v1.y : v1.x;
                            bs = tau ? v2.x : sgn2 * v2.x + v2.y;
                            bd = tau ? v2.y : 3.0f * v2.x;
                        }
                        Ds = __builtin_amdgcn_mfma_f32_16x16x4f32(as, bs, Ds, 0, 0, 0);
                        Dd = __builtin_amdgcn_mfma_f32_16x16x4f32(as, bd, Dd, 0, 0, 0);
                    }
                } else {
                    // t (flow_su2_t.jl:56-74), the batch's nodes on the inner index: two local terms
                    // with the site-0 vertices v3 (ν) and v4 (ν'), then one term per overlap row
                    // (j1, j2, m) of site i, whose factor -2 S² m sits beside the row; the ν'
                    // vertices carry -tp, and node slots beyond the batch hold zeros.  The nodes
                    // go through the rows eight at a time (two groups of four).
                    const int ng = (nb + 3) / 4;  // groups of four nodes holding the batch
                    // one accumulator pair per group: four independent product chains
                    sh_v4f Dsg[2] = {Ds, Ds}, Ddg[2] = {Dd, Dd};
                    const int r0 = __builtin_amdgcn_readfirstlane(s_ovoff[i]);
                    const int r1 = __builtin_amdgcn_readfirstlane(s_ovoff[i + 1]);
                    for (int g0 = 0; g0 < ng; g0 += 2) {
                        const int gn = (ng - g0) < 2 ? ng - g0 : 2;
                        const float2 *V0[2], *V1[2];
#pragma unroll
                        for (int g = 0; g < 2; ++g) {
                            const int kn = 4 * (g0 + g) + kq;
                            V0[g] = Rb + kn * Rs;
                            V1[g] = Cb + kn * Cs;
                            if (g < gn) {
                                const float2 v1i = V0[g][i], v2i = V1[g][i];
                                const float2 v3 = vr >= 0 ? s_loc[kn * nAS + vr] : float2{0.f, 0.f};
                                const float2 v4 = vc >= 0 ? s_loc[kn * nAS + nA2 + vc] : float2{0.f, 0.f};
                                Dsg[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(v1i.x, -1.0f * v4.x + v4.y, Dsg[g], 0, 0, 0);
                                Ddg[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(v1i.y, 3.0f * v4.x + v4.y, Ddg[g], 0, 0, 0);
                                Dsg[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(-1.0f * v3.x + v3.y, v2i.x, Dsg[g], 0, 0, 0);
                                Ddg[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(3.0f * v3.x + v3.y, v2i.y, Ddg[g], 0, 0, 0);
                            }
                        }
                        // the rows, 64 at a time, one per lane; each step reads its row from a
                        // lane (v_readlane), so that only the vertex reads wait on LDS
                        float bx[2] = {0.f, 0.f}, by[2] = {0.f, 0.f};  // the current j1 run's ν' sums
                        for (int cb = r0; cb < r1; cb += 64) {
                            const int nr = (r1 - cb) < 64 ? r1 - cb : 64;
                            const uint2 rl = (lane < nr) ? s_ov[cb + lane] : uint2{0u, 0u};
                            const int rj = (int)rl.x, rm = (int)rl.y;
                            // rows in j1 runs: the ν' factors m v(j2) of a run are summed on the
                            // vector unit, then one product per run and group of four nodes with
                            // the ν vector of j1; the vertex reads of four rows are issued together
                            auto rows = [&](auto ngc) {
                                constexpr int G = decltype(ngc)::value;
                                auto step = [&](uint32_t jj, float mm, const float2 *f2) {
#pragma unroll
                                    for (int g = 0; g < G; ++g) {
                                        bx[g] += mm * f2[g].x;
                                        by[g] += mm * f2[g].y;
                                    }
                                    if (jj >> 31) {
#pragma unroll
                                        for (int g = 0; g < G; ++g) {
                                            const float2 e1 = V0[g][jj & 255];
                                            Dsg[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(e1.x, bx[g], Dsg[g], 0, 0, 0);
                                            Ddg[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(e1.y, by[g], Ddg[g], 0, 0, 0);
                                            bx[g] = 0.f;
                                            by[g] = 0.f;
                                        }
                                    }
                                };
                                int r = 0;
                                for (; r + 4 <= nr; r += 4) {
                                    uint32_t jj[4];
                                    float mm[4];
                                    float2 f2[4][G];
#pragma unroll
                                    for (int u = 0; u < 4; ++u) {
                                        jj[u] = (uint32_t)__builtin_amdgcn_readlane(rj, r + u);
                                        mm[u] = __int_as_float(__builtin_amdgcn_readlane(rm, r + u));
#pragma unroll
                                        for (int g = 0; g < G; ++g) f2[u][g] = V1[g][(jj[u] >> 8) & 255];
                                    }
#pragma unroll
                                    for (int u = 0; u < 4; ++u) step(jj[u], mm[u], f2[u]);
                                }
                                for (; r < nr; ++r) {
                                    const uint32_t jj = (uint32_t)__builtin_amdgcn_readlane(rj, r);
                                    const float mm = __int_as_float(__builtin_amdgcn_readlane(rm, r));
                                    float2 f2[G];
#pragma unroll
                                    for (int g = 0; g < G; ++g) f2[g] = V1[g][(jj >> 8) & 255];
                                    step(jj, mm, f2);
                                }
                            };
                            if (gn == 1) rows(std::integral_constant<int, 1>{});
                            else rows(std::integral_constant<int, 2>{});
                        }
                    }
                    Ds = Dsg[0] + Dsg[1];
                    Dd = Ddg[0] + Ddg[1];
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    b2[(s * 2 + 0) * 4 + r] += wq * (double)Ds[r];
                    b2[(s * 2 + 1) * 4 + r] += wq * (double)Dd[r];
                    if (addp) {
                        Pp[(s * 2 + 0) * 4 + r] += Ds[r];
                        Pp[(s * 2 + 1) * 4 + r] += Dd[r];
                    }
                }
            }
            if (wv == 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (act[r]) my_nodes += nb;
            }
        }
        __syncthreads();  // the descriptors and site vectors are rewritten by the next batch
#if defined(PFFRG_SHARED_PROFILE) && defined(PFFRG_SHARED_PROFILE_CELLS)
        // contraction cycles and batches with at most 4 active rows and 4 active columns
        if (t == 0 && A.count && nA2 <= 4 && nA3 <= 4) {
            s_prof[2 * PR_N + 11] += clock64() - _tb;
            s_prof[2 * PR_N + 12] += 1;
        }
#endif
        SH_PROF_LAP(PR_CONTRACT);
#undef SH_PROF_LAP
    };

    // active rows and columns of the tile: those holding at least one unconverged element,
    // compacted by the first wave with a ballot (s_act2/s_act3: tile positions of the active
    // rows/columns in order; s_rmap/s_cmap: compact index of each position, -1 if inactive), and
    // the level's batch size: the most nodes (a power of two, at most SH_KMAX) whose site
    // vectors fit the batch's capacity of shared_batch_nodes(spt) nodes of 32 vectors
    auto set_active = [&]() {
        SH_PROF_T0();
        if (t < SH_T) { s_rowf[t] = 0; s_colf[t] = 0; }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (act[r]) {
                s_rowf[r2 + r] = 1;
                s_colf[e3] = 1;
            }
        __syncthreads();
        if (t < 64) {
            const int q = t & 31;
            const bool isrow = t < 32;
            const bool on = q < SH_T && (isrow ? s_rowf[q] : s_colf[q]) != 0;
            const unsigned long long bal = __ballot(on);
            const unsigned long long mine = isrow ? (bal & 0xffffffffull) : (bal >> 32);
            const int p = __popcll(mine & ((1ull << q) - 1ull));
            if (on) (isrow ? s_act2 : s_act3)[p] = q;
            if (q < SH_T) (isrow ? s_rmap : s_cmap)[q] = on ? p : -1;
            if (t == 0) {
                const int nas = __popcll(bal);
                s_nact[0] = __popcll(bal & 0xffffffffull);
                int kb = 1;
                while (2 * kb <= A.kmax && 2 * kb * nas <= Kcap) kb *= 2;
                s_kb = kb;
            }
            if (t == 32) s_nact[1] = __popcll(bal >> 32);
        }
        __syncthreads();
        SH_PROF_ADD(PR_ACTIVE);
    };

    // Convergence test of every active element over all its sites (quadrature.jl:54-65): each
    // wave's partial sums of squares go through LDS and are added in wave order, the same on
    // every wave; a converged element adds its Richardson result to its partial sum.
    auto check_level = [&](long nn_) {
        SH_PROF_T0();
        double q[4][3];
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r][0] = q[r][1] = q[r][2] = 0.0;
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            const int s = o >> 3, r = o & 3;
            if (s < nsw) {
                double R = (double)b1[o * NTp] * (-1.0 / 15.0);
                R += 16.0 / 15.0 * b2[o];
                q[r][0] += R * R;
                q[r][1] += b2[o] * b2[o];
                const double dd = R - b2[o];
                q[r][2] += dd * dd;
            }
        }
        if (wv < nw) {  // the problem's own waves (a batched launch's block may have more)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = (r2 + r) * SH_T + e3;
                double *d = s_red + ((long)wv * SH_T * SH_T + e) * 3;
                d[0] = q[r][0];
                d[1] = q[r][1];
                d[2] = q[r][2];
            }
        }
        __syncthreads();
        // each element's test once: thread e adds the waves' partial sums in wave order
        if (t < SH_T * SH_T) {
            double q1 = 0.0, q2 = 0.0, q3 = 0.0;
            for (int w = 0; w < nw; ++w) {
                const double *d = s_red + ((long)w * SH_T * SH_T + t) * 3;
                q1 += d[0];
                q2 += d[1];
                q3 += d[2];
            }
            const double n1 = sqrt(q1), n2 = sqrt(q2), adiff = sqrt(q3);
            const double rdiff = adiff / fmax(n1, n2);
            s_conv[t] = (adiff < P.atol || rdiff < P.rtol || nn_ > (1L << 20)) ? 1 : 0;
        }
        __syncthreads();
        bool any = false;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = (r2 + r) * SH_T + e3;
            const bool conv = act[r] && s_conv[e] != 0;
            if (conv) {
                // Richardson-extrapolated result (quadrature.jl:72-88), added in sub-interval order
                double *pp = part_at(dom, r);
#pragma unroll
                for (int s = 0; s < SPT; ++s) {
                    if (s >= nsw) break;
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const int o = (s * 2 + c) * 4 + r;
                        double R = (double)b1[o * NTp] * (-1.0 / 15.0);
                        R += 16.0 / 15.0 * b2[o];
                        pp[c * y.ns + msite[s]] += (R - b2[o]) + b2[o];
                        // one read-modify-write at a time (once per element and sub-interval): keeps
                        // the compiler from holding every load of the loop in registers at once
                        asm volatile("" ::: "memory");
                    }
                }
                act[r] = false;
            }
            any = any || act[r];
        }
        // the next level: S1 <- S2, S2 <- S2 / 2 (quadrature.jl:74-76)
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            if ((o >> 3) < nsw) b1[o * NTp] = (float)b2[o];
            b2[o] = 0.5 * b2[o];
        }
        const bool more = __syncthreads_or(any ? 1 : 0);
        SH_PROF_ADD(PR_CHECK);
        return more;
    };

    // Quadrature of the unit's sub-intervals in order, then (last domain) the two tail nodes.
    // One preparation and one evaluation call site serve every step: the three nodes of the
    // initial Simpson rule (a and b, then the midpoint, whose weights differ), the new nodes of
    // each level (SH_NODES prepared at a time, K evaluated at a time), and the tail nodes
    // (one at a time: their weights differ).
    enum { ST_SUB, ST_INIT, ST_LEVEL, ST_TAIL };
    int st = ST_SUB, kk = -1, n = 0, q0 = 0;
    long nlev = 0, base = 0;
    double a = 0.0, b = 0.0, sg = 1.0, h = 0.0;
    bool need_prep = false;
    auto start_level = [&]() {
        // residual! (quadrature.jl:2-24): nlev new nodes a + (2i - 1) h with weight 4h/3; the
        // previous level's new nodes get -2h/3, applied to their kept sum
        h = 0.5 * (b - a) / nlev;
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            b2[o] += -2.0 * h / 3.0 * (double)Pp[o];
            Pp[o] = 0.f;
        }
        base = 0;
        n = (int)(nlev < SH_NODES ? nlev : SH_NODES);
        need_prep = true;
    };
    while (true) {
        if (st == ST_SUB) {
            ++kk;
            if (kk > P.eval || (kk == P.eval && dom != 3)) break;
#pragma unroll
            for (int r = 0; r < 4; ++r) act[r] = valid[r];
#pragma unroll
            for (int o = 0; o < NO; ++o) { b2[o] = 0.0; Pp[o] = 0.f; }
            if (kk == P.eval) {
                // boundary corrections (flow_s.jl:37-38): -val and +val times the tail factors
                st = ST_TAIL;
                n = 2;
            } else {
                if (!subinterval(dom * P.eval + kk, P.eval, ref, logpts, a, b, sg)) continue;
                a = wave_uniform(a);
                b = wave_uniform(b);
                sg = wave_uniform(sg);
                // simps! (quadrature.jl:57-62): a and b with weight (b-a)/6, the midpoint with 2(b-a)/3
                st = ST_INIT;
                n = 3;
            }
            set_active();
            need_prep = true;
        }
        if (need_prep) {
            prep_nodes(n, [&](int q, double &x, double &w, int &rw) {
                if (st == ST_INIT) {
                    x = sg * (q == 0 ? a : (q == 1 ? b : 0.5 * (a + b)));
                    w = (q == 2) ? 2.0 * (b - a) / 3.0 : 1.0 * (b - a) / 6.0;
                    rw = (q == 2);  // the midpoint is the node the first level reweights
                } else if (st == ST_TAIL) {
                    x = (q == 0) ? -val : val;
                    w = P.corrs[q + 2 * (ch + 3 * w1)];
                    rw = 0;
                } else {
                    x = sg * node_at(a, 2.0 * (double)(base + q + 1) - 1.0, h);
                    w = 4.0 * h / 3.0;
                    rw = 1;
                }
            });
            need_prep = false;
            q0 = 0;
        }
        const int kbu = __builtin_amdgcn_readfirstlane(s_kb);
        const int nb = (st == ST_INIT) ? (q0 == 0 ? 2 : 1) : (st == ST_TAIL ? 1 : ((n - q0) < kbu ? n - q0 : kbu));
        eval_batch(q0, nb);
        q0 += nb;
        if (q0 < n) continue;
        if (st == ST_INIT) {
            // S1, then S2 starts from S1 / 2 (quadrature.jl:62-66)
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                if ((o >> 3) < nsw) b1[o * NTp] = (float)b2[o];
                b2[o] = 0.5 * b2[o];
            }
            nlev = 2;
            st = ST_LEVEL;
            start_level();
            continue;
        }
        if (st == ST_TAIL) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (valid[r]) {
                    double *pt = part_at(4, r);
#pragma unroll
                    for (int s = 0; s < SPT; ++s) {
                        if (s >= nsw) break;
                        pt[msite[s]] = b2[(s * 2 + 0) * 4 + r];
                        pt[y.ns + msite[s]] = b2[(s * 2 + 1) * 4 + r];
                    }
                }
            break;
        }
        // ST_LEVEL: the next chunk of the level's nodes, or the level's convergence test
        base += n;
        if (base < nlev) {
            n = (int)((nlev - base) < SH_NODES ? nlev - base : SH_NODES);
            need_prep = true;
            continue;
        }
        if (!check_level(2 * nlev)) {
            st = ST_SUB;
            continue;
        }
        set_active();
        nlev *= 2;
        start_level();
    }
    // the unit's run time in constant-rate clock ticks, which orders the next launch's units
    if (cost_slot && t == 0) {
        const unsigned long long dt = wall_clock64() - rt_unit;
        *cost_slot = dt < 0xffffffffull ? (uint32_t)dt : 0xffffffffu;
    }
    if (A.count) {
        if (my_nodes) atomicAdd(&s_cnt[0], my_nodes);
        __syncthreads();
        if (t == 0) {
            atomicAdd(A.count, s_cnt[0]);
            atomicAdd(A.count + 1, s_cnt[1]);
#ifdef PFFRG_SHARED_PROFILE
            s_prof[prof_base + PR_TOTAL] = clock64() - t_begin;
            for (int k = 0; k < 2 * PR_N + 13; ++k) atomicAdd(A.count + 2 + k, s_prof[k]);
            // the block's duration in constant-rate (wall clock) ticks
            const unsigned long long rt_end = wall_clock64();
            atomicAdd(A.count + 31, rt_end - rt_begin);
#ifdef PFFRG_SHARED_PROFILE_UNITS
            const unsigned long long r = atomicAdd(A.count + N_COUNTERS, 1ull);
            if (r < 16384) {
                unsigned long long *rec = A.count + N_COUNTERS + 1 + 4 * r;
                rec[0] = (unsigned long long)uc | ((unsigned long long)PASS << 32);
                rec[1] = rt_begin;
                rec[2] = rt_end;
                rec[3] = (unsigned long long)__smid() | ((unsigned long long)blockIdx.x << 32);
            }
#endif
#endif
        }
    }
}

template <int PASS, int SPT>
__global__ __launch_bounds__(SH_MAXT) void k_shared_pass(SharedArgs A) {
    shared_pass_body<PASS, SPT>(A, __builtin_amdgcn_readfirstlane(A.units[blockIdx.x]),
                                A.cost ? A.cost + blockIdx.x : nullptr);
}

// The same pass of several problems in one launch: unit code bits 26-31 select the problem's
// arguments in tab; the block has the threads of the problem that needs most, the others idle.
// cost (may be null) receives each unit's run time at the unit's place in units.
template <int PASS, int SPT>
__global__ __launch_bounds__(SH_MAXT) void k_shared_pass_batch(const SharedArgs *__restrict__ tab,
                                                                const uint32_t *__restrict__ units, uint32_t *cost,
                                                                int n) {
    if ((int)blockIdx.x >= n) return;
    const uint32_t uc = __builtin_amdgcn_readfirstlane(units[blockIdx.x]);
    shared_pass_body<PASS, SPT>(tab[uc >> 26], uc & 0x03ffffffu, cost ? cost + blockIdx.x : nullptr);
}

// Sum of an element's partial results in the reference's order (the four domains, then the two
// tail nodes) into the output vertex at the element's place (flow_s.jl:41-51).
template <int PASS>
__global__ __launch_bounds__(256) void k_shared_finish(SharedArgs A) {
    const Sys &y = A.P.y;
    const long total = (long)A.P.n_elems * 2 * y.ns;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const int i = (int)(q % y.ns);
        const int comp = (int)((q / y.ns) & 1);
        const long el = q / (2L * y.ns);
        const uint32_t code = A.P.elems[el];
        int ch, kernel, w1, w2, w3;
        decode(code, ch, kernel, w1, w2, w3);
        const int nn = y.nn;
        int w2x, w3x;
        if (A.mode == 2) {
            w2x = 0;
            w3x = (kernel == 1) ? nn : w3;
        } else {
            w2x = (kernel == 2 || kernel == 4) ? w2 : nn;
            w3x = (kernel == 3 || kernel == 4) ? w3 : nn;
        }
        double acc = 0.0;
        for (int d = 0; d < SH_NPART; ++d)
            acc += A.part[shared_part_index(d, ch, w1, w2x, w3x, y.nO, A.NX2, A.NX3, y.ns) + (long)comp * y.ns + i];
        long base;
        if (kernel == 1) base = (long)y.ns * w1;
        else if (kernel == 2) base = y.q21 + (long)y.ns * (w1 + (long)y.nO * w2);
        else if (kernel == 3) base = y.q22 + (long)y.ns * (w1 + (long)y.nO * w3);
        else base = y.q3 + (long)y.ns * (w1 + (long)y.nO * (w2 + (long)y.nn * w3));
        A.P.out[(long)comp * y.complen + (long)ch * y.chlen + base + i] = acc;
    }
}

// threads of a block: nw waves, each holding the whole tile on its share of the sites
__host__ __device__ inline int shared_threads(const SharedArgs &A) { return 64 * A.nw; }

template <int PASS, int SPT>
hipError_t launch_shared_pass_spt(const SharedArgs &A, size_t lds, hipStream_t stream) {
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_shared_pass<PASS, SPT>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_shared_pass<PASS, SPT>), dim3(A.n_units), dim3(shared_threads(A)), lds, stream, A);
    return hipGetLastError();
}

template <int PASS>
// the work units of one pass; launch_shared_finish then writes the outputs from the partial sums
hipError_t launch_shared_pass(const SharedArgs &A, size_t lds, hipStream_t stream) {
    if (A.n_units <= 0) return hipSuccess;
    switch (A.spt) {
    case 1: return launch_shared_pass_spt<PASS, 1>(A, lds, stream);
    case 2: return launch_shared_pass_spt<PASS, 2>(A, lds, stream);
    case 3: return launch_shared_pass_spt<PASS, 3>(A, lds, stream);
    default: return hipErrorInvalidValue;
    }
}

template <int PASS, int SPT>
hipError_t launch_shared_pass_batch_spt(const SharedArgs *tab, const uint32_t *units, uint32_t *cost, int n,
                                        int nthreads, size_t lds, hipStream_t stream) {
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_shared_pass_batch<PASS, SPT>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_shared_pass_batch<PASS, SPT>), dim3(n), dim3(nthreads), lds, stream, tab, units, cost, n);
    return hipGetLastError();
}

// the units of all problems of a batch (problem index in bits 26-31), nthreads per block, the
// per-lane state sized for spt sites per wave (the most any problem of the batch has); the
// partial sums of each problem are then summed by launch_shared_finish on the same stream
template <int PASS>
hipError_t launch_shared_pass_batch(const SharedArgs *tab, const uint32_t *units, uint32_t *cost, int n, int nthreads,
                                    int spt, size_t lds, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    switch (spt) {
    case 1: return launch_shared_pass_batch_spt<PASS, 1>(tab, units, cost, n, nthreads, lds, stream);
    case 2: return launch_shared_pass_batch_spt<PASS, 2>(tab, units, cost, n, nthreads, lds, stream);
    case 3: return launch_shared_pass_batch_spt<PASS, 3>(tab, units, cost, n, nthreads, lds, stream);
    default: return hipErrorInvalidValue;
    }
}

template <int PASS>
hipError_t launch_shared_finish(const SharedArgs &A, hipStream_t stream) {
    long total = (long)A.P.n_elems * 2 * A.P.y.ns;
    long blocks = (total + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_shared_finish<PASS>), dim3((unsigned)blocks), dim3(256), 0, stream, A);
    return hipGetLastError();
}

#define PFFRG_SHARED_INSTANCES(KW, PASS)                                                                        \
    KW template hipError_t launch_shared_pass<PASS>(const SharedArgs &, size_t, hipStream_t);                  \
    KW template hipError_t launch_shared_pass_batch<PASS>(const SharedArgs *, const uint32_t *, uint32_t *, int,  \
                                                          int, int, size_t, hipStream_t);                      \
    KW template hipError_t launch_shared_finish<PASS>(const SharedArgs &, hipStream_t);

PFFRG_SHARED_INSTANCES(extern, P_KAT)
PFFRG_SHARED_INSTANCES(extern, P_LEFT)
PFFRG_SHARED_INSTANCES(extern, P_CENTRAL)
PFFRG_SHARED_INSTANCES(extern, P_BSE)
PFFRG_SHARED_INSTANCES(extern, P_REDUCED)

}  // namespace pffrg
