/* pffrg.h -- C ABI of the MI355X (gfx950) pf-FRG flow right-hand side.
 *
 * Drop-in replacement for the internal Julia functions of PFFRGSolver.jl that
 * evaluate the flow right-hand side, the parquet equations and the ODE stage
 * algebra.  All entry points are plain C (ccall / ctypes / cgo compatible):
 * pointers, sizes and doubles, no C++ types, no exceptions.
 *
 * Conventions
 *  - An opaque context owns one HIP stream and a fixed set of device "slots".
 *    Each slot holds one SU(2) action: Σ[n_sigma], bare[2][n_sites] and the
 *    packed vertex  [comp spin,dens][channel s,t,u][q1 | q2_1 | q2_2 | q3]
 *    with column-major arrays (site fastest), i.e. exactly the memory of the
 *    reference's Action_su2 / Channel arrays and of its HDF5 datasets.
 *  - Calls are synchronous with respect to the host unless stated otherwise.
 *  - Every int-returning call returns 0 on success and a negative code on
 *    failure; pffrg_last_error() then describes the failure.
 *  - Tolerances are (atol, rtol) pairs as in the reference's NTuple{2,Float64}.
 *  - Λ is in units of the normalised coupling (Launcher.jl:647-650).
 */
#ifndef PFFRG_H
#define PFFRG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFFRG_OK 0
#define PFFRG_ERR_ARG -1
#define PFFRG_ERR_HIP -2
#define PFFRG_ERR_STATE -3
#define PFFRG_ERR_LIMIT -4

/* arithmetic of the vertex passes */
#define PFFRG_FP64 0        /* FP64 storage, interpolation, products and accumulation */
#define PFFRG_FP32_MIXED 1  /* FP32 vertex copy, interpolation and products; FP64 quadrature sums,
                               Simpson sums and convergence tests; the elements of one (channel, Ω)
                               share node evaluations (shared-node pass) */
#define PFFRG_FP32_ELEMENT 2 /* the arithmetic of PFFRG_FP32_MIXED in the per-element pass
                               (measurement and comparison) */

typedef struct pffrg_ctx pffrg_ctx;

typedef struct {
    int32_t n_sites;  /* irreducible sites, 1..64 */
    int32_t n_sigma;  /* self-energy mesh points (num_σ + 1), <= 64 */
    int32_t n_Omega;  /* bosonic mesh points (num_Ω + 1), <= 64 */
    int32_t n_nu;     /* fermionic mesh points (num_ν + 1), <= 64 */
    double S;         /* spin length (Action_su2.S) */
    int32_t n_slots;  /* number of device action slots */
    int32_t device;   /* HIP device ordinal */
} pffrg_dims;

/* Action symmetries (get_action_empty, Action/Action.jl:564-577): the SU(2) action with two vertex
 * components (spin, density; action_lib/action_su2.jl:9-13) and the U(1)-symmetric DM action with
 * six (xx, zz, DM, dd, zd, dz; action_lib/action_u1_dm.jl:1-11).  Slot arrays hold n_components
 * blocks: bare[n_components * n_sites], gamma[n_components * 3 * channel length]. */
#define PFFRG_SU2 0
#define PFFRG_U1_DM 1

/* Context lifetime.  Replaces the per-run allocations of launch_1l! (launcher_1l.jl:38-51);
 * pffrg_create is pffrg_create_symmetry with PFFRG_SU2.  The batched entry points and the
 * shared-node FP32 pass serve the SU(2) action; a U(1)-DM context runs the FP32-mixed mode on
 * the per-element pass. */
int pffrg_create(const pffrg_dims *dims, pffrg_ctx **out);
int pffrg_create_symmetry(const pffrg_dims *dims, int32_t symmetry, pffrg_ctx **out);
int32_t pffrg_n_components(const pffrg_ctx *ctx);
void pffrg_destroy(pffrg_ctx *ctx);
const char *pffrg_last_error(const pffrg_ctx *ctx);

/* Frequency meshes of the Mesh struct (Frequency/mesh.jl:16-27): σ, Ωs, νs, Ωt, νt. */
int pffrg_set_mesh(pffrg_ctx *ctx, const double *sigma, const double *Os, const double *ns, const double *Ot,
                   const double *nt);

/* Reduced_lattice tables (Lattice/reduced.jl:15-25): exchange[n_sites], mult[n_sites] (the Int64
 * orbit sizes, >= 1), overlap as CSR: ov_off[n_sites+1] row offsets, ov_rows[3*R] = (j1, j2,
 * multiplicity).  index_base = 1 for the reference's 1-based Julia tables, 0 for 0-based ones.
 * A rejected call (bad index, or a device error) keeps the previous lattice. */
int pffrg_set_lattice(pffrg_ctx *ctx, const int64_t *exchange, const int64_t *mult, const int64_t *ov_off,
                      const int64_t *ov_rows, int32_t index_base);

/* Host <-> device transfer of one action slot; NULL skips a part. */
int pffrg_upload_action(pffrg_ctx *ctx, int32_t slot, const double *Sigma, const double *bare, const double *gamma);
int pffrg_download_action(pffrg_ctx *ctx, int32_t slot, double *Sigma, double *bare, double *gamma);
int64_t pffrg_gamma_length(const pffrg_ctx *ctx);

int pffrg_set_precision(pffrg_ctx *ctx, int32_t mode);

/* compute_dΣ!  (Flow/flow_self.jl:2-20): da.Σ[2:end] from a. */
int pffrg_compute_dsigma(pffrg_ctx *ctx, double Lambda, int32_t a, int32_t da, const double *sigma_tol);

/* compute_dΣ_corr!  (Flow/flow_self.jl:23-50): multiloop self-energy corrections. */
int pffrg_compute_dsigma_corr(pffrg_ctx *ctx, double Lambda, int32_t a, int32_t da, int32_t da_sigma,
                              const double *sigma_tol);

/* compute_dΓ_1l!  (Flow/compute.jl:57-112): Katanin one-loop vertex flow into da
 * (reads da.Σ, so compute_dsigma must run first). */
int pffrg_compute_dgamma_1l(pffrg_ctx *ctx, double Lambda, int32_t a, int32_t da, int32_t eval,
                            const double *gamma_tol);

/* Launch shape of the FP32-mixed shared-node vertex pass (no reference counterpart: the
 * reference has no kernels): shape[0] waves per block (sites split over them), shape[1] the
 * batch capacity in nodes of 32 site vectors (4 or 8), shape[2] the most quadrature nodes per
 * batch (16..64).  Every shape gives results within FP32 rounding of the others; a context keeps
 * one shape for all its shared-node passes.  pffrg_set_shared_shape refuses a shape the problem
 * cannot take (at most 3 sites per wave, 160 KB of LDS).
 *
 * pffrg_tune_shared is the autotuner: it times this problem's own Katanin pass (compute_dΣ! and
 * compute_dΓ_1l! of slot a into slot da at Lambda; slot da is overwritten, and the runs count in
 * the statistics) for each candidate of a fixed list of shapes, keeps the fastest for the
 * context and returns it in best[3]; ms[k] is candidate k's time (-1 where infeasible), for at
 * most *n_cand entries, and *n_cand is set to the list's length. */
int pffrg_set_shared_shape(pffrg_ctx *ctx, int32_t waves, int32_t batch_nodes, int32_t max_nodes);
int pffrg_get_shared_shape(pffrg_ctx *ctx, int32_t *shape);
int pffrg_tune_shared(pffrg_ctx *ctx, double Lambda, int32_t a, int32_t da, int32_t eval, const double *gamma_tol,
                      int32_t *best, double *ms, int32_t *n_cand);

/* compute_dΓ_2l!  (Flow/compute.jl:115-174). */
int pffrg_compute_dgamma_2l(pffrg_ctx *ctx, double Lambda, int32_t a, int32_t da, int32_t da_l, int32_t eval,
                            const double *gamma_tol);

/* compute_dΓ_ml!  (Flow/compute.jl:177-286). */
int pffrg_compute_dgamma_ml(pffrg_ctx *ctx, double Lambda, int32_t loops, int32_t a, int32_t da, int32_t da_l,
                            int32_t da_c, int32_t da_temp, int32_t da_sigma, int32_t eval, const double *gamma_tol);

/* compute_Σ!  (Flow/SDE.jl:58-96): Schwinger-Dyson self energy of a1 into a2 (and the
 * reduced s bubble into a2's s channel q1/q2_2). */
int pffrg_compute_sigma(pffrg_ctx *ctx, double Lambda, int32_t a1, int32_t a2, int32_t eval, const double *gamma_tol,
                        const double *sigma_tol);

/* compute_Γ!  (Flow/compute.jl:2-54): Bethe-Salpeter right side of a1 into a2; uses the
 * boundary corrections left by the preceding pffrg_compute_sigma, as the reference does. */
int pffrg_compute_gamma(pffrg_ctx *ctx, double Lambda, int32_t a1, int32_t a2, int32_t eval,
                        const double *gamma_tol);

/* symmetrize! / symmetrize_add_to!  (Action/action_lib/action_su2.jl:141-250). */
int pffrg_symmetrize(pffrg_ctx *ctx, int32_t slot);
int pffrg_symmetrize_add_to(pffrg_ctx *ctx, int32_t a_l, int32_t a);

/* Action algebra (Action/Action.jl:180-337): vertex_only = 1 for the *_Γ! variants.
 * replace_with!(dst, src); mult_with!(slot, fac); mult_with_add_to!(src, fac, dst): dst += fac*src. */
int pffrg_replace_with(pffrg_ctx *ctx, int32_t dst, int32_t src, int32_t vertex_only);
int pffrg_mult_with(pffrg_ctx *ctx, int32_t slot, double fac, int32_t vertex_only);
int pffrg_mult_with_add_to(pffrg_ctx *ctx, int32_t src, double fac, int32_t dst, int32_t vertex_only);

/* get_abs_max  (Action/Action.jl:346-359): max |q3| over components and channels, in FP64. */
int pffrg_abs_max(pffrg_ctx *ctx, int32_t slot, double *out);

/* Σ of one slot (n_sigma doubles) for the host-side parquet damping (parquet.jl:57-61). */
int pffrg_get_sigma(pffrg_ctx *ctx, int32_t slot, double *Sigma);
int pffrg_set_sigma(pffrg_ctx *ctx, int32_t slot, const double *Sigma);

/* Measurement hooks (no reference counterpart; bench and tests only). */
int pffrg_synchronize(pffrg_ctx *ctx);
int pffrg_reset_stats(pffrg_ctx *ctx);
/* stats[0] = quadrature nodes evaluated by vertex passes, stats[1] = vertex-pass elements,
 * stats[2] = vertex-pass launches, stats[3] = vertex bytes gathered by the interpolations */
int pffrg_get_stats(pffrg_ctx *ctx, int64_t *stats);
/* the context's n <= 32 raw device counters since pffrg_reset_stats: [0] quadrature nodes
 * (element-nodes), [1] gathered bytes; [2..9] and [10..17] cycle counts of the shared-node pass's
 * phases (active set, node preparation, descriptors, gathers, contraction, convergence tests,
 * batches, total) in the units of the s and u channels and of the t channel, [18..22] batches by
 * active elements of the tile (1-4, 5-16, 17-64, 65-128, 129-256), [23] active element-nodes,
 * [24] active row and column vectors per node, in a build with PFFRG_SHARED_PROFILE defined
 * (zero otherwise); with PFFRG_SHARED_PROFILE_CELLS also defined, [25] wave steps of the gather
 * loop (one node of one site vector per lane), [26] those in which a lane loaded corners, [27] their
 * cycles, [28] the cycles of the others, [29] contraction cycles of batches with at most 4 active
 * rows and columns, [30] those batches; with PFFRG_SHARED_PROFILE_PAIRS instead, [25] consecutive node pairs of one batch and
 * vector, [26] those reading the same corners in all three channels, [27] channels reading the same
 * corners */
int pffrg_get_counters(pffrg_ctx *ctx, int64_t *out, int32_t n);
/* boundary-correction factors corrs[Ω][channel][side] (2*3*n_Omega doubles) of the last pass */
int pffrg_get_corrs(pffrg_ctx *ctx, double *corrs);
/* integrand of one output element (pass 0..4 = Katanin, left, central, BSE, reduced SDE; channel
 * 0..2 = s,t,u; kernel 1..4; 0-based w indices, -1 for unused) at n loop frequencies v, without
 * quadrature: out[k][comp][site] (n*2*n_sites doubles) */
int pffrg_probe_integrand(pffrg_ctx *ctx, int32_t pass, double Lambda, int32_t a, int32_t da, int32_t ch,
                          int32_t kernel, int32_t w1, int32_t w2, int32_t w3, int32_t n, const double *v,
                          double *out);
/* arm a one-shot recording for the next vertex pass: per (element, sub-interval) quadrature
 * node counts and the site-0 spin result, written to the host arrays when that pass ends
 * (capacity entries each, indexed element * 4*eval + sub-interval); diagnostics only */
int pffrg_debug_record(pffrg_ctx *ctx, int32_t *sub_nodes, double *sub_res, int64_t capacity);
/* enable = 1: time every vertex-pass launch with HIP events on the context stream */
int pffrg_set_timing(pffrg_ctx *ctx, int32_t enable);
/* total milliseconds and launch count of the vertex-pass kernel since the last reset */
int pffrg_get_timing(pffrg_ctx *ctx, double *total_ms, int64_t *launches);
/* Practical ceiling of the vertex pass's gathers on this device: random trilinear q3 lookups
 * (8 corners x 3 channels per lookup, the vertex pass's lane layout) on the slot's gather copy
 * in the given precision, with no other work.  waves_per_simd = 0 leaves occupancy to the
 * hardware, 1..8 caps it with reserved LDS; footprint = 0 draws from the whole q3 block,
 * footprint > 0 from as few ν' planes as fit that many bytes (at least 2 planes); n_iter
 * lookups per site group; n_rep timed launches after one untimed one.  Returns the gathered
 * GB/s and the mean launch time. */
int pffrg_gather_ceiling(pffrg_ctx *ctx, int32_t slot, int32_t precision, int32_t waves_per_simd, int64_t footprint,
                         int32_t n_iter, int32_t n_rep, double *gbps, double *ms_per_launch);

/* compute_χ!  (Observable/correlation_lib/correlation_su2.jl:58-86): static real-space
 * correlations chi[i + n_sites*k] of the irreducible sites i at the n_chi frequencies
 * chi_mesh[k] (Mesh.χ), from slot a at Lambda, to the tolerance (atol, rtol) of χ_tol.  The
 * nine boxes of integrate_χ_boxes (correlation_lib.jl:2-30) are integrated by adaptive 2-D
 * cubature with hcubature_v's schedule (degree-7 Genz-Malik rule, several regions split per
 * round, at most 10^7 evaluations per box, as the reference's call); nreg (optional,
 * 9*n_sites*n_chi entries) receives the final region count of every box.  A U(1)-DM context
 * (compute_χ!, correlation_u1_dm.jl:138-190) writes three blocks of n_sites*n_chi, χxx, χzz and
 * χxy, each integrated on its own (and three blocks of nreg). */
int pffrg_compute_chi(pffrg_ctx *ctx, double Lambda, int32_t a, const double *chi_mesh, int32_t n_chi,
                      const double *chi_tol, double *chi, int32_t *nreg);

/* ---- Mesh rescanning (resample_from_to, Action/Action.jl:483-557), in three parts:
 * pffrg_scan_cuts gathers what scan_channel reads, pffrg_scan is the scan heuristic, and
 * pffrg_resample moves an action onto the new meshes. */

/* scan_channel (Action.jl:449-480), device part: for slot a and every (component c, channel
 * ch), k = 3*c + ch (k < 3 n_components), the first position of max |q3| in column-major order (Julia's argmax),
 * 0-based (site, w, v, v') into idx[4k .. 4k+3], and the four cuts through it into
 * cuts[k*(n_Omega + 3*n_nu) ...]:
 *   q3[i, :, v, v'] (n_Omega values),
 *   q3[i, w, x, x]  - q3[i, w, end, end],
 *   q3[i, w, x, v'] - q3[i, w, end, v'],
 *   q3[i, w, v, x]  - q3[i, w, v, end]   (n_nu values each, x = 0 .. n_nu-1). */
int pffrg_scan_cuts(pffrg_ctx *ctx, int32_t a, int32_t *idx, double *cuts);

/* pffrg_scan_cuts without device allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and outputs as pffrg_scan_cuts; its argmax partials and
 * cuts live in one static 54.5 KB scratch per GPU (held until the process ends) instead of the
 * context's lazily allocated one.  Every context on a GPU shares that scratch, so calls that
 * overlap on one GPU (several host threads or streams) race; each call overwrites what it reads.
 * Rejects (PFFRG_ERR_LIMIT) meshes whose cuts exceed the scratch, which the n_Omega, n_nu <= 64
 * limits of pffrg_create already exclude. */
int pffrg_scan_cuts_noalloc(pffrg_ctx *ctx, int32_t a, int32_t *idx, double *cuts);

/* scan (Action.jl:369-446), host only: the new linear extent of a mesh x[0..n-1] (built by
 * get_mesh with linear fraction p[0]) for the cut y[0..n-1], from the peak and origin
 * heuristics with bounds p[1] < Δ < p[2] and p[3] <= δ <= p[4] (p[3], p[4] already
 * scaled by Λ).  Needs no device. */
int pffrg_scan(const double *x, const double *y, int32_t n, const double *p, double *delta);

/* resample_from_to (Action.jl:531-556; channel.jl:416-470; vertex.jl:148-160): slot src,
 * held on the context's meshes, is interpolated onto the new meshes (sigma, Os, ns, Ot, nt;
 * same lengths) into slot dst: Σ by get_Σ (Action.jl:8-26), every q1/q2/q3 entry by the
 * reference's multilinear rule (get_param clamps beyond the old mesh), the bare couplings
 * copied.  The new meshes then replace the context's, as pffrg_set_mesh; the other slots
 * keep their data.  src != dst. */
int pffrg_resample(pffrg_ctx *ctx, int32_t src, int32_t dst, const double *sigma, const double *Os,
                   const double *ns, const double *Ot, const double *nt);

/* ---- Batched problems.  A batch groups contexts of one device (up to 64: different lattices,
 * truncations, couplings, meshes and Λ) whose 1-loop right-hand sides run together: the
 * self-energy and boundary-correction kernels per problem, then ONE vertex-pass launch over the
 * elements of all problems, then the per-problem symmetrisation.  It replaces a sequence of
 * compute_dΣ! + compute_dΓ_1l! calls (flow_self.jl:2-20, compute.jl:57-112), one per problem,
 * for callers that advance many flows at once (coupling sweeps; launch_1l!, launcher_1l.jl:85-109,
 * run once per problem).  The contexts stay owned by the caller and must outlive the batch. */
typedef struct pffrg_batch pffrg_batch;
int pffrg_batch_create(pffrg_ctx *const *ctxs, int32_t n, pffrg_batch **out);
void pffrg_batch_destroy(pffrg_batch *batch);
const char *pffrg_batch_last_error(const pffrg_batch *batch);
/* per problem k (caller's order): dΣ and dΓ_1l of slot a[k] into slot da[k] at Lambda[k] with
 * eval[k]; all contexts must use the same precision mode */
int pffrg_batch_rhs_1l(pffrg_batch *batch, const double *Lambda, const int32_t *a, const int32_t *da,
                       const int32_t *eval, const double *sigma_tol, const double *gamma_tol);

/* replace_with! / mult_with_add_to! / get_abs_max (Action.jl:181-359) on the same slots of every
 * problem of the batch, with fac[k] for problem k (caller's order) and out[k] its |q3| maximum:
 * the BS3 stage algebra of launch_1l! (launcher_1l.jl:84-116) for all flows of a batch in one
 * call each.  Each problem's result equals the per-context call's bitwise. */
int pffrg_batch_replace_with(pffrg_batch *batch, int32_t dst, int32_t src, int32_t vertex_only);
int pffrg_batch_mult_with_add_to(pffrg_batch *batch, int32_t src, const double *fac, int32_t dst, int32_t vertex_only);
int pffrg_batch_abs_max(pffrg_batch *batch, int32_t slot, double *out);

/* compute_χ! (correlation_su2.jl:58-86) of every problem k of the batch (caller's order): slot
 * a[k] at Lambda[k], χ frequencies chi_mesh[k][0 .. n_chi[k]-1], result chi[k][i + n_sites*w] as
 * pffrg_compute_chi; all problems' cubature tasks run through one set of rounds.  Each
 * problem's result equals its pffrg_compute_chi bitwise.  n_chi[k] = 0 leaves problem k out
 * (its chi_mesh[k] and chi[k] may then be NULL). */
int pffrg_batch_compute_chi(pffrg_batch *batch, const double *Lambda, const int32_t *a, const double *const *chi_mesh,
                            const int32_t *n_chi, const double *chi_tol, double *const *chi);

/* ---- Learned warm start of the parquet iteration (no reference counterpart: the reference
 * starts launch_parquet! from the bare action, Launcher.jl:676-678 and parquet.jl:54).  A
 * multilayer perceptron maps (couplings, Λ) features to coefficients of a fixed basis of
 * parquet solutions; pffrg_decoder expands coefficients into a vertex.  Products run on MFMA
 * with BF16 operands and FP32 accumulation; weights, moments and the decoder are FP32.
 * Layers: n_in -> width (n_layers - 1 times, SiLU) -> n_out (linear). */
typedef struct pffrg_mlp pffrg_mlp;
int pffrg_mlp_create(int32_t n_in, int32_t width, int32_t n_layers, int32_t n_out, uint64_t seed, int32_t device,
                     pffrg_mlp **out);
void pffrg_mlp_destroy(pffrg_mlp *mlp);
const char *pffrg_mlp_last_error(const pffrg_mlp *mlp);
int64_t pffrg_mlp_num_params(const pffrg_mlp *mlp);
/* parameters: per layer W (out x in, row-major) then b (out), FP32 */
int pffrg_mlp_get_params(pffrg_mlp *mlp, float *params);
int pffrg_mlp_set_params(pffrg_mlp *mlp, const float *params);
/* y[batch][n_out] = mlp(x[batch][n_in]) */
int pffrg_mlp_forward(pffrg_mlp *mlp, const float *x, int32_t batch, float *y);
/* loss = mean (y - target)^2 and its gradient w.r.t. every parameter (kept on the device; also
 * copied to grads when it is not NULL) */
int pffrg_mlp_backward(pffrg_mlp *mlp, const float *x, const float *target, int32_t batch, float *loss, float *grads);
/* one Adam update with the last gradient (step counts from 1) */
int pffrg_mlp_adam(pffrg_mlp *mlp, float lr, float beta1, float beta2, float eps, int32_t step);

typedef struct pffrg_decoder pffrg_decoder;
/* out[j] = mean[j] + Σ_k coef[k] basis[k*length + j] */
int pffrg_decoder_create(int64_t length, int32_t n_coef, const float *mean, const float *basis, int32_t device,
                         pffrg_decoder **out);
void pffrg_decoder_destroy(pffrg_decoder *dec);
int pffrg_decoder_apply(pffrg_decoder *dec, const float *coef, double *out);

#ifdef __cplusplus
}
#endif

#endif /* PFFRG_H */
