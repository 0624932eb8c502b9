/*
 * gfhip -- MI355X (gfx950) implementation of the guarded-floor GMG-FGMRES
 * solver path of solver-admissibility-and-guarded-floor-selection.
 *
 * C ABI: plain pointers and sizes.  Every pointer argument named d_* (or a
 * vector argument of a compute call) is DEVICE memory; `stream` is a
 * hipStream_t (NULL = default stream).  Every function returns 0 on success
 * and a negative code on failure (-1 HIP error or an expired grid-wide
 * wait, -2 bad argument, -3 not set up, -4 nothing to continue);
 * gf_last_error() returns a description of the last failure on the
 * calling thread.  Calls on one object are not re-entrant.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   gf_ctx_*            MatrixFreeKff.__init__            src/gpu_fem/solver_v2.py:1153-1175
 *   gf_matvec_f64/f32   MatrixFreeKff.matvec              src/gpu_fem/solver_v2.py:1182-1222
 *   gf_matvec_f32, gf_matvec_full_f32, gf_matvec_full_bf16
 *                       FusedMatvec.matvec / .matvec_full  cuda_fused_matvec.py:343-416
 *   gf_diag_f64         MatrixFreeKff.extract_diagonal    src/gpu_fem/solver_v2.py:1228-1251
 *   gf_elem_sensitivity_f64  compliance sensitivity of SolverV4 src/gpu_fem/solver_v4.py:258-265
 *   gf_nodeblock_*      NodeBlockJacobiMF.update / apply   src/gpu_fem/solver_v2.py:1381-1451
 *   gf_gmg_create       GalerkinMatFreeGMG.__init__       src/gpu_fem/multigrid_v4.py:546-770
 *   gf_gmg_setup        GalerkinMatFreeGMG.setup          src/gpu_fem/multigrid_v4.py:1119-1341
 *   gf_gmg_apply        GalerkinMatFreeGMG.apply          src/gpu_fem/multigrid_v4.py:1759-1767
 *   gf_gmg_vcycle       GalerkinMatFreeGMG._apply_base_cycle multigrid_v4.py:1737-1740
 *   gf_gmg_level_*      GalerkinMatFreeGMG._K_gpu / _diag_inv_gpu (read-out for checks)
 *   gf_gmg_solve        _cupy_fgmres(A=MatrixFreeKff.matvec(., E), M=gmg.apply)
 *                                                         src/gpu_fem/multigrid_v4.py:416-534
 *   gf_gmg_solve_continue  the kept floor's re-solve after the probe
 *                       experiments/phase5/run_gmg_floor_detector_prospective.py:263-277, 313-351
 *   gf_gmg_set_wait_flag  none (test hook for the wait-expiry error path)
 *   gf_dot_f64 / gf_nrm2_f64  cp.dot / cp.linalg.norm used by _cupy_fgmres / _cupy_pcg
 *   gf_cnn_*            learned floor predictor: no reference counterpart (the
 *                       reference's policy always probes, run_gmg_floor_detector_
 *                       prospective.py:263-290); its proposals feed the same guard
 */
#ifndef GFHIP_H
#define GFHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gf_ctx gf_ctx;
typedef struct gf_gmg gf_gmg;

/* GalerkinMatFreeGMG keyword arguments (multigrid_v4.py:546-717).  Strings of
 * the reference are encoded as small integers:
 *   fine_smoother: 0 "fp64", 1 "fp32", 2 "bf16"
 *   level_prec[i]: -1 default schedule, 0 "fp64", 1 "fp32", 2 "bf16"
 *   smoother_type: 0 "chebyshev", 1 "jacobi"
 *   cycle_type:    0 "v", 1 "w"
 *   coarse_correction_policy: 0 "none", 1 "residual_line_search"
 *   root_enrichment_mode: 0 "none", 1 "weighted_residual_second",
 *       2 "weighted_gap_second", 3 "void_residual_second", 4 "void_gap_second"
 *   root_local_correction_mode: 0 "none", 1 "node_block_line_search",
 *       2 "node_block_inner_fgmres"                                        */
typedef struct gf_gmg_params {
  int32_t n_levels;
  int32_t coarse_smooth_iters;
  int32_t fine_smoother;
  int32_t fine_smoother_degree;
  double omega;
  int32_t dense_chol_max;
  int32_t coarse_pcg_iters;
  int32_t smoother_type;
  double cheb_lower_frac;
  int32_t level_prec[8];
  int32_t cycle_type;
  int32_t coarse_correction_policy;
  double coarse_correction_acceptance_ratio;
  double coarse_correction_max_scale;
  int32_t root_enrichment_mode;
  double root_enrichment_weight_floor;
  double root_enrichment_weight_power;
  int32_t root_local_correction_mode;
  double root_local_correction_max_scale;
  int32_t root_local_correction_inner_steps;
  int32_t root_local_correction_inner_restart;
  int32_t inner_krylov_steps;
  int32_t inner_krylov_restart;
  /* coarsest dense solve (not a reference option): 0 = explicit inverse
   * L^-T L^-1 formed once per set-up and applied as one parallel GEMV
   * (default), 1 = the two triangular solves of the reference */
  int32_t coarsest_solve;
  /* 1 = record the whole preconditioner application as a hipGraph after each
   * set-up and replay it (default); 0 = launch kernel by kernel */
  int32_t graph_apply;
  /* outer FGMRES orthogonalisation (gf_gmg_solve): 1 = the whole modified
   * Gram-Schmidt sweep of an Arnoldi step in one persistent launch with the
   * vector held in registers and bounded grid-wide waits (default; falls
   * back when n is too large; also used by the inner and root-local FGMRES
   * recorded in the preconditioner graph), 2 = as 1 but the outer sweeps go
   * through hipLaunchCooperativeKernel, 0 = one launch per projection.  Same
   * arithmetic order in every mode up to the partial-sum grouping. */
  int32_t mgs_sweep;
  /* iterative coarsest solve (more free DOFs than dense_chol_max): 0 = the
   * fixed-count Jacobi PCG in two launches per iteration, 1 = in one
   * persistent launch with bounded grid-wide exchanges through uncached
   * memory where its grid fits the co-residency rule, else in two launches
   * (default; 13.3 vs 15.1 us per iteration at 1M elements).
   * The same recurrence; the dot products differ only in partial-sum
   * grouping. */
  int32_t coarse_pcg_persistent;
} gf_gmg_params;

/* State read back after gf_gmg_setup (properties of GalerkinMatFreeGMG,
 * multigrid_v4.py:771-836). */
typedef struct gf_gmg_info {
  int32_t n_levels;
  int32_t n_free[8];
  int32_t dims[8][3];
  double lambda_max[8];
  double omega_used;
  double contrast_ratio;
  double void_frac;
  int32_t is_high_contrast;
  int32_t fine_degree_adapt;
  int32_t coarse_iters_adapt;
  int32_t coarse_chol_mode; /* 0 none, 1 dense, 2 iterative */
  int32_t coarse_chol_regularised;
  int32_t level_prec[8];
  int32_t setup_count;
} gf_gmg_info;

int32_t gf_version(void);
const char* gf_last_error(void);

/* ---- fine operator ---------------------------------------------------- */
/* Structured hex grid nelx x nely x nelz with the reference node numbering
 * (pub_simp_solver.py:196-207).  KE: 576 doubles (host), row-major 24x24.
 * free: n_free sorted full-DOF indices (host).  ndof = 3*(nelx+1)(nely+1)(nelz+1). */
int32_t gf_ctx_create(gf_ctx** out, int32_t nelx, int32_t nely, int32_t nelz,
                      const double* KE, const int32_t* free_dofs, int32_t n_free, int32_t ndof);
int32_t gf_ctx_destroy(gf_ctx* ctx);
/* Launch-shape autotuner of the fine operator (no reference counterpart):
 * times the fine-operator kernels on this grid -- shape 1 one lane per node,
 * 4 four waves per 64 nodes, 8 and 16 the same with the node and element
 * windows staged in LDS for an 8x8 / 4x16 (y,z) patch, 72 and 80 those tiles
 * with the FP64 element-row contraction on the matrix cores
 * (v_mfma_f64_4x4x4f64; the FP32 operator runs the VALU tile there), 48 and
 * 50 a node per lane with the u and E windows of a 4x8x8 or 2x8x16 (x,y,z)
 * node box staged in LDS, 44 a 2x8x8 box with two lanes per node (one per
 * x side of the node's elements) -- and keeps the fastest for every later call on ctx (each shape's best of three
 * interleaved timing rounds; the FP32 operator's shape is chosen the same
 * way from its own timings).  All shapes give bitwise-identical
 * results.  us_out (nullable, host): 9 floats of FP64 timings in that order;
 * shape_out (nullable, host): the chosen shape.  Synchronises. */
int32_t gf_ctx_autotune(gf_ctx* ctx, float* us_out, int32_t* shape_out, void* stream);
/* Force a launch shape (FP64 and FP32 operators): 0 = process default, 1, 4,
 * 8, 16, 72, 80, 48, 50 or 44. */
int32_t gf_ctx_set_shape(gf_ctx* ctx, int32_t shape);
/* Process-wide launch form of the coarse-level (27-point block stencil)
 * operator kernels: 1 = node-per-lane workgroups of eight waves with
 * coalesced coefficient reads, 64, 32 or 16 nodes per workgroup by level
 * size (default); 64, 32 or 16 = that form with the node count fixed;
 * 0 = one lane per (node, slot).  All fold the slot products in the same
 * order: identical results. */
int32_t gf_set_coarse_form(int32_t form);
/* Process-wide form of the residual line search of a coarse correction
 * (multigrid_v4.py:867-931; no reference counterpart as a switch): 1 = one
 * persistent launch for the statistics, the trial norm and the apply where
 * its grid fits the co-residency rule (default), 0 = three launches.  Both
 * use the same reduction trees: identical results. */
int32_t gf_set_ls_form(int32_t form);
/* Process-wide form of the persistent Gram-Schmidt sweep (the modified
 * Gram-Schmidt of multigrid_v4.py:474-482 in one launch per Arnoldi step; no
 * reference counterpart as a switch): 1 = the sweep that prefetches the next
 * basis vector where its grid fits the co-residency rule, else the one
 * without that prefetch (fewer registers: a larger grid fits), else one
 * launch per projection (default); 2 = the sweep without the prefetch
 * wherever it fits.  The sweeps give identical results. */
int32_t gf_set_sweep_form(int32_t form);
/* y_free = K_ff(E) u_free. E: n_elem moduli; u,y: n_free (device). */
int32_t gf_matvec_f64(gf_ctx* ctx, const double* E, const double* u, double* y, void* stream);
int32_t gf_matvec_f32(gf_ctx* ctx, const float* E, const float* u, float* y, void* stream);
/* y = b - K_ff(E) u  (fused residual) */
int32_t gf_residual_f64(gf_ctx* ctx, const double* E, const double* b, const double* u,
                        double* y, void* stream);
/* Full-length (ndof) variant: u_full must be zero on fixed DOFs; y_full gets
 * the unrestricted K u on every DOF (FusedMatvec.matvec_full contract,
 * cuda_fused_matvec.py:379-391). */
int32_t gf_matvec_full_f32(gf_ctx* ctx, const float* E, const float* u_full, float* y_full,
                           void* stream);
/* BF16 matrix-core form of gf_matvec_full_f32 (FusedMatvec dtype="bf16",
 * cuda_fused_matvec.py:138-246, 393-416): KE and u rounded to bf16, products
 * accumulated in FP32 on v_mfma_f32_16x16x32_bf16, E applied in FP32,
 * element rows summed per node in FP32 in a fixed order (deterministic).
 * One kernel; element results stay in LDS.  Uploads the 32x32 bf16 KE to ctx
 * on first use.  The same operator smooths level 0 of a hierarchy created
 * with fine_smoother = 2 ("bf16") and a reduced-precision level 0
 * (multigrid_v4.py:1765-1780). */
int32_t gf_matvec_full_bf16(gf_ctx* ctx, const float* E, const float* u_full, float* y_full,
                            void* stream);
int32_t gf_diag_f64(gf_ctx* ctx, const double* E, double* diag, void* stream);
/* Element strain energy and SIMP sensitivity of a free-DOF displacement u
 * (SolverV4._solve_cupy_matfree, solver_v4.py:258-265):
 *   ce[e] = u_e^T KE u_e (unit modulus), dc[e] = -penal*scale*rho[e]^(penal-1)*ce[e]
 * with scale = E0 - Emin; ce or dc may be NULL (rho may be NULL when dc is). */
int32_t gf_elem_sensitivity_f64(gf_ctx* ctx, const double* u, const double* rho, double penal,
                                double scale, double* ce, double* dc, void* stream);
int32_t gf_nodeblock_update(gf_ctx* ctx, const double* E, const double* scalar_diag /*nullable*/,
                            void* stream);
int32_t gf_nodeblock_apply(gf_ctx* ctx, const double* r, double* z, void* stream);

/* ---- BLAS-1 (deterministic fixed-order reductions) -------------------- */
/* result written to a device double (d_out); ctx may be NULL (a per-thread
 * scratch: calls from one thread must not overlap on two streams) */
int32_t gf_dot_f64(gf_ctx* ctx, int64_t n, const double* x, const double* y, double* d_out,
                   void* stream);
int32_t gf_nrm2_f64(gf_ctx* ctx, int64_t n, const double* x, double* d_out, void* stream);
/* The recomputed-residual guard's two norms without any allocation:
 * d_out[0] = ||b - K_ff(E) u||_2, d_out[1] = ||b||_2 (device, 2 doubles),
 * bitwise what gf_residual_f64 followed by gf_nrm2_f64 on the residual and
 * on b give with this context.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc): the residual and the
 * partial sums live in a file-scope device array sized for at most 4,194,304
 * free DOFs (32 MiB, held by every GPU the library runs on until the process
 * ends); a larger context returns -2.  Calls on one GPU share that array, so
 * calls that overlap on one GPU race.  No host sync. */
int32_t gf_residual_norm_f64_noalloc(gf_ctx* ctx, const double* E, const double* b,
                                     const double* u, double* d_out, void* stream);
/* y = a*x + b*y (host scalars); the vector updates of _cupy_fgmres/_cupy_pcg
 * when they run with caller-supplied operators */
int32_t gf_axpby_f64(int64_t n, double a, const double* x, double b, double* y, void* stream);

/* ---- Galerkin GMG hierarchy ------------------------------------------- */
void gf_gmg_default_params(gf_gmg_params* p);
int32_t gf_gmg_create(gf_gmg** out, gf_ctx* ctx, const gf_gmg_params* params);
int32_t gf_gmg_destroy(gf_gmg* g);
/* E: n_elem FP64 moduli (device); fine_diag: n_free or NULL. Synchronises. */
int32_t gf_gmg_setup(gf_gmg* g, const double* E, const double* fine_diag, void* stream);
int32_t gf_gmg_info_get(gf_gmg* g, gf_gmg_info* info);
/* out = M^{-1} rhs (the full nonlinear preconditioner).  Synchronises the
 * stream once at the end to read the wait-expiry flags: when a grid-wide
 * wait of a Gram-Schmidt sweep (inner or root-local FGMRES), of the
 * persistent coarsest PCG or of a one-launch line search gave up, the call
 * returns -1, the error names the wait, and the flags, the exchange states
 * and the recorded graph are reset (every such call, and gf_gmg_solve,
 * reports and clears them). */
int32_t gf_gmg_apply(gf_gmg* g, const double* rhs, double* out, void* stream);
/* out = one base cycle (V or W) on rhs, no wrapper; wait flags as apply. */
int32_t gf_gmg_vcycle(gf_gmg* g, const double* rhs, double* out, void* stream);
/* No counterpart in the reference (its waits are host synchronisations that
 * cannot expire): sets wait-expiry flag `which` (0 sweep, 1 persistent PCG,
 * 2 one-launch line search)
 * from the host, so the error path of apply / vcycle / cycle_trace / solve
 * can be exercised without any kernel actually giving up. */
int32_t gf_gmg_set_wait_flag(gf_gmg* g, int32_t which, void* stream);
/* y = K_l x at level l (l = 0: fine operator with the set-up moduli). */
int32_t gf_gmg_level_matvec(gf_gmg* g, int32_t level, const double* x, double* y, void* stream);
/* copy the level's inverse diagonal (n_free[level] doubles) */
int32_t gf_gmg_level_dinv(gf_gmg* g, int32_t level, double* out, void* stream);
/* dense copy (row-major n_free x n_free, device) of the level operator, lv >= 1 */
int32_t gf_gmg_level_dense(gf_gmg* g, int32_t level, double* out, void* stream);
/* restriction / prolongation between level and level+1 */
int32_t gf_gmg_restrict(gf_gmg* g, int32_t level, const double* rf, double* rc, void* stream);
int32_t gf_gmg_prolong(gf_gmg* g, int32_t level, const double* ec, double* ef, void* stream);
/* free-DOF list of a level (host ints, n_free[level] entries) */
int32_t gf_gmg_level_free(gf_gmg* g, int32_t level, int32_t* out_host);

/* ---- diagnostics: traced cycle (GalerkinMatFreeGMG.cycle_trace,
 * multigrid_v4.py:1361-1728) ------------------------------------------- */
#define GF_TRACE_SLOTS 32
/* One level visit of a traced cycle, in visit-id (pre-order) order.  v holds
 * the visit's norms and line-search scalars:
 *   0 |rhs|  1 |D^-1 rhs|  2 |x after pre-smoothing|  3 |pre-smoothed
 *   residual r|  4 |R r|  5 |P R r|  6 |r - P R r|  7 |e_c|  8 |P e_c|
 *   9 r.(P e_c)  10 |residual after the first correction|  11-17 first
 *   correction scalars [c.c, r.r, Ac.Ac, r.Ac, applied scale, degenerate
 *   flag, |trial residual|] (line search only)  18 |second coarse rhs|
 *   19 |e_c2|  20 |P e_c2|  21 r1.(P e_c2)  22 |residual after the second
 *   correction|  23-29 second correction scalars  30 |x after
 *   post-smoothing|  31 |rhs - K x_post|.
 * A coarsest visit fills 0, 1, 7 (|x|) and 10 (|rhs - K x|). */
typedef struct gf_trace_visit {
  int32_t visit_id;
  int32_t level;
  int32_t is_coarsest;
  int32_t n_coarse_corrections;
  int32_t enrichment;            /* second correction from root enrichment */
  int32_t extra_fine_postsmooth; /* high-contrast second fine post-smooth */
  double v[GF_TRACE_SLOTS];
} gf_trace_visit;
/* One traced cycle on rhs (cycle_type -1 = configured, 0 = V, 1 = W)
 * followed by the configured root-local correction; out receives the same
 * iterate as gf_gmg_apply with inner_krylov_steps = 0.  visits: host array
 * of capacity cap (n_visits receives the count, also when cap is too
 * small).  summary (host, 16 doubles): 0 |rhs|  1 |D0^-1 rhs|  2 |out|
 * 3 rhs.out  4 |pre-local residual|  5-8 root-local scalars (inner FGMRES:
 * [c.c, r.r, t.t]; line search: [c.c, r.r, Ac.Ac, r.Ac])  9 |line-search
 * trial residual|.  Allocates a small device buffer; synchronises. */
int32_t gf_gmg_cycle_trace(gf_gmg* g, const double* rhs, double* out, int32_t cycle_type,
                           gf_trace_visit* visits, int32_t cap, int32_t* n_visits,
                           double* summary, void* stream);

/* FGMRES(restart) with A = K_ff(E of the last setup), M = gf_gmg_apply.
 * b, x: n_free device.  x0_given: use x as start vector.  history: host
 * array receiving the relative residual history (capacity hist_cap). */
int32_t gf_gmg_solve(gf_gmg* g, const double* b, double* x, int32_t x0_given, double tol,
                     int32_t maxiter, int32_t restart, double* history, int32_t hist_cap,
                     int32_t* n_hist, int32_t* iters, int32_t* converged, void* stream);
/* Continue the last gf_gmg_solve on this hierarchy to a larger maxiter, as if
 * that solve had been called with it: the policy's kept floor re-solves the
 * probe's system from x = 0 with restart >= probe iterations, and the first
 * probe_iters Arnoldi steps of that solve are the probe's, bit for bit
 * (run_gmg_floor_detector_prospective.py:263-277 then 313-351 re-run them).
 * Valid when that solve started from x = 0 (x0_given = 0) with the same b
 * pointer, tol and restart, ran one cycle that converged or stopped at its
 * maxiter before restart, and no gf_gmg_setup came since; otherwise returns
 * -4 and changes nothing.  x receives the iterate of the longer solve;
 * history receives only the entries after the first solve's (iters and
 * converged are those of the whole solve). */
int32_t gf_gmg_solve_continue(gf_gmg* g, const double* b, double* x, double tol,
                              int32_t maxiter, int32_t restart, double* history, int32_t hist_cap,
                              int32_t* n_hist, int32_t* iters, int32_t* converged, void* stream);

/* ---- learned floor predictor (no reference counterpart: the reference's
 * policy always runs the probe; the learned form proposes the floor and r50
 * from the density and the recomputed-residual guard judges the proposal).
 * A 3D CNN on the density average-pooled to 16x8x8 plus three grid
 * features (log2(dim) - 5): conv 3^3 1->16 ReLU, conv 3^3 16->32 ReLU,
 * global average pool, dense 35->32 ReLU, dense 32->4 (output 0 regresses
 * the normalised log10 r50, outputs 1..3 are floor-class logits).  BF16
 * operands with FP32 accumulation in every product, FP32 elsewhere.
 * theta: n_params floats [W1 16x27 | b1 16 | W2 32x27x16 | b2 32 |
 * W3 32x35 | b3 32 | W4 4x32 | b4 4].  work: batch * work_floats_per_sample
 * floats of device scratch owned by the caller.  All pointers are device
 * pointers; no host syncs; deterministic. */
int32_t gf_cnn_sizes(int32_t* n_params, int64_t* work_floats_per_sample);
/* x_out[16*8*8] = adaptive average pool of rho (element order
 * e = (ix*ny + iy)*nz + iz) over bins [floor(i n/m), ceil((i+1) n/m)). */
int32_t gf_cnn_pool_f64(const double* rho, int32_t nx, int32_t ny, int32_t nz, float* x_out,
                        void* stream);
/* out[batch][4] for x[batch][1024], feat[batch][3]. */
int32_t gf_cnn_forward(const float* theta, const float* x, const float* feat, int32_t batch,
                       float* out, float* work, void* stream);
/* Mean loss (w_reg (out0 - y_reg)^2 + cross-entropy(out1..3, y_cls)) into
 * loss_out[1] (nullable) and its gradient into grad_out[n_params]; out
 * (nullable) receives the forward outputs. */
int32_t gf_cnn_grad(const float* theta, const float* x, const float* feat, const float* y_reg,
                    const int32_t* y_cls, int32_t batch, float w_reg, float* loss_out,
                    float* grad_out, float* out, float* work, void* stream);
/* One Adam step (step counts from 1) on n parameters, FP32. */
int32_t gf_cnn_adam(float* theta, float* m, float* v, const float* grad, int32_t n, float lr,
                    float beta1, float beta2, float eps, int32_t step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GFHIP_H */
