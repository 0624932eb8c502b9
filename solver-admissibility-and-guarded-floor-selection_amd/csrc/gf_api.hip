// extern "C" entry points (declared in include/gfhip.h).
#include "gf_reduce.h"
#include "gf_solver.h"

#include <mutex>
#include <new>

using namespace gf;

struct gf_ctx {
  Ctx c;
};
struct gf_gmg {
  Gmg g;
};

static inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

#define CHECK_PTR(p)                                  \
  do {                                                \
    if (!(p)) {                                       \
      set_error(std::string("null argument: ") + #p); \
      return -2;                                      \
    }                                                 \
  } while (0)

// ---------------------------------------------------------------------------
// Recomputed-residual guard without allocation (gf_residual_norm_f64_noalloc).
// Device scratch of a fixed maximum size in a file-scope array: the residual
// (kNoAllocMaxFree doubles) and the per-block partial sums (kRedBlocks).
// Every GPU the library runs on holds its own copy until the process ends;
// host code uses the address hipGetSymbolAddress returns for the current
// device, never the array's name.  Calls on one GPU share that GPU's copy,
// so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {
constexpr long kNoAllocMaxFree = 1L << 22;  // 4,194,304 free DOFs (32 MiB)
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocMaxFree + kRedBlocks];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
static int scratch_ptr(double** out) {
  int dev = 0;
  GF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error("gf_residual_norm_f64_noalloc: device ordinal out of range");
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("gf_residual_norm_f64_noalloc: hipGetSymbolAddress: ") +
                hipGetErrorString(e));
      return -1;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Block partials of x.x: the loop and block tree of k_dot (same bits).
__global__ __launch_bounds__(kBlock) void k_sq_partials(long n, const double* __restrict__ x,
                                                        double* __restrict__ partials) {
  __shared__ double lds[4];
  const double bs = block_sum(dot_thread<double>(n, x, x), lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// sqrt of the partials folded in block order: the fold of k_dot's last block.
__global__ __launch_bounds__(kBlock) void k_fold_sqrt(int nb, const double* __restrict__ partials,
                                                      double* __restrict__ out) {
  __shared__ double lds[4];
  const double s = fold_thread_first(nb, partials, lds);
  if (threadIdx.x == 0) *out = sqrt(s);
}

extern "C" {

int32_t gf_version(void) { return 10000; }

const char* gf_last_error(void) { return last_error(); }

int32_t gf_ctx_create(gf_ctx** out, int32_t nelx, int32_t nely, int32_t nelz, const double* KE,
                      const int32_t* free_dofs, int32_t n_free, int32_t ndof) {
  CHECK_PTR(out);
  CHECK_PTR(KE);
  CHECK_PTR(free_dofs);
  gf_ctx* c = new (std::nothrow) gf_ctx();
  if (!c) {
    set_error("gf_ctx_create: out of host memory");
    return -1;
  }
  int rc = c->c.init(nelx, nely, nelz, KE, free_dofs, n_free, ndof);
  if (rc) {
    c->c.release();
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int32_t gf_ctx_autotune(gf_ctx* ctx, float* us_out, int32_t* shape_out, void* stream) {
  CHECK_PTR(ctx);
  int rc = ctx->c.autotune(S(stream), us_out);
  if (rc) return rc;
  if (shape_out) *shape_out = ctx->c.g.shape;
  return 0;
}

int32_t gf_ctx_set_shape(gf_ctx* ctx, int32_t shape) {
  CHECK_PTR(ctx);
  if (shape != 0 && shape != 1 && shape != 4 && shape != 8 && shape != 16 &&
      shape != kShapeMfma8 && shape != kShapeMfma16 && shape != kShapeBox488 &&
      shape != kShapeBox2816 && shape != kShapeBox288) {
    set_error("gf_ctx_set_shape: shape must be 0, 1, 4, 8, 16, 72, 80, 48, 50 or 44");
    return -2;
  }
  ctx->c.g.shape = shape;
  ctx->c.g.shape32 = 0;  // a forced shape holds for both precisions
  return 0;
}

int32_t gf_set_coarse_form(int32_t form) {
  if (form != 0 && form != 1 && form != 16 && form != 32 && form != 64) {
    set_error("gf_set_coarse_form: form must be 0, 1, 16, 32 or 64");
    return -2;
  }
  set_stencil_form(form != 0);
  set_stencil_npb(form > 1 ? form : 0);
  return 0;
}

int32_t gf_set_ls_form(int32_t form) {
  if (form != 0 && form != 1) {
    set_error("gf_set_ls_form: form must be 0 or 1");
    return -2;
  }
  set_ls_form(form);
  return 0;
}

int32_t gf_set_sweep_form(int32_t form) {
  if (form != 1 && form != 2) {
    set_error("gf_set_sweep_form: form must be 1 or 2");
    return -2;
  }
  set_sweep_form(form);
  return 0;
}

int32_t gf_ctx_destroy(gf_ctx* ctx) {
  if (!ctx) return 0;
  ctx->c.release();
  delete ctx;
  return 0;
}

int32_t gf_matvec_f64(gf_ctx* ctx, const double* E, const double* u, double* y, void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(E);
  CHECK_PTR(u);
  CHECK_PTR(y);
  return fine_matvec<double>(ctx->c.g, ctx->c.map(), ctx->c.Ke, E, u, y, EPI_PLAIN,
                             EpiArgs<double>{}, nullptr, S(stream));
}

int32_t gf_matvec_f32(gf_ctx* ctx, const float* E, const float* u, float* y, void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(E);
  CHECK_PTR(u);
  CHECK_PTR(y);
  return fine_matvec<float>(ctx->c.g, ctx->c.map(), ctx->c.Ke32, E, u, y, EPI_PLAIN,
                            EpiArgs<float>{}, nullptr, S(stream));
}

int32_t gf_residual_f64(gf_ctx* ctx, const double* E, const double* b, const double* u,
                        double* y, void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(E);
  CHECK_PTR(b);
  CHECK_PTR(u);
  CHECK_PTR(y);
  EpiArgs<double> ep;
  ep.b = b;
  return fine_matvec<double>(ctx->c.g, ctx->c.map(), ctx->c.Ke, E, u, y, EPI_RESID, ep, nullptr,
                             S(stream));
}

int32_t gf_matvec_full_f32(gf_ctx* ctx, const float* E, const float* u_full, float* y_full,
                           void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(E);
  CHECK_PTR(u_full);
  CHECK_PTR(y_full);
  // full layout = identity DOF map: y = K u over all DOFs (fixed rows give
  // the reaction K u there), the FusedMatvec.matvec_full contract
  DofMap ident{0, nullptr};
  return fine_matvec<float>(ctx->c.g, ident, ctx->c.Ke32, E, u_full, y_full, EPI_PLAIN,
                            EpiArgs<float>{}, nullptr, S(stream));
}

int32_t gf_matvec_full_bf16(gf_ctx* ctx, const float* E, const float* u_full, float* y_full,
                            void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(E);
  CHECK_PTR(u_full);
  CHECK_PTR(y_full);
  return matvec_full_bf16(ctx->c, E, u_full, y_full, S(stream));
}

int32_t gf_elem_sensitivity_f64(gf_ctx* ctx, const double* u, const double* rho, double penal,
                                 double scale, double* ce, double* dc, void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(u);
  if (!ce && !dc) {
    set_error("gf_elem_sensitivity_f64: ce and dc are both NULL");
    return -2;
  }
  if (dc && !rho) {
    set_error("gf_elem_sensitivity_f64: dc needs rho");
    return -2;
  }
  return elem_sensitivity(ctx->c.g, ctx->c.map(), ctx->c.Ke, u, rho, penal, scale, ce, dc,
                          S(stream));
}

int32_t gf_diag_f64(gf_ctx* ctx, const double* E, double* diag, void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(E);
  CHECK_PTR(diag);
  return fine_diag(ctx->c.g, ctx->c.map(), ctx->c.Ke, E, diag, S(stream));
}

int32_t gf_nodeblock_update(gf_ctx* ctx, const double* E, const double* scalar_diag,
                            void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(E);
  int rc = ctx->c.nodeblock_alloc();
  if (rc) return rc;
  ctx->c.nb_has_pinv = scalar_diag ? 1 : 0;
  return nodeblock_update(ctx->c.g, ctx->c.map(), ctx->c.Ke, E, scalar_diag, ctx->c.nb_inv9,
                          ctx->c.nb_pinv, S(stream));
}

int32_t gf_nodeblock_apply(gf_ctx* ctx, const double* r, double* z, void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(r);
  CHECK_PTR(z);
  if (!ctx->c.nb_inv9) {
    set_error("gf_nodeblock_apply: update() has not been called");
    return -3;
  }
  return nodeblock_apply(ctx->c.g, ctx->c.map(), ctx->c.nb_inv9, ctx->c.nb_pinv,
                         ctx->c.nb_has_pinv, r, z, nullptr, S(stream));
}

// reductions without a context use a per-thread scratch (one stream at a
// time per thread)
static Work* blas_work(gf_ctx* ctx) {
  if (ctx) return &ctx->c.work;
  static thread_local Work w;
  if (w.init() != 0) return nullptr;
  return &w;
}

int32_t gf_dot_f64(gf_ctx* ctx, int64_t n, const double* x, const double* y, double* d_out,
                   void* stream) {
  CHECK_PTR(x);
  CHECK_PTR(y);
  CHECK_PTR(d_out);
  Work* w = blas_work(ctx);
  if (!w) return -1;
  return dot_f64(*w, n, x, y, d_out, false, nullptr, S(stream));
}

int32_t gf_nrm2_f64(gf_ctx* ctx, int64_t n, const double* x, double* d_out, void* stream) {
  CHECK_PTR(x);
  CHECK_PTR(d_out);
  Work* w = blas_work(ctx);
  if (!w) return -1;
  return dot_f64(*w, n, x, x, d_out, true, nullptr, S(stream));
}

int32_t gf_residual_norm_f64_noalloc(gf_ctx* ctx, const double* E, const double* b,
                                     const double* u, double* d_out, void* stream) {
  CHECK_PTR(ctx);
  CHECK_PTR(E);
  CHECK_PTR(b);
  CHECK_PTR(u);
  CHECK_PTR(d_out);
  const long n = ctx->c.n_free;
  if (n > kNoAllocMaxFree) {
    set_error("gf_residual_norm_f64_noalloc: " + std::to_string(n) +
              " free DOFs exceed the scratch maximum of " + std::to_string(kNoAllocMaxFree));
    return -2;
  }
  double* base = nullptr;
  GF_TRY(scratch_ptr(&base));
  double* r = base;
  double* partials = base + kNoAllocMaxFree;
  const hipStream_t st = S(stream);
  EpiArgs<double> ep;
  ep.b = b;
  GF_TRY(fine_matvec<double>(ctx->c.g, ctx->c.map(), ctx->c.Ke, E, u, r, EPI_RESID, ep, nullptr,
                             st));
  // every partial the fold reads is written by the kernel before it on this
  // stream, so nothing in the scratch has to start at zero
  const int nb = red_grid(n);
  const double* vec[2] = {r, b};
  for (int k = 0; k < 2; ++k) {
    hipLaunchKernelGGL(k_sq_partials, dim3(nb), dim3(kBlock), 0, st, n, vec[k], partials);
    GF_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fold_sqrt, dim3(1), dim3(kBlock), 0, st, nb, partials, d_out + k);
    GF_CHECK_LAUNCH();
  }
  return 0;
}

__global__ void k_axpby_host(long n, double a, const double* __restrict__ x, double b,
                             double* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] = (b == 1.0 ? y[i] : b * y[i]) + a * x[i];
}

int32_t gf_axpby_f64(int64_t n, double a, const double* x, double b, double* y, void* stream) {
  CHECK_PTR(x);
  CHECK_PTR(y);
  if (n <= 0) return 0;
  int grid = cdiv(n, kBlock);
  if (grid > 4096) grid = 4096;
  hipLaunchKernelGGL(k_axpby_host, dim3(grid), dim3(kBlock), 0, S(stream), (long)n, a, x, b, y);
  GF_CHECK_LAUNCH();
  return 0;
}

void gf_gmg_default_params(gf_gmg_params* p) {
  if (!p) return;
  p->n_levels = 4;
  p->coarse_smooth_iters = 2;
  p->fine_smoother = 1;
  p->fine_smoother_degree = 2;
  p->omega = 0.5;
  p->dense_chol_max = 5000;
  p->coarse_pcg_iters = 80;
  p->smoother_type = 0;
  p->cheb_lower_frac = 1.0 / 30.0;
  for (int i = 0; i < 8; ++i) p->level_prec[i] = -1;
  p->cycle_type = 0;
  p->coarse_correction_policy = 0;
  p->coarse_correction_acceptance_ratio = 1.0;
  p->coarse_correction_max_scale = 1.0;
  p->root_enrichment_mode = 0;
  p->root_enrichment_weight_floor = 1e-6;
  p->root_enrichment_weight_power = 0.5;
  p->root_local_correction_mode = 0;
  p->root_local_correction_max_scale = 1.0;
  p->root_local_correction_inner_steps = 10;
  p->root_local_correction_inner_restart = 0;
  p->inner_krylov_steps = 0;
  p->inner_krylov_restart = 0;
  p->coarsest_solve = 0;
  p->graph_apply = 1;
  p->mgs_sweep = 1;
  p->coarse_pcg_persistent = 1;
}

int32_t gf_gmg_create(gf_gmg** out, gf_ctx* ctx, const gf_gmg_params* params) {
  CHECK_PTR(out);
  CHECK_PTR(ctx);
  CHECK_PTR(params);
  if (params->inner_krylov_restart > 1024 || params->root_local_correction_inner_restart > 1024) {
    set_error("gf_gmg_create: inner restarts are limited to 1024");
    return -2;
  }
  gf_gmg* g = new (std::nothrow) gf_gmg();
  if (!g) {
    set_error("gf_gmg_create: out of host memory");
    return -1;
  }
  int rc = g->g.create(&ctx->c, params);
  if (rc) {
    g->g.release();
    delete g;
    return rc;
  }
  *out = g;
  return 0;
}

int32_t gf_gmg_destroy(gf_gmg* g) {
  if (!g) return 0;
  g->g.release();
  delete g;
  return 0;
}

int32_t gf_gmg_setup(gf_gmg* g, const double* E, const double* fine_diag, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(E);
  return g->g.setup(E, fine_diag, S(stream));
}

int32_t gf_gmg_info_get(gf_gmg* g, gf_gmg_info* info) {
  CHECK_PTR(g);
  CHECK_PTR(info);
  Gmg& G = g->g;
  *info = gf_gmg_info{};
  info->n_levels = G.L;
  for (int l = 0; l < G.L; ++l) {
    info->n_free[l] = G.lv[l].n_free;
    info->dims[l][0] = G.lv[l].g.nx;
    info->dims[l][1] = G.lv[l].g.ny;
    info->dims[l][2] = G.lv[l].g.nz;
    info->lambda_max[l] = l < (int)G.lam.size() ? G.lam[l] : 0.0;
    info->level_prec[l] = G.lv[l].prec;
  }
  info->omega_used = G.omega_used;
  info->contrast_ratio = G.contrast;
  info->void_frac = G.void_frac;
  info->is_high_contrast = G.high;
  info->fine_degree_adapt = G.fine_deg;
  info->coarse_iters_adapt = G.coarse_it;
  info->coarse_chol_mode = G.chol_mode;
  info->coarse_chol_regularised = G.chol_regularised;
  info->setup_count = G.setup_count;
  return 0;
}

int32_t gf_gmg_apply(gf_gmg* g, const double* rhs, double* out, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(rhs);
  CHECK_PTR(out);
  const int rc = g->g.apply(rhs, out, S(stream));
  if (rc) return rc;
  return g->g.take_wait_error("gf_gmg_apply", S(stream));
}

int32_t gf_gmg_vcycle(gf_gmg* g, const double* rhs, double* out, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(rhs);
  CHECK_PTR(out);
  if (!g->g.is_setup) {
    set_error("GalerkinMatFreeGMG.setup() must be called before apply().");
    return -3;
  }
  const int rc = g->g.base_cycle(rhs, out, nullptr, S(stream));
  if (rc) return rc;
  return g->g.take_wait_error("gf_gmg_vcycle", S(stream));
}

int32_t gf_gmg_cycle_trace(gf_gmg* g, const double* rhs, double* out, int32_t cycle_type,
                           gf_trace_visit* visits, int32_t cap, int32_t* n_visits,
                           double* summary, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(rhs);
  CHECK_PTR(out);
  CHECK_PTR(visits);
  const int rc =
      g->g.cycle_trace(rhs, out, cycle_type, visits, cap, n_visits, summary, S(stream));
  if (rc) return rc;
  return g->g.take_wait_error("gf_gmg_cycle_trace", S(stream));
}

int32_t gf_gmg_set_wait_flag(gf_gmg* g, int32_t which, void* stream) {
  CHECK_PTR(g);
  if (which < 0 || which >= kWaitFlags) {
    set_error("gf_gmg_set_wait_flag: which must be 0 (sweep), 1 (persistent PCG) or 2 (line "
              "search)");
    return -2;
  }
  const double one = 1.0;
  GF_HIP(hipMemcpyAsync(g->g.d_err + which, &one, sizeof(double), hipMemcpyHostToDevice,
                        S(stream)));
  GF_HIP(hipStreamSynchronize(S(stream)));
  return 0;
}

static int check_level(gf_gmg* g, int level, bool need_coarse) {
  if (!g->g.is_setup) {
    set_error("GalerkinMatFreeGMG.setup() must be called first");
    return -3;
  }
  if (level < (need_coarse ? 1 : 0) || level >= g->g.L) {
    set_error("level out of range");
    return -2;
  }
  return 0;
}

int32_t gf_gmg_level_matvec(gf_gmg* g, int32_t level, const double* x, double* y, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(x);
  CHECK_PTR(y);
  int rc = check_level(g, level, false);
  if (rc) return rc;
  return g->g.level_op(level, x, y, EPI_PLAIN, EpiArgs<double>{}, nullptr, S(stream));
}

int32_t gf_gmg_level_dinv(gf_gmg* g, int32_t level, double* out, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(out);
  int rc = check_level(g, level, false);
  if (rc) return rc;
  GF_HIP(hipMemcpyAsync(out, g->g.lv[level].dinv, sizeof(double) * g->g.lv[level].n_free,
                        hipMemcpyDeviceToDevice, S(stream)));
  return 0;
}

int32_t gf_gmg_level_dense(gf_gmg* g, int32_t level, double* out, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(out);
  int rc = check_level(g, level, true);
  if (rc) return rc;
  Level& v = g->g.lv[level];
  return stencil_to_dense(v.g, v.map(), v.A, v.n_free, out, S(stream));
}

int32_t gf_gmg_restrict(gf_gmg* g, int32_t level, const double* rf, double* rc_, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(rf);
  CHECK_PTR(rc_);
  if (level < 0 || level + 1 >= g->g.L) {
    set_error("level out of range");
    return -2;
  }
  Level& f = g->g.lv[level];
  Level& c = g->g.lv[level + 1];
  return restrict_(f.g, f.map(), c.g, c.map(), rf, rc_, nullptr, S(stream));
}

int32_t gf_gmg_prolong(gf_gmg* g, int32_t level, const double* ec, double* ef, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(ec);
  CHECK_PTR(ef);
  if (level < 0 || level + 1 >= g->g.L) {
    set_error("level out of range");
    return -2;
  }
  Level& f = g->g.lv[level];
  Level& c = g->g.lv[level + 1];
  return prolong(f.g, f.map(), c.g, c.map(), ec, ef, nullptr, S(stream));
}

int32_t gf_gmg_level_free(gf_gmg* g, int32_t level, int32_t* out_host) {
  CHECK_PTR(g);
  CHECK_PTR(out_host);
  if (level < 0 || level >= g->g.L) {
    set_error("level out of range");
    return -2;
  }
  const std::vector<int>& f = g->g.lv[level].h_free;
  for (size_t i = 0; i < f.size(); ++i) out_host[i] = f[i];
  return 0;
}

int32_t gf_gmg_solve(gf_gmg* g, const double* b, double* x, int32_t x0_given, double tol,
                     int32_t maxiter, int32_t restart, double* history, int32_t hist_cap,
                     int32_t* n_hist, int32_t* iters, int32_t* converged, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(b);
  CHECK_PTR(x);
  CHECK_PTR(n_hist);
  CHECK_PTR(iters);
  CHECK_PTR(converged);
  int nh = 0, it = 0, cv = 0;
  int rc = g->g.outer_solve(b, x, x0_given, tol, maxiter, restart, history, hist_cap, &nh, &it,
                            &cv, S(stream));
  *n_hist = nh;
  *iters = it;
  *converged = cv;
  return rc;
}

int32_t gf_gmg_solve_continue(gf_gmg* g, const double* b, double* x, double tol,
                              int32_t maxiter, int32_t restart, double* history, int32_t hist_cap,
                              int32_t* n_hist, int32_t* iters, int32_t* converged, void* stream) {
  CHECK_PTR(g);
  CHECK_PTR(b);
  CHECK_PTR(x);
  CHECK_PTR(n_hist);
  CHECK_PTR(iters);
  CHECK_PTR(converged);
  int nh = 0, it = 0, cv = 0;
  int rc = g->g.outer_solve(b, x, 0, tol, maxiter, restart, history, hist_cap, &nh, &it, &cv,
                            S(stream), true);
  *n_hist = nh;
  *iters = it;
  *converged = cv;
  return rc;
}

}  // extern "C"
