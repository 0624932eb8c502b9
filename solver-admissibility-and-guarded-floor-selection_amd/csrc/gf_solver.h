// Solver objects behind the C ABI: the fine-operator context, the Galerkin
// GMG hierarchy with its nonlinear wrapper, and device-resident FGMRES.
#pragma once

#include "gf_internal.h"
#include "../../include/gfhip.h"

#include <functional>
#include <vector>

namespace gf {

// Fine operator: structured hex grid, free-DOF layout, unit stiffness.
struct Ctx {
  Grid g{};
  int n_free = 0;
  int off = -1;                 // offset-map start or -1 (table mode)
  int* d_f2f = nullptr;         // ndof ints
  std::vector<int> h_free;
  std::vector<int> h_f2f;
  double h_Ke[576];
  double* Ke = nullptr;
  float* Ke32 = nullptr;
  Work work;
  // node-block Jacobi state (solver_v2.py NodeBlockJacobiMF)
  double* nb_inv9 = nullptr;
  double* nb_pinv = nullptr;
  int nb_has_pinv = 0;
  // bf16 matrix-core operator (allocated on first use)
  unsigned short* KEb = nullptr;  // 32x32 bf16 bits, zero padded
  int bf16_alloc();
  DofMap map() const { return DofMap{off, d_f2f}; }
  int init(int nelx, int nely, int nelz, const double* Ke_host, const int* free_host, int nfree,
           int ndof);
  int nodeblock_alloc();
  // time the fine-operator launch shapes on this grid, keep the fastest
  int autotune(hipStream_t st, float* us_out);
  void release();
};

// y_full = sum_e E_e KE_bf16 bf16(u_e) over all DOFs (FP32 accumulation)
int matvec_full_bf16(Ctx& c, const float* E, const float* u_full, float* y_full, hipStream_t st);
// The same operator on a DOF layout with an operator epilogue (the bf16 fine
// smoother's A_op); KEb from Ctx::bf16_alloc.
int fine_matvec_bf16(const Grid& g, const DofMap& map, const unsigned short* KEb, const float* E,
                     const float* u, float* y, int epi, const EpiArgs<float>& ep, const int* gate,
                     hipStream_t st);

// One device-resident flexible GMRES instance (fixed maximum restart).
struct Krylov {
  int m = 0;
  long n = 0;
  double* V = nullptr;   // (m+1) x n
  double* Z = nullptr;   // m x n
  double* H = nullptr;   // (m+1) x m, row-major
  double* cs = nullptr;
  double* sn = nullptr;
  double* g = nullptr;   // m+1
  double* y = nullptr;   // m
  double* sc = nullptr;  // scalars: [0]=bnorm [1]=beta [2]=hist value [3]=h_{j+1,j} [4]=sweep error
  int* fl = nullptr;     // [0]=step gate [1]=run gate [2]=jfinal [3]=converged [4]=total [5]=err
  int alloc(int m_, long n_);
  void release();
};

struct Level {
  Grid g{};
  int n_free = 0;
  int off = -1;
  int* d_f2f = nullptr;
  std::vector<int> h_free;
  double* A = nullptr;     // 27x9 block stencil (lv >= 1)
  float* A32 = nullptr;
  double* KP = nullptr;    // R(KP) scratch on the fine side of this level's product
  double* diag = nullptr;
  double* dinv = nullptr;
  float* dinv32 = nullptr;
  int prec = 0;            // 0 fp64, 1 fp32
  // work vectors
  double* X[2] = {nullptr, nullptr};
  double *RHS = nullptr, *R = nullptr, *PE = nullptr, *AC = nullptr, *DV = nullptr,
         *T = nullptr, *R2 = nullptr;
  float *X32[2] = {nullptr, nullptr};
  float *RHS32 = nullptr, *DV32 = nullptr;
  double* st = nullptr;    // 16 device scalars for the safeguarded correction
  DofMap map() const { return DofMap{off, d_f2f}; }
};

struct Gmg {
  Ctx* ctx = nullptr;
  gf_gmg_params p{};
  int L = 1;
  std::vector<Level> lv;
  // set-up state
  double* E64 = nullptr;
  float* E32 = nullptr;
  double* Q = nullptr;        // 8 x 576 sub-element Galerkin blocks
  bool uploaded = false;
  double lam0 = NAN;          // fine lambda_max estimate
  double lam_E_max = NAN;
  std::vector<double> lam;    // per level
  double omega_used = 0.5;
  double contrast = NAN, void_frac = NAN, E_max = NAN, E_min = NAN;
  int high = 0, fine_deg = 2, coarse_it = 2;
  bool fine_bf16 = false;     // level 0 smooths with the bf16 matrix-core operator
  int chol_mode = 0;          // 0 none, 1 dense, 2 iterative
  int chol_regularised = 0;
  double* CH = nullptr;       // dense L (row-major)
  double* CHt = nullptr;      // its transpose
  double* CHr = nullptr;      // reciprocal pivots 1/L_ii
  double* CHi = nullptr;      // K^-1 = L^-T L^-1 (coarsest_solve == 0)
  double* CHy = nullptr;      // scratch for the inverse
  int ch_n = 0;
  int* d_fail = nullptr;
  double* pcg_work = nullptr;
  void* pcg_uc = nullptr;     // uncached buffers of the persistent coarsest PCG
  size_t pcg_uc_bytes = 0;
  double* re_w = nullptr;     // root-enrichment weights
  double* nb_inv9 = nullptr;  // root-local node-block Jacobi (own copy)
  double* nb_pinv = nullptr;
  bool is_setup = false;
  int setup_count = 0;
  // wrapper work
  Krylov ik, rl;
  double *ap_x = nullptr, *ap_res = nullptr, *ap_corr = nullptr, *ap_trial = nullptr,
         *ap_st = nullptr;
  // outer solve
  Krylov outer;
  double* hs = nullptr;       // pinned host mirror for outer control
  double* d_err = nullptr;    // kWaitFlags slots: set to 1 by a sweep / persistent PCG whose
                              // grid-wide wait expired
  // scratch for power iterations
  double *pw_v = nullptr, *pw_d = nullptr, *pw_kv = nullptr, *pw_av = nullptr, *pw_s = nullptr;
  Work work;

  // hipGraph of one preconditioner application (rhs in gr_in, result in gr_out)
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  double *gr_in = nullptr, *gr_out = nullptr;
  int graph_builds = 0;

  int create(Ctx* c, const gf_gmg_params* params);
  int setup(const double* E, const double* fine_diag, hipStream_t st);
  int apply(const double* rhs, double* out, hipStream_t st);
  int apply_impl(const double* rhs, double* out, hipStream_t st);
  int root_local(const double* rhs, const double* xp, double* out, hipStream_t st);
  // traced cycle (diagnostics): slots of one visit are GF_TRACE_SLOTS doubles
  struct TraceMeta {
    int level, is_coarsest, n_corr, enrich, extra;
  };
  struct TraceState {
    std::vector<TraceMeta> meta;
    double* slots = nullptr;  // device, cap * GF_TRACE_SLOTS
    const double* one = nullptr;
    int cap = 0;
  };
  int trace_visit(int l, const double* rhs, double* out, bool wcyc, hipStream_t st,
                  double** result, TraceState& ts);
  int cycle_trace(const double* rhs, double* out, int cycle_type, gf_trace_visit* visits, int cap,
                  int* n_visits, double* summary, hipStream_t st);
  int capture();
  void drop_graph();
  int base_cycle(const double* rhs, double* out, const int* gate, hipStream_t st);
  int outer_solve(const double* b, double* x, int x0_given, double tol, int maxiter, int restart,
                  double* history, int hist_cap, int* n_hist, int* iters, int* converged,
                  hipStream_t st, bool resume = false);
  // where the last outer solve stopped, for gf_gmg_solve_continue
  struct Resume {
    bool valid = false;
    const double* b = nullptr;
    double tol = 0.0, bnorm = 0.0;
    int restart = 0, total = 0;
    long n = 0;
    bool conv = false;
  } rs;
  void release();

  // internals
  int level_op(int l, const double* x, double* y, int epi, const EpiArgs<double>& ep,
               const int* gate, hipStream_t st);
  int smooth(int l, const double* x_in, const double* rhs, double* out, int sweeps_scale,
             const int* gate, hipStream_t st, double** result);
  int cheb(int l, const double* x_in, const double* rhs, double* out, const int* gate,
           hipStream_t st, double** result);
  int cycle(int l, const double* rhs, double* out, const int* gate, hipStream_t st,
            double** result);
  int correct(int l, double* x, const double* residual, const double* corr, double* residual_out,
              const int* gate, hipStream_t st);
  int solve_coarsest(const double* rhs, double* out, const int* gate, hipStream_t st,
                     double** result);
  int enrichment_rhs(const double* residual, double* out, const int* gate, hipStream_t st,
                     const ZeroStart* zs = nullptr);
  const ZeroStart* fuse_zero_start(int lc);
  ZeroStart zs_{};
  int zs_level = -1;  // level whose next zero-start step a restriction applied
  int fgmres_dev(Krylov& K, const std::function<int(const double*, double*, const int*)>& M,
                 const double* b, double* x, double tol, int maxiter, int restart,
                 hipStream_t st);
  int estimate_fine_lambda(hipStream_t st, double* out);
  int estimate_level_lambda(int l, hipStream_t st, double* out);
  int read_scalars(const double* dptr, double* host, int n, hipStream_t st);
  int take_wait_error(const char* where, hipStream_t st);
  int wait_failed(const double* flags, const char* where, hipStream_t st);
};

}  // namespace gf
