// Internal interfaces between the gfhip translation units.
#pragma once

#include "gf_common.h"

#include <utility>
#include <vector>

namespace gf {

// Reduction scratch shared by the kernels of one stream (fixed-size, owned
// by the object that issues the launches).
struct Work {
  double* partials = nullptr;  // 4 * kRedBlocks doubles
  int* ticket = nullptr;
  int* flags = nullptr;        // kSweepMaxBlocks ring positions of the MGS sweep
  double* ring = nullptr;      // 3 x kSweepMaxBlocks exchange slots of the MGS sweep
  int sweep_G = -1;            // grid of the last sweep launched on flags (-1: unknown)
  // one-launch line search: 5 x kRedBlocks block partials, and an arrival
  // counter with the value it holds when a launch starts
  double* lsparts = nullptr;
  unsigned long long* lscnt = nullptr;
  int init();
  void release();
};

// Blocks of one cooperative MGS sweep (bounded by the ring-position array and
// by the exchange ring, three slots per block).
constexpr int kSweepMaxBlocks = 1024;
// Bit pattern of an empty exchange slot (a NaN no arithmetic produces).
constexpr unsigned long long kRingEmpty = 0x7FF4DEADBEEFCAFEull;

// Hardware queues of this process (GPU_MAX_HW_QUEUES, HIP's default 4,
// clamped to [1, 32]).  Kernels whose blocks wait for each other (the
// Gram-Schmidt sweep, the persistent coarsest PCG) may each run on a
// different queue at the same moment, so one such grid may use at most
// 1/hw_queues() of the device's co-resident capacity: then all of them fit
// at once and every wait ends.
int hw_queues();

// Slots of the wait-expiry flag (three doubles per hierarchy): a grid-wide
// wait of a Gram-Schmidt sweep, of the persistent coarsest PCG, of the
// one-launch line search.
enum WaitFlag { WAIT_SWEEP = 0, WAIT_PCG = 1, WAIT_LS = 2, kWaitFlags = 3 };

// Fine-operator launch shapes timed by the autotuner, in this order: 1 lane
// per node, 4 waves per 64 nodes, LDS tiles 8x8 and 4x16, and the same tiles
// with the FP64 contraction on the matrix cores (codes 72 and 80; the FP32
// operator runs the VALU tile of the same patch there), then LDS boxes of
// 4x8x8 and 2x8x16 nodes with one lane per node (codes 48, 50) and of 2x8x8
// nodes with two lanes per node (code 44).
constexpr int kShapeMfma8 = 72, kShapeMfma16 = 80;
constexpr int kShapeBox488 = 48, kShapeBox2816 = 50, kShapeBox288 = 44;

// Largest coarsest system of the two-triangular-solve form (its right-hand
// side is held in LDS); the explicit-inverse form has no such bound.
constexpr int kTrsvMaxN = 5000;
constexpr int kFineShapes = 9;

// Epilogue applied by the operator kernels to each output DOF i:
//   PLAIN : y_i = (A x)_i
//   RESID : y_i = b_i - (A x)_i
//   CHEB0 : dv_i = dinv_i (b_i - (A x)_i);         xo_i = x_i + a dv_i
//   CHEB  : dv_i = a (dinv_i (b_i - (A x)_i) + c dv_i); xo_i = x_i + dv_i
//   JACOBI: xo_i = x_i + (a dinv_i) (b_i - (A x)_i)
enum Epilogue { EPI_PLAIN = 0, EPI_RESID = 1, EPI_CHEB0 = 2, EPI_CHEB = 3, EPI_JACOBI = 4 };

template <typename T>
struct EpiArgs {
  const T* b = nullptr;
  const T* dinv = nullptr;
  T* dv = nullptr;
  T* xo = nullptr;
  T a = 0;
  T c = 0;
};

// Explicitly rounded arithmetic for the operator epilogues: every product
// and sum is one IEEE operation (no contraction left to the compiler), so
// every kernel that applies an epilogue produces the same bits whatever code
// surrounds it.
__device__ __forceinline__ double rn_mul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float rn_mul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double rn_add(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ float rn_add(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ double rn_sub(double a, double b) { return __dsub_rn(a, b); }
__device__ __forceinline__ float rn_sub(float a, float b) { return __fsub_rn(a, b); }
__device__ __forceinline__ double rn_fma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float rn_fma(float a, float b, float c) { return fmaf(a, b, c); }

// The operator epilogue at one output index i from its loaded operands
// (multigrid_v4.py:1783-1812 smoothing steps, solver_v2.py residuals):
//   PLAIN  y = A x               RESID  y = b - A x
//   CHEB0  z = D^-1 (b - A x), dv = z, xo = x + a z
//   CHEB   z = D^-1 (b - A x), d = a (z + c dv), dv = d, xo = x + d
//   JACOBI xo = x + (a D^-1)(b - A x)
template <typename T, int EPI>
__device__ __forceinline__ void epi_store(int i, T ax, T b, T di, T dv, T xv, T* __restrict__ y,
                                          const EpiArgs<T>& ep) {
  if (EPI == EPI_PLAIN) {
    y[i] = ax;
  } else if (EPI == EPI_RESID) {
    y[i] = rn_sub(b, ax);
  } else if (EPI == EPI_CHEB0) {
    const T z = rn_mul(di, rn_sub(b, ax));
    ep.dv[i] = z;
    ep.xo[i] = rn_fma(ep.a, z, xv);
  } else if (EPI == EPI_CHEB) {
    const T z = rn_mul(di, rn_sub(b, ax));
    const T d = rn_mul(ep.a, rn_fma(ep.c, dv, z));
    ep.dv[i] = d;
    ep.xo[i] = rn_add(xv, d);
  } else {
    ep.xo[i] = rn_fma(rn_mul(ep.a, di), rn_sub(b, ax), xv);
  }
}

template <int EPI>
constexpr bool epi_reads_x() {
  return EPI == EPI_CHEB0 || EPI == EPI_CHEB || EPI == EPI_JACOBI;
}

// Epilogue at one output index (i >= 0).
template <typename T, int EPI>
__device__ __forceinline__ void epilogue1(int i, T ax, const T* __restrict__ x,
                                          T* __restrict__ y, const EpiArgs<T>& ep) {
  const T b = EPI != EPI_PLAIN ? ep.b[i] : T(0);
  const T di = epi_reads_x<EPI>() ? ep.dinv[i] : T(0);
  const T xv = epi_reads_x<EPI>() ? x[i] : T(0);
  const T dv = EPI == EPI_CHEB ? ep.dv[i] : T(0);
  epi_store<T, EPI>(i, ax, b, di, dv, xv, y, ep);
}

// Epilogue of the three DOFs of one node (o[q] < 0: fixed DOF, skipped).
// Every operand is loaded before the first store: a store through one
// pointer may alias a later load through another as far as the compiler
// knows, which otherwise serialises the three DOFs' loads.  The indices of
// one node are distinct, so the early loads read what epilogue1 would.
template <typename T>
struct EpiOps3 {
  T b[3], di[3], dv[3], xv[3];
};

// The epilogue operands of a node's three DOFs; they do not depend on the
// operator result, so a kernel may load them before its arithmetic and let
// their latency pass under it.
template <typename T, int EPI>
__device__ __forceinline__ void epi_load3(const int (&o)[3], const T* __restrict__ x,
                                          const EpiArgs<T>& ep, EpiOps3<T>& v) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int i = o[q] >= 0 ? o[q] : 0;
    v.b[q] = EPI != EPI_PLAIN ? ep.b[i] : T(0);
    v.di[q] = epi_reads_x<EPI>() ? ep.dinv[i] : T(0);
    v.xv[q] = epi_reads_x<EPI>() ? x[i] : T(0);
    v.dv[q] = EPI == EPI_CHEB ? ep.dv[i] : T(0);
  }
}

template <typename T, int EPI>
__device__ __forceinline__ void epi_store3(const int (&o)[3], const T (&ax)[3],
                                           const EpiOps3<T>& v, T* __restrict__ y,
                                           const EpiArgs<T>& ep) {
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (o[q] >= 0) epi_store<T, EPI>(o[q], ax[q], v.b[q], v.di[q], v.dv[q], v.xv[q], y, ep);
}

template <typename T, int EPI>
__device__ __forceinline__ void epilogue3(const int (&o)[3], const T (&ax)[3],
                                          const T* __restrict__ x, T* __restrict__ y,
                                          const EpiArgs<T>& ep) {
  EpiOps3<T> v;
  epi_load3<T, EPI>(o, x, ep, v);
  epi_store3<T, EPI>(o, ax, v, y, ep);
}

// ---- BLAS-1 (gf_blas.hip) --------------------------------------------------
int dot_f64(Work& w, long n, const double* x, const double* y, double* out, bool sqrt_,
            const int* gate, hipStream_t st);
int dot_f32(Work& w, long n, const float* x, const float* y, double* out, bool sqrt_,
            const int* gate, hipStream_t st);
int dot4_ls(Work& w, long n, const double* c, const double* r, const double* Ac, double* out4,
            const int* gate, hipStream_t st);
// dot4_ls and, in the same launch, the residual line-search step from those
// statistics (ls_scale): out6[0..3] statistics, out6[4] alpha, out6[5]
// rejection flag.
int dot4_ls_step(Work& w, long n, const double* c, const double* r, const double* Ac,
                 double* out6, double max_scale, const int* gate, hipStream_t st);
int mgs_axpy_dot(Work& w, long n, const double* v, double* wv, const double* h,
                 const double* next, double* out, bool sqrt_, const int* gate, hipStream_t st);
// One Arnoldi orthogonalisation by modified Gram-Schmidt in a single
// cooperative launch: for s = 0..j, h_s = V_s . w, w -= h_s V_s; then
// h_{j+1} = ||w|| and w /= h_{j+1} when h_{j+1} > 1e-14.  H column j
// (row-major, leading dimension m) receives h_0..h_{j+1}.  *err is set to 1
// if a grid barrier wait exceeds its bound.  Skipped when *gate == 0.  Safe to
// record in a hipGraph (no launch-time state).  Returns 1 when the sweep was
// launched, 0 when n is outside the register-resident range (the caller then
// uses mgs_axpy_dot per projection), negative on error.
// Device state of one FGMRES instance that the Givens update of Arnoldi step
// j reads and writes (Krylov::cs, sn, g, sc, fl; see givens_step).
struct GivensArgs {
  double tol;
  double* cs;
  double* sn;
  double* g;
  double* sc;
  int* fl;
};

// Givens update of Hessenberg column j (multigrid_v4.py:495-515): apply the
// previous rotations, form the new one (cs = H_jj/rho, sn = H_{j+1,j}/rho),
// rotate g, record the projected residual |g_{j+1}|/||b|| and h_{j+1,j} in
// sc[2], sc[3], count the step in fl[4], fl[2] = j+1, and on convergence or
// breakdown (projected residual <= tol or h_{j+1,j} <= 1e-14) set fl[3] and
// clear the step gate fl[0].  One thread; skipped while fl[0] == 0.
__device__ __forceinline__ void givens_step(int m, int j, const GivensArgs& a, double* H) {
  int* fl = a.fl;
  if (!fl[0]) return;
  double* cs = a.cs;
  double* sn = a.sn;
  double* g = a.g;
  const double hn = H[(j + 1) * m + j];
  for (int i = 0; i < j; ++i) {
    const double hij = H[i * m + j], hi1 = H[(i + 1) * m + j];
    const double tmp = cs[i] * hij + sn[i] * hi1;
    H[(i + 1) * m + j] = -sn[i] * hij + cs[i] * hi1;
    H[i * m + j] = tmp;
  }
  const double hjj = H[j * m + j], hj1 = H[(j + 1) * m + j];
  const double denom = sqrt(hjj * hjj + hj1 * hj1);
  double c, s;
  if (denom < 1e-300) {
    c = 1.0;
    s = 0.0;
  } else {
    c = hjj / denom;
    s = hj1 / denom;
  }
  cs[j] = c;
  sn[j] = s;
  H[j * m + j] = c * hjj + s * hj1;
  H[(j + 1) * m + j] = 0.0;
  const double gj = g[j];
  g[j + 1] = -s * gj;
  g[j] = c * gj;
  fl[4] += 1;
  fl[2] = j + 1;
  const double res = fabs(g[j + 1]) / a.sc[0];
  a.sc[2] = res;
  a.sc[3] = hn;
  if (res <= a.tol || hn <= 1e-14) {
    fl[3] = 1;
    fl[0] = 0;
  }
}

// When gv is non-null the sweep also performs the Givens update of step j in
// its last block-0 thread (givens_step), in place of a separate launch.
int mgs_sweep(Work& w, long n, const double* V, int j, int m, double* wv, double* H, double* err,
              const int* gate, int cooperative, const GivensArgs* gv, hipStream_t st);
int trial_norm(Work& w, long n, const double* b, const double* x, const double* s, double* y,
               double* out, const int* gate, hipStream_t st);
// The residual line search of one coarse correction in one persistent
// launch: the statistics of dot4_ls_step with the step (into st[0..5]), the
// trial residual y = r - st[4] Ac (y nullable) and its norm (st[6]), and the
// accept/apply of k_ls_apply (x += s c; y = r when rejected) -- the same
// reduction trees and arithmetic as those three launches, so the same bits.
// Returns 1 when launched, 0 when the grid does not fit the co-residency
// rule (the caller then takes the three launches), < 0 on error.
int ls_fused(Work& w, long n, const double* c, const double* r, const double* Ac, double* x,
             double* y, double* st, double max_scale, double ratio, double* err,
             const int* gate, hipStream_t stream);
// Clears the one-launch line search's arrival counter (after an expired wait).
int ls_reset(Work& w, hipStream_t stream);
// Line-search form: 1 = one launch where it fits (default), 0 = three launches.
void set_ls_form(int form);
int get_ls_form();
// Gram-Schmidt sweep variant: 1 = prefetching where it fits, else the lean
// one; 2 = the lean one wherever it fits (same bits)
void set_sweep_form(int form);
int axpby_dev(long n, const double* a, double asign, const double* x, const double* b, double* y,
              const int* gate, hipStream_t st);
int scale_inv(long n, const double* x, const double* h, double thresh, double* y,
              const int* gate, hipStream_t st);
int combine(long n, const double* Z, long ldz, const double* y, const int* jcount, bool acc,
            double* x, const int* gate, hipStream_t st);
int cast_f64_to_f32(long n, const double* x, float* y, hipStream_t st);
int cast_f32_to_f64(long n, const float* x, double* y, hipStream_t st);
int sub_f64(long n, const double* a, const double* b, double* y, hipStream_t st);
int minmax_f64(Work& w, long n, const double* E, double* out2, hipStream_t st);
int count_gt(Work& w, long n, const double* E, const double* thr_src, double scale, double* out,
             hipStream_t st);

// ---- fine level (gf_fine.hip) ----------------------------------------------
// y = K_ff(E) u over the structured hex grid `g` (free layout via `map`).
template <typename T>
int fine_matvec(const Grid& g, const DofMap& map, const T* Ke, const T* E, const T* u, T* y,
                int epi, const EpiArgs<T>& ep, const int* gate, hipStream_t st);
// Launch shape of the fine operator kernel: 1 = one lane per node, 4 = four
// waves per 64 nodes; both give bitwise-identical results.  Process-wide,
// chosen by the autotuner.
void set_fine_split(int split_f64, int split_f32);
int get_fine_split(int is_f64);
// Zero-start first Chebyshev/Jacobi step: dv = dinv*b; x = a*dv (dv may be null).
template <typename T>
int zero_start_step(long n, const T* b, const T* dinv, T a, T* dv, T* x, const int* gate,
                    hipStream_t st);
// ce_e = u_e^T KE u_e and dc_e = -penal*scale*rho_e^(penal-1)*ce_e (either output nullable)
int elem_sensitivity(const Grid& g, const DofMap& map, const double* Ke, const double* u,
                     const double* rho, double penal, double scale, double* ce, double* dc,
                     hipStream_t st);
int fine_diag(const Grid& g, const DofMap& map, const double* Ke, const double* E, double* diag,
              hipStream_t st);
// diag -> 1/where(|d|>1e-12, d, 1) (also fp32 copy when d32 != null)
int safe_inverse(long n, const double* d, double* dinv, float* dinv32, hipStream_t st);
int nodeblock_update(const Grid& g, const DofMap& map, const double* Ke, const double* E,
                     const double* scalar_diag, double* inv9, double* pinv, hipStream_t st);
int nodeblock_apply(const Grid& g, const DofMap& map, const double* inv9, const double* pinv,
                    int has_pinv, const double* r, double* z, const int* gate, hipStream_t st);

// ---- coarse levels (gf_coarse.hip) -----------------------------------------
int level1_assemble(const Grid& gc, const DofMap& mapc, const Grid& gf, const double* Q,
                    const double* E, double* A, hipStream_t st);
template <typename T>
int stencil_matvec(const Grid& g, const DofMap& map, const T* A, const T* x, T* y, int epi,
                   const EpiArgs<T>& ep, const int* gate, hipStream_t st);
// 0 = slot-per-lane, 1 = node-per-lane coarse operator kernel (same bits)
void set_stencil_form(int form);
// Nodes per workgroup of the node-per-lane coarse operator (16, 32 or 64;
// 0 = by level size).
void set_stencil_npb(int npb);
int stencil_diag(const Grid& g, const DofMap& map, const double* A, double* diag, hipStream_t st);
int prolong(const Grid& gf, const DofMap& mapf, const Grid& gc, const DofMap& mapc,
            const double* ec, double* yf, const int* gate, hipStream_t st);
// First smoothing step from a zero start on the coarse level, applied by the
// restriction that produces that level's right-hand side b (one launch
// fewer per coarse level visit): mode 1 Chebyshev, z = D^-1 b, dv = z,
// xo = 0 + a z; mode 2 Jacobi, xo = 0 + (a D^-1) b.  The same expressions as
// zero_start_step, so the bits are the same.
struct ZeroStart {
  int mode = 0;
  const double* dinv = nullptr;
  double* dv = nullptr;
  double* xo = nullptr;
  double a = 0.0;
};
int restrict_(const Grid& gf, const DofMap& mapf, const Grid& gc, const DofMap& mapc,
              const double* rf, double* yc, const int* gate, hipStream_t st,
              const ZeroStart* zs = nullptr);
int rap(const Grid& gf, const DofMap& mapf, const Grid& gc, const DofMap& mapc,
        const double* Af, double* KP, double* Ac, hipStream_t st);
int stencil_to_dense(const Grid& g, const DofMap& map, const double* A, int n_free, double* D,
                     hipStream_t st);
int dense_cholesky(int n, double* D, double* Lt, int* fail, hipStream_t st);
int dense_add_diag(int n, double* D, double eps, hipStream_t st);
int dense_chol_solve(int n, const double* L, const double* Lt, const double* rdiag,
                     const double* b, double* x, const int* gate, hipStream_t st);
int dense_recip_diag(int n, const double* L, double* rdiag, hipStream_t st);
int dense_inverse_from_chol(int n, const double* L, double* Y, double* Yt, double* Kinv,
                            hipStream_t st);
int dense_gemv(int n, const double* K, const double* b, double* x, const int* gate,
               hipStream_t st);
// Fixed-count Jacobi PCG (zero start) on a stencil level, one launch per
// vector phase over the whole grid; work: 4 * n_free + 4 doubles.
int fixed_pcg_grid(Work& w, const Grid& g, const DofMap& map, const double* A,
                   const double* dinv, int n_free, int iters, const double* b, double* x,
                   double* work, const int* gate, hipStream_t st);
// The same PCG in one persistent launch (grid-wide exchanges through the
// uncached buffer from pcg_buffers_acquire; *err = 1 if a bounded wait
// expires).  Returns 1 when launched, 0 when the level needs more than
// 1024 blocks (the caller then uses fixed_pcg_grid), negative on error.
int fixed_pcg_persistent(const Grid& g, const DofMap& map, const double* A, const double* dinv,
                         int n_free, int iters, const double* b, double* x, void* ucbuf,
                         double* err, const int* gate, hipStream_t st);
// Uncached scratch of the persistent PCG (exchange ring, r and two
// direction buffers), recycled by size and never returned to the driver.
int pcg_buffers_acquire(long n_free, void** out, size_t* bytes_out);
void pcg_buffers_release(void* p, size_t bytes);
int f64_to_f32_block(long n, const double* A, float* A32, hipStream_t st);

}  // namespace gf
