// Fine-operator context, Galerkin GMG hierarchy (set-up + nonlinear
// preconditioner) and flexible GMRES, orchestrated on one HIP stream.
//
// The preconditioner application never synchronises with the host: every
// scalar decision of the reference (safeguarded line search accept/reject,
// root-local accept/reject, Givens rotations, lucky breakdown, inner-Krylov
// stopping) is evaluated by single-thread kernels on device scalars, and the
// kernels that follow read the outcome through `gate` flags.  Only set-up
// (spectral estimates, contrast statistics, factorisation status) and the
// outer FGMRES convergence test read scalars back to the host.
#include "gf_solver.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace gf {

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__global__ void k_fill(long n, double* x, double v, const int* gate) {
  if (gate && *gate == 0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    x[i] = v;
}

__global__ void k_copy(long n, const double* __restrict__ a, double* __restrict__ b,
                       const int* gate) {
  if (gate && *gate == 0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_add(long n, double* __restrict__ x, const double* __restrict__ a,
                      const int* gate) {
  if (gate && *gate == 0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    x[i] = x[i] + a[i];
}

// elementwise helpers for the power iterations
__global__ void k_recip(long n, const double* a, double* b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    b[i] = 1.0 / a[i];
}
__global__ void k_mul(long n, const double* a, const double* b, double* c) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    c[i] = a[i] * b[i];
}
__global__ void k_div_host(long n, const double* a, double s, double* b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    b[i] = a[i] / s;
}
__global__ void k_div_dev(long n, double* a, const double* s) {
  const double v = *s;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    a[i] = a[i] / v;
}
// out = sum_i (a_i * w_i) * b_i     (mode 0)
// out = sum_i a_i * (a_i / w_i)     (mode 1)
__global__ __launch_bounds__(kBlock) void k_wdot(long n, const double* a, const double* w,
                                                 const double* b, int mode, double* partials,
                                                 double* out) {
  __shared__ double lds[kBlock];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    acc += mode == 0 ? (a[i] * w[i]) * b[i] : a[i] * (a[i] / w[i]);
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = lds[0];
}
__global__ void k_fold(int nb, const double* partials, double* out, int post) {
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += partials[i];
  *out = post ? sqrt(s) : s;
}

// ---- FGMRES control ---------------------------------------------------------
// fl: [0]=step gate [1]=run gate [2]=jfinal [3]=converged [4]=total [5]=err
__global__ void k_fgm_begin(int* fl, const double* sc) {
  const int run = sc[0] > 1e-30;
  fl[0] = run;
  fl[1] = run;
  fl[2] = 0;
  fl[3] = 0;
  fl[4] = 0;
  fl[5] = 0;
}

// cycle start: beta = sc[1]; stop if converged earlier or beta/bnorm <= tol
__global__ void k_cycle_start(int* fl, double* sc, double* g, double tol) {
  if (fl[3]) {
    fl[0] = 0;
    fl[1] = 0;
    return;
  }
  if (!fl[1]) {
    fl[0] = 0;
    return;
  }
  const double beta = sc[1];
  if (beta / sc[0] <= tol) {
    fl[3] = 1;
    fl[0] = 0;
    fl[1] = 0;
    return;
  }
  g[0] = beta;
  fl[0] = 1;
  fl[2] = 0;
}

__global__ void k_copy_scalar(const double* a, double* b) { *b = *a; }

// Givens update of column j (multigrid_v4.py:485-517).  sc[2] <- |g_{j+1}|/bnorm,
// sc[3] <- h_{j+1,j} before rotation.
__global__ void k_givens(int m, int j, double tol, double* H, double* cs, double* sn, double* g,
                         double* sc, int* fl) {
  givens_step(m, j, GivensArgs{tol, cs, sn, g, sc, fl}, H);
}

// y = R^{-1} g over the leading jfinal x jfinal triangle (column sweep,
// descending columns as LAPACK's upper-triangular solve).
__global__ __launch_bounds__(256) void k_backsolve(int m, const double* H, const double* g,
                                                   double* y, int* fl) {
  if (!fl[1]) return;
  const int n = fl[2];
  __shared__ double v[1024];
  for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = g[i];
  __syncthreads();
  for (int l = n - 1; l >= 0; --l) {
    const double d = H[l * m + l];
    double yl = 0.0;
    if (d != 0.0) yl = v[l] / d;
    else if (threadIdx.x == 0) fl[5] = 1;
    __syncthreads();
    if (threadIdx.x == 0) v[l] = yl;
    for (int k = threadIdx.x; k < l; k += blockDim.x) v[k] -= yl * H[k * m + l];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) y[i] = v[i];
}

// ---- safeguarded corrections -----------------------------------------------
// st: [0]=c.c [1]=r.r [2]=Ac.Ac [3]=r.Ac [4]=scale [5]=degenerate [6]=trial norm
__global__ void k_ls_apply(long n, double* __restrict__ x, const double* __restrict__ c,
                           const double* st, double ratio, const double* __restrict__ r,
                           double* __restrict__ r2, const int* gate) {
  if (gate && *gate == 0) return;
  const double rn = sqrt(st[1]);
  const bool acc = (st[5] == 0.0) && !(st[6] > ratio * rn);
  const double s = acc ? st[4] : 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    x[i] = x[i] + s * c[i];
    if (r2 && !acc) r2[i] = r[i];
  }
}

// root-local inner FGMRES acceptance (multigrid_v4.py:969-1004)
// st: [0]=corr.corr [1]=res.res [2]=trial.trial
__global__ void k_rl_accept(long n, const double* __restrict__ x, const double* __restrict__ c,
                            const double* st, double* __restrict__ out) {
  const double cn = sqrt(st[0]), rn = sqrt(st[1]), tn = sqrt(st[2]);
  const bool acc = !(cn <= 1e-300 || rn <= 1e-300) && !(tn > rn);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = acc ? x[i] + c[i] : x[i];
}

// root-local node-block line search (multigrid_v4.py:1006-1055)
// st: [0]=c.c [1]=r.r [2]=Ac.Ac [3]=r.Ac
__global__ void k_rl_ls(long n, const double* __restrict__ x, const double* __restrict__ c,
                        const double* st, double max_scale, double* __restrict__ out) {
  const double cn = sqrt(st[0]), rn = sqrt(st[1]), den = st[2];
  double s = 0.0;
  if (!(cn <= 1e-300 || rn <= 1e-300 || den <= 1e-300)) {
    const double raw = st[3] / den;
    s = fmin(max_scale, fmax(0.0, raw));
    if (s <= 1e-12) s = 0.0;
  }
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = s != 0.0 ? x[i] + s * c[i] : x[i];
}

// root-enrichment weighting (multigrid_v4.py:938-962)
__global__ void k_enrich(long n, const double* w, int void_mode, double wfloor,
                         const double* res, const double* proj, double* out, const int* gate) {
  if (gate && *gate == 0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    double wi = w[i];
    if (void_mode) wi = fmin(fmax(1.0 - wi, wfloor), 1.0);
    const double v = proj ? res[i] - proj[i] : res[i];
    out[i] = wi * v;
  }
}

__global__ void k_enrich_weights(long n, const double* d, const double* mx, double wfloor,
                                 double power, double* w) {
  const double m = fmax(*mx, 1e-300);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    double v = fabs(d[i]) / m;
    v = fmin(fmax(v, wfloor), 1.0);
    if (fabs(power - 1.0) > 1e-12) v = pow(v, power);
    w[i] = v;
  }
}

__global__ void k_safe_diag(long n, double* d) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    d[i] = fabs(d[i]) > 1e-12 ? d[i] : 1.0;
}

static inline int vgrid(long n) {
  int g = cdiv(n, kBlock);
  return g < 1 ? 1 : (g > 4096 ? 4096 : g);
}

#define LAUNCH(k, n, st, ...)                                                  \
  do {                                                                         \
    hipLaunchKernelGGL(k, dim3(vgrid(n)), dim3(kBlock), 0, st, __VA_ARGS__);   \
    GF_CHECK_LAUNCH();                                                         \
  } while (0)

#define LAUNCH1(k, st, ...)                                                    \
  do {                                                                         \
    hipLaunchKernelGGL(k, dim3(1), dim3(1), 0, st, __VA_ARGS__);               \
    GF_CHECK_LAUNCH();                                                         \
  } while (0)

template <typename T>
static int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  GF_HIP(hipMalloc((void**)p, sizeof(T) * n));
  return 0;
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// ---------------------------------------------------------------------------
// Ctx
// ---------------------------------------------------------------------------
int Ctx::init(int nelx, int nely, int nelz, const double* Ke_host, const int* free_host,
              int nfree, int ndof) {
  GF_DEVICE_LOCK();
  if (nelx < 1 || nely < 1 || nelz < 1) {
    set_error("gf_ctx_create: grid dimensions must be >= 1");
    return -2;
  }
  g = make_grid(nelx, nely, nelz);
  if (ndof != g.ndof) {
    set_error("gf_ctx_create: ndof does not match the grid");
    return -2;
  }
  if (nfree < 1 || nfree > ndof) {
    set_error("gf_ctx_create: n_free out of range");
    return -2;
  }
  h_free.assign(free_host, free_host + nfree);
  h_f2f.assign(ndof, -1);
  for (int i = 0; i < nfree; ++i) {
    const int d = h_free[i];
    if (d < 0 || d >= ndof || (i > 0 && d <= h_free[i - 1])) {
      set_error("gf_ctx_create: free DOFs must be sorted, unique and in range");
      return -2;
    }
    h_f2f[d] = i;
  }
  n_free = nfree;
  off = (h_free[0] + nfree == ndof) ? h_free[0] : -1;  // sorted+unique => contiguous tail
  std::memcpy(h_Ke, Ke_host, sizeof(double) * 576);
  GF_TRY(dalloc(&d_f2f, ndof));
  GF_TRY(upload_now(d_f2f, h_f2f.data(), sizeof(int) * ndof));
  GF_TRY(dalloc(&Ke, 576));
  GF_TRY(dalloc(&Ke32, 576));
  float k32[576];
  for (int i = 0; i < 576; ++i) k32[i] = (float)h_Ke[i];
  GF_TRY(upload_now(Ke, h_Ke, sizeof(double) * 576));
  GF_TRY(upload_now(Ke32, k32, sizeof(float) * 576));
  GF_TRY(work.init());
  return 0;
}

int Ctx::nodeblock_alloc() {
  GF_DEVICE_LOCK();
  if (nb_inv9) return 0;
  GF_TRY(dalloc(&nb_inv9, 9L * g.n_nodes));
  GF_TRY(dalloc(&nb_pinv, n_free));
  GF_TRY(zero_now(nb_inv9, sizeof(double) * 9L * g.n_nodes));
  GF_TRY(zero_now(nb_pinv, sizeof(double) * n_free));
  return 0;
}

int Ctx::autotune(hipStream_t st, float* us_out) {
  GF_DEVICE_LOCK();
  const long ne = (long)g.nx * g.ny * g.nz;
  // timing buffers and events, released on every return path
  struct Scratch {
    double *E = nullptr, *u = nullptr, *y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scratch() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      dfree(E);
      dfree(u);
      dfree(y);
    }
  } sc;
  GF_TRY(dalloc(&sc.E, ne));
  GF_TRY(dalloc(&sc.u, n_free));
  GF_TRY(dalloc(&sc.y, n_free));
  double *E = sc.E, *u = sc.u, *y = sc.y;
  hipLaunchKernelGGL(k_fill, dim3(256), dim3(kBlock), 0, st, ne, E, 1.0, (const int*)nullptr);
  hipLaunchKernelGGL(k_fill, dim3(256), dim3(kBlock), 0, st, (long)n_free, u, 1.0,
                     (const int*)nullptr);
  GF_HIP(hipEventCreate(&sc.e0));
  GF_HIP(hipEventCreate(&sc.e1));
  hipEvent_t e0 = sc.e0, e1 = sc.e1;
  // the shapes are timed in three interleaved rounds and each keeps its best
  // round, so a burst of other work on the device (concurrent instances set
  // up at the same time) does not decide the choice
  const int shapes[kFineShapes] = {1,
                                   4,
                                   8,
                                   16,
                                   kShapeMfma8,
                                   kShapeMfma16,
                                   kShapeBox488,
                                   kShapeBox2816,
                                   kShapeBox288};
  float tmin[kFineShapes];
  for (int si = 0; si < kFineShapes; ++si) tmin[si] = 1e30f;
  for (int round = 0; round < 3; ++round) {
    for (int si = 0; si < kFineShapes; ++si) {
      Grid gg = g;
      gg.shape = shapes[si];
      for (int w = 0; w < 3; ++w)
        GF_TRY(fine_matvec<double>(gg, map(), Ke, E, u, y, EPI_PLAIN, EpiArgs<double>{}, nullptr,
                                   st));
      GF_HIP(hipEventRecord(e0, st));
      const int reps = 20;
      for (int r = 0; r < reps; ++r)
        GF_TRY(fine_matvec<double>(gg, map(), Ke, E, u, y, EPI_PLAIN, EpiArgs<double>{}, nullptr,
                                   st));
      GF_HIP(hipEventRecord(e1, st));
      GF_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      GF_HIP(hipEventElapsedTime(&ms, e0, e1));
      tmin[si] = std::min(tmin[si], ms * 1000.f / reps);
    }
  }
  float best = 1e30f;
  int best_shape = 4;
  for (int si = 0; si < kFineShapes; ++si) {
    if (us_out) us_out[si] = tmin[si];
    if (tmin[si] < best) {
      best = tmin[si];
      best_shape = shapes[si];
    }
  }
  g.shape = best_shape;
  // the FP32 operator (multigrid smoothers of the FP32 stacks) is timed on
  // its own: its best shape can differ (one lane per node in a 4x8x8 box at
  // 1M elements); the double buffers serve as float buffers of the same sizes
  {
    const float* E32 = reinterpret_cast<const float*>(E);
    const float* u32 = reinterpret_cast<const float*>(u);
    float* y32 = reinterpret_cast<float*>(y);
    float t32[kFineShapes];
    for (int si = 0; si < kFineShapes; ++si) t32[si] = 1e30f;
    for (int round = 0; round < 3; ++round) {
      for (int si = 0; si < kFineShapes; ++si) {
        Grid gg = g;
        gg.shape = shapes[si];
        gg.shape32 = 0;
        for (int w = 0; w < 3; ++w)
          GF_TRY(fine_matvec<float>(gg, map(), Ke32, E32, u32, y32, EPI_PLAIN, EpiArgs<float>{},
                                    nullptr, st));
        GF_HIP(hipEventRecord(e0, st));
        const int reps = 20;
        for (int r = 0; r < reps; ++r)
          GF_TRY(fine_matvec<float>(gg, map(), Ke32, E32, u32, y32, EPI_PLAIN, EpiArgs<float>{},
                                    nullptr, st));
        GF_HIP(hipEventRecord(e1, st));
        GF_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        GF_HIP(hipEventElapsedTime(&ms, e0, e1));
        t32[si] = std::min(t32[si], ms * 1000.f / reps);
      }
    }
    int best32 = best_shape;
    float b32 = 1e30f;
    for (int si = 0; si < kFineShapes; ++si)
      if (t32[si] < b32) {
        b32 = t32[si];
        best32 = shapes[si];
      }
    g.shape32 = best32;
  }
  return 0;
}

void Ctx::release() {
  GF_DEVICE_LOCK();
  dfree(d_f2f);
  dfree(Ke);
  dfree(Ke32);
  dfree(KEb);
  dfree(nb_inv9);
  dfree(nb_pinv);
  work.release();
}

// ---------------------------------------------------------------------------
// Krylov
// ---------------------------------------------------------------------------
int Krylov::alloc(int m_, long n_) {
  GF_DEVICE_LOCK();
  if (V && m_ == m && n_ == n) return 0;
  release();
  m = m_;
  n = n_;
  GF_TRY(dalloc(&V, (size_t)(m + 1) * n));
  GF_TRY(dalloc(&Z, (size_t)m * n));
  GF_TRY(dalloc(&H, (size_t)(m + 1) * m));
  GF_TRY(dalloc(&cs, m));
  GF_TRY(dalloc(&sn, m));
  GF_TRY(dalloc(&g, m + 1));
  GF_TRY(dalloc(&y, m));
  GF_TRY(dalloc(&sc, 8));
  GF_TRY(dalloc(&fl, 8));
  GF_TRY(zero_now(fl, sizeof(int) * 8));
  GF_TRY(zero_now(sc, sizeof(double) * 8));
  return 0;
}

void Krylov::release() {
  GF_DEVICE_LOCK();
  dfree(V);
  dfree(Z);
  dfree(H);
  dfree(cs);
  dfree(sn);
  dfree(g);
  dfree(y);
  dfree(sc);
  dfree(fl);
  m = 0;
  n = 0;
}

// ---------------------------------------------------------------------------
// hierarchy construction (multigrid_v4.py:546-770, solver_v2.py:62-149/716-763)
// ---------------------------------------------------------------------------
static void local_prolongation(int s, int t, int u, double P[8][8]) {
  static const int pos[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const int stu[3] = {s, t, u};
  for (int fn = 0; fn < 8; ++fn) {
    double cc[3];
    for (int a = 0; a < 3; ++a) cc[a] = (stu[a] + pos[fn][a]) / 2.0;
    for (int cn = 0; cn < 8; ++cn) {
      double w = 1.0;
      for (int a = 0; a < 3; ++a) w *= pos[cn][a] == 0 ? (1.0 - cc[a]) : cc[a];
      P[fn][cn] = w;
    }
  }
}

static void galerkin_q(const double* Ke, std::vector<double>& Q) {
  Q.assign(8 * 576, 0.0);
  int idx = 0;
  for (int s = 0; s < 2; ++s)
    for (int t = 0; t < 2; ++t)
      for (int u = 0; u < 2; ++u, ++idx) {
        double P[8][8];
        local_prolongation(s, t, u, P);
        // P24 = kron(P, I3): P24[3f+a][3c+b] = P[f][c] * (a==b)
        // (P24^T Ke) first, then times P24 -- P24 = kron(P, I3)
        double KP[24][24];
        double PtK[24][24];
        for (int i = 0; i < 24; ++i)
          for (int k = 0; k < 24; ++k) {
            double acc = 0.0;
            for (int r = 0; r < 24; ++r) {
              const double p = (r % 3 == i % 3) ? P[r / 3][i / 3] : 0.0;
              acc += p * Ke[r * 24 + k];
            }
            PtK[i][k] = acc;
          }
        for (int i = 0; i < 24; ++i)
          for (int j = 0; j < 24; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 24; ++k) {
              const double p = (k % 3 == j % 3) ? P[k / 3][j / 3] : 0.0;
              acc += PtK[i][k] * p;
            }
            KP[i][j] = acc;
          }
        for (int i = 0; i < 24; ++i)
          for (int j = 0; j < 24; ++j) Q[idx * 576 + i * 24 + j] = KP[i][j];
      }
}

static std::vector<int> coarse_free_injection(const Grid& gf_, const Grid& gc,
                                              const std::vector<int>& free_f) {
  std::vector<char> mask_f(gf_.ndof, 0);
  for (int d : free_f) mask_f[d] = 1;
  std::vector<char> is_free(gc.ndof, 0);
  int cn = 0;
  for (int ix = 0; ix < gc.NX; ++ix)
    for (int iy = 0; iy < gc.NY; ++iy)
      for (int iz = 0; iz < gc.NZ; ++iz, ++cn) {
        const long fine_node = 2L * ix * gf_.NY * gf_.NZ + 2L * iy * gf_.NZ + 2L * iz;
        for (int d = 0; d < 3; ++d) {
          const long fd = 3 * fine_node + d;
          if (fd < gf_.ndof) is_free[3 * cn + d] = mask_f[fd];
        }
      }
  std::vector<int> out;
  for (int i = 0; i < gc.ndof; ++i)
    if (is_free[i]) out.push_back(i);
  return out;
}

// Level precision code (0 fp64, 1 fp32, 2 bf16).  An explicit list is padded
// with fp64; otherwise the default schedule of multigrid_v4.py:1135-1144.
static int level_prec_code(const gf_gmg_params& p, int l) {
  if (p.level_prec[0] >= 0) return p.level_prec[l] >= 0 ? p.level_prec[l] : 0;
  if (p.fine_smoother == 2) return l == 0 ? 2 : (l == 1 ? 1 : 0);
  return l == 0 ? p.fine_smoother : 0;
}

int Gmg::create(Ctx* c, const gf_gmg_params* params) {
  GF_DEVICE_LOCK();
  ctx = c;
  p = *params;
  if (p.smoother_type < 0 || p.smoother_type > 1 || p.cycle_type < 0 || p.cycle_type > 1 ||
      p.coarse_correction_policy < 0 || p.coarse_correction_policy > 1 ||
      p.root_enrichment_mode < 0 || p.root_enrichment_mode > 4 ||
      p.root_local_correction_mode < 0 || p.root_local_correction_mode > 2) {
    set_error("gf_gmg_create: invalid option code");
    return -2;
  }
  if (p.coarse_correction_acceptance_ratio <= 0.0 || p.coarse_correction_max_scale <= 0.0 ||
      p.root_local_correction_max_scale <= 0.0 || p.root_local_correction_inner_steps <= 0 ||
      p.inner_krylov_steps < 0 || !(p.root_enrichment_weight_floor > 0.0 &&
                                    p.root_enrichment_weight_floor <= 1.0) ||
      p.root_enrichment_weight_power <= 0.0) {
    set_error("gf_gmg_create: invalid numeric option");
    return -2;
  }
  if (p.root_local_correction_inner_restart <= 0)
    p.root_local_correction_inner_restart = std::max(1, p.root_local_correction_inner_steps);
  if (p.inner_krylov_restart <= 0) p.inner_krylov_restart = std::max(1, p.inner_krylov_steps);
  p.fine_smoother_degree = std::max(1, p.fine_smoother_degree);
  p.coarse_smooth_iters = std::max(1, p.coarse_smooth_iters);
  L = std::max(1, p.n_levels);
  if (L > 8) {
    set_error("gf_gmg_create: at most 8 levels");
    return -2;
  }
  lv.resize(L);
  lv[0].g = ctx->g;
  lv[0].n_free = ctx->n_free;
  lv[0].off = ctx->off;
  lv[0].d_f2f = ctx->d_f2f;
  lv[0].h_free = ctx->h_free;
  for (int l = 1; l < L; ++l) {
    const Grid& gp = lv[l - 1].g;
    lv[l].g = make_grid(std::max(1, gp.nx / 2), std::max(1, gp.ny / 2), std::max(1, gp.nz / 2));
    lv[l].h_free = coarse_free_injection(gp, lv[l].g, lv[l - 1].h_free);
    lv[l].n_free = (int)lv[l].h_free.size();
    if (lv[l].n_free == 0) {
      set_error("gf_gmg_create: a coarse level has no free DOFs");
      return -2;
    }
    const int nf = lv[l].n_free;
    lv[l].off = (lv[l].h_free[0] + nf == lv[l].g.ndof) ? lv[l].h_free[0] : -1;
    std::vector<int> f2f(lv[l].g.ndof, -1);
    for (int i = 0; i < nf; ++i) f2f[lv[l].h_free[i]] = i;
    GF_TRY(dalloc(&lv[l].d_f2f, lv[l].g.ndof));
    GF_TRY(upload_now(lv[l].d_f2f, f2f.data(), sizeof(int) * f2f.size()));
    GF_TRY(dalloc(&lv[l].A, 243L * lv[l].g.n_nodes));
  }
  for (int l = 0; l < L; ++l) {
    Level& v = lv[l];
    const long n = v.n_free;
    // coarse levels smooth in FP32 only for "fp32" (multigrid_v4.py:1878);
    // level 0 smooths in reduced precision for anything but "fp64", through
    // the fine_smoother's operator (multigrid_v4.py:1769-1780, 1851-1864)
    const int code = level_prec_code(p, l);
    v.prec = (l == 0) ? (code == 0 ? 0 : 1) : (code == 1 ? 1 : 0);
    if (l == 0 && v.prec == 1 && p.fine_smoother != 1 && p.fine_smoother != 2) {
      set_error("Unsupported fine smoother for a reduced-precision level 0");
      return -2;
    }
    if (l == 0 && v.prec == 1 && p.fine_smoother == 2) {
      GF_TRY(ctx->bf16_alloc());
      fine_bf16 = true;
    }
    GF_TRY(dalloc(&v.diag, n));
    GF_TRY(dalloc(&v.dinv, n));
    GF_TRY(dalloc(&v.dinv32, n));
    for (int k = 0; k < 2; ++k) {
      GF_TRY(dalloc(&v.X[k], n));
      GF_TRY(dalloc(&v.X32[k], n));
    }
    GF_TRY(dalloc(&v.RHS, n));
    GF_TRY(dalloc(&v.R, n));
    GF_TRY(dalloc(&v.R2, n));
    GF_TRY(dalloc(&v.PE, n));
    GF_TRY(dalloc(&v.AC, n));
    GF_TRY(dalloc(&v.DV, n));
    GF_TRY(dalloc(&v.T, n));
    GF_TRY(dalloc(&v.RHS32, n));
    GF_TRY(dalloc(&v.DV32, n));
    GF_TRY(dalloc(&v.st, 16));
    if (l >= 1 && l + 1 < L) GF_TRY(dalloc(&v.KP, 243L * v.g.n_nodes));
    if (l >= 1 && v.prec == 1) GF_TRY(dalloc(&v.A32, 243L * v.g.n_nodes));
  }
  const long n0 = lv[0].n_free;
  const long ne = (long)ctx->g.nx * ctx->g.ny * ctx->g.nz;
  GF_TRY(dalloc(&E64, ne));
  GF_TRY(dalloc(&E32, ne));
  std::vector<double> Qh;
  galerkin_q(ctx->h_Ke, Qh);
  GF_TRY(dalloc(&Q, Qh.size()));
  GF_TRY(upload_now(Q, Qh.data(), sizeof(double) * Qh.size()));
  GF_TRY(dalloc(&d_fail, 4));
  GF_TRY(dalloc(&ap_x, n0));
  GF_TRY(dalloc(&ap_res, n0));
  GF_TRY(dalloc(&ap_corr, n0));
  GF_TRY(dalloc(&ap_trial, n0));
  GF_TRY(dalloc(&ap_st, 16));
  GF_TRY(dalloc(&gr_in, n0));
  GF_TRY(dalloc(&gr_out, n0));
  GF_TRY(dalloc(&pw_v, n0));
  GF_TRY(dalloc(&pw_d, n0));
  GF_TRY(dalloc(&pw_kv, n0));
  GF_TRY(dalloc(&pw_av, n0));
  GF_TRY(dalloc(&pw_s, 1024));
  GF_HIP(hipHostMalloc((void**)&hs, sizeof(double) * 16, hipHostMallocDefault));
  GF_TRY(dalloc(&d_err, kWaitFlags));
  GF_TRY(zero_now(d_err, sizeof(double) * kWaitFlags));
  if (p.root_enrichment_mode != 0) GF_TRY(dalloc(&re_w, n0));
  if (p.inner_krylov_steps > 0) GF_TRY(ik.alloc(p.inner_krylov_restart, n0));
  if (p.root_local_correction_mode == 2)
    GF_TRY(rl.alloc(p.root_local_correction_inner_restart, n0));
  if (p.root_local_correction_mode != 0) {
    GF_TRY(dalloc(&nb_inv9, 9L * ctx->g.n_nodes));
    GF_TRY(dalloc(&nb_pinv, n0));
    GF_TRY(zero_now(nb_inv9, sizeof(double) * 9L * ctx->g.n_nodes));
    GF_TRY(zero_now(nb_pinv, sizeof(double) * n0));
  }
  lam.assign(L, NAN);
  GF_TRY(work.init());
  return 0;
}

void Gmg::release() {
  GF_DEVICE_LOCK();
  dfree(d_err);
  for (int l = 0; l < (int)lv.size(); ++l) {
    Level& v = lv[l];
    if (l > 0) dfree(v.d_f2f);
    dfree(v.A);
    dfree(v.A32);
    dfree(v.KP);
    dfree(v.diag);
    dfree(v.dinv);
    dfree(v.dinv32);
    for (int k = 0; k < 2; ++k) {
      dfree(v.X[k]);
      dfree(v.X32[k]);
    }
    dfree(v.RHS);
    dfree(v.R);
    dfree(v.R2);
    dfree(v.PE);
    dfree(v.AC);
    dfree(v.DV);
    dfree(v.T);
    dfree(v.RHS32);
    dfree(v.DV32);
    dfree(v.st);
  }
  dfree(E64);
  dfree(E32);
  dfree(Q);
  dfree(CH);
  dfree(CHt);
  dfree(CHr);
  dfree(CHi);
  dfree(CHy);
  dfree(d_fail);
  dfree(pcg_work);
  pcg_buffers_release(pcg_uc, pcg_uc_bytes);
  pcg_uc = nullptr;
  pcg_uc_bytes = 0;
  dfree(re_w);
  dfree(nb_inv9);
  dfree(nb_pinv);
  dfree(ap_x);
  dfree(ap_res);
  dfree(ap_corr);
  dfree(ap_trial);
  dfree(ap_st);
  drop_graph();
  if (cap_stream) (void)hipStreamDestroy(cap_stream);
  cap_stream = nullptr;
  dfree(gr_in);
  dfree(gr_out);
  dfree(pw_v);
  dfree(pw_d);
  dfree(pw_kv);
  dfree(pw_av);
  dfree(pw_s);
  if (hs) (void)hipHostFree(hs);
  hs = nullptr;
  ik.release();
  rl.release();
  outer.release();
  work.release();
}

int Gmg::read_scalars(const double* dptr, double* host, int n, hipStream_t st) {
  GF_HIP(hipMemcpyAsync(hs, dptr, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  GF_HIP(hipStreamSynchronize(st));
  std::memcpy(host, hs, sizeof(double) * n);
  return 0;
}

// A grid-wide wait that expired leaves its slot of d_err set (the blocks
// stopped waiting, so the result of that launch is not valid).  The flags
// are cleared, the exchange state of the sweep is reset (its ring positions
// may disagree) and the recorded preconditioner graph is dropped, so the
// next call starts clean; the error names the wait that gave up.
int Gmg::wait_failed(const double* flags, const char* where, hipStream_t st) {
  work.sweep_G = -1;
  drop_graph();
  LAUNCH(k_fill, kWaitFlags, st, (long)kWaitFlags, d_err, 0.0, (const int*)nullptr);
  (void)ls_reset(work, st);
  (void)hipStreamSynchronize(st);
  std::string what;
  if (flags[WAIT_SWEEP] != 0.0) what = "the Gram-Schmidt sweep";
  if (flags[WAIT_PCG] != 0.0)
    what += std::string(what.empty() ? "" : " and ") + "the persistent coarsest PCG";
  if (flags[WAIT_LS] != 0.0)
    what += std::string(what.empty() ? "" : " and ") + "the one-launch line search";
  set_error(std::string(where) + ": a grid-wide wait of " + what +
            " exceeded its bound; the result is not valid");
  return -1;
}

// Synchronise st and report (and clear) an expired grid-wide wait of any
// launch issued so far; the ABI calls that run the preconditioner or the
// coarsest PCG outside the outer solve end with this check.
int Gmg::take_wait_error(const char* where, hipStream_t st) {
  GF_HIP(hipMemcpyAsync(hs + 8, d_err, sizeof(double) * kWaitFlags, hipMemcpyDeviceToHost, st));
  GF_HIP(hipStreamSynchronize(st));
  double f[kWaitFlags];
  std::memcpy(f, hs + 8, sizeof(f));
  for (int k = 0; k < kWaitFlags; ++k)
    if (f[k] != 0.0) return wait_failed(f, where, st);
  return 0;
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
int Gmg::level_op(int l, const double* x, double* y, int epi, const EpiArgs<double>& ep,
                  const int* gate, hipStream_t st) {
  if (l == 0)
    return fine_matvec<double>(ctx->g, ctx->map(), ctx->Ke, E64, x, y, epi, ep, gate, st);
  return stencil_matvec<double>(lv[l].g, lv[l].map(), lv[l].A, x, y, epi, ep, gate, st);
}

template <typename T>
static int op_T(Gmg& G, int l, const T* x, T* y, int epi, const EpiArgs<T>& ep, const int* gate,
                hipStream_t st);
template <>
int op_T<double>(Gmg& G, int l, const double* x, double* y, int epi, const EpiArgs<double>& ep,
                 const int* gate, hipStream_t st) {
  return G.level_op(l, x, y, epi, ep, gate, st);
}
template <>
int op_T<float>(Gmg& G, int l, const float* x, float* y, int epi, const EpiArgs<float>& ep,
                const int* gate, hipStream_t st) {
  if (l == 0 && G.fine_bf16)
    return fine_matvec_bf16(G.ctx->g, G.ctx->map(), G.ctx->KEb, G.E32, x, y, epi, ep, gate, st);
  if (l == 0)
    return fine_matvec<float>(G.ctx->g, G.ctx->map(), G.ctx->Ke32, G.E32, x, y, epi, ep, gate,
                              st);
  return stencil_matvec<float>(G.lv[l].g, G.lv[l].map(), G.lv[l].A32, x, y, epi, ep, gate, st);
}

__global__ void k_cast_gated_f2d(long n, const float* a, double* b, const int* gate) {
  if (gate && *gate == 0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}
__global__ void k_cast_gated_d2f(long n, const double* a, float* b, const int* gate) {
  if (gate && *gate == 0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// Chebyshev coefficients of level l in precision T (multigrid_v4.py:1783-1812):
// returns the first step's alpha and sets sigma, delta.
template <typename T>
static T cheb_coeffs(const Gmg& G, int l, T& sigma, T& delta) {
  double lamd = l == 0 ? G.lam0 : G.lam[l];
  if (l > 0 && !std::isfinite(lamd)) lamd = 1.0;
  double lfd = G.p.cheb_lower_frac;
  if (G.high) lfd = std::min(G.p.cheb_lower_frac, 1.0 / 100.0);
  const T lam = (T)lamd, lf = (T)lfd;
  const T lam_min = lf * lam;
  sigma = (lam + lam_min) * (T)0.5;
  delta = (lam - lam_min) * (T)0.5;
  return (T)2.0 / (lam + lam_min);
}

// The zero-start first smoothing step of coarse level lc, for the
// restriction that produces its right-hand side to apply (FP64 levels that
// are not the coarsest; the pre-smoother from zero is the first thing the
// level's cycle does).  Marks the level so the smoother skips that step.
const ZeroStart* Gmg::fuse_zero_start(int lc) {
  if (lc <= 0 || lc >= L - 1 || lv[lc].prec != 0) return nullptr;
  Level& v = lv[lc];
  zs_ = ZeroStart{};
  zs_.dinv = v.dinv;
  zs_.xo = v.X[0];
  if (p.smoother_type == 0) {
    double sigma, delta;
    zs_.mode = 1;
    zs_.dv = v.DV;
    zs_.a = cheb_coeffs<double>(*this, lc, sigma, delta);
  } else {
    zs_.mode = 2;
    zs_.a = omega_used;
  }
  zs_level = lc;
  return &zs_;
}

// Generic smoother body in precision T.  xs[2] are the ping-pong buffers,
// x_in (nullable = zero start) must not alias them unless it is xs[cur].
template <typename T>
static int smooth_T(Gmg& G, int l, const T* x_in, const T* rhs, const T* dinv, T* dv, T* xs0,
                    T* xs1, T* out_last, const int* gate, hipStream_t st, T** result) {
  Level& v = G.lv[l];
  const long n = v.n_free;
  const int deg = l == 0 ? G.fine_deg : G.coarse_it;
  T* bufs[2] = {xs0, xs1};
  int nb = (x_in == xs0) ? 1 : 0;
  const T* x = x_in;
  auto dest = [&](int step) -> T* {
    if (step == deg - 1 && out_last) return out_last;
    T* d = bufs[nb];
    nb ^= 1;
    return d;
  };
  // a zero-start first step already applied by the restriction that
  // produced rhs (Gmg::fuse_zero_start) is not applied again
  const bool started = !x && G.zs_level == l;
  if (started) G.zs_level = -1;
  if (G.p.smoother_type == 0) {
    T sigma, delta;
    T alpha = cheb_coeffs<T>(G, l, sigma, delta);
    T* xo = dest(0);
    if (!x) {
      if (!started) GF_TRY(zero_start_step<T>(n, rhs, dinv, alpha, dv, xo, gate, st));
    } else {
      EpiArgs<T> ep;
      ep.b = rhs;
      ep.dinv = dinv;
      ep.dv = dv;
      ep.xo = xo;
      ep.a = alpha;
      GF_TRY(op_T<T>(G, l, x, xo, EPI_CHEB0, ep, gate, st));
    }
    x = xo;
    for (int k = 1; k < deg; ++k) {
      const T c = ((delta * delta) * alpha) * (T)0.25;
      const T alpha_new = (T)1.0 / (sigma - c);
      T* xn = dest(k);
      EpiArgs<T> ep;
      ep.b = rhs;
      ep.dinv = dinv;
      ep.dv = dv;
      ep.xo = xn;
      ep.a = alpha_new;
      ep.c = c;
      GF_TRY(op_T<T>(G, l, x, xn, EPI_CHEB, ep, gate, st));
      x = xn;
      alpha = alpha_new;
    }
  } else {
    const T om = (T)G.omega_used;
    for (int k = 0; k < deg; ++k) {
      T* xn = dest(k);
      if (!x) {
        if (!started) GF_TRY(zero_start_step<T>(n, rhs, dinv, om, (T*)nullptr, xn, gate, st));
      } else {
        EpiArgs<T> ep;
        ep.b = rhs;
        ep.dinv = dinv;
        ep.xo = xn;
        ep.a = om;
        GF_TRY(op_T<T>(G, l, x, xn, EPI_JACOBI, ep, gate, st));
      }
      x = xn;
    }
  }
  *result = const_cast<T*>(x);
  return 0;
}

int Gmg::smooth(int l, const double* x_in, const double* rhs, double* out, int,
                const int* gate, hipStream_t st, double** result) {
  Level& v = lv[l];
  const long n = v.n_free;
  if (v.prec == 0) {
    return smooth_T<double>(*this, l, x_in, rhs, v.dinv, v.DV, v.X[0], v.X[1], out, gate, st,
                            result);
  }
  // reduced-precision smoother: cast in, smooth in FP32, cast result to FP64
  float* x32 = nullptr;
  if (x_in) {
    LAUNCH(k_cast_gated_d2f, n, st, n, x_in, v.X32[0], gate);
    x32 = v.X32[0];
  }
  LAUNCH(k_cast_gated_d2f, n, st, n, rhs, v.RHS32, gate);
  float* r32 = nullptr;
  GF_TRY(smooth_T<float>(*this, l, x32, v.RHS32, v.dinv32, v.DV32, v.X32[0], v.X32[1], nullptr,
                         gate, st, &r32));
  double* dst = out;
  if (!dst) dst = (x_in == v.X[0]) ? v.X[1] : v.X[0];
  LAUNCH(k_cast_gated_f2d, n, st, n, r32, dst, gate);
  *result = dst;
  return 0;
}

int Gmg::solve_coarsest(const double* rhs, double* out, const int* gate, hipStream_t st,
                        double** result) {
  Level& c = lv[L - 1];
  double* dst = out ? out : c.X[0];
  if (L == 1) return smooth(0, nullptr, rhs, out, 1, gate, st, result);
  if (chol_mode == 1) {
    if (p.coarsest_solve == 0)
      GF_TRY(dense_gemv(ch_n, CHi, rhs, dst, gate, st));
    else
      GF_TRY(dense_chol_solve(ch_n, CH, CHt, CHr, rhs, dst, gate, st));
  } else {
    int it = p.coarse_pcg_iters;
    if (high) it = std::max(p.coarse_pcg_iters, 200);
    int launched = 0;
    if (pcg_uc) {
      launched = fixed_pcg_persistent(c.g, c.map(), c.A, c.dinv, c.n_free, it, rhs, dst, pcg_uc,
                                      d_err + WAIT_PCG, gate, st);
      if (launched < 0) return launched;
    }
    if (!launched)
      GF_TRY(fixed_pcg_grid(work, c.g, c.map(), c.A, c.dinv, c.n_free, it, rhs, dst, pcg_work,
                            gate, st));
  }
  *result = dst;
  return 0;
}

// x <- x + accepted correction; residual_out (nullable) <- corrected residual
int Gmg::correct(int l, double* x, const double* residual, const double* corr,
                 double* residual_out, const int* gate, hipStream_t st) {
  Level& v = lv[l];
  const long n = v.n_free;
  if (p.coarse_correction_policy == 0) {
    LAUNCH(k_add, n, st, n, x, corr, gate);
    if (residual_out) {
      EpiArgs<double> ep;
      ep.b = residual;
      GF_TRY(level_op(l, corr, residual_out, EPI_RESID, ep, gate, st));
    }
    return 0;
  }
  GF_TRY(level_op(l, corr, v.AC, EPI_PLAIN, EpiArgs<double>{}, gate, st));
  // the statistics, trial norm and apply in one launch where its grid fits,
  // else in three (identical bits)
  const int fused = ls_fused(work, n, corr, residual, v.AC, x, residual_out, v.st,
                             p.coarse_correction_max_scale,
                             p.coarse_correction_acceptance_ratio, d_err + WAIT_LS, gate, st);
  if (fused < 0) return fused;
  if (fused) return 0;
  GF_TRY(dot4_ls_step(work, n, corr, residual, v.AC, v.st, p.coarse_correction_max_scale, gate,
                      st));
  GF_TRY(trial_norm(work, n, residual, v.AC, v.st + 4, residual_out, v.st + 6, gate, st));
  LAUNCH(k_ls_apply, n, st, n, x, corr, v.st, p.coarse_correction_acceptance_ratio, residual,
         residual_out, gate);
  return 0;
}

int Gmg::enrichment_rhs(const double* residual, double* out, const int* gate, hipStream_t st,
                        const ZeroStart* zs) {
  Level& f = lv[0];
  Level& c = lv[1];
  const long n = f.n_free;
  const int mode = p.root_enrichment_mode;
  const bool gap = (mode == 2 || mode == 4);
  const bool vm = (mode == 3 || mode == 4);
  if (gap) {
    GF_TRY(restrict_(f.g, f.map(), c.g, c.map(), residual, c.T, gate, st));
    GF_TRY(prolong(f.g, f.map(), c.g, c.map(), c.T, f.T, gate, st));
  }
  LAUNCH(k_enrich, n, st, n, re_w, vm ? 1 : 0, p.root_enrichment_weight_floor, residual,
         gap ? f.T : nullptr, f.AC, gate);
  return restrict_(f.g, f.map(), c.g, c.map(), f.AC, out, gate, st, zs);
}

int Gmg::cycle(int l, const double* rhs, double* out, const int* gate, hipStream_t st,
               double** result) {
  if (l == L - 1) return solve_coarsest(rhs, out, gate, st, result);
  Level& v = lv[l];
  Level& c = lv[l + 1];
  const long n = v.n_free;
  const bool wcyc = p.cycle_type == 1;
  const bool enrich = (l == 0 && p.root_enrichment_mode != 0);
  double* x = nullptr;
  GF_TRY(smooth(l, nullptr, rhs, nullptr, 1, gate, st, &x));
  {
    EpiArgs<double> ep;
    ep.b = rhs;
    GF_TRY(level_op(l, x, v.R, EPI_RESID, ep, gate, st));
  }
  GF_TRY(restrict_(v.g, v.map(), c.g, c.map(), v.R, c.RHS, gate, st, fuse_zero_start(l + 1)));
  double* ec = nullptr;
  GF_TRY(cycle(l + 1, c.RHS, nullptr, gate, st, &ec));
  GF_TRY(prolong(v.g, v.map(), c.g, c.map(), ec, v.PE, gate, st));
  if (!wcyc) {
    GF_TRY(correct(l, x, v.R, v.PE, enrich ? v.R2 : nullptr, gate, st));
    if (enrich) {
      GF_TRY(enrichment_rhs(v.R2, c.RHS, gate, st, fuse_zero_start(l + 1)));
      double* e2 = nullptr;
      GF_TRY(cycle(l + 1, c.RHS, nullptr, gate, st, &e2));
      GF_TRY(prolong(v.g, v.map(), c.g, c.map(), e2, v.PE, gate, st));
      GF_TRY(correct(l, x, v.R2, v.PE, nullptr, gate, st));
    }
  } else {
    if (p.coarse_correction_policy == 0) {
      LAUNCH(k_add, n, st, n, x, v.PE, gate);
      EpiArgs<double> ep;
      ep.b = rhs;
      GF_TRY(level_op(l, x, v.R2, EPI_RESID, ep, gate, st));
    } else {
      GF_TRY(correct(l, x, v.R, v.PE, v.R2, gate, st));
    }
    if (enrich) {
      GF_TRY(enrichment_rhs(v.R2, c.RHS, gate, st, fuse_zero_start(l + 1)));
    } else {
      GF_TRY(restrict_(v.g, v.map(), c.g, c.map(), v.R2, c.RHS, gate, st,
                       fuse_zero_start(l + 1)));
    }
    double* e2 = nullptr;
    GF_TRY(cycle(l + 1, c.RHS, nullptr, gate, st, &e2));
    GF_TRY(prolong(v.g, v.map(), c.g, c.map(), e2, v.PE, gate, st));
    GF_TRY(correct(l, x, v.R2, v.PE, nullptr, gate, st));
  }
  const bool extra = (l == 0 && high);
  GF_TRY(smooth(l, x, rhs, extra ? nullptr : out, 1, gate, st, &x));
  if (extra) GF_TRY(smooth(l, x, rhs, out, 1, gate, st, &x));
  *result = x;
  return 0;
}

int Gmg::base_cycle(const double* rhs, double* out, const int* gate, hipStream_t st) {
  double* res = nullptr;
  GF_TRY(cycle(0, rhs, out, gate, st, &res));
  if (res != out) LAUNCH(k_copy, lv[0].n_free, st, lv[0].n_free, res, out, gate);
  return 0;
}

// ---------------------------------------------------------------------------
// device-resident FGMRES (x0 = 0), multigrid_v4.py:416-534
// ---------------------------------------------------------------------------
int Gmg::fgmres_dev(Krylov& K, const std::function<int(const double*, double*, const int*)>& M,
                    const double* b, double* x, double tol, int maxiter, int restart,
                    hipStream_t st) {
  const long n = K.n;
  const int m = K.m;
  int* step = K.fl;
  int* run = K.fl + 1;
  LAUNCH(k_fill, n, st, n, x, 0.0, (const int*)nullptr);
  GF_TRY(dot_f64(work, n, b, b, K.sc, true, nullptr, st));
  LAUNCH1(k_fgm_begin, st, K.fl, K.sc);
  int total = 0, cyc = 0;
  while (total < maxiter) {
    const int inner = std::min(restart, maxiter - total);
    const double* r = b;
    if (cyc == 0) {
      LAUNCH1(k_copy_scalar, st, K.sc, K.sc + 1);
    } else {
      EpiArgs<double> ep;
      ep.b = b;
      GF_TRY(level_op(0, x, K.V, EPI_RESID, ep, run, st));
      GF_TRY(dot_f64(work, n, K.V, K.V, K.sc + 1, true, run, st));
      r = K.V;
    }
    LAUNCH1(k_cycle_start, st, K.fl, K.sc, K.g, tol);
    GF_TRY(scale_inv(n, r, K.sc + 1, 0.0, K.V, step, st));
    for (int j = 0; j < inner; ++j) {
      double* Vj = K.V + (long)j * n;
      double* Zj = K.Z + (long)j * n;
      double* w = K.V + (long)(j + 1) * n;
      GF_TRY(M(Vj, Zj, step));
      GF_TRY(level_op(0, Zj, w, EPI_PLAIN, EpiArgs<double>{}, step, st));
      int swept = 0;
      if (p.mgs_sweep) {  // plain launch: this sequence is recorded in a graph
        const GivensArgs gv{tol, K.cs, K.sn, K.g, K.sc, K.fl};
        swept = mgs_sweep(work, n, K.V, j, m, w, K.H, d_err + WAIT_SWEEP, step, 0, &gv, st);
        if (swept < 0) return swept;
      }
      if (!swept) {
        GF_TRY(dot_f64(work, n, K.V, w, K.H + j, false, step, st));
        for (int i = 0; i < j; ++i)
          GF_TRY(mgs_axpy_dot(work, n, K.V + (long)i * n, w, K.H + (long)i * m + j,
                              K.V + (long)(i + 1) * n, K.H + (long)(i + 1) * m + j, false,
                              step, st));
        GF_TRY(mgs_axpy_dot(work, n, Vj, w, K.H + (long)j * m + j, nullptr,
                            K.H + (long)(j + 1) * m + j, true, step, st));
        GF_TRY(scale_inv(n, w, K.H + (long)(j + 1) * m + j, 1e-14, w, step, st));
      }
      if (!swept) LAUNCH1(k_givens, st, m, j, tol, K.H, K.cs, K.sn, K.g, K.sc, K.fl);
    }
    hipLaunchKernelGGL(k_backsolve, dim3(1), dim3(256), 0, st, m, K.H, K.g, K.y, K.fl);
    GF_CHECK_LAUNCH();
    GF_TRY(combine(n, K.Z, n, K.y, K.fl + 2, true, x, run, st));
    total += inner;
    ++cyc;
  }
  return 0;
}

void Gmg::drop_graph() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
  graph_valid = false;
}

// Record apply_impl(gr_in -> gr_out) on a private stream.  The sequence is
// static for a given set-up (all decisions are device-side), so one graph is
// replayed until the next set-up.  It contains kernels only.
int Gmg::capture() {
  GF_DEVICE_LOCK();
  drop_graph();
  if (!cap_stream) GF_HIP(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
  GF_HIP(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
  const int rc = apply_impl(gr_in, gr_out, cap_stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(cap_stream, &g);
  if (rc != 0 || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc == 0) set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return rc ? rc : -1;
  }
  graph = g;
  GF_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  graph_valid = true;
  ++graph_builds;
  return 0;
}

int Gmg::apply(const double* rhs, double* out, hipStream_t st) {
  if (!is_setup) {
    set_error("GalerkinMatFreeGMG.setup() must be called before apply().");
    return -3;
  }
  if (!p.graph_apply) return apply_impl(rhs, out, st);
  if (!graph_valid) GF_TRY(capture());
  const size_t bytes = sizeof(double) * lv[0].n_free;
  GF_HIP(hipMemcpyAsync(gr_in, rhs, bytes, hipMemcpyDeviceToDevice, st));
  GF_HIP(hipGraphLaunch(gexec, st));
  GF_HIP(hipMemcpyAsync(out, gr_out, bytes, hipMemcpyDeviceToDevice, st));
  return 0;
}

int Gmg::apply_impl(const double* rhs, double* out, hipStream_t st) {
  const long n = lv[0].n_free;
  const double* xp;
  if (p.inner_krylov_steps > 0) {
    auto M = [&](const double* in, double* o, const int* gate) {
      return base_cycle(in, o, gate, st);
    };
    GF_TRY(fgmres_dev(ik, M, rhs, ap_x, 0.0, p.inner_krylov_steps, p.inner_krylov_restart, st));
    xp = ap_x;
  } else {
    GF_TRY(base_cycle(rhs, p.root_local_correction_mode == 0 ? out : ap_x, nullptr, st));
    if (p.root_local_correction_mode == 0) return 0;
    xp = ap_x;
  }
  if (p.root_local_correction_mode == 0) {
    LAUNCH(k_copy, n, st, n, xp, out, (const int*)nullptr);
    return 0;
  }
  return root_local(rhs, xp, out, st);
}

// Root-local correction of x_p against rhs (multigrid_v4.py:964-1055): the
// pre-local residual goes to ap_res, the decision scalars to ap_st, the
// corrected iterate to out.  Inner FGMRES mode: ap_st = [c.c, r.r, t.t] with
// the trial residual t in ap_trial; line-search mode: ap_st = [c.c, r.r,
// Ac.Ac, r.Ac] with Ac in ap_trial.
int Gmg::root_local(const double* rhs, const double* xp, double* out, hipStream_t st) {
  const long n = lv[0].n_free;
  EpiArgs<double> ep;
  ep.b = rhs;
  GF_TRY(level_op(0, xp, ap_res, EPI_RESID, ep, nullptr, st));
  if (p.root_local_correction_mode == 2) {
    auto M = [&](const double* in, double* o, const int* gate) {
      return nodeblock_apply(ctx->g, ctx->map(), nb_inv9, nb_pinv, 1, in, o, gate, st);
    };
    GF_TRY(fgmres_dev(rl, M, ap_res, ap_corr, 0.0, p.root_local_correction_inner_steps,
                      p.root_local_correction_inner_restart, st));
    GF_TRY(dot_f64(work, n, ap_corr, ap_corr, ap_st, false, nullptr, st));
    GF_TRY(dot_f64(work, n, ap_res, ap_res, ap_st + 1, false, nullptr, st));
    EpiArgs<double> e2;
    e2.b = ap_res;
    GF_TRY(level_op(0, ap_corr, ap_trial, EPI_RESID, e2, nullptr, st));
    GF_TRY(dot_f64(work, n, ap_trial, ap_trial, ap_st + 2, false, nullptr, st));
    LAUNCH(k_rl_accept, n, st, n, xp, ap_corr, ap_st, out);
    return 0;
  }
  GF_TRY(nodeblock_apply(ctx->g, ctx->map(), nb_inv9, nb_pinv, 1, ap_res, ap_corr, nullptr, st));
  GF_TRY(level_op(0, ap_corr, ap_trial, EPI_PLAIN, EpiArgs<double>{}, nullptr, st));
  GF_TRY(dot4_ls(work, n, ap_corr, ap_res, ap_trial, ap_st, nullptr, st));
  LAUNCH(k_rl_ls, n, st, n, xp, ap_corr, ap_st, p.root_local_correction_max_scale, out);
  return 0;
}

// ---------------------------------------------------------------------------
// set-up (multigrid_v4.py:1119-1341)
// ---------------------------------------------------------------------------
static int wdot_host(Work& w, long n, const double* a, const double* wt, const double* b,
                     int mode, double* dout, hipStream_t st) {
  const int nb = std::min(kRedBlocks, std::max(1, cdiv(n, kBlock * 4)));
  hipLaunchKernelGGL(k_wdot, dim3(nb), dim3(kBlock), 0, st, n, a, wt, b, mode, w.partials, dout);
  GF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(1), 0, st, nb, w.partials, dout, 0);
  GF_CHECK_LAUNCH();
  return 0;
}

int Gmg::estimate_fine_lambda(hipStream_t st, double* out) {
  const long n = lv[0].n_free;
  double h[2];
  LAUNCH(k_recip, n, st, n, lv[0].dinv, pw_d);
  LAUNCH(k_fill, n, st, n, pw_v, 1.0, (const int*)nullptr);
  GF_TRY(wdot_host(work, n, pw_v, pw_d, pw_v, 0, pw_s, st));
  LAUNCH1(k_fold, st, 1, pw_s, pw_s + 1, 1);  // sqrt on device
  LAUNCH(k_div_dev, n, st, n, pw_v, pw_s + 1);
  double lamv = 1.0;
  for (int it = 0; it < 20; ++it) {
    GF_TRY(level_op(0, pw_v, pw_kv, EPI_PLAIN, EpiArgs<double>{}, nullptr, st));
    GF_TRY(wdot_host(work, n, pw_v, pw_d, pw_v, 0, pw_s, st));
    GF_TRY(dot_f64(work, n, pw_v, pw_kv, pw_s + 1, false, nullptr, st));
    GF_TRY(read_scalars(pw_s, h, 2, st));
    lamv = h[1] / std::max(h[0], 1e-300);
    LAUNCH(k_mul, n, st, n, lv[0].dinv, pw_kv, pw_av);
    GF_TRY(wdot_host(work, n, pw_av, pw_d, pw_av, 0, pw_s, st));
    GF_TRY(read_scalars(pw_s, h, 1, st));
    const double nrm = std::sqrt(h[0]);
    if (nrm <= 1e-300) break;
    LAUNCH(k_div_host, n, st, n, pw_av, nrm, pw_v);
  }
  *out = lamv;
  return 0;
}

int Gmg::estimate_level_lambda(int l, hipStream_t st, double* out) {
  Level& v = lv[l];
  const long n = v.n_free;
  double h[2];
  LAUNCH(k_fill, n, st, n, pw_v, 1.0, (const int*)nullptr);
  GF_TRY(dot_f64(work, n, pw_v, pw_v, pw_s, true, nullptr, st));
  GF_TRY(read_scalars(pw_s, h, 1, st));
  LAUNCH(k_div_host, n, st, n, pw_v, h[0], pw_v);
  double lamv = 1.0;
  for (int it = 0; it < 10; ++it) {
    GF_TRY(level_op(l, pw_v, pw_kv, EPI_PLAIN, EpiArgs<double>{}, nullptr, st));
    LAUNCH(k_mul, n, st, n, v.dinv, pw_kv, pw_av);
    GF_TRY(dot_f64(work, n, pw_av, pw_av, pw_s, true, nullptr, st));
    GF_TRY(read_scalars(pw_s, h, 1, st));
    const double nrm = h[0];
    if (nrm <= 1e-300) break;
    GF_TRY(dot_f64(work, n, pw_v, pw_kv, pw_s, false, nullptr, st));
    GF_TRY(wdot_host(work, n, pw_v, v.dinv, pw_v, 1, pw_s + 1, st));
    GF_TRY(read_scalars(pw_s, h, 2, st));
    lamv = h[0] / std::max(h[1], 1e-300);
    LAUNCH(k_div_host, n, st, n, pw_av, nrm, pw_v);
  }
  *out = std::max(lamv, 1e-6);
  return 0;
}

int Gmg::setup(const double* E, const double* fine_diag, hipStream_t st) {
  rs.valid = false;
  drop_graph();
  const long ne = (long)ctx->g.nx * ctx->g.ny * ctx->g.nz;
  const long n0 = lv[0].n_free;
  GF_HIP(hipMemcpyAsync(E64, E, sizeof(double) * ne, hipMemcpyDeviceToDevice, st));
  GF_TRY(cast_f64_to_f32(ne, E64, E32, st));
  Level& f = lv[0];
  if (fine_diag) {
    GF_HIP(hipMemcpyAsync(f.diag, fine_diag, sizeof(double) * n0, hipMemcpyDeviceToDevice, st));
  } else {
    GF_TRY(gf::fine_diag(ctx->g, ctx->map(), ctx->Ke, E64, f.diag, st));
  }
  LAUNCH(k_safe_diag, n0, st, n0, f.diag);
  GF_TRY(safe_inverse(n0, f.diag, f.dinv, f.dinv32, st));
  if (p.root_enrichment_mode != 0) {
    GF_TRY(minmax_f64(work, n0, f.diag, pw_s, st));
    LAUNCH(k_enrich_weights, n0, st, n0, f.diag, pw_s, p.root_enrichment_weight_floor,
           p.root_enrichment_weight_power, re_w);
  }
  if (p.root_local_correction_mode != 0) {
    GF_TRY(nodeblock_update(ctx->g, ctx->map(), ctx->Ke, E64, f.diag, nb_inv9, nb_pinv, st));
  }
  double h[4];
  GF_TRY(minmax_f64(work, ne, E64, pw_s + 8, st));
  GF_TRY(read_scalars(pw_s + 8, h, 2, st));
  E_max = h[0];
  E_min = h[1];
  ++setup_count;
  const double prev = lam_E_max;
  const bool reest = std::isnan(prev) || std::fabs(E_max - prev) / std::max(std::fabs(prev), 1e-12) > 0.10;
  if (reest) {
    GF_TRY(estimate_fine_lambda(st, &lam0));
    lam_E_max = E_max;
  }
  omega_used = std::min(p.omega, 0.9 * 2.0 / std::max(lam0, 1e-6));
  if ((int)lam.size() != L) lam.assign(L, NAN);
  lam[0] = lam0;
  contrast = E_max / std::max(E_min, 1e-300);
  GF_TRY(count_gt(work, ne, E64, pw_s + 8, 1e-8, pw_s + 10, st));
  GF_TRY(read_scalars(pw_s + 10, h, 1, st));
  void_frac = 1.0 - h[0] / (double)std::max(ne, 1L);
  const int bc = p.coarse_smooth_iters, bf = p.fine_smoother_degree;
  if (contrast > 1e5 || void_frac > 0.8) {
    coarse_it = std::max(bc, 6); fine_deg = std::max(bf, 4); high = 1;
  } else if (contrast > 100 || void_frac > 0.5) {
    coarse_it = std::max(bc, 4); fine_deg = std::max(bf, 3); high = 1;
  } else if (contrast > 10) {
    coarse_it = std::max(bc, 3); fine_deg = bf; high = 0;
  } else {
    coarse_it = bc; fine_deg = bf; high = 0;
  }
  for (int l = 1; l < L; ++l) {
    Level& v = lv[l];
    if (l == 1) {
      GF_TRY(level1_assemble(v.g, v.map(), ctx->g, Q, E64, v.A, st));
    } else {
      Level& u = lv[l - 1];
      GF_TRY(rap(u.g, u.map(), v.g, v.map(), u.A, u.KP, v.A, st));
    }
    GF_TRY(stencil_diag(v.g, v.map(), v.A, v.diag, st));
    LAUNCH(k_safe_diag, v.n_free, st, (long)v.n_free, v.diag);
    GF_TRY(safe_inverse(v.n_free, v.diag, v.dinv, v.dinv32, st));
    if (reest || std::isnan(lam[l])) GF_TRY(estimate_level_lambda(l, st, &lam[l]));
    if (v.prec == 1) GF_TRY(f64_to_f32_block(243L * v.g.n_nodes, v.A, v.A32, st));
  }
  chol_mode = 0;
  chol_regularised = 0;
  if (L > 1) {
    Level& c = lv[L - 1];
    const int n = c.n_free;
    // dense Cholesky when n <= dense_chol_max (multigrid_v4.py:1308); the
    // two triangular solves keep their right-hand side in LDS (at most
    // kTrsvMaxN entries), the explicit-inverse GEMV has no such bound
    if (n <= p.dense_chol_max && (p.coarsest_solve == 0 || n <= kTrsvMaxN)) {
      if (ch_n != n) {
        dfree(CH);
        dfree(CHt);
        dfree(CHr);
        dfree(CHi);
        dfree(CHy);
        GF_TRY(dalloc(&CH, (size_t)n * n));
        GF_TRY(dalloc(&CHt, (size_t)n * n));
        GF_TRY(dalloc(&CHr, (size_t)n));
        if (p.coarsest_solve == 0) {
          GF_TRY(dalloc(&CHi, (size_t)n * n));
          GF_TRY(dalloc(&CHy, (size_t)n * n));
        }
        ch_n = n;
      }
      GF_TRY(stencil_to_dense(c.g, c.map(), c.A, n, CH, st));
      GF_TRY(dense_cholesky(n, CH, CHt, d_fail, st));
      int fail = 0;
      GF_HIP(hipMemcpyAsync(&fail, d_fail, sizeof(int), hipMemcpyDeviceToHost, st));
      GF_HIP(hipStreamSynchronize(st));
      if (!fail) {
        chol_mode = 1;
      } else {
        // eps = max(mean(diag) * 1e-8, 1e-14) of the unregularised matrix
        GF_TRY(stencil_to_dense(c.g, c.map(), c.A, n, CH, st));
        std::vector<double> dg(n);
        std::vector<double> rawd(n);
        GF_TRY(stencil_diag(c.g, c.map(), c.A, pw_v, st));
        GF_HIP(hipMemcpyAsync(rawd.data(), pw_v, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += rawd[i];
        const double eps = std::max(s / n * 1e-8, 1e-14);
        GF_TRY(dense_add_diag(n, CH, eps, st));
        GF_TRY(dense_cholesky(n, CH, CHt, d_fail, st));
        GF_HIP(hipMemcpyAsync(&fail, d_fail, sizeof(int), hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        if (!fail) {
          chol_mode = 1;
          chol_regularised = 1;
        } else {
          chol_mode = 2;
        }
      }
    } else {
      chol_mode = 2;
    }
    if (chol_mode == 1) {
      GF_TRY(dense_recip_diag(n, CH, CHr, st));
      // Y = L^-1 in CHy, its transpose in CHt (the triangular path is unused here)
      if (p.coarsest_solve == 0) GF_TRY(dense_inverse_from_chol(n, CH, CHy, CHt, CHi, st));
    }
    if (chol_mode == 2 && !pcg_work) GF_TRY(dalloc(&pcg_work, 4L * c.n_free + 4));
    if (chol_mode == 2 && p.coarse_pcg_persistent && !pcg_uc)
      GF_TRY(pcg_buffers_acquire(c.n_free, &pcg_uc, &pcg_uc_bytes));
  }
  GF_HIP(hipStreamSynchronize(st));
  is_setup = true;
  return 0;
}

// ---------------------------------------------------------------------------
// outer FGMRES: host-driven restart/convergence logic on device Hessenberg
// ---------------------------------------------------------------------------
__global__ void k_outer_cycle(int* fl, double* g, const double* sc) {
  g[0] = sc[1];
  fl[0] = 1;
  fl[1] = 1;
  fl[2] = 0;
  fl[3] = 0;
}

int Gmg::outer_solve(const double* b, double* x, int x0_given, double tol, int maxiter,
                     int restart, double* history, int hist_cap, int* n_hist, int* iters,
                     int* converged, hipStream_t st, bool resume) {
  if (!is_setup) {
    set_error("GalerkinMatFreeGMG.setup() must be called before apply().");
    return -3;
  }
  if (restart < 1 || restart > 1024 || maxiter < 0) {
    set_error("gf_gmg_solve: restart must be in [1, 1024] and maxiter >= 0");
    return -2;
  }
  const long n = lv[0].n_free;
  if (resume) {
    if (!rs.valid || rs.b != b || rs.tol != tol || rs.restart != restart || rs.n != n ||
        maxiter < rs.total) {
      set_error("gf_gmg_solve_continue: no resumable solve (the last gf_gmg_solve on this "
                "hierarchy must have started from x = 0, with the same b, tol and restart, and "
                "stopped at its maxiter inside its first cycle or converged there; no set-up "
                "since, and maxiter not below the iterations already done)");
      return -4;
    }
  } else {
    rs.valid = false;
    GF_TRY(outer.alloc(restart, n));
  }
  Krylov& K = outer;
  const int m = K.m;
  int nh = 0;
  auto push = [&](double v) {
    if (history && nh < hist_cap) history[nh] = v;
    ++nh;
  };
  // a continued solve is the same first cycle, from x = 0
  if (!x0_given || resume) LAUNCH(k_fill, n, st, n, x, 0.0, (const int*)nullptr);
  double h[4];
  double bnorm;
  if (resume) {
    bnorm = rs.bnorm;
  } else {
    GF_TRY(dot_f64(work, n, b, b, K.sc, true, nullptr, st));
    GF_TRY(read_scalars(K.sc, h, 1, st));
    bnorm = h[0];
  }
  *iters = 0;
  *converged = 0;
  if (bnorm <= 1e-30) {
    *converged = 1;
    *n_hist = 0;
    return 0;
  }
  int total = resume ? rs.total : 0;
  bool conv = resume && rs.conv;
  int j0 = resume ? rs.total : 0;  // Arnoldi step to continue from (first cycle)
  int cycles = 0;
  bool cut_in_cycle = false;       // the last cycle stopped at maxiter before restart
  if (resume && conv) {
    // the first solve converged: its iterate, recombined from the same
    // Hessenberg solution and basis
    GF_TRY(combine(n, K.Z, n, K.y, K.fl + 2, true, x, nullptr, st));
    cycles = 1;
  }
  while (!conv && total < maxiter) {
    const int cycle_start = total - j0;
    if (j0 == 0) {
      EpiArgs<double> ep;
      ep.b = b;
      GF_TRY(level_op(0, x, K.V, EPI_RESID, ep, nullptr, st));
      GF_TRY(dot_f64(work, n, K.V, K.V, K.sc + 1, true, nullptr, st));
      GF_TRY(read_scalars(K.sc + 1, h, 1, st));
      const double beta = h[0];
      push(beta / bnorm);
      if (beta / bnorm <= tol) {
        conv = true;
        break;
      }
      LAUNCH1(k_outer_cycle, st, K.fl, K.g, K.sc);
      GF_TRY(scale_inv(n, K.V, K.sc + 1, -1.0, K.V, nullptr, st));
    }
    const int inner = std::min(restart, maxiter - cycle_start);
    for (int j = j0; j < inner; ++j) {
      double* Vj = K.V + (long)j * n;
      double* Zj = K.Z + (long)j * n;
      double* w = K.V + (long)(j + 1) * n;
      GF_TRY(apply(Vj, Zj, st));
      GF_TRY(level_op(0, Zj, w, EPI_PLAIN, EpiArgs<double>{}, nullptr, st));
      int swept = 0;
      if (p.mgs_sweep) {
        const GivensArgs gv{tol, K.cs, K.sn, K.g, K.sc, K.fl};
        swept = mgs_sweep(work, n, K.V, j, m, w, K.H, d_err + WAIT_SWEEP, nullptr,
                          p.mgs_sweep == 2, &gv, st);
        if (swept < 0) return swept;
      }
      if (!swept) {
        GF_TRY(dot_f64(work, n, K.V, w, K.H + j, false, nullptr, st));
        for (int i = 0; i < j; ++i)
          GF_TRY(mgs_axpy_dot(work, n, K.V + (long)i * n, w, K.H + (long)i * m + j,
                              K.V + (long)(i + 1) * n, K.H + (long)(i + 1) * m + j, false,
                              nullptr, st));
        GF_TRY(mgs_axpy_dot(work, n, Vj, w, K.H + (long)j * m + j, nullptr,
                            K.H + (long)(j + 1) * m + j, true, nullptr, st));
        GF_TRY(scale_inv(n, w, K.H + (long)(j + 1) * m + j, 1e-14, w, nullptr, st));
      }
      if (!swept) LAUNCH1(k_givens, st, m, j, tol, K.H, K.cs, K.sn, K.g, K.sc, K.fl);
      // projected residual, h_{j+1,j} and the wait-expiry flags in one wait
      GF_HIP(hipMemcpyAsync(hs, K.sc + 2, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
      GF_HIP(hipMemcpyAsync(hs + 2, d_err, sizeof(double) * kWaitFlags, hipMemcpyDeviceToHost,
                            st));
      GF_HIP(hipStreamSynchronize(st));
      std::memcpy(h, hs, sizeof(double) * 2);
      bool expired = false;
      for (int k = 0; k < kWaitFlags; ++k) expired |= hs[2 + k] != 0.0;
      if (expired) {
        rs.valid = false;
        return wait_failed(hs + 2, "gf_gmg_solve", st);
      }
      ++total;
      push(h[0]);
      if (h[0] <= tol || h[1] <= 1e-14) {
        conv = true;
        break;
      }
    }
    cut_in_cycle = !conv && (total - cycle_start) < restart;
    j0 = 0;
    ++cycles;
    hipLaunchKernelGGL(k_backsolve, dim3(1), dim3(256), 0, st, m, K.H, K.g, K.y, K.fl);
    GF_CHECK_LAUNCH();
    GF_TRY(combine(n, K.Z, n, K.y, K.fl + 2, true, x, nullptr, st));
    if (conv) break;
  }
  GF_HIP(hipStreamSynchronize(st));
  // resumable: one cycle from x = 0 that converged or stopped at maxiter
  // before its restart length (V, Z, H, the Givens rotations and g then hold
  // exactly the state a longer solve has at that step)
  const bool zero_start = resume || !x0_given;
  rs.valid = zero_start && cycles == 1 && (conv || cut_in_cycle);
  rs.b = b;
  rs.tol = tol;
  rs.restart = restart;
  rs.n = n;
  rs.total = total;
  rs.conv = conv;
  rs.bnorm = bnorm;
  *iters = total;
  *converged = conv ? 1 : 0;
  *n_hist = nh;
  return 0;
}

}  // namespace gf

namespace gf {

// ---------------------------------------------------------------------------
// traced cycle (multigrid_v4.py:1361-1728): one V or W cycle on rhs with the
// kernels of cycle() in the same order (the iterate, and so the output, is
// the one cycle() produces), plus the norms and dot products of the
// reference's per-visit record, written to device slots and read back once.
// Slot layout of one visit (gf_trace_visit.v):
//   0 |rhs|  1 |D^-1 rhs|  2 |x_pre|  3 |r_pre|  4 |R r_pre|  5 |P R r_pre|
//   6 |r_pre - P R r_pre|  7 |e_c|  8 |P e_c|  9 r_pre.(P e_c)
//   10 |r after the first correction|  11-17 first line-search scalars
//      [c.c, r.r, Ac.Ac, r.Ac, applied scale, degenerate, |trial|]
//   18 |second coarse rhs|  19 |e_c2|  20 |P e_c2|  21 r_1.(P e_c2)
//   22 |r after the second correction|  23-29 second line-search scalars
//   30 |x_post|  31 |rhs - K x_post|
// Coarsest visit: 0, 1, 7 = |x|, 10 = |rhs - K x|.
// Summary (16 doubles): 0 |rhs|  1 |D0^-1 rhs|  2 |out|  3 rhs.out
//   4 |pre-local residual|  5-8 root-local scalars (ap_st)  9 |line-search
//   trial residual|.
// ---------------------------------------------------------------------------
namespace {
enum : int {
  TV_RHS = 0, TV_JAC = 1, TV_PRE_X = 2, TV_PRE_R = 3, TV_RC = 4, TV_PROJ = 5, TV_GAP = 6,
  TV_EC = 7, TV_PEC = 8, TV_R_PEC = 9, TV_CORR_R = 10, TV_ST1 = 11, TV_RC2 = 18, TV_EC2 = 19,
  TV_PEC2 = 20, TV_R_PEC2 = 21, TV_CORR_R2 = 22, TV_ST2 = 23, TV_POST_X = 30, TV_POST_R = 31
};
static_assert(TV_POST_R < GF_TRACE_SLOTS, "trace slots");
}  // namespace

// applied root-local line-search scale (multigrid_v4.py:1035-1045)
__global__ void k_rl_scale(const double* st, double max_scale, double* s_out) {
  const double cn = sqrt(st[0]), rn = sqrt(st[1]), den = st[2];
  double s = 0.0;
  if (!(cn <= 1e-300 || rn <= 1e-300 || den <= 1e-300)) {
    s = fmin(max_scale, fmax(0.0, st[3] / den));
    if (s <= 1e-12) s = 0.0;
  }
  *s_out = s;
}

static int trace_nrm(Work& w, long n, const double* x, double* out, hipStream_t st) {
  return dot_f64(w, n, x, x, out, true, nullptr, st);
}

int Gmg::trace_visit(int l, const double* rhs, double* out, bool wcyc, hipStream_t st,
                     double** result, TraceState& ts) {
  const int id = (int)ts.meta.size();
  if (id >= ts.cap) {
    set_error("cycle_trace: more visits than slots");
    return -2;
  }
  ts.meta.push_back(TraceMeta{l, l == L - 1 ? 1 : 0, 0, 0, 0});
  double* s = ts.slots + (long)id * GF_TRACE_SLOTS;
  Level& v = lv[l];
  const long n = v.n_free;
  GF_TRY(trace_nrm(work, n, rhs, s + TV_RHS, st));
  LAUNCH(k_mul, n, st, n, v.dinv, rhs, v.T);
  GF_TRY(trace_nrm(work, n, v.T, s + TV_JAC, st));
  if (l == L - 1) {
    double* x = nullptr;
    GF_TRY(solve_coarsest(rhs, out, nullptr, st, &x));
    EpiArgs<double> ep;
    ep.b = rhs;
    GF_TRY(level_op(l, x, v.T, EPI_RESID, ep, nullptr, st));
    GF_TRY(trace_nrm(work, n, x, s + TV_EC, st));
    GF_TRY(trace_nrm(work, n, v.T, s + TV_CORR_R, st));
    *result = x;
    return 0;
  }
  Level& c = lv[l + 1];
  const long nc = c.n_free;
  const bool enrich = (l == 0 && p.root_enrichment_mode != 0);
  const bool ls = p.coarse_correction_policy != 0;
  double* x = nullptr;
  GF_TRY(smooth(l, nullptr, rhs, nullptr, 1, nullptr, st, &x));
  {
    EpiArgs<double> ep;
    ep.b = rhs;
    GF_TRY(level_op(l, x, v.R, EPI_RESID, ep, nullptr, st));
  }
  GF_TRY(trace_nrm(work, n, x, s + TV_PRE_X, st));
  GF_TRY(trace_nrm(work, n, v.R, s + TV_PRE_R, st));
  GF_TRY(restrict_(v.g, v.map(), c.g, c.map(), v.R, c.RHS, nullptr, st));
  GF_TRY(trace_nrm(work, nc, c.RHS, s + TV_RC, st));
  GF_TRY(prolong(v.g, v.map(), c.g, c.map(), c.RHS, v.T, nullptr, st));
  GF_TRY(trace_nrm(work, n, v.T, s + TV_PROJ, st));
  GF_TRY(trial_norm(work, n, v.R, v.T, ts.one, nullptr, s + TV_GAP, nullptr, st));
  double* ec = nullptr;
  GF_TRY(trace_visit(l + 1, c.RHS, nullptr, wcyc, st, &ec, ts));
  GF_TRY(trace_nrm(work, nc, ec, s + TV_EC, st));
  GF_TRY(prolong(v.g, v.map(), c.g, c.map(), ec, v.PE, nullptr, st));
  GF_TRY(trace_nrm(work, n, v.PE, s + TV_PEC, st));
  GF_TRY(dot_f64(work, n, v.R, v.PE, s + TV_R_PEC, false, nullptr, st));
  const size_t st_bytes = 7 * sizeof(double);
  int n_corr = 1;
  bool second = false;
  if (!wcyc) {
    GF_TRY(correct(l, x, v.R, v.PE, v.R2, nullptr, st));
    if (ls) GF_HIP(hipMemcpyAsync(s + TV_ST1, v.st, st_bytes, hipMemcpyDeviceToDevice, st));
    GF_TRY(trace_nrm(work, n, v.R2, s + TV_CORR_R, st));
    if (enrich) {
      GF_TRY(enrichment_rhs(v.R2, c.RHS, nullptr, st));
      second = true;
    }
  } else {
    if (!ls) {
      LAUNCH(k_add, n, st, n, x, v.PE, (const int*)nullptr);
      EpiArgs<double> ep;
      ep.b = rhs;
      GF_TRY(level_op(l, x, v.R2, EPI_RESID, ep, nullptr, st));
    } else {
      GF_TRY(correct(l, x, v.R, v.PE, v.R2, nullptr, st));
      GF_HIP(hipMemcpyAsync(s + TV_ST1, v.st, st_bytes, hipMemcpyDeviceToDevice, st));
    }
    GF_TRY(trace_nrm(work, n, v.R2, s + TV_CORR_R, st));
    if (enrich)
      GF_TRY(enrichment_rhs(v.R2, c.RHS, nullptr, st));
    else
      GF_TRY(restrict_(v.g, v.map(), c.g, c.map(), v.R2, c.RHS, nullptr, st));
    second = true;
  }
  if (second) {
    GF_TRY(trace_nrm(work, nc, c.RHS, s + TV_RC2, st));
    double* e2 = nullptr;
    GF_TRY(trace_visit(l + 1, c.RHS, nullptr, wcyc, st, &e2, ts));
    GF_TRY(trace_nrm(work, nc, e2, s + TV_EC2, st));
    GF_TRY(prolong(v.g, v.map(), c.g, c.map(), e2, v.PE, nullptr, st));
    GF_TRY(trace_nrm(work, n, v.PE, s + TV_PEC2, st));
    GF_TRY(dot_f64(work, n, v.R2, v.PE, s + TV_R_PEC2, false, nullptr, st));
    GF_TRY(correct(l, x, v.R2, v.PE, v.T, nullptr, st));
    if (ls) GF_HIP(hipMemcpyAsync(s + TV_ST2, v.st, st_bytes, hipMemcpyDeviceToDevice, st));
    GF_TRY(trace_nrm(work, n, v.T, s + TV_CORR_R2, st));
    n_corr = 2;
  }
  const bool extra = (l == 0 && high);
  GF_TRY(smooth(l, x, rhs, extra ? nullptr : out, 1, nullptr, st, &x));
  if (extra) GF_TRY(smooth(l, x, rhs, out, 1, nullptr, st, &x));
  {
    EpiArgs<double> ep;
    ep.b = rhs;
    GF_TRY(level_op(l, x, v.T, EPI_RESID, ep, nullptr, st));
  }
  GF_TRY(trace_nrm(work, n, x, s + TV_POST_X, st));
  GF_TRY(trace_nrm(work, n, v.T, s + TV_POST_R, st));
  TraceMeta& m = ts.meta[id];
  m.n_corr = n_corr;
  m.enrich = enrich ? 1 : 0;
  m.extra = extra ? 1 : 0;
  *result = x;
  return 0;
}

static int trace_visit_count(int l, int L, bool wcyc, bool enrich) {
  if (l == L - 1) return 1;
  const int k = (wcyc || (l == 0 && enrich)) ? 2 : 1;
  return 1 + k * trace_visit_count(l + 1, L, wcyc, enrich);
}

int Gmg::cycle_trace(const double* rhs, double* out, int cycle_type, gf_trace_visit* visits,
                     int cap, int* n_visits, double* summary, hipStream_t st) {
  if (!is_setup) {
    set_error("GalerkinMatFreeGMG.setup() must be called before cycle_trace().");
    return -3;
  }
  if (cycle_type < -1 || cycle_type > 1) {
    set_error("cycle_trace: cycle_type must be -1 (configured), 0 (V) or 1 (W)");
    return -2;
  }
  const bool wcyc = cycle_type < 0 ? p.cycle_type == 1 : cycle_type == 1;
  const int nv = trace_visit_count(0, L, wcyc, p.root_enrichment_mode != 0);
  if (n_visits) *n_visits = nv;
  if (nv > cap) {
    set_error("cycle_trace: visit capacity too small");
    return -2;
  }
  const long n = lv[0].n_free;
  const long nslots = (long)nv * GF_TRACE_SLOTS + 16 + 1;
  double* dbuf = nullptr;
  GF_TRY(dalloc(&dbuf, nslots));
  TraceState ts;
  ts.slots = dbuf;
  ts.cap = nv;
  double* sm = dbuf + (long)nv * GF_TRACE_SLOTS;
  double* one = sm + 16;
  ts.one = one;
  int rc = 0;
  do {
    LAUNCH(k_fill, nslots, st, nslots, dbuf, 0.0, (const int*)nullptr);
    LAUNCH(k_fill, 1, st, 1, one, 1.0, (const int*)nullptr);
    const int rl = p.root_local_correction_mode;
    double* target = rl == 0 ? out : ap_x;
    double* xr = nullptr;
    if ((rc = trace_visit(0, rhs, target, wcyc, st, &xr, ts))) break;
    if (xr != target) LAUNCH(k_copy, n, st, n, xr, target, (const int*)nullptr);
    if (rl == 0) {
      EpiArgs<double> ep;
      ep.b = rhs;
      if ((rc = level_op(0, out, ap_res, EPI_RESID, ep, nullptr, st))) break;
    } else {
      if ((rc = root_local(rhs, ap_x, out, st))) break;
      GF_HIP(hipMemcpyAsync(sm + 5, ap_st, 4 * sizeof(double), hipMemcpyDeviceToDevice, st));
      if (rl == 1) {
        hipLaunchKernelGGL(k_rl_scale, dim3(1), dim3(1), 0, st, ap_st,
                           p.root_local_correction_max_scale, sm + 10);
        GF_CHECK_LAUNCH();
        if ((rc = trial_norm(work, n, ap_res, ap_trial, sm + 10, nullptr, sm + 9, nullptr, st)))
          break;
      }
    }
    if ((rc = trace_nrm(work, n, ap_res, sm + 4, st))) break;
    if ((rc = trace_nrm(work, n, rhs, sm + 0, st))) break;
    LAUNCH(k_mul, n, st, n, lv[0].dinv, rhs, lv[0].T);
    if ((rc = trace_nrm(work, n, lv[0].T, sm + 1, st))) break;
    if ((rc = trace_nrm(work, n, out, sm + 2, st))) break;
    if ((rc = dot_f64(work, n, rhs, out, sm + 3, false, nullptr, st))) break;
    std::vector<double> h(nslots);
    GF_HIP(hipMemcpyAsync(h.data(), dbuf, sizeof(double) * nslots, hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nv; ++i) {
      const TraceMeta& m = ts.meta[i];
      gf_trace_visit& o = visits[i];
      o.visit_id = i;
      o.level = m.level;
      o.is_coarsest = m.is_coarsest;
      o.n_coarse_corrections = m.n_corr;
      o.enrichment = m.enrich;
      o.extra_fine_postsmooth = m.extra;
      std::memcpy(o.v, h.data() + (long)i * GF_TRACE_SLOTS, sizeof(double) * GF_TRACE_SLOTS);
    }
    if (summary) std::memcpy(summary, h.data() + (long)nv * GF_TRACE_SLOTS, sizeof(double) * 16);
  } while (0);
  (void)hipStreamSynchronize(st);
  dfree(dbuf);
  return rc;
}

}  // namespace gf
