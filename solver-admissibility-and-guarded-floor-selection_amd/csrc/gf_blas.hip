// Deterministic BLAS-1 building blocks with device-resident scalars.
//
// Every reduction uses a fixed grid of at most kRedBlocks blocks, a fixed
// in-block tree and a ticket so that the last block to finish folds the block
// partials in index order: the result is bitwise reproducible for a given n
// and never needs a host round trip.  All kernels accept an optional `gate`
// (device int); when it points at 0 the kernel is a no-op.  The inner Krylov
// loops use this to stop after a breakdown without host synchronisation.
#include "gf_reduce.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <mutex>

namespace gf {

std::string& err_slot() {
  static thread_local std::string s;
  return s;
}
void set_error(const std::string& msg) { err_slot() = msg; }
const char* last_error() { return err_slot().c_str(); }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(long n, const T* __restrict__ x,
                                                const T* __restrict__ y, double* partials,
                                                int* ticket, double* out, int post,
                                                const int* gate) {
  if (gate && *gate == 0) return;
  double a[1] = {dot_thread<T>(n, x, y)};
  finish_reduce<1>(a, partials, ticket, out, post);
}

// out[0]=x.x  out[1]=y.y  out[2]=z.z  out[3]=y.z  (line-search statistics);
// with `step` set the folding block also forms the line-search step
// (ls_scale) into out[4], out[5]
__global__ __launch_bounds__(kBlock) void k_dot4_ls(long n, const double* __restrict__ c,
                                                    const double* __restrict__ r,
                                                    const double* __restrict__ Ac,
                                                    double* partials, int* ticket, double* out,
                                                    int step, double max_scale,
                                                    const int* gate) {
  if (gate && *gate == 0) return;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  // four strides' loads in flight, sums in the plain loop's order
  const long st = (long)gridDim.x * blockDim.x;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + 3 * st < n; i += 4 * st) {
    double cv[4], rv[4], av[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      cv[u] = c[i + u * st];
      rv[u] = r[i + u * st];
      av[u] = Ac[i + u * st];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[0] += cv[u] * cv[u];
      a[1] += rv[u] * rv[u];
      a[2] += av[u] * av[u];
      a[3] += rv[u] * av[u];
    }
  }
  for (; i < n; i += st) {
    double cv = c[i], rv = r[i], av = Ac[i];
    a[0] += cv * cv;
    a[1] += rv * rv;
    a[2] += av * av;
    a[3] += rv * av;
  }
  if (finish_reduce<4>(a, partials, ticket, out, 0) && step && threadIdx.x == 0)
    ls_scale(out, max_scale);
}

// w -= h*v (h = *hptr) and out = next . w  (fused MGS step)
__global__ __launch_bounds__(kBlock) void k_mgs_axpy_dot(long n, const double* __restrict__ v,
                                                         double* __restrict__ w,
                                                         const double* hptr,
                                                         const double* __restrict__ next,
                                                         double* partials, int* ticket,
                                                         double* out, int post, const int* gate) {
  if (gate && *gate == 0) return;
  const double h = *hptr;
  double a[1] = {0.0};
  // four strides' loads in flight, sums in the plain loop's order
  const long st = (long)gridDim.x * blockDim.x;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + 3 * st < n; i += 4 * st) {
    double vv[4], ww[4], nx[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      vv[u] = v[i + u * st];
      ww[u] = w[i + u * st];
      nx[u] = next ? next[i + u * st] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double wv = ww[u] - h * vv[u];
      w[i + u * st] = wv;
      a[0] += (next ? nx[u] : wv) * wv;
    }
  }
  for (; i < n; i += st) {
    double wv = w[i] - h * v[i];
    w[i] = wv;
    a[0] += (next ? next[i] : wv) * wv;
  }
  finish_reduce<1>(a, partials, ticket, out, post);
}


// ---------------------------------------------------------------------------
// Modified Gram-Schmidt sweep of one Arnoldi step in one persistent launch.
// w stays in registers (PER elements per thread); each projection needs one
// grid-wide exchange of block partials, through a ring of three slots per
// block in uncached memory.  The value in a slot is its own arrival signal:
// an empty slot holds the bit pattern kRingEmpty, a block publishes its
// partial of round s into slot (r0+s) mod 3, and every block polls the G
// slots of that round until none is empty and folds the values it read in
// block order, so every block derives the same h_s bit for bit (the same
// values and order as separate flag and partial reads).  Once a block has
// read round s, every block has published round s and so finished reading
// round s-1: the block then empties its own slot of round s-1, which is
// written next at round s+2, and waits for that store's acknowledgement
// before it publishes round s+1, so no reader of round s+2 can see the
// stale value.  flags[b] carries the ring position r0 from sweep to sweep.
// ---------------------------------------------------------------------------
constexpr int kSweepBlock = 256;
constexpr long kSweepSpinCap = 1L << 20;

// The variant without the prefetch asks for three waves per SIMD (at most
// 168 registers a lane): 768 co-resident blocks, so the 755-block grid of a
// 1M-element solve fits a process with one hardware queue.
template <int PER, bool PF>
__global__ __launch_bounds__(kSweepBlock, PF ? 1 : 3) void k_mgs_sweep(long n, int j, int m,
                                                          const double* __restrict__ V,
                                                          double* __restrict__ w, double* H,
                                                          double* ring, int* flags, double* err,
                                                          const int* gate, GivensArgs gv) {
  const bool open = gate_open(gate);  // tested before the first exchange
  __shared__ double lds[4];
  __shared__ double hsh;
  __shared__ unsigned f0s;
  const int G = gridDim.x, b = blockIdx.x;
  const long stride = (long)G * kSweepBlock;
  const long base = (long)b * kSweepBlock + threadIdx.x;
  // PF: V_{s+1} is prefetched while round s's exchange is in flight (a
  // third register array); without it the same loads follow the exchange,
  // with the same values and order (identical bits) and fewer registers
  double wr[PER], vr[PER], vn[PF ? PER : 1];
  bool broken = false;  // wave 0 only
  // w and V_0 are requested before the ring position, so the two load
  // latencies overlap at the barrier below
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const long i = base + k * stride;
    wr[k] = i < n ? w[i] : 0.0;
    vr[k] = i < n ? V[i] : 0.0;
  }
  // Every complete sweep advances every block's ring position by j+2 (mod 3),
  // so all positions are equal when a sweep starts.
  if (threadIdx.x == 0)
    f0s = (unsigned)__hip_atomic_load(&flags[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (!open) return;  // uniform over the grid: no block enters the exchange
  const int r0 = (int)(f0s % 3u);
  for (int s = 0; s <= j + 1; ++s) {
    double a = 0.0;
    if (s <= j) {
#pragma unroll
      for (int k = 0; k < PER; ++k) a += wr[k] * vr[k];
    } else {
#pragma unroll
      for (int k = 0; k < PER; ++k) a += wr[k] * wr[k];
    }
    const double bs = block_sum(a, lds);
    const int r = (r0 + s) % 3;
    double* cur = ring + (long)r * kSweepMaxBlocks;
    if (threadIdx.x == 0) {
      store_ack_wait();  // the previous round's slot emptying has landed
      __hip_atomic_store(&cur[b], bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (PF && s < j) {  // prefetch V_{s+1} while the exchange is in flight
      const double* vs = V + (long)(s + 1) * n;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const long i = base + k * stride;
        vn[PF ? k : 0] = i < n ? vs[i] : 0.0;
      }
    }
    if (threadIdx.x < 64) {
      // bounded wait; after one expired wait the block stops waiting (the
      // host sees *err and discards the result), so a sweep always ends
      long spins = 0;
      double t = 0.0;
      while (!broken) {
        int ok = 1;
        t = 0.0;
        for (int q = threadIdx.x; q < G; q += 64) {
          const double v = __hip_atomic_load(&cur[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok &= (int)((unsigned long long)__double_as_longlong(v) != kRingEmpty);
          t += v;
        }
        if (__all(ok)) break;
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSweepSpinCap) {
          broken = true;
          if (threadIdx.x == 0) *err = 1.0;
        }
      }
      t = wave_sum(t);
      if (threadIdx.x == 0) {
        hsh = t;
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(
                               &ring[(long)((r + 2) % 3) * kSweepMaxBlocks + b]),
                           kRingEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __syncthreads();
    const double h = hsh;
    if (s <= j) {
#pragma unroll
      for (int k = 0; k < PER; ++k) wr[k] -= h * vr[k];
      if (b == 0 && threadIdx.x == 0) H[(long)s * m + j] = h;
      if (s < j) {
        const double* vs = V + (long)(s + 1) * n;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
          const long i = base + k * stride;
          vr[k] = PF ? vn[PF ? k : 0] : (i < n ? vs[i] : 0.0);
        }
      }
    } else {
      const double nrm = sqrt(h);
      if (b == 0 && threadIdx.x == 0) {
        H[(long)(j + 1) * m + j] = nrm;
        // the Givens update of this step, in place of a separate launch
        if (gv.fl) givens_step(m, j, gv, H);
      }
      const bool scale = nrm > 1e-14;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const long i = base + k * stride;
        if (i < n) w[i] = scale ? wr[k] / nrm : wr[k];
      }
    }
  }
  // the ring position after this sweep's j+2 rounds
  if (threadIdx.x == 0)
    __hip_atomic_store(&flags[b], (r0 + j + 2) % 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_fill_int(int n, int* p, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_u64(int n, unsigned long long* p, unsigned long long v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

int hw_queues() {
  static std::atomic<int> q{0};
  if (q.load() == 0) {
    int v = 4;
    if (const char* s = std::getenv("GPU_MAX_HW_QUEUES")) {
      const int e = std::atoi(s);
      if (e > 0) v = e;
    }
    q.store(std::min(32, std::max(1, v)));
  }
  return q.load();
}

namespace {
// A sweep's blocks spin until every block of its grid has arrived, so the
// whole grid must be resident at once.  Solver instances on other streams
// may be running sweeps (or persistent PCGs) at the same moment, one per
// hardware queue, so a grid may use at most 1/hw_queues() of the device's
// co-resident capacity; larger orthogonalisations use one launch per
// projection.
template <int PER, bool PF>
int sweep_fits(int G) {
  static std::atomic<int> max_blocks{-1};  // co-resident blocks of this instantiation
  if (max_blocks.load() < 0) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mgs_sweep<PER, PF>, kSweepBlock,
                                                     0) != hipSuccess)
      return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    max_blocks.store(per_cu * cus);
  }
  return (long)hw_queues() * G <= max_blocks.load();
}

template <int PER, bool PF>
int launch_sweep_v(Work& wk, long n, const double* V, int j, int m, double* wv, double* H,
                   double* err, const int* gate, int cooperative, const GivensArgs* gvp, int G,
                   hipStream_t st) {
  GivensArgs gv = gvp ? *gvp : GivensArgs{0.0, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (wk.sweep_G != G) {
    // the sweep needs all ring positions of its grid equal and the ring
    // empty at entry; state left by a different grid (or an unknown history)
    // is reset first -- plain kernels, so this also holds where the sequence
    // is recorded in a graph
    hipLaunchKernelGGL(k_fill_int, dim3(4), dim3(256), 0, st, kSweepMaxBlocks, wk.flags, 0);
    GF_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fill_u64, dim3(12), dim3(256), 0, st, 3 * kSweepMaxBlocks,
                       reinterpret_cast<unsigned long long*>(wk.ring), kRingEmpty);
    GF_CHECK_LAUNCH();
    wk.sweep_G = G;
  }
  if (cooperative) {
    void* args[] = {&n,   &j,   &m,       (void*)&V,  &wv, &H, &wk.ring, &wk.flags, &err,
                    (void*)&gate, &gv};
    GF_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_mgs_sweep<PER, PF>),
                                      dim3(G), dim3(kSweepBlock), args, 0, st));
  } else {
    // G is far below the device's co-resident capacity (checked above), so
    // every block is resident once the kernels ahead of it drain; the waits
    // are bounded regardless
    hipLaunchKernelGGL((k_mgs_sweep<PER, PF>), dim3(G), dim3(kSweepBlock), 0, st, n, j, m, V,
                       wv, H, wk.ring, wk.flags, err, gate, gv);
    GF_CHECK_LAUNCH();
  }
  return 1;
}

// The prefetching sweep where its grid fits the co-residency rule, else the
// one without the prefetch (fewer registers, more co-resident blocks: at 1M
// elements its 755 blocks fit a process with one hardware queue), else none
// (the caller takes one launch per projection).  Both give the same bits.
std::atomic<int> g_sweep_form{1};
template <int PER>
int launch_sweep(Work& wk, long n, const double* V, int j, int m, double* wv, double* H,
                 double* err, const int* gate, int cooperative, const GivensArgs* gvp, int G,
                 hipStream_t st) {
  if (g_sweep_form.load() == 1 && sweep_fits<PER, true>(G))
    return launch_sweep_v<PER, true>(wk, n, V, j, m, wv, H, err, gate, cooperative, gvp, G, st);
  if (sweep_fits<PER, false>(G))
    return launch_sweep_v<PER, false>(wk, n, V, j, m, wv, H, err, gate, cooperative, gvp, G, st);
  return 0;
}
}  // namespace

void set_sweep_form(int form) { g_sweep_form = form == 2 ? 2 : 1; }

int mgs_sweep(Work& wk, long n, const double* V, int j, int m, double* wv, double* H, double* err,
              const int* gate, int cooperative, const GivensArgs* gv, hipStream_t st) {
  constexpr int kPerMax = 16;
  int G = cdiv(n, (long)kSweepBlock * kPerMax);
  if (G < 1) G = 1;
  if (G > kSweepMaxBlocks) return 0;
  const long per = (n + (long)G * kSweepBlock - 1) / ((long)G * kSweepBlock);
  if (per <= 1) return launch_sweep<1>(wk, n, V, j, m, wv, H, err, gate, cooperative, gv, G, st);
  if (per <= 2) return launch_sweep<2>(wk, n, V, j, m, wv, H, err, gate, cooperative, gv, G, st);
  if (per <= 4) return launch_sweep<4>(wk, n, V, j, m, wv, H, err, gate, cooperative, gv, G, st);
  if (per <= 8) return launch_sweep<8>(wk, n, V, j, m, wv, H, err, gate, cooperative, gv, G, st);
  return launch_sweep<16>(wk, n, V, j, m, wv, H, err, gate, cooperative, gv, G, st);
}

// y = b - s*x ; out = ||y||  with s = *sptr (trial residual norm, y may be null)
__global__ __launch_bounds__(kBlock) void k_trial_norm(long n, const double* __restrict__ b,
                                                       const double* __restrict__ x,
                                                       const double* sptr, double* y,
                                                       double* partials, int* ticket,
                                                       double* out, const int* gate) {
  if (gate && *gate == 0) return;
  const double s = *sptr;
  double a[1] = {0.0};
  // four strides' loads in flight, sums in the plain loop's order
  const long st = (long)gridDim.x * blockDim.x;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + 3 * st < n; i += 4 * st) {
    double bv[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      bv[u] = b[i + u * st];
      xv[u] = x[i + u * st];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double t = bv[u] - s * xv[u];
      if (y) y[i + u * st] = t;
      a[0] += t * t;
    }
  }
  for (; i < n; i += st) {
    double t = b[i] - s * x[i];
    if (y) y[i] = t;
    a[0] += t * t;
  }
  finish_reduce<1>(a, partials, ticket, out, 1);
}

// ---------------------------------------------------------------------------
// Residual line search of one coarse correction (multigrid_v4.py:867-931) in
// one persistent launch of red_grid(n) blocks.  Phase 1 forms the four
// statistics of k_dot4_ls, phase 2 the trial residual and its norm as
// k_trial_norm, phase 3 the accept/apply of k_ls_apply; between phases every
// block publishes its partial sums and all blocks wait until every block
// has arrived.  Each block then folds all partials in block order with the
// folding block's code of finish_reduce, so every block derives the
// statistics, the step and the trial norm bit for bit as the three launches
// do.  Phases 2 and 3 give a thread the same indices, so y (written by both)
// needs no ordering between blocks.
//
// Arrival: cnt[0] counts block arrivals over all launches; cnt[1] holds its
// value when a launch starts (launches on one Work are ordered by their
// stream), so round k of this launch is complete at cnt[1] + k*G.  Block 0
// advances cnt[1] once round 2 is complete, when every block has read it.
// The waits are bounded: an expired wait raises *err and the block goes on.
// ---------------------------------------------------------------------------
constexpr long kLsSpinCap = 1L << 20;

__device__ __forceinline__ void ls_arrive_wait(unsigned long long* cnt, unsigned long long target,
                                               double* err) {
  store_ack_wait();  // this block's partials have landed
  __hip_atomic_fetch_add(&cnt[0], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  long spins = 0;
  while (__hip_atomic_load(&cnt[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins > kLsSpinCap) {
      *err = 1.0;
      break;
    }
  }
}

__device__ __forceinline__ double ls_fold(const double* parts, int G, double* lds) {
  double a = 0.0;
  for (int i = threadIdx.x; i < G; i += blockDim.x)
    a += __hip_atomic_load(&parts[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return block_sum(a, lds);
}

// ls_fold of NP partial arrays (kRedBlocks apart) at once; results in thread 0
template <int NP>
__device__ __forceinline__ void ls_fold_n(const double* parts, int G, double (&out)[NP],
                                          double* lds) {
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    out[p] = 0.0;
    for (int i = threadIdx.x; i < G; i += blockDim.x)
      out[p] += __hip_atomic_load(&parts[p * kRedBlocks + i], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum_n<NP>(out, lds);
}

__global__ __launch_bounds__(kBlock) void k_ls_fused(long n, const double* __restrict__ c,
                                                     const double* __restrict__ r,
                                                     const double* __restrict__ Ac,
                                                     double* __restrict__ x, double* y,
                                                     double* stat, double max_scale, double ratio,
                                                     double* parts, unsigned long long* cnt,
                                                     double* err, const int* gate) {
  const bool open = gate_open(gate);  // tested before the first exchange
  __shared__ double lds[16];
  __shared__ double sst[8];
  __shared__ unsigned long long base_s;
  const int G = gridDim.x, b = blockIdx.x;
  if (threadIdx.x == 0)
    base_s = __hip_atomic_load(&cnt[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const long st = (long)G * blockDim.x;
  const long i0 = b * (long)blockDim.x + threadIdx.x;
  // phase 1: c.c, r.r, Ac.Ac, r.Ac (k_dot4_ls's loop and order)
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  {
    long i = i0;
    for (; i + 3 * st < n; i += 4 * st) {
      double cv[4], rv[4], av[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cv[u] = c[i + u * st];
        rv[u] = r[i + u * st];
        av[u] = Ac[i + u * st];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[0] += cv[u] * cv[u];
        a[1] += rv[u] * rv[u];
        a[2] += av[u] * av[u];
        a[3] += rv[u] * av[u];
      }
    }
    for (; i < n; i += st) {
      double cv = c[i], rv = r[i], av = Ac[i];
      a[0] += cv * cv;
      a[1] += rv * rv;
      a[2] += av * av;
      a[3] += rv * av;
    }
  }
  block_sum_n<4>(a, lds);
  if (!open) return;  // uniform over the grid: no block enters the exchange
  const unsigned long long base = base_s;  // written by thread 0 before block_sum_n's barriers
  if (threadIdx.x == 0) {
#pragma unroll
    for (int p = 0; p < 4; ++p)
      __hip_atomic_store(&parts[p * kRedBlocks + b], a[p], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    ls_arrive_wait(cnt, base + (unsigned long long)G, err);
  }
  __syncthreads();
  {
    double t[4];
    ls_fold_n<4>(parts, G, t, lds);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int p = 0; p < 4; ++p) sst[p] = t[p];
      ls_scale(sst, max_scale);
    }
  }
  __syncthreads();
  // phase 2: y = r - alpha Ac and ||y||^2 (k_trial_norm's loop and order)
  const double s = sst[4];
  double a2 = 0.0;
  {
    long i = i0;
    for (; i + 3 * st < n; i += 4 * st) {
      double bv[4], xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        bv[u] = r[i + u * st];
        xv[u] = Ac[i + u * st];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        double t = bv[u] - s * xv[u];
        if (y) y[i + u * st] = t;
        a2 += t * t;
      }
    }
    for (; i < n; i += st) {
      double t = r[i] - s * Ac[i];
      if (y) y[i] = t;
      a2 += t * t;
    }
  }
  const double bs2 = block_sum(a2, lds);
  if (threadIdx.x == 0) {
    __hip_atomic_store(&parts[4 * kRedBlocks + b], bs2, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    ls_arrive_wait(cnt, base + 2ull * (unsigned long long)G, err);
  }
  __syncthreads();
  {
    const double t = ls_fold(parts + 4 * kRedBlocks, G, lds);
    if (threadIdx.x == 0) sst[6] = sqrt(t);
  }
  __syncthreads();
  if (b == 0 && threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) stat[k] = sst[k];
    __hip_atomic_store(&cnt[1], base + 2ull * (unsigned long long)G, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  // phase 3: accept or reject, x += s c (k_ls_apply)
  const double rn = sqrt(sst[1]);
  const bool acc = (sst[5] == 0.0) && !(sst[6] > ratio * rn);
  const double sc = acc ? sst[4] : 0.0;
  for (long i = i0; i < n; i += st) {
    x[i] = x[i] + sc * c[i];
    if (y && !acc) y[i] = r[i];
  }
}

// y = a*x + b*y  with scalars read from device (aptr/bptr may be null = 0)
__global__ __launch_bounds__(kBlock) void k_axpby_dev(long n, const double* aptr, double asign,
                                                      const double* __restrict__ x,
                                                      const double* bptr, double* __restrict__ y,
                                                      const int* gate) {
  if (gate && *gate == 0) return;
  const double a = asign * *aptr;
  const double bb = bptr ? *bptr : 1.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] = (bptr ? bb * y[i] : y[i]) + a * x[i];
}

// y = x / h  if h > thresh  (Arnoldi normalisation; otherwise y untouched)
__global__ __launch_bounds__(kBlock) void k_scale_inv(long n, const double* __restrict__ x,
                                                      const double* hptr, double thresh,
                                                      double* __restrict__ y, const int* gate) {
  if (gate && *gate == 0) return;
  const double h = *hptr;
  if (!(h > thresh)) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] = x[i] / h;
}

// x = sum_j y[j] * Z[j]   (j < *jcount), accumulated in j order from 0
__global__ __launch_bounds__(kBlock) void k_combine(long n, const double* __restrict__ Z,
                                                    long ldz, const double* ycoef,
                                                    const int* jcount, int accumulate,
                                                    double* __restrict__ x, const int* gate) {
  if (gate && *gate == 0) return;
  const int m = *jcount;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    double acc = accumulate ? x[i] : 0.0;
    // eight basis vectors' loads in flight, accumulated in j order
    int j = 0;
    for (; j + 8 <= m; j += 8) {
      double z[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) z[u] = Z[(long)(j + u) * ldz + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = acc + ycoef[j + u] * z[u];
    }
    for (; j < m; ++j) acc = acc + ycoef[j] * Z[(long)j * ldz + i];
    x[i] = acc;
  }
}

template <typename TO, typename TI>
__global__ __launch_bounds__(kBlock) void k_cast(long n, const TI* __restrict__ x,
                                                 TO* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] = (TO)x[i];
}

__global__ __launch_bounds__(kBlock) void k_sub(long n, const double* __restrict__ a,
                                                const double* __restrict__ b,
                                                double* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] = a[i] - b[i];
}

// E statistics: out[0]=max(E) out[1]=min(E) (exact, order-free)
__global__ __launch_bounds__(kBlock) void k_minmax(long n, const double* __restrict__ E,
                                                   double* out2, double* partials, int* ticket) {
  __shared__ double smax[kBlock], smin[kBlock];
  __shared__ bool last;
  double mx = -INFINITY, mn = INFINITY;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    mx = fmax(mx, E[i]);
    mn = fmin(mn, E[i]);
  }
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partials[blockIdx.x], smax[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&partials[kRedBlocks + blockIdx.x], smin[0], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    store_ack_wait();
    last = (__hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
            (int)gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  if (threadIdx.x == 0) {
    double a = -INFINITY, b = INFINITY;
    for (int i = 0; i < (int)gridDim.x; ++i) {
      a = fmax(a, __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      b = fmin(b, __hip_atomic_load(&partials[kRedBlocks + i], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT));
    }
    out2[0] = a;
    out2[1] = b;
    __hip_atomic_store(ticket, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// count of E > thr   (thr read from device: *thr_src * scale)
__global__ __launch_bounds__(kBlock) void k_count_gt(long n, const double* __restrict__ E,
                                                     const double* thr_src, double scale,
                                                     double* partials, int* ticket, double* out) {
  const double thr = (double)(*thr_src * scale);
  double a[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    a[0] += (E[i] > thr) ? 1.0 : 0.0;
  finish_reduce<1>(a, partials, ticket, out, 0);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
int dot_f64(Work& w, long n, const double* x, const double* y, double* out, bool sqrt_,
            const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_dot<double>, dim3(red_grid(n)), dim3(kBlock), 0, st, n, x, y,
                     w.partials, w.ticket, out, sqrt_ ? 1 : 0, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int dot_f32(Work& w, long n, const float* x, const float* y, double* out, bool sqrt_,
            const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_dot<float>, dim3(red_grid(n)), dim3(kBlock), 0, st, n, x, y, w.partials,
                     w.ticket, out, sqrt_ ? 1 : 0, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int dot4_ls(Work& w, long n, const double* c, const double* r, const double* Ac, double* out4,
            const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_dot4_ls, dim3(red_grid(n)), dim3(kBlock), 0, st, n, c, r, Ac, w.partials,
                     w.ticket, out4, 0, 0.0, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int dot4_ls_step(Work& w, long n, const double* c, const double* r, const double* Ac,
                 double* out6, double max_scale, const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_dot4_ls, dim3(red_grid(n)), dim3(kBlock), 0, st, n, c, r, Ac, w.partials,
                     w.ticket, out6, 1, max_scale, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int mgs_axpy_dot(Work& w, long n, const double* v, double* wv, const double* h,
                 const double* next, double* out, bool sqrt_, const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_mgs_axpy_dot, dim3(red_grid(n)), dim3(kBlock), 0, st, n, v, wv, h, next,
                     w.partials, w.ticket, out, sqrt_ ? 1 : 0, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int trial_norm(Work& w, long n, const double* b, const double* x, const double* s, double* y,
               double* out, const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_trial_norm, dim3(red_grid(n)), dim3(kBlock), 0, st, n, b, x, s, y,
                     w.partials, w.ticket, out, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

static std::atomic<int> g_ls_form{1};
void set_ls_form(int form) { g_ls_form = form ? 1 : 0; }
int get_ls_form() { return g_ls_form.load(); }

// The line search's blocks wait for each other like a sweep's, so its grid
// follows the same co-residency rule (1/hw_queues() of the capacity).
static int ls_fits(int G) {
  static std::atomic<int> max_blocks{-1};
  if (max_blocks.load() < 0) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ls_fused, kBlock, 0) !=
        hipSuccess)
      return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    max_blocks.store(per_cu * cus);
  }
  return (long)hw_queues() * G <= max_blocks.load();
}

int ls_fused(Work& w, long n, const double* c, const double* r, const double* Ac, double* x,
             double* y, double* st, double max_scale, double ratio, double* err,
             const int* gate, hipStream_t stream) {
  if (!get_ls_form()) return 0;
  const int G = red_grid(n);
  if (!ls_fits(G)) return 0;
  hipLaunchKernelGGL(k_ls_fused, dim3(G), dim3(kBlock), 0, stream, n, c, r, Ac, x, y, st,
                     max_scale, ratio, w.lsparts, w.lscnt, err, gate);
  GF_CHECK_LAUNCH();
  return 1;
}

int ls_reset(Work& w, hipStream_t stream) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, stream, 2, w.lscnt, 0ull);
  GF_CHECK_LAUNCH();
  return 0;
}

int axpby_dev(long n, const double* a, double asign, const double* x, const double* b, double* y,
              const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby_dev, dim3(red_grid(n)), dim3(kBlock), 0, st, n, a, asign, x, b, y,
                     gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int scale_inv(long n, const double* x, const double* h, double thresh, double* y,
              const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_inv, dim3(red_grid(n)), dim3(kBlock), 0, st, n, x, h, thresh, y,
                     gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int combine(long n, const double* Z, long ldz, const double* y, const int* jcount, bool acc,
            double* x, const int* gate, hipStream_t st) {
  // one element per thread up to 4096 blocks (no cross-thread sum: the grid
  // does not change the result)
  const int grid = (int)std::min<long>(4096, std::max<long>(1, cdiv(n, kBlock)));
  hipLaunchKernelGGL(k_combine, dim3(grid), dim3(kBlock), 0, st, n, Z, ldz, y, jcount,
                     acc ? 1 : 0, x, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int cast_f64_to_f32(long n, const double* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL((k_cast<float, double>), dim3(red_grid(n)), dim3(kBlock), 0, st, n, x, y);
  GF_CHECK_LAUNCH();
  return 0;
}

int cast_f32_to_f64(long n, const float* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL((k_cast<double, float>), dim3(red_grid(n)), dim3(kBlock), 0, st, n, x, y);
  GF_CHECK_LAUNCH();
  return 0;
}

int sub_f64(long n, const double* a, const double* b, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_sub, dim3(red_grid(n)), dim3(kBlock), 0, st, n, a, b, y);
  GF_CHECK_LAUNCH();
  return 0;
}

int minmax_f64(Work& w, long n, const double* E, double* out2, hipStream_t st) {
  hipLaunchKernelGGL(k_minmax, dim3(red_grid(n)), dim3(kBlock), 0, st, n, E, out2, w.partials,
                     w.ticket);
  GF_CHECK_LAUNCH();
  return 0;
}

int count_gt(Work& w, long n, const double* E, const double* thr_src, double scale, double* out,
             hipStream_t st) {
  hipLaunchKernelGGL(k_count_gt, dim3(red_grid(n)), dim3(kBlock), 0, st, n, E, thr_src, scale,
                     w.partials, w.ticket, out);
  GF_CHECK_LAUNCH();
  return 0;
}

// Reduction scratch is uncached device memory, recycled through a process-
// wide free list and never returned to the driver (uncached allocations that
// kernels have read are kept until the process ends).
namespace {
struct Pooled {
  double* partials;
  int* ticket;
  int* flags;
  double* ring;
  double* lsparts;
  unsigned long long* lscnt;
};
struct UncachedPool {
  std::mutex mu;
  std::vector<Pooled> free_list;
} g_pool;
}  // namespace

std::recursive_mutex& device_mutex() {
  static std::recursive_mutex mu;
  return mu;
}

static hipStream_t init_stream(int* rc) {
  static hipStream_t s = nullptr;  // guarded by the device lock
  *rc = 0;
  if (!s && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreateWithFlags failed for the set-up stream");
    *rc = -1;
  }
  return s;
}

int upload_now(void* dst, const void* src, size_t bytes) {
  GF_DEVICE_LOCK();
  int rc = 0;
  hipStream_t s = init_stream(&rc);
  if (rc) return rc;
  GF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  GF_HIP(hipStreamSynchronize(s));
  return 0;
}

int zero_now(void* dst, size_t bytes) {
  GF_DEVICE_LOCK();
  int rc = 0;
  hipStream_t s = init_stream(&rc);
  if (rc) return rc;
  GF_HIP(hipMemsetAsync(dst, 0, bytes, s));
  GF_HIP(hipStreamSynchronize(s));
  return 0;
}

int Work::init() {
  if (partials) return 0;
  GF_DEVICE_LOCK();
  std::lock_guard<std::mutex> lock(g_pool.mu);
  if (!g_pool.free_list.empty()) {
    const Pooled& e = g_pool.free_list.back();
    partials = e.partials;
    ticket = e.ticket;
    flags = e.flags;
    ring = e.ring;
    lsparts = e.lsparts;
    lscnt = e.lscnt;
    sweep_G = -1;
    g_pool.free_list.pop_back();
    // the previous owner may have left the counter mid-launch (an expired wait)
    return zero_now(lscnt, sizeof(unsigned long long) * 2);
  }
  GF_HIP(hipExtMallocWithFlags((void**)&partials, sizeof(double) * 4 * kRedBlocks,
                               hipDeviceMallocUncached));
  GF_HIP(hipExtMallocWithFlags((void**)&ticket, sizeof(int) * 64, hipDeviceMallocUncached));
  GF_TRY(zero_now(ticket, sizeof(int) * 64));
  GF_HIP(hipExtMallocWithFlags((void**)&flags, sizeof(int) * kSweepMaxBlocks,
                               hipDeviceMallocUncached));
  GF_TRY(zero_now(flags, sizeof(int) * kSweepMaxBlocks));
  GF_HIP(hipExtMallocWithFlags((void**)&ring, sizeof(double) * 3 * kSweepMaxBlocks,
                               hipDeviceMallocUncached));
  {
    std::vector<unsigned long long> empty(3 * kSweepMaxBlocks, kRingEmpty);
    GF_TRY(upload_now(ring, empty.data(), sizeof(unsigned long long) * empty.size()));
  }
  GF_HIP(hipExtMallocWithFlags((void**)&lsparts, sizeof(double) * 5 * kRedBlocks,
                               hipDeviceMallocUncached));
  GF_HIP(hipExtMallocWithFlags((void**)&lscnt, sizeof(unsigned long long) * 2,
                               hipDeviceMallocUncached));
  GF_TRY(zero_now(lscnt, sizeof(unsigned long long) * 2));
  sweep_G = -1;
  return 0;
}

void Work::release() {
  GF_DEVICE_LOCK();
  std::lock_guard<std::mutex> lock(g_pool.mu);
  if (partials)
    g_pool.free_list.push_back(Pooled{partials, ticket, flags, ring, lsparts, lscnt});
  partials = nullptr;
  ticket = nullptr;
  flags = nullptr;
  ring = nullptr;
  lsparts = nullptr;
  lscnt = nullptr;
}

}  // namespace gf
