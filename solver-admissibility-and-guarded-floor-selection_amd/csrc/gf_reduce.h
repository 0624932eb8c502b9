// Deterministic reductions shared by the kernel translation units: a fixed
// in-block tree, then the last block to arrive (ticket) folds the per-block
// partials in block order, so a sum is bitwise reproducible for a given n and
// grid and lands in device memory without a host round trip.
#pragma once

#include "gf_internal.h"

namespace gf {

// ---------------------------------------------------------------------------
// block reduction helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum over a 256-thread block; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return r;
}

// NP sums over a 256-thread block at once, each by block_sum's tree (the
// same bits as NP block_sum calls), with one pair of barriers for all of
// them; lds holds 4 * NP doubles.  Results valid in thread 0 (v[p]).
template <int NP>
__device__ __forceinline__ void block_sum_n(double (&v)[NP], double* lds) {
#pragma unroll
  for (int p = 0; p < NP; ++p) v[p] = wave_sum(v[p]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int p = 0; p < NP; ++p) lds[4 * p + wid] = v[p];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int p = 0; p < NP; ++p)
      v[p] = (lds[4 * p] + lds[4 * p + 1]) + (lds[4 * p + 2] + lds[4 * p + 3]);
  }
  __syncthreads();
}

// One thread's share of sum_i x_i y_i over a grid-stride loop: four
// strides' loads issue before the first is used, and the sum is taken in the
// same order as the plain stride loop (identical bits).
template <typename T>
__device__ __forceinline__ double dot_thread(long n, const T* __restrict__ x,
                                             const T* __restrict__ y) {
  double a = 0.0;
  const long st = (long)gridDim.x * blockDim.x;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + 3 * st < n; i += 4 * st) {
    T xv[4], yv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xv[u] = x[i + u * st];
      yv[u] = y[i + u * st];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) a += (double)xv[u] * (double)yv[u];
  }
  for (; i < n; i += st) a += (double)x[i] * (double)y[i];
  return a;
}

// Sum of nb block partials in block order over one 256-thread block (each
// thread takes i, i + 256, ..., then the block tree); result in thread 0.
__device__ __forceinline__ double fold_thread_first(int nb, const double* partials, double* lds) {
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) a += partials[i];
  return block_sum(a, lds);
}

// Fold NP partial sums of every block into out[0..NP) (last block only).
// Returns true in the block that performed the fold.
template <int NP>
__device__ __forceinline__ bool finish_reduce(const double (&v)[NP], double* partials,
                                              int* ticket, double* out, int post) {
  __shared__ double lds[4 * NP];
  __shared__ bool last;
  double bs[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) bs[p] = v[p];
  block_sum_n<NP>(bs, lds);
  // partials and ticket live in uncached device memory, so a partial is
  // visible to every XCD once its store is acknowledged.  Waiting for that
  // acknowledgement (vmcnt 0) before the ticket increment is the part of an
  // agent-scope release that matters here; the L2 write-back that such a
  // release would add is not needed for memory that L2 does not hold.
  if (threadIdx.x == 0) {
#pragma unroll
    for (int p = 0; p < NP; ++p)
      __hip_atomic_store(&partials[p * kRedBlocks + blockIdx.x], bs[p], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    store_ack_wait();
    const int t = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == (int)gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  const int nb = gridDim.x;
  double a[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    a[p] = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
      a[p] += __hip_atomic_load(&partials[p * kRedBlocks + i], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum_n<NP>(a, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int p = 0; p < NP; ++p) out[p] = post == 1 ? sqrt(a[p]) : a[p];
  }
  if (threadIdx.x == 0) __hip_atomic_store(ticket, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Residual line-search step (multigrid_v4.py:867-931) from the statistics
// st[0] = c.c, st[1] = r.r, st[2] = Ac.Ac, st[3] = r.Ac: st[4] = alpha =
// clip((r.Ac)/(Ac.Ac), 0, max_scale), st[5] = 1 when the correction is
// rejected outright (a zero norm), else 0.  One thread.
__device__ __forceinline__ void ls_scale(double* st, double max_scale) {
  const double cn = sqrt(st[0]), rn = sqrt(st[1]), den = st[2];
  if (cn <= 1e-300 || rn <= 1e-300 || den <= 1e-300) {
    st[4] = 0.0;
    st[5] = 1.0;
    return;
  }
  const double raw = st[3] / den;
  st[4] = fmin(max_scale, fmax(0.0, raw));
  st[5] = 0.0;
}

// Reduction grid: ~4096 elements per block, between 16 and kRedBlocks blocks
// (measured on MI355X: fewer, fuller blocks shorten the ticket tail).
inline int red_grid(long n) {
  int g = cdiv(n, 4096);
  return g < 16 ? 16 : (g > kRedBlocks ? kRedBlocks : g);
}

}  // namespace gf
