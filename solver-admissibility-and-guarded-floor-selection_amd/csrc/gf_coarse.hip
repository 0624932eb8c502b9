// Coarse levels: element-wise level-1 Galerkin assembly, Galerkin R(KP)
// products, block-stencil operators, trilinear transfers, the dense coarsest
// factorisation and the fixed-count coarsest PCG.
//
// Every coarse operator is a 27-point stencil of 3x3 blocks on its node grid
// (SoA layout, see gf_common.h).  Entries whose row or column DOF is fixed are
// stored as exact zeros, which reproduces the reference's free-restricted
// CSR matrices value for value.
#include "gf_reduce.h"

#include <atomic>
#include <vector>

namespace gf {

__host__ __device__ constexpr int lidx_c(int px, int py, int pz) {
  return pz * 4 + (py ? (px ? 2 : 3) : (px ? 1 : 0));
}

__device__ __forceinline__ void node_xyz(const Grid& g, int n, int& ix, int& iy, int& iz) {
  iz = n % g.NZ;
  const int t = n / g.NZ;
  iy = t % g.NY;
  ix = t / g.NY;
}

// ---------------------------------------------------------------------------
// level-1 assembly: one thread per (coarse node I, stencil slot k)
//   K1[I, I+d_k] = sum_{coarse elements C containing both, ascending}
//                  sum_{stu in s-major order} E[fine(C,stu)] Q_stu[(I in C), (J in C)]
// fine(C,stu) = min(2C+stu, nel_f - 1) per axis (odd fine dims drop the last
// fine layer, exactly as the reference's element-wise construction).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_level1(Grid gc, DofMap mapc, Grid gf,
                                                   const double* __restrict__ Q,
                                                   const double* __restrict__ E,
                                                   double* __restrict__ A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (n >= gc.n_nodes) return;
  const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
  int ix, iy, iz;
  node_xyz(gc, n, ix, iy, iz);
  const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
  const bool jin = (jx >= 0) & (jx < gc.NX) & (jy >= 0) & (jy < gc.NY) & (jz >= 0) &
                   (jz < gc.NZ);
  const int m = (jx * gc.NY + jy) * gc.NZ + jz;
  int orow[3], ocol[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    orow[q] = mapc(3 * n + q);
    ocol[q] = jin ? mapc(3 * m + q) : -1;
  }
  double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (jin) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          // element C = (ix-1+a, iy-1+b, iz-1+c); J's local position 1+d-a must be 0/1
          const int pjx = 1 + dx - a, pjy = 1 + dy - b, pjz = 1 + dz - c;
          if (pjx < 0 || pjx > 1 || pjy < 0 || pjy > 1 || pjz < 0 || pjz > 1) continue;
          const int cx = ix - 1 + a, cy = iy - 1 + b, cz = iz - 1 + c;
          if (cx < 0 || cx >= gc.nx || cy < 0 || cy >= gc.ny || cz < 0 || cz >= gc.nz) continue;
          const int li = lidx_c(1 - a, 1 - b, 1 - c);
          const int lj = lidx_c(pjx, pjy, pjz);
          double Ef[8];
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            const int su = s >> 2, tu = (s >> 1) & 1, uu = s & 1;
            const int fx = min(2 * cx + su, gf.nx - 1);
            const int fy = min(2 * cy + tu, gf.ny - 1);
            const int fz = min(2 * cz + uu, gf.nz - 1);
            Ef[s] = E[(fx * gf.ny + fy) * gf.nz + fz];
          }
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              double raw = 0.0;
#pragma unroll
              for (int s = 0; s < 8; ++s)
                raw += Ef[s] * Q[s * 576 + (3 * li + r) * 24 + 3 * lj + q];
              blk[r * 3 + q] += raw;
            }
        }
  }
  const long nn = gc.n_nodes;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const bool keep = (orow[r] >= 0) & (ocol[q] >= 0);
      A[(long)(k * 9 + r * 3 + q) * nn + n] = keep ? blk[r * 3 + q] : 0.0;
    }
}

int level1_assemble(const Grid& gc, const DofMap& mapc, const Grid& gf, const double* Q,
                    const double* E, double* A, hipStream_t st) {
  hipLaunchKernelGGL(k_level1, dim3(cdiv(gc.n_nodes, kBlock), 27), dim3(kBlock), 0, st, gc, mapc,
                     gf, Q, E, A);
  GF_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// block-stencil operator with the shared epilogues
// ---------------------------------------------------------------------------
// One thread per (node, stencil slot): a block covers 8 nodes x 32 lanes
// (27 slots used); slot products Sum_q A[k][r][q] x[J_k, q] meet in LDS and
// lane r < 3 of each node adds them in slot order before the epilogue.  The
// small coarse grids get 27 times the parallelism of a thread-per-node map.
template <typename T, int EPI, bool OFF>
__global__ __launch_bounds__(kBlock) void k_stencil(Grid g, DofMap map, const T* __restrict__ A,
                                                    const T* __restrict__ x, T* __restrict__ y,
                                                    EpiArgs<T> ep, const int* gate) {
  if (gate && *gate == 0) return;
  __shared__ T part[8][27][3];
  const int k = threadIdx.x & 31, ln = threadIdx.x >> 5;
  const int n = xcd_tile(blockIdx.x, gridDim.x) * 8 + ln;
  const bool valid = n < g.n_nodes;
  if (valid && k < 27) {
    int ix, iy, iz;
    node_xyz(g, n, ix, iy, iz);
    const int jx = ix + k / 9 - 1, jy = iy + (k / 3) % 3 - 1, jz = iz + k % 3 - 1;
    const bool in = (jx >= 0) & (jx < g.NX) & (jy >= 0) & (jy < g.NY) & (jz >= 0) & (jz < g.NZ);
    const int m = (jx * g.NY + jy) * g.NZ + jz;
    const long nn = g.n_nodes;
    T s0 = 0, s1 = 0, s2 = 0;
    // masked loads read 0 (branch-free); a zero x contributes exact zeros, as
    // a skipped term does
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T xv = map_load<OFF>(map, x, 3 * m + q, in);
      s0 += A[(k * 9 + 0 * 3 + q) * nn + n] * xv;
      s1 += A[(k * 9 + 1 * 3 + q) * nn + n] * xv;
      s2 += A[(k * 9 + 2 * 3 + q) * nn + n] * xv;
    }
    part[ln][k][0] = s0;
    part[ln][k][1] = s1;
    part[ln][k][2] = s2;
  }
  __syncthreads();
  if (!valid || k >= 3) return;
  const int o = map_idx<OFF>(map, 3 * n + k);
  if (o < 0) return;
  T acc = 0;
#pragma unroll
  for (int kk = 0; kk < 27; ++kk) acc += part[ln][kk][k];
  epilogue1<T, EPI>(o, acc, x, y, ep);
}

// Node-per-lane form of k_stencil with coalesced coefficient reads: a block
// is NPB consecutive nodes x NW waves; lane l of wave w takes node l % NPB
// and evaluates the slots k = w*S + l/NPB, + NW*S, ... (S = 64/NPB sub-slots
// per wave), so each coefficient plane A[(k*9 + r*3 + q) * n_nodes + n] is
// read as runs of NPB consecutive values, storing the slot products
// t_k = sum_q A[k][r][q] x[J_k, q] in LDS.  Lanes l < NPB of wave 0 then fold
// them in slot order 0..26 exactly as k_stencil does, so every form and
// every NPB gives the same bits.  Smaller NPB spreads a small level over
// more workgroups.  Wave 0 loads its epilogue operands (b, D^-1, x, dv of
// its node's three DOFs) before its slot loads, so that round trip overlaps
// the coefficient stream instead of following the fold.
template <typename T, int EPI, bool OFF, int NW, int NPB>
__global__ __launch_bounds__(64 * NW) void k_stencil_nodes(Grid g, DofMap map,
                                                           const T* __restrict__ A,
                                                           const T* __restrict__ x,
                                                           T* __restrict__ y, EpiArgs<T> ep,
                                                           const int* gate) {
  constexpr int S = 64 / NPB;                   // sub-slots per wave and round
  constexpr int R = (27 + NW * S - 1) / (NW * S);  // slot rounds per wave
  const bool open = gate_open(gate);  // tested after the slot loads
  __shared__ T tk[27][3][NPB];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ln = lane % NPB, sub = lane / NPB;
  const int n = xcd_tile(blockIdx.x, gridDim.x) * NPB + ln;
  const bool valid = n < g.n_nodes;
  const int nn_ = valid ? n : 0;
  int ix, iy, iz;
  node_xyz(g, nn_, ix, iy, iz);
  const long nn = g.n_nodes;
  const bool folder = (wv == 0) & (sub == 0);
  // epilogue operands of the folding lane's node (read before any slot load)
  int o[3] = {-1, -1, -1};
  T eb[3] = {0, 0, 0}, ed[3] = {0, 0, 0}, ex[3] = {0, 0, 0}, ev[3] = {0, 0, 0};
  if (folder) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      o[r] = valid ? map_idx<OFF>(map, 3 * n + r) : -1;
      const int i = o[r] >= 0 ? o[r] : 0;
      eb[r] = EPI != EPI_PLAIN ? ep.b[i] : T(0);
      ed[r] = epi_reads_x<EPI>() ? ep.dinv[i] : T(0);
      ex[r] = epi_reads_x<EPI>() ? x[i] : T(0);
      ev[r] = EPI == EPI_CHEB ? ep.dv[i] : T(0);
    }
  }
  // rounds with a compile-time trip count (slots past 26 masked rather than
  // branched over) let a lane's loads be issued together
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int kk = wv * S + sub + NW * S * t;
    const bool kin = kk < 27;
    const int k = kin ? kk : 0;
    const int jx = ix + k / 9 - 1, jy = iy + (k / 3) % 3 - 1, jz = iz + k % 3 - 1;
    const bool in = kin & valid & (jx >= 0) & (jx < g.NX) & (jy >= 0) & (jy < g.NY) &
                    (jz >= 0) & (jz < g.NZ);
    const int m = (jx * g.NY + jy) * g.NZ + jz;
    T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T xv = map_load<OFF>(map, x, 3 * m + q, in);
      s0 += A[(k * 9 + 0 * 3 + q) * nn + nn_] * xv;
      s1 += A[(k * 9 + 1 * 3 + q) * nn + nn_] * xv;
      s2 += A[(k * 9 + 2 * 3 + q) * nn + nn_] * xv;
    }
    if (kin) {
      tk[k][0][ln] = s0;
      tk[k][1][ln] = s1;
      tk[k][2][ln] = s2;
    }
  }
  __syncthreads();
  if (!open || !folder || !valid) return;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    if (o[r] < 0) continue;
    T acc = 0;
#pragma unroll
    for (int k = 0; k < 27; ++k) acc += tk[k][r][ln];
    epi_store<T, EPI>(o[r], acc, eb[r], ed[r], ev[r], ex[r], y, ep);
  }
}

// Nodes per workgroup of k_stencil_nodes: 64 when the level fills the
// device with 64-node blocks, else 32 or 16 (more, smaller workgroups).
static std::atomic<int> g_stencil_npb{0};  // 0 = by level size
void set_stencil_npb(int npb) { g_stencil_npb = (npb == 16 || npb == 32 || npb == 64) ? npb : 0; }
static int stencil_npb(const Grid& g) {
  const int f = g_stencil_npb.load();
  if (f) return f;
  if (cdiv(g.n_nodes, 64) >= 1024) return 64;
  if (cdiv(g.n_nodes, 32) >= 512) return 32;
  return 16;
}

template <typename T, bool OFF, bool NODES>
static void launch_stencil(const Grid& g, const DofMap& map, const T* A, const T* x, T* y, int epi,
                           const EpiArgs<T>& ep, const int* gate, hipStream_t st) {
  const int npb = NODES ? stencil_npb(g) : 8;
  const dim3 grid(cdiv(g.n_nodes, npb)), block(kBlock);
  const dim3 block8(8 * 64);
#define GF_STENCIL_LAUNCH(E)                                                                   \
  if (!NODES)                                                                                  \
    hipLaunchKernelGGL((k_stencil<T, E, OFF>), grid, block, 0, st, g, map, A, x, y, ep, gate); \
  else if (npb == 64)                                                                          \
    hipLaunchKernelGGL((k_stencil_nodes<T, E, OFF, 8, 64>), grid, block8, 0, st, g, map, A, x, \
                       y, ep, gate);                                                           \
  else if (npb == 32)                                                                          \
    hipLaunchKernelGGL((k_stencil_nodes<T, E, OFF, 8, 32>), grid, block8, 0, st, g, map, A, x, \
                       y, ep, gate);                                                           \
  else                                                                                         \
    hipLaunchKernelGGL((k_stencil_nodes<T, E, OFF, 8, 16>), grid, block8, 0, st, g, map, A, x, \
                       y, ep, gate);
  switch (epi) {
    case EPI_PLAIN: GF_STENCIL_LAUNCH(EPI_PLAIN) break;
    case EPI_RESID: GF_STENCIL_LAUNCH(EPI_RESID) break;
    case EPI_CHEB0: GF_STENCIL_LAUNCH(EPI_CHEB0) break;
    case EPI_CHEB: GF_STENCIL_LAUNCH(EPI_CHEB) break;
    default: GF_STENCIL_LAUNCH(EPI_JACOBI) break;
  }
#undef GF_STENCIL_LAUNCH
}

// Coarse-operator launch form: 0 = slot-per-lane (k_stencil), 1 = node-per-lane
// (k_stencil_nodes, default).  Bitwise-identical results.
static std::atomic<int> g_stencil_form{1};
void set_stencil_form(int form) { g_stencil_form = form ? 1 : 0; }

template <typename T>
int stencil_matvec(const Grid& g, const DofMap& map, const T* A, const T* x, T* y, int epi,
                   const EpiArgs<T>& ep, const int* gate, hipStream_t st) {
  if (epi < EPI_PLAIN || epi > EPI_JACOBI) {
    set_error("stencil_matvec: bad epilogue");
    return -2;
  }
  const bool nodes = g_stencil_form.load() != 0;
  if (map.off >= 0) {
    if (nodes) launch_stencil<T, true, true>(g, map, A, x, y, epi, ep, gate, st);
    else launch_stencil<T, true, false>(g, map, A, x, y, epi, ep, gate, st);
  } else {
    if (nodes) launch_stencil<T, false, true>(g, map, A, x, y, epi, ep, gate, st);
    else launch_stencil<T, false, false>(g, map, A, x, y, epi, ep, gate, st);
  }
  GF_CHECK_LAUNCH();
  return 0;
}
template int stencil_matvec<double>(const Grid&, const DofMap&, const double*, const double*,
                                    double*, int, const EpiArgs<double>&, const int*,
                                    hipStream_t);
template int stencil_matvec<float>(const Grid&, const DofMap&, const float*, const float*,
                                   float*, int, const EpiArgs<float>&, const int*, hipStream_t);

__global__ __launch_bounds__(kBlock) void k_stencil_diag(Grid g, DofMap map,
                                                         const double* __restrict__ A,
                                                         double* __restrict__ d) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const long nn = g.n_nodes;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int f = map(3 * n + q);
    if (f >= 0) d[f] = A[(13 * 9 + q * 4) * nn + n];
  }
}

int stencil_diag(const Grid& g, const DofMap& map, const double* A, double* diag,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_stencil_diag, dim3(cdiv(g.n_nodes, kBlock)), dim3(kBlock), 0, st, g, map,
                     A, diag);
  GF_CHECK_LAUNCH();
  return 0;
}

__global__ __launch_bounds__(kBlock) void k_to_f32(long n, const double* __restrict__ a,
                                                   float* __restrict__ b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

int f64_to_f32_block(long n, const double* A, float* A32, hipStream_t st) {
  int grid = cdiv(n, kBlock);
  if (grid > 8192) grid = 8192;
  hipLaunchKernelGGL(k_to_f32, dim3(grid), dim3(kBlock), 0, st, n, A, A32);
  GF_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// trilinear transfers (row-normalised scalar weights, applied per component)
// ---------------------------------------------------------------------------
// Row-normalised trilinear weight of coarse parent I for fine index i >= 0:
// a parent's 1-D weight (1 for the one parent of an even index, 1/2 for each
// of the two of an odd one) divided by the row's sum, in closed form.  An
// odd index past the last coarse node keeps one parent, 1/2 / 1/2 = 1.
// Every quotient is exact (1, 1/2 or 0), so no division is needed.
__device__ __forceinline__ double pweight(int i, int I, int ncn) {
  const int h = i >> 1;
  const bool lo = h < ncn, hi = h + 1 < ncn;
  if ((i & 1) == 0) return (I == h && lo) ? 1.0 : 0.0;
  if (I == h && lo) return hi ? 0.5 : 1.0;
  return (I == h + 1 && hi) ? 0.5 : 0.0;
}

// Parents of fine index i >= 0 along one axis with their normalised weights
// (pweight of each), in ascending index order; the count is 0, 1 or 2.
__device__ __forceinline__ int parents_norm(int i, int ncn, int* p, double* w) {
  const int h = i >> 1;
  const bool lo = h < ncn, hi = h + 1 < ncn;
  p[0] = h;
  p[1] = h + 1;
  if ((i & 1) == 0) {
    w[0] = 1.0;
    w[1] = 0.0;
    return lo ? 1 : 0;
  }
  w[0] = hi ? 0.5 : 1.0;
  w[1] = 0.5;
  return lo ? (hi ? 2 : 1) : 0;
}

// Prolongation: fine node n gathers its up to 2x2x2 coarse parents.  The
// map modes of both levels are template parameters (OFF: free index = DOF -
// off), so an offset-mapped level costs no table load and no branch.
template <bool OFFF, bool OFFC>
__global__ __launch_bounds__(kBlock) void k_prolong(Grid gf, DofMap mapf, Grid gc, DofMap mapc,
                                                    const double* __restrict__ ec,
                                                    double* __restrict__ yf, const int* gate) {
  const bool open = gate_open(gate);  // tested after the parent loads
  const int n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= gf.n_nodes) return;
  const int o0 = map_idx<OFFF>(mapf, 3 * n), o1 = map_idx<OFFF>(mapf, 3 * n + 1),
            o2 = map_idx<OFFF>(mapf, 3 * n + 2);
  if ((o0 & o1 & o2) < 0) return;
  int ix, iy, iz;
  node_xyz(gf, n, ix, iy, iz);
  int px[2], py[2], pz[2];
  double wx[2], wy[2], wz[2];
  const int cx = parents_norm(ix, gc.NX, px, wx);
  const int cy = parents_norm(iy, gc.NY, py, wy);
  const int cz = parents_norm(iz, gc.NZ, pz, wz);
  double acc[3] = {0.0, 0.0, 0.0};
  // all 2x2x2 parent slots with a fixed trip count (slots past a parent
  // count are masked and read index 0), so the up to 24 loads issue
  // together; terms are added in (a, b, c) order.  Every weight is a power
  // of two, so ((wx wy) wz) is the exact normalised product.
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const bool in = (a < cx) & (b < cy) & (c < cz);
        const double w = (wx[a] * wy[b]) * wz[c];
        const int m = in ? (px[a] * gc.NY + py[b]) * gc.NZ + pz[c] : 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int f = in ? map_idx<OFFC>(mapc, 3 * m + q) : -1;
          const double v = ec[f >= 0 ? (unsigned)f : 0u];
          if (f >= 0) acc[q] += w * v;
        }
      }
  if (!open) return;
  if (o0 >= 0) yf[o0] = acc[0];
  if (o1 >= 0) yf[o1] = acc[1];
  if (o2 >= 0) yf[o2] = acc[2];
}

// Restriction R = P^T: coarse node I gathers its (up to 27) fine children in
// ascending fine-node order.  64 coarse nodes x 4 child groups per block,
// partial sums combined in LDS in group order.  Map modes as for k_prolong.
template <bool OFFF, bool OFFC>
__global__ __launch_bounds__(kBlock) void k_restrict(Grid gf, DofMap mapf, Grid gc, DofMap mapc,
                                                     const double* __restrict__ rf,
                                                     double* __restrict__ yc, ZeroStart zs,
                                                     const int* gate) {
  const bool open = gate_open(gate);  // tested after the child loads
  __shared__ double part[3][3][64];
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int n = xcd_tile(blockIdx.x, gridDim.x) * 64 + lane;
  const bool valid = n < gc.n_nodes;
  double acc[3] = {0.0, 0.0, 0.0};
  int o[3] = {-1, -1, -1};
  double dinv[3] = {0.0, 0.0, 0.0};
  if (valid) {
    o[0] = map_idx<OFFC>(mapc, 3 * n);
    o[1] = map_idx<OFFC>(mapc, 3 * n + 1);
    o[2] = map_idx<OFFC>(mapc, 3 * n + 2);
    // the zero start's diagonal is loaded before the children, so its
    // latency passes under theirs
    if (zs.mode != 0 && grp == 0) {
#pragma unroll
      for (int q = 0; q < 3; ++q) dinv[q] = zs.dinv[o[q] >= 0 ? (unsigned)o[q] : 0u];
    }
    int ix, iy, iz;
    node_xyz(gc, n, ix, iy, iz);
    double wx[3], wy[3], wz[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int fx = 2 * ix + a - 1, fy = 2 * iy + a - 1, fz = 2 * iz + a - 1;
      wx[a] = (fx >= 0 && fx < gf.NX) ? pweight(fx, ix, gc.NX) : 0.0;
      wy[a] = (fy >= 0 && fy < gf.NY) ? pweight(fy, iy, gc.NY) : 0.0;
      wz[a] = (fz >= 0 && fz < gf.NZ) ? pweight(fz, iz, gc.NZ) : 0.0;
    }
    // the group's children c27 = grp, grp+4, ..., grp+24 (the last for
    // grp < 3 only) with a fixed trip count, absent children masked and
    // reading index 0, so the loads issue together; terms are added in
    // ascending child order
#pragma unroll
    for (int t = 0; t < 7; ++t) {
      const bool kin = (t < 6) | (grp < 3);
      const int c27 = kin ? grp + 4 * t : 0;
      const int a = c27 / 9, b = (c27 / 3) % 3, c = c27 % 3;
      const bool in = kin & (wx[a] != 0.0) & (wy[b] != 0.0) & (wz[c] != 0.0);
      const double w = (wx[a] * wy[b]) * wz[c];
      const int m =
          in ? ((2 * ix + a - 1) * gf.NY + (2 * iy + b - 1)) * gf.NZ + (2 * iz + c - 1) : 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int f = in ? map_idx<OFFF>(mapf, 3 * m + q) : -1;
        const double v = rf[f >= 0 ? (unsigned)f : 0u];
        if (f >= 0) acc[q] += w * v;
      }
    }
  }
  if (grp > 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) part[grp - 1][q][lane] = acc[q];
  }
  __syncthreads();
  if (!open || grp != 0 || !valid) return;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double v = ((acc[q] + part[0][q][lane]) + part[1][q][lane]) + part[2][q][lane];
    const int i = o[q];
    if (i >= 0) {
      yc[i] = v;
      if (zs.mode == 1) {
        const double z = rn_mul(dinv[q], v);
        zs.dv[i] = z;
        zs.xo[i] = rn_add(0.0, rn_mul(zs.a, z));
      } else if (zs.mode == 2) {
        zs.xo[i] = rn_add(0.0, rn_mul(rn_mul(zs.a, dinv[q]), v));
      }
    }
  }
}

#define GF_TRANSFER(KERNEL, GRID, ...)                                                     \
  do {                                                                                     \
    const bool of_ = mapf.off >= 0, oc_ = mapc.off >= 0;                                   \
    if (of_ && oc_)                                                                        \
      hipLaunchKernelGGL((KERNEL<true, true>), GRID, dim3(kBlock), 0, st, __VA_ARGS__);    \
    else if (of_)                                                                          \
      hipLaunchKernelGGL((KERNEL<true, false>), GRID, dim3(kBlock), 0, st, __VA_ARGS__);   \
    else if (oc_)                                                                          \
      hipLaunchKernelGGL((KERNEL<false, true>), GRID, dim3(kBlock), 0, st, __VA_ARGS__);   \
    else                                                                                   \
      hipLaunchKernelGGL((KERNEL<false, false>), GRID, dim3(kBlock), 0, st, __VA_ARGS__);  \
  } while (0)

int prolong(const Grid& gf, const DofMap& mapf, const Grid& gc, const DofMap& mapc,
            const double* ec, double* yf, const int* gate, hipStream_t st) {
  GF_TRANSFER(k_prolong, dim3(cdiv(gf.n_nodes, kBlock)), gf, mapf, gc, mapc, ec, yf, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int restrict_(const Grid& gf, const DofMap& mapf, const Grid& gc, const DofMap& mapc,
              const double* rf, double* yc, const int* gate, hipStream_t st,
              const ZeroStart* zs) {
  const ZeroStart z = zs ? *zs : ZeroStart{};
  GF_TRANSFER(k_restrict, dim3(cdiv(gc.n_nodes, 64)), gf, mapf, gc, mapc, rf, yc, z, gate);
  GF_CHECK_LAUNCH();
  return 0;
}
#undef GF_TRANSFER

// ---------------------------------------------------------------------------
// Galerkin product K_c = R (K_f P), in two stages with the association of the
// reference's sparse products (K_f P first, then R times that; every sum runs
// in ascending column/row order).
//   stage 1: KP[a, slot q] (3x3) = sum_b K_f[a,b] P[b, J_q],
//            J_q = ceil((a-2)/2) + q per axis, q in {0,1,2}^3
//   stage 2: K_c[I, I+d_k] = sum_{a child of I, ascending} P[a,I] KP[a, J]
// ---------------------------------------------------------------------------
__device__ __forceinline__ int jbase(int a) { return (a - 1) >> 1; }  // ceil((a-2)/2)

__global__ __launch_bounds__(kBlock) void k_rap_kp(Grid gf, DofMap mapf, Grid gc, DofMap mapc,
                                                   const double* __restrict__ Af,
                                                   double* __restrict__ KP) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (n >= gf.n_nodes) return;
  int ax, ay, az;
  node_xyz(gf, n, ax, ay, az);
  const int Jx = jbase(ax) + q / 9, Jy = jbase(ay) + (q / 3) % 3, Jz = jbase(az) + q % 3;
  const long nnf = gf.n_nodes;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const bool Jin = (Jx >= 0) & (Jx < gc.NX) & (Jy >= 0) & (Jy < gc.NY) & (Jz >= 0) &
                   (Jz < gc.NZ);
  int cfree[3] = {-1, -1, -1};
  if (Jin) {
    const int mJ = (Jx * gc.NY + Jy) * gc.NZ + Jz;
#pragma unroll
    for (int c = 0; c < 3; ++c) cfree[c] = mapc(3 * mJ + c);
    for (int k = 0; k < 27; ++k) {
      const int bx = ax + k / 9 - 1, by = ay + (k / 3) % 3 - 1, bz = az + k % 3 - 1;
      if (bx < 0 || bx >= gf.NX || by < 0 || by >= gf.NY || bz < 0 || bz >= gf.NZ) continue;
      const double w = (pweight(bx, Jx, gc.NX) * pweight(by, Jy, gc.NY)) * pweight(bz, Jz, gc.NZ);
      if (w == 0.0) continue;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          acc[r * 3 + c] += Af[(long)(k * 9 + r * 3 + c) * nnf + n] * w;
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      KP[(long)(q * 9 + r * 3 + c) * nnf + n] = (cfree[c] >= 0) ? acc[r * 3 + c] : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_rap_r(Grid gf, DofMap mapf, Grid gc, DofMap mapc,
                                                  const double* __restrict__ KP,
                                                  double* __restrict__ Ac) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (n >= gc.n_nodes) return;
  int ix, iy, iz;
  node_xyz(gc, n, ix, iy, iz);
  const int Jx = ix + k / 9 - 1, Jy = iy + (k / 3) % 3 - 1, Jz = iz + k % 3 - 1;
  const bool Jin = (Jx >= 0) & (Jx < gc.NX) & (Jy >= 0) & (Jy < gc.NY) & (Jz >= 0) &
                   (Jz < gc.NZ);
  const long nnf = gf.n_nodes, nnc = gc.n_nodes;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (Jin) {
    for (int a = -1; a <= 1; ++a) {
      const int fx = 2 * ix + a;
      if (fx < 0 || fx >= gf.NX) continue;
      const double wxv = pweight(fx, ix, gc.NX);
      if (wxv == 0.0) continue;
      const int qx = Jx - jbase(fx);
      if (qx < 0 || qx > 2) continue;
      for (int b = -1; b <= 1; ++b) {
        const int fy = 2 * iy + b;
        if (fy < 0 || fy >= gf.NY) continue;
        const double wyv = pweight(fy, iy, gc.NY);
        if (wyv == 0.0) continue;
        const int qy = Jy - jbase(fy);
        if (qy < 0 || qy > 2) continue;
        for (int c = -1; c <= 1; ++c) {
          const int fz = 2 * iz + c;
          if (fz < 0 || fz >= gf.NZ) continue;
          const double wzv = pweight(fz, iz, gc.NZ);
          if (wzv == 0.0) continue;
          const int qz = Jz - jbase(fz);
          if (qz < 0 || qz > 2) continue;
          const double w = (wxv * wyv) * wzv;
          const int a_node = (fx * gf.NY + fy) * gf.NZ + fz;
          const int q = qx * 9 + qy * 3 + qz;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            if (mapf(3 * a_node + r) < 0) continue;
#pragma unroll
            for (int c = 0; c < 3; ++c)
              acc[r * 3 + c] += w * KP[(long)(q * 9 + r * 3 + c) * nnf + a_node];
          }
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const bool rf = mapc(3 * n + r) >= 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      Ac[(long)(k * 9 + r * 3 + c) * nnc + n] = rf ? acc[r * 3 + c] : 0.0;
  }
}

int rap(const Grid& gf, const DofMap& mapf, const Grid& gc, const DofMap& mapc, const double* Af,
        double* KP, double* Ac, hipStream_t st) {
  hipLaunchKernelGGL(k_rap_kp, dim3(cdiv(gf.n_nodes, kBlock), 27), dim3(kBlock), 0, st, gf, mapf,
                     gc, mapc, Af, KP);
  GF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_rap_r, dim3(cdiv(gc.n_nodes, kBlock), 27), dim3(kBlock), 0, st, gf, mapf,
                     gc, mapc, KP, Ac);
  GF_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// dense coarsest: gather, Cholesky (one work-group), two triangular solves
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_to_dense(Grid g, DofMap map,
                                                     const double* __restrict__ A, int nf,
                                                     double* __restrict__ D) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (n >= g.n_nodes) return;
  int ix, iy, iz;
  node_xyz(g, n, ix, iy, iz);
  const int jx = ix + k / 9 - 1, jy = iy + (k / 3) % 3 - 1, jz = iz + k % 3 - 1;
  if (jx < 0 || jx >= g.NX || jy < 0 || jy >= g.NY || jz < 0 || jz >= g.NZ) return;
  const int m = (jx * g.NY + jy) * g.NZ + jz;
  const long nn = g.n_nodes;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int fr = map(3 * n + r);
    if (fr < 0) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int fc = map(3 * m + c);
      if (fc < 0) continue;
      D[(long)fr * nf + fc] = A[(long)(k * 9 + r * 3 + c) * nn + n];
    }
  }
}

int stencil_to_dense(const Grid& g, const DofMap& map, const double* A, int n_free, double* D,
                     hipStream_t st) {
  GF_HIP(hipMemsetAsync(D, 0, sizeof(double) * (size_t)n_free * n_free, st));
  hipLaunchKernelGGL(k_to_dense, dim3(cdiv(g.n_nodes, kBlock), 27), dim3(kBlock), 0, st, g, map,
                     A, n_free, D);
  GF_CHECK_LAUNCH();
  return 0;
}

// Blocked right-looking Cholesky (lower factor, row-major) over
// 64-wide panels, many work-groups per panel.  Every entry receives exactly
// the updates of the unblocked right-looking algorithm, in the same order:
// for pivot k ascending, D[i][j] -= D[i][k] * D[j][k] (k < j <= i), then the
// column below a pivot is divided by it.  Per panel [kb, kb+m):
//   k_potrf_diag  factors the m x m diagonal block in LDS (one work-group);
//   k_trsm_rows   solves the panel rows below it, one row per thread, by
//                 forward substitution against the diagonal block (the
//                 updates from this panel's pivots, ascending, then the
//                 division);
//   k_syrk_tiles  applies this panel's pivots, ascending, to the trailing
//                 lower triangle in 64x64 tiles.
// A non-positive or non-finite pivot sets *fail and every later kernel of
// the factorisation returns at once.
constexpr int kCB = 64;

__global__ __launch_bounds__(256) void k_potrf_diag(int n, double* __restrict__ D, int kb,
                                                    int* fail) {
  __shared__ double T[kCB][kCB + 1];
  __shared__ double piv;
  __shared__ int bad;
  if (*fail) return;
  const int m = min(kCB, n - kb);
  const int tid = threadIdx.x;
  for (int e = tid; e < m * m; e += blockDim.x) {
    const int r = e / m, c = e % m;
    T[r][c] = c <= r ? D[(long)(kb + r) * n + kb + c] : 0.0;
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    if (tid == 0) {
      const double d = T[k][k];
      if (!(d > 0.0) || !isfinite(d)) bad = 1;
      piv = sqrt(d);
      T[k][k] = piv;
    }
    __syncthreads();
    if (bad) break;
    const double p = piv;
    for (int i = k + 1 + tid; i < m; i += blockDim.x) T[i][k] /= p;
    __syncthreads();
    const int mm = m - k - 1;
    for (int e = tid; e < mm * mm; e += blockDim.x) {
      const int i = k + 1 + e / mm, j = k + 1 + e % mm;
      if (j <= i) T[i][j] -= T[i][k] * T[j][k];
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0) *fail = 1;
    return;
  }
  for (int e = tid; e < m * m; e += blockDim.x) {
    const int r = e / m, c = e % m;
    if (c <= r) D[(long)(kb + r) * n + kb + c] = T[r][c];
  }
}

__global__ __launch_bounds__(kCB) void k_trsm_rows(int n, double* __restrict__ D, int kb,
                                                   const int* fail) {
  __shared__ double L11[kCB][kCB + 1];
  if (*fail) return;
  const int m = min(kCB, n - kb);
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int r = e / m, c = e % m;
    L11[r][c] = c <= r ? D[(long)(kb + r) * n + kb + c] : 0.0;
  }
  __syncthreads();
  const int i = kb + m + blockIdx.x * kCB + threadIdx.x;
  if (i >= n) return;
  double* row = D + (long)i * n + kb;
  double x[kCB];
#pragma unroll
  for (int c = 0; c < kCB; ++c) x[c] = c < m ? row[c] : 0.0;
#pragma unroll
  for (int c = 0; c < kCB; ++c) {
    if (c < m) {
#pragma unroll
      for (int l = 0; l < c; ++l) x[c] -= x[l] * L11[c][l];
      x[c] /= L11[c][c];
    }
  }
#pragma unroll
  for (int c = 0; c < kCB; ++c)
    if (c < m) row[c] = x[c];
}

// Tiles (ti, tj), tj <= ti, of the trailing block [t0, n)^2, t0 = kb + m;
// tile index q enumerates the lower triangle row by row.  Each thread owns a
// 4x4 set of entries; the panel columns of the tile's rows and columns are
// staged in LDS.
__global__ __launch_bounds__(256) void k_syrk_tiles(int n, double* __restrict__ D, int kb,
                                                    const int* fail) {
  __shared__ double Ai[kCB][kCB + 1];
  __shared__ double Aj[kCB][kCB + 1];
  if (*fail) return;
  const int m = min(kCB, n - kb);
  const int t0 = kb + m;
  const int q = blockIdx.x;
  int ti = (int)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= q) ++ti;
  while (ti * (ti + 1) / 2 > q) --ti;
  const int tj = q - ti * (ti + 1) / 2;
  const int r0 = t0 + ti * kCB, c0 = t0 + tj * kCB;
  const int tid = threadIdx.x;
  for (int e = tid; e < kCB * kCB; e += blockDim.x) {
    const int r = e / kCB, l = e % kCB;
    Ai[r][l] = (r0 + r < n && l < m) ? D[(long)(r0 + r) * n + kb + l] : 0.0;
    Aj[r][l] = (c0 + r < n && l < m) ? D[(long)(c0 + r) * n + kb + l] : 0.0;
  }
  __syncthreads();
  const int rr = (tid / 16) * 4, cc = (tid % 16) * 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = r0 + rr + a, j = c0 + cc + b;
      acc[a][b] = (i < n && j < n && j <= i) ? D[(long)i * n + j] : 0.0;
    }
  for (int l = 0; l < m; ++l) {
    double av[4], bv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) av[a] = Ai[rr + a][l];
#pragma unroll
    for (int b = 0; b < 4; ++b) bv[b] = Aj[cc + b][l];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] -= av[a] * bv[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = r0 + rr + a, j = c0 + cc + b;
      if (i < n && j < n && j <= i) D[(long)i * n + j] = acc[a][b];
    }
}

// upper triangle zeroed, Lt = L^T (row-major); skipped after a failure
__global__ __launch_bounds__(kBlock) void k_chol_finish(int n, double* __restrict__ D,
                                                        double* __restrict__ Lt,
                                                        const int* fail) {
  if (*fail) return;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  const double v = j > i ? 0.0 : D[e];
  D[e] = v;
  Lt[(long)j * n + i] = v;
}

__global__ void k_clear_flag(int* fail) { *fail = 0; }

int dense_cholesky(int n, double* D, double* Lt, int* fail, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, st, fail);
  GF_CHECK_LAUNCH();
  for (int kb = 0; kb < n; kb += kCB) {
    const int m = std::min(kCB, n - kb);
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(256), 0, st, n, D, kb, fail);
    GF_CHECK_LAUNCH();
    const int rows = n - kb - m;
    if (rows <= 0) break;
    hipLaunchKernelGGL(k_trsm_rows, dim3(cdiv(rows, kCB)), dim3(kCB), 0, st, n, D, kb, fail);
    GF_CHECK_LAUNCH();
    const int T = cdiv(rows, kCB);
    hipLaunchKernelGGL(k_syrk_tiles, dim3(T * (T + 1) / 2), dim3(256), 0, st, n, D, kb, fail);
    GF_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_chol_finish, dim3(cdiv((long)n * n, kBlock)), dim3(kBlock), 0, st, n, D,
                     Lt, fail);
  GF_CHECK_LAUNCH();
  return 0;
}

__global__ __launch_bounds__(kBlock) void k_add_diag(int n, double* D, double eps) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    D[(long)i * n + i] += eps;
}

int dense_add_diag(int n, double* D, double eps, hipStream_t st) {
  hipLaunchKernelGGL(k_add_diag, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, n, D, eps);
  GF_CHECK_LAUNCH();
  return 0;
}

// x = L^-T L^-1 b by blocked substitution in one work-group.  The working
// vector lives in LDS; 64x64 diagonal blocks are staged in LDS and solved by
// one wavefront (lane r owns row r, column sweep with the stored reciprocal
// pivots); the off-diagonal updates are 64-term row dot products streamed by
// all threads (row-major L for the forward sweep, its transpose Lt for the
// backward sweep, so every row segment is contiguous).
constexpr int kTrsvThreads = 256;
constexpr int kTB = 64;

__global__ __launch_bounds__(kTrsvThreads) void k_chol_solve(int n, const double* __restrict__ L,
                                                             const double* __restrict__ Lt,
                                                             const double* __restrict__ rdiag,
                                                             const double* __restrict__ b,
                                                             double* __restrict__ x,
                                                             const int* gate) {
  if (gate && *gate == 0) return;
  __shared__ double v[kTrsvMaxN];
  __shared__ double D[kTB * (kTB + 1)];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int i = tid; i < n; i += blockDim.x) v[i] = b[i];
  __syncthreads();
  // forward: L y = b
  for (int kb = 0; kb < n; kb += kTB) {
    const int m = min(kTB, n - kb);
    for (int idx = tid; idx < m * m; idx += blockDim.x) {
      const int r = idx / m, c = idx % m;
      D[r * (kTB + 1) + c] = (c <= r) ? L[(long)(kb + r) * n + kb + c] : 0.0;
    }
    __syncthreads();
    if (wid == 0) {
      double val = lane < m ? v[kb + lane] : 0.0;
      const double rd = lane < m ? rdiag[kb + lane] : 0.0;
      for (int c = 0; c < m; ++c) {
        const double yc = __shfl(val, c, 64) * __shfl(rd, c, 64);
        if (lane == c) val = yc;
        else if (lane > c) val -= D[lane * (kTB + 1) + c] * yc;
      }
      if (lane < m) v[kb + lane] = val;
    }
    __syncthreads();
    for (int i = kb + m + tid; i < n; i += blockDim.x) {
      const double* row = L + (long)i * n + kb;
      double acc = 0.0;
      for (int c = 0; c < m; ++c) acc += row[c] * v[kb + c];
      v[i] -= acc;
    }
    __syncthreads();
  }
  // backward: L^T x = y  (block rows from the bottom)
  const int nb = (n + kTB - 1) / kTB;
  for (int bi = nb - 1; bi >= 0; --bi) {
    const int kb = bi * kTB;
    const int m = min(kTB, n - kb);
    // U = (L block)^T: U[r][c] = L[kb+c][kb+r] = Lt[(kb+r)*n + kb+c], c >= r
    for (int idx = tid; idx < m * m; idx += blockDim.x) {
      const int r = idx / m, c = idx % m;
      D[r * (kTB + 1) + c] = (c >= r) ? Lt[(long)(kb + r) * n + kb + c] : 0.0;
    }
    __syncthreads();
    if (wid == 0) {
      double val = lane < m ? v[kb + lane] : 0.0;
      const double rd = lane < m ? rdiag[kb + lane] : 0.0;
      for (int c = m - 1; c >= 0; --c) {
        const double xc = __shfl(val, c, 64) * __shfl(rd, c, 64);
        if (lane == c) val = xc;
        else if (lane < c) val -= D[lane * (kTB + 1) + c] * xc;
      }
      if (lane < m) v[kb + lane] = val;
    }
    __syncthreads();
    for (int i = tid; i < kb; i += blockDim.x) {
      const double* row = Lt + (long)i * n + kb;
      double acc = 0.0;
      for (int c = 0; c < m; ++c) acc += row[c] * v[kb + c];
      v[i] -= acc;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += blockDim.x) x[i] = v[i];
}

int dense_chol_solve(int n, const double* L, const double* Lt, const double* rdiag,
                     const double* b, double* x, const int* gate, hipStream_t st) {
  if (n > kTrsvMaxN) {
    set_error("dense_chol_solve: n exceeds the dense coarsest limit");
    return -2;
  }
  hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(kTrsvThreads), 0, st, n, L, Lt, rdiag, b, x,
                     gate);
  GF_CHECK_LAUNCH();
  return 0;
}

__global__ void k_recip_diag(int n, const double* __restrict__ L, double* __restrict__ r) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    r[i] = 1.0 / L[(long)i * n + i];
}

// Y = L^{-1} (row-major, lower) by right-looking blocked forward substitution
// of the identity over 64-row blocks.  Every entry receives the updates of
// column-wise substitution, Y[i][j] = (delta_ij - sum_{k=j}^{i-1} L[i][k]
// Y[k][j]) / L[i][i], in the same ascending k order (terms with k < j are
// exact zeros), so the result does not depend on the blocking:
//   k_inv_diag_rows  finishes the block's rows: thread j owns column j and
//                    applies the block's own pivots (L block in LDS);
//   k_inv_update     subtracts the finished block's contribution from every
//                    later row, in 64x64 tiles.
__global__ __launch_bounds__(kBlock) void k_eye(long nn, int n, double* __restrict__ Y) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  Y[e] = (e / n == e % n) ? 1.0 : 0.0;
}

__global__ __launch_bounds__(kCB) void k_inv_diag_rows(int n, const double* __restrict__ L,
                                                       double* __restrict__ Y, int ib) {
  __shared__ double Lb[kCB][kCB + 1];
  const int m = min(kCB, n - ib);
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int r = e / m, c = e % m;
    Lb[r][c] = c <= r ? L[(long)(ib + r) * n + ib + c] : 0.0;
  }
  __syncthreads();
  const int j = blockIdx.x * kCB + threadIdx.x;  // columns 0 .. ib+m-1
  if (j >= ib + m) return;
  double x[kCB];
#pragma unroll
  for (int r = 0; r < kCB; ++r) x[r] = r < m ? Y[(long)(ib + r) * n + j] : 0.0;
#pragma unroll
  for (int r = 0; r < kCB; ++r) {
    if (r < m) {
#pragma unroll
      for (int l = 0; l < r; ++l) x[r] -= Lb[r][l] * x[l];
      x[r] /= Lb[r][r];
    }
  }
#pragma unroll
  for (int r = 0; r < kCB; ++r)
    if (r < m) Y[(long)(ib + r) * n + j] = x[r];
}

// tile q = (row tile R of the rows below the block, column tile C <= the
// block's): Y[R][C] -= L[R][block] * Y[block][C]
__global__ __launch_bounds__(256) void k_inv_update(int n, const double* __restrict__ L,
                                                    double* __restrict__ Y, int ib, int ncol) {
  __shared__ double Lr[kCB][kCB + 1];
  __shared__ double Yb[kCB][kCB + 1];
  const int m = min(kCB, n - ib);
  const int R = blockIdx.x / ncol, C = blockIdx.x % ncol;
  const int r0 = ib + m + R * kCB, c0 = C * kCB;
  const int tid = threadIdx.x;
  for (int e = tid; e < kCB * kCB; e += blockDim.x) {
    const int r = e / kCB, l = e % kCB;
    Lr[r][l] = (r0 + r < n && l < m) ? L[(long)(r0 + r) * n + ib + l] : 0.0;
    Yb[r][l] = (r < m && c0 + l < n) ? Y[(long)(ib + r) * n + c0 + l] : 0.0;
  }
  __syncthreads();
  const int rr = (tid / 16) * 4, cc = (tid % 16) * 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = r0 + rr + a, j = c0 + cc + b;
      acc[a][b] = (i < n && j < n) ? Y[(long)i * n + j] : 0.0;
    }
  for (int l = 0; l < m; ++l) {
    double av[4], bv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) av[a] = Lr[rr + a][l];
#pragma unroll
    for (int b = 0; b < 4; ++b) bv[b] = Yb[l][cc + b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] -= av[a] * bv[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = r0 + rr + a, j = c0 + cc + b;
      if (i < n && j < n) Y[(long)i * n + j] = acc[a][b];
    }
}

// Kinv = Y^T Y over lower 64x64 tiles (I >= J), mirrored into the upper
// triangle: Kinv[i][j] = sum_{k >= max(i,j)} Y[k][i] Y[k][j], k ascending
// from +0 (rows k < max(i,j) add exact zeros).
__global__ __launch_bounds__(256) void k_yty_tiles(int n, const double* __restrict__ Y,
                                                   double* __restrict__ K) {
  __shared__ double Ai[kCB][kCB + 1];
  __shared__ double Aj[kCB][kCB + 1];
  const int q = blockIdx.x;
  int ti = (int)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= q) ++ti;
  while (ti * (ti + 1) / 2 > q) --ti;
  const int tj = q - ti * (ti + 1) / 2;
  const int i0 = ti * kCB, j0 = tj * kCB;
  const int tid = threadIdx.x;
  const int rr = (tid / 16) * 4, cc = (tid % 16) * 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int k0 = i0; k0 < n; k0 += kCB) {
    __syncthreads();
    for (int e = tid; e < kCB * kCB; e += blockDim.x) {
      const int l = e / kCB, c = e % kCB;
      const bool in = k0 + l < n;
      Ai[l][c] = (in && i0 + c < n) ? Y[(long)(k0 + l) * n + i0 + c] : 0.0;
      Aj[l][c] = (in && j0 + c < n) ? Y[(long)(k0 + l) * n + j0 + c] : 0.0;
    }
    __syncthreads();
    const int lm = min(kCB, n - k0);
    for (int l = 0; l < lm; ++l) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = Ai[l][rr + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Aj[l][cc + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = i0 + rr + a, j = j0 + cc + b;
      if (i < n && j < n) {
        K[(long)i * n + j] = acc[a][b];
        K[(long)j * n + i] = acc[a][b];
      }
    }
}

// Kinv = L^{-T} L^{-1} from the lower Cholesky factor L (row-major); Y is an
// n*n scratch buffer (the caller's second scratch buffer Yt is unused).
int dense_inverse_from_chol(int n, const double* L, double* Y, double* /*Yt*/, double* Kinv,
                            hipStream_t st) {
  const long nn = (long)n * n;
  hipLaunchKernelGGL(k_eye, dim3(cdiv(nn, kBlock)), dim3(kBlock), 0, st, nn, n, Y);
  GF_CHECK_LAUNCH();
  for (int ib = 0; ib < n; ib += kCB) {
    const int m = std::min(kCB, n - ib);
    hipLaunchKernelGGL(k_inv_diag_rows, dim3(cdiv(ib + m, kCB)), dim3(kCB), 0, st, n, L, Y, ib);
    GF_CHECK_LAUNCH();
    const int rows = n - ib - m;
    if (rows <= 0) break;
    const int ncol = cdiv(ib + m, kCB);
    hipLaunchKernelGGL(k_inv_update, dim3(cdiv(rows, kCB) * ncol), dim3(256), 0, st, n, L, Y,
                       ib, ncol);
    GF_CHECK_LAUNCH();
  }
  const int T = cdiv(n, kCB);
  hipLaunchKernelGGL(k_yty_tiles, dim3(T * (T + 1) / 2), dim3(256), 0, st, n, Y, Kinv);
  GF_CHECK_LAUNCH();
  return 0;
}

// x = Kinv b: one wavefront per row, lanes stride the row, fixed-order
// wave reduction.
__global__ __launch_bounds__(kBlock) void k_dense_gemv(int n, const double* __restrict__ K,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ x, const int* gate) {
  const bool open = gate_open(gate);  // tested after the row's loads
  const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double* r = K + (long)row * n;
  double acc = 0.0;
  // four columns' loads in flight before the first product, the products
  // added in ascending column order (the plain stride loop's order)
  int c = lane;
  for (; c + 192 < n; c += 256) {
    double rv[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      rv[u] = r[c + 64 * u];
      bv[u] = b[c + 64 * u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += rv[u] * bv[u];
  }
  for (; c < n; c += 64) acc += r[c] * b[c];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if (lane == 0 && open) x[row] = acc;
}

int dense_gemv(int n, const double* K, const double* b, double* x, const int* gate,
               hipStream_t st) {
  hipLaunchKernelGGL(k_dense_gemv, dim3(cdiv(n, kBlock / 64)), dim3(kBlock), 0, st, n, K, b, x,
                     gate);
  GF_CHECK_LAUNCH();
  return 0;
}

int dense_recip_diag(int n, const double* L, double* rdiag, hipStream_t st) {
  hipLaunchKernelGGL(k_recip_diag, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, n, L, rdiag);
  GF_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// Fixed-count Jacobi PCG on the coarsest stencil, zero start, no stopping
// test (the reference's device-scalar loop, solver_v2.py:781-810), in two
// grid-wide launches per iteration with every scalar in device memory (the
// loop is stream-ordered and is recorded in the preconditioner graph):
//   k_pcg_spmv  : p_it = dinv r + beta p_{it-1} (beta from s), formed on the
//                 fly for every stencil neighbour and stored for owned DOFs
//                 in the buffer of this iteration's parity; Ap = A p_it;
//                 s[2] = p_it . Ap
//   k_pcg_update: alpha = s[cur] / (s[2] + 1e-300); x += alpha p_it;
//                 r -= alpha Ap; s[cur^1] = r . (dinv r)
// The direction is one expression (pcg_dir) evaluated identically wherever a
// DOF's value is needed, so every block sees the same bits.  s[0], s[1] hold
// r.z of alternate iterations.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pcg_dir(double dinv, double r, double beta, double pold) {
  return __fma_rn(beta, pold, __dmul_rn(dinv, r));
}

__global__ __launch_bounds__(kBlock) void k_pcg_init(long n, const double* __restrict__ b,
                                                     const double* __restrict__ dinv,
                                                     double* __restrict__ x,
                                                     double* __restrict__ r, double* partials,
                                                     int* ticket, double* s, const int* gate) {
  if (gate && *gate == 0) return;
  double a[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    a[0] += bi * __dmul_rn(dinv[i], bi);
  }
  finish_reduce<1>(a, partials, ticket, s, 0);
}

// FIRST: p_0 = dinv b (= dinv r_0, beta = 0).  64 nodes x 4 waves per block
// as k_stencil_nodes; the slot products are folded in slot order 0..26.
template <bool OFF, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(Grid g, DofMap map,
                                                     const double* __restrict__ A,
                                                     const double* __restrict__ dinv,
                                                     const double* __restrict__ r,
                                                     const double* __restrict__ pold,
                                                     double* __restrict__ pnew,
                                                     double* __restrict__ Ap, double* partials,
                                                     int* ticket, double* s, int cur,
                                                     const int* gate) {
  if (gate && *gate == 0) return;
  __shared__ double tk[27][3][64];
  const double beta = FIRST ? 0.0 : s[cur] / (s[cur ^ 1] + 1e-300);
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = blockIdx.x * 64 + lane;
  const bool valid = n < g.n_nodes;
  const int nn_ = valid ? n : 0;
  int ix, iy, iz;
  node_xyz(g, nn_, ix, iy, iz);
  const long nn = g.n_nodes;
  for (int k = wv; k < 27; k += 4) {
    const int jx = ix + k / 9 - 1, jy = iy + (k / 3) % 3 - 1, jz = iz + k % 3 - 1;
    const bool in = valid & (jx >= 0) & (jx < g.NX) & (jy >= 0) & (jy < g.NY) & (jz >= 0) &
                    (jz < g.NZ);
    const int m = (jx * g.NY + jy) * g.NZ + jz;
    double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int f = in ? map_idx<OFF>(map, 3 * m + q) : -1;
      const int fi = f >= 0 ? f : 0;
      const double dv = dinv[fi], rv = r[fi];
      const double pv = FIRST ? __dmul_rn(dv, rv) : pcg_dir(dv, rv, beta, pold[fi]);
      const double xv = f >= 0 ? pv : 0.0;
      s0 += A[(k * 9 + 0 * 3 + q) * nn + nn_] * xv;
      s1 += A[(k * 9 + 1 * 3 + q) * nn + nn_] * xv;
      s2 += A[(k * 9 + 2 * 3 + q) * nn + nn_] * xv;
    }
    tk[k][0][lane] = s0;
    tk[k][1][lane] = s1;
    tk[k][2][lane] = s2;
  }
  __syncthreads();
  double a[1] = {0.0};
  if (wv == 0 && valid) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int o = map_idx<OFF>(map, 3 * n + q);
      if (o < 0) continue;
      double acc = 0;
#pragma unroll
      for (int kk = 0; kk < 27; ++kk) acc += tk[kk][q][lane];
      const double pv = FIRST ? __dmul_rn(dinv[o], r[o]) : pcg_dir(dinv[o], r[o], beta, pold[o]);
      pnew[o] = pv;
      Ap[o] = acc;
      a[0] += pv * acc;
    }
  }
  finish_reduce<1>(a, partials, ticket, s + 2, 0);
}

__global__ __launch_bounds__(kBlock) void k_pcg_update(long n, const double* __restrict__ dinv,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ Ap,
                                                       double* __restrict__ x,
                                                       double* __restrict__ r, double* partials,
                                                       int* ticket, double* s, int cur,
                                                       const int* gate) {
  if (gate && *gate == 0) return;
  const double alpha = s[cur] / (s[2] + 1e-300);
  double a[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    x[i] = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    a[0] += ri * __dmul_rn(dinv[i], ri);
  }
  finish_reduce<1>(a, partials, ticket, s + (cur ^ 1), 0);
}

int fixed_pcg_grid(Work& w, const Grid& g, const DofMap& map, const double* A,
                   const double* dinv, int n_free, int iters, const double* b, double* x,
                   double* work, const int* gate, hipStream_t st) {
  const long n = n_free;
  double* r = work;
  double* pb[2] = {work + n, work + 2 * n};
  double* Ap = work + 3 * n;
  double* s = work + 4 * n;
  const int G = red_grid(n);
  const int Gs = cdiv(g.n_nodes, 64);
  if (Gs > kRedBlocks) {
    set_error("fixed_pcg_grid: coarsest grid too large for one reduction grid");
    return -2;
  }
  hipLaunchKernelGGL(k_pcg_init, dim3(G), dim3(kBlock), 0, st, n, b, dinv, x, r, w.partials,
                     w.ticket, s, gate);
  GF_CHECK_LAUNCH();
  for (int it = 0; it < iters; ++it) {
    // r.z of this iteration sits in s[cur]; the previous one in s[cur ^ 1]
    const int cur = it & 1;
    double* pn = pb[cur];
    const double* po = pb[cur ^ 1];
#define GF_PCG_SPMV(OFF, FIRST)                                                                \
  hipLaunchKernelGGL((k_pcg_spmv<OFF, FIRST>), dim3(Gs), dim3(kBlock), 0, st, g, map, A, dinv, \
                     r, po, pn, Ap, w.partials, w.ticket, s, cur, gate)
    if (map.off >= 0) {
      if (it == 0) GF_PCG_SPMV(true, true); else GF_PCG_SPMV(true, false);
    } else {
      if (it == 0) GF_PCG_SPMV(false, true); else GF_PCG_SPMV(false, false);
    }
#undef GF_PCG_SPMV
    GF_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pcg_update, dim3(G), dim3(kBlock), 0, st, n, dinv, pn, Ap, x, r,
                       w.partials, w.ticket, s, cur, gate);
    GF_CHECK_LAUNCH();
  }
  return 0;
}

// ---------------------------------------------------------------------------
// The same PCG in ONE persistent launch (k_pcg_persistent): a block owns 64
// consecutive coarsest nodes (4 waves share the 27 stencil slots as in
// k_pcg_spmv; wave 0 keeps its nodes' x, r, p, Ap in registers).  The two
// dot products of an iteration are grid-wide exchanges of block partials
// through a ring of three slots per block in uncached memory, as in the
// Gram-Schmidt sweep: the value in a slot is its own arrival signal (an empty
// slot holds the bit pattern kRingEmpty), exchange e uses slot e mod 3, and
// every block polls the G slots of its round until none is empty and folds
// the values in block order (identical bits everywhere).  A block that has
// read round e empties its slot of round e-1 (every block is then past it),
// and that store is acknowledged before the block publishes round e+1, so a
// reader of round e+2 never sees the stale value.  The vectors other blocks
// read (r and the two direction buffers) also live in uncached memory and are
// written and read with agent-scope atomic accesses (which bypass the CU's
// L1); their stores are acknowledged before the block's partial is published,
// so they are visible to every block that has seen the partial -- no L2
// write-back is needed.  The ring is emptied by a fill kernel before every
// launch.  Waits are bounded: an expired wait sets *err (the outer solve
// reports it) and the block stops waiting, so the launch always ends.
// ---------------------------------------------------------------------------
constexpr long kPcgSpinCap = 1L << 20;
constexpr int kPcgMaxBlocks = 1024;

// Block-sum v, publish it as round e of the ring, wait for every block's
// round-e partial, fold them in block order.  All threads of the block must
// call it.
__device__ double pcg_exchange(double v, int e, double* ring, double* err, bool& broken,
                               double* lds, double* bc) {
  const int G = gridDim.x, b = blockIdx.x;
  const double bs = block_sum(v, lds);
  double* cur = ring + (long)(e % 3) * kPcgMaxBlocks;
  if (threadIdx.x == 0) {
    // wave 0's vector stores of this iteration and the previous round's slot
    // emptying have landed before the partial can be seen
    store_ack_wait();
    __hip_atomic_store(&cur[b], bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x < 64) {
    long spins = 0;
    double t = 0.0;
    while (!broken) {
      int ok = 1;
      t = 0.0;
      for (int q = threadIdx.x; q < G; q += 64) {
        const double pv = __hip_atomic_load(&cur[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok &= (int)((unsigned long long)__double_as_longlong(pv) != kRingEmpty);
        t += pv;
      }
      if (__all(ok)) break;
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kPcgSpinCap) {
        broken = true;
        if (threadIdx.x == 0) *err = 1.0;
      }
    }
    t = wave_sum(t);
    if (threadIdx.x == 0) {
      *bc = t;
      __hip_atomic_store(reinterpret_cast<unsigned long long*>(
                             &ring[(long)((e + 2) % 3) * kPcgMaxBlocks + b]),
                         kRingEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  return *bc;
}

template <bool OFF>
__global__ __launch_bounds__(kBlock) void k_pcg_persistent(Grid g, DofMap map,
                                                           const double* __restrict__ A,
                                                           const double* __restrict__ dinv,
                                                           int iters, const double* __restrict__ b,
                                                           double* __restrict__ x, double* r,
                                                           double* p0, double* p1, double* ring,
                                                           double* err, const int* gate) {
  if (gate && *gate == 0) return;
  __shared__ double tk[27][3][64];
  __shared__ double lds[4];
  __shared__ double bc;
  bool broken = false;  // wave 0 only
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = blockIdx.x * 64 + lane;
  const bool valid = n < g.n_nodes;
  const int nn_ = valid ? n : 0;
  int ix, iy, iz;
  node_xyz(g, nn_, ix, iy, iz);
  const long nn = g.n_nodes;
  int o[3];
  double xr[3] = {0, 0, 0}, rr[3] = {0, 0, 0}, dr[3] = {0, 0, 0}, pr[3] = {0, 0, 0},
         ar[3] = {0, 0, 0};
  double a = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    o[q] = (valid && wv == 0) ? map_idx<OFF>(map, 3 * n + q) : -1;
    if (o[q] >= 0) {
      const double bi = b[o[q]];
      dr[q] = dinv[o[q]];
      rr[q] = bi;
      __hip_atomic_store(&r[o[q]], bi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      a += bi * __dmul_rn(dr[q], bi);
    }
  }
  int e = 0;
  double rz = pcg_exchange(a, e++, ring, err, broken, lds, &bc);
  double rz_prev = 1.0;
  for (int it = 0; it < iters; ++it) {
    double* pn = (it & 1) ? p1 : p0;
    const double* po = (it & 1) ? p0 : p1;
    const double beta = it == 0 ? 0.0 : rz / (rz_prev + 1e-300);
    // A wave's 7 slot rounds (the last masked off for wave 3) with a
    // compile-time trip count and the LDS writes after them, so the
    // uncached r and p loads of every round are in flight together.
    double sv[7][3];
#pragma unroll
    for (int t = 0; t < 7; ++t) {
      const bool kin = (t < 6) | (wv < 3);
      const int k = kin ? wv + 4 * t : 0;
      const int jx = ix + k / 9 - 1, jy = iy + (k / 3) % 3 - 1, jz = iz + k % 3 - 1;
      const bool in = kin & valid & (jx >= 0) & (jx < g.NX) & (jy >= 0) & (jy < g.NY) &
                      (jz >= 0) & (jz < g.NZ);
      const int m = (jx * g.NY + jy) * g.NZ + jz;
      double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int f = in ? map_idx<OFF>(map, 3 * m + q) : -1;
        const int fi = f >= 0 ? f : 0;
        // r and the previous direction change during this launch (written by
        // other blocks): agent-scope loads, which do not hit the CU's L1
        const double dv = dinv[fi];
        const double rv = __hip_atomic_load(&r[fi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const double pv =
            it == 0 ? __dmul_rn(dv, rv)
                    : pcg_dir(dv, rv, beta,
                              __hip_atomic_load(&po[fi], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT));
        const double xv = f >= 0 ? pv : 0.0;
        s0 += A[(k * 9 + 0 * 3 + q) * nn + nn_] * xv;
        s1 += A[(k * 9 + 1 * 3 + q) * nn + nn_] * xv;
        s2 += A[(k * 9 + 2 * 3 + q) * nn + nn_] * xv;
      }
      sv[t][0] = s0;
      sv[t][1] = s1;
      sv[t][2] = s2;
    }
#pragma unroll
    for (int t = 0; t < 7; ++t) {
      if ((t < 6) | (wv < 3)) {
        const int k = wv + 4 * t;
        tk[k][0][lane] = sv[t][0];
        tk[k][1][lane] = sv[t][1];
        tk[k][2][lane] = sv[t][2];
      }
    }
    __syncthreads();
    a = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (o[q] < 0) continue;
      double acc = 0;
#pragma unroll
      for (int kk = 0; kk < 27; ++kk) acc += tk[kk][q][lane];
      pr[q] = it == 0 ? __dmul_rn(dr[q], rr[q]) : pcg_dir(dr[q], rr[q], beta, pr[q]);
      __hip_atomic_store(&pn[o[q]], pr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ar[q] = acc;
      a += pr[q] * acc;
    }
    const double pAp = pcg_exchange(a, e++, ring, err, broken, lds, &bc);
    const double alpha = rz / (pAp + 1e-300);
    a = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (o[q] < 0) continue;
      xr[q] = xr[q] + alpha * pr[q];
      rr[q] = rr[q] - alpha * ar[q];
      __hip_atomic_store(&r[o[q]], rr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      a += rr[q] * __dmul_rn(dr[q], rr[q]);
    }
    rz_prev = rz;
    rz = pcg_exchange(a, e++, ring, err, broken, lds, &bc);
  }
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (o[q] >= 0) x[o[q]] = xr[q];
}

__global__ void k_empty_ring(int n, unsigned long long* p) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = kRingEmpty;
}

// Uncached buffers of the persistent PCG: handed out by size and recycled,
// never returned to the driver (uncached memory that kernels have read is kept
// until the process ends).
namespace {
struct UcBlock {
  char* p;
  size_t bytes;
};
std::vector<UcBlock>& uc_free_list() {
  static std::vector<UcBlock> v;
  return v;
}
}  // namespace

int pcg_buffers_acquire(long n_free, void** out, size_t* bytes_out) {
  GF_DEVICE_LOCK();
  const size_t bytes = sizeof(double) * 3 * kPcgMaxBlocks + sizeof(double) * 3 * (size_t)n_free +
                       256;
  auto& fl = uc_free_list();
  for (size_t i = 0; i < fl.size(); ++i) {
    if (fl[i].bytes >= bytes) {
      *out = fl[i].p;
      *bytes_out = fl[i].bytes;
      fl.erase(fl.begin() + (long)i);
      return 0;
    }
  }
  GF_HIP(hipExtMallocWithFlags(out, bytes, hipDeviceMallocUncached));
  *bytes_out = bytes;
  return 0;
}

void pcg_buffers_release(void* p, size_t bytes) {
  if (!p) return;
  GF_DEVICE_LOCK();
  uc_free_list().push_back(UcBlock{static_cast<char*>(p), bytes});
}

// Co-resident blocks of the k_pcg_persistent variant that will be launched
// (the kernel's blocks wait for each other, so its grid must fit at once,
// with room for the waiting grids of other hardware queues: at most
// 1/hw_queues() of the capacity, as for the sweep).
template <bool OFF>
static int pcg_capacity() {
  static std::atomic<int> cap{-1};
  if (cap.load() < 0) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<OFF>, kBlock, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    cap.store(per_cu * cus);
  }
  return cap.load();
}

int fixed_pcg_persistent(const Grid& g, const DofMap& map, const double* A, const double* dinv,
                         int n_free, int iters, const double* b, double* x, void* ucbuf,
                         double* err, const int* gate, hipStream_t st) {
  const int G = cdiv(g.n_nodes, 64);
  const int cap = map.off >= 0 ? pcg_capacity<true>() : pcg_capacity<false>();
  if (G > kPcgMaxBlocks || (long)hw_queues() * G > cap) return 0;
  double* ring = static_cast<double*>(ucbuf);
  double* r = ring + 3 * kPcgMaxBlocks;
  double* p0 = r + n_free;
  double* p1 = p0 + n_free;
  hipLaunchKernelGGL(k_empty_ring, dim3(12), dim3(256), 0, st, 3 * kPcgMaxBlocks,
                     reinterpret_cast<unsigned long long*>(ring));
  GF_CHECK_LAUNCH();
  if (map.off >= 0)
    hipLaunchKernelGGL(k_pcg_persistent<true>, dim3(G), dim3(kBlock), 0, st, g, map, A, dinv,
                       iters, b, x, r, p0, p1, ring, err, gate);
  else
    hipLaunchKernelGGL(k_pcg_persistent<false>, dim3(G), dim3(kBlock), 0, st, g, map, A, dinv,
                       iters, b, x, r, p0, p1, ring, err, gate);
  GF_CHECK_LAUNCH();
  return 1;
}

}  // namespace gf
