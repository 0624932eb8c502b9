// Fine-level matrix-free hex8 operator K_ff(E) u and its set-up companions.
//
// Owner-computes formulation: one thread per grid node gathers the 27
// neighbouring node displacements (3 components each) and the moduli of its
// (up to) 8 adjacent elements into registers, then forms
//     y_n = sum_{e adjacent} E_e * (Ke[rows of n in e, :] . u_e)
// with the element loop in ascending element index and the 24-term row
// products in column order.  No atomics, no element->DOF table (indices are
// computed from the structured grid), and every result is bitwise
// reproducible.  Ke is read with wave-uniform addresses, so it travels through
// the scalar cache and feeds the FMAs as scalar operands.
#include "gf_internal.h"

#include <atomic>
#include <type_traits>

namespace gf {

template <typename T>
struct Acc {
  using type = T;
};
// FP32 operator: element products in FP32, node sums in FP64, result rounded
// once to FP32 (the reference's FP32 path accumulates its scatter in FP64).
template <>
struct Acc<float> {
  using type = double;
};

// Element contributions are summed with one fixed tree for every launch
// shape:  P_ab = E_ab0*t_ab0 + E_ab1*t_ab1  (element (a,b,c) = (ix-1+a,
// iy-1+b, iz-1+c)),  Q_a = P_a0 + P_a1,  y = Q_0 + Q_1,  each t the 24-term
// row product in column order.  SPLIT lanes cooperate on one node (SPLIT=1:
// a whole node per lane; 2: one lane per a; 4: one lane per (a,b)) and
// combine their partial trees with shuffles, so all shapes are bitwise
// identical and the shape can be autotuned freely.
template <typename T>
__device__ __forceinline__ void elem_rows(const T* __restrict__ Ke, const T (&un)[27][3], int a,
                                          int b, int c, T& t0, T& t1, T& t2) {
  const int L = lidx(1 - a, 1 - b, 1 - c);
  t0 = 0;
  t1 = 0;
  t2 = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = (a + lpos(j, 0)) * 9 + (b + lpos(j, 1)) * 3 + (c + lpos(j, 2));
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T uu = un[k][q];
      t0 += Ke[(3 * L + 0) * 24 + 3 * j + q] * uu;
      t1 += Ke[(3 * L + 1) * 24 + 3 * j + q] * uu;
      t2 += Ke[(3 * L + 2) * 24 + 3 * j + q] * uu;
    }
  }
}

// SPLIT = 1: one lane per node, the whole tree in registers.
template <typename T, int EPI, bool OFF>
__global__ __launch_bounds__(kBlock) void k_fine_matvec_node(Grid g, DofMap map,
                                                             const T* __restrict__ Ke,
                                                             const T* __restrict__ E,
                                                             const T* __restrict__ u,
                                                             T* __restrict__ y, EpiArgs<T> ep,
                                                             const int* gate) {
  using A = typename Acc<T>::type;
  if (gate && *gate == 0) return;
  const int n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int o0 = map_idx<OFF>(map, 3 * n), o1 = map_idx<OFF>(map, 3 * n + 1),
            o2 = map_idx<OFF>(map, 3 * n + 2);
  if ((o0 & o1 & o2) < 0) return;
  const int iz = n % g.NZ, t = n / g.NZ, iy = t % g.NY, ix = t / g.NY;
  T un[27][3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
      for (int dz = 0; dz < 3; ++dz) {
        const int jx = ix + dx - 1, jy = iy + dy - 1, jz = iz + dz - 1;
        const bool in = (jx >= 0) & (jx < g.NX) & (jy >= 0) & (jy < g.NY) & (jz >= 0) &
                        (jz < g.NZ);
        const int m = (jx * g.NY + jy) * g.NZ + jz;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          un[dx * 9 + dy * 3 + dz][c] = map_load<OFF>(map, u, 3 * m + c, in);
      }
  A Q[2][3];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    A P[2][3];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      P[b][0] = 0;
      P[b][1] = 0;
      P[b][2] = 0;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int ex = ix - 1 + a, ey = iy - 1 + b, ez = iz - 1 + c;
        const bool in = (ex >= 0) & (ex < g.nx) & (ey >= 0) & (ey < g.ny) & (ez >= 0) &
                        (ez < g.nz);
        const T ev = E[in ? (ex * g.ny + ey) * g.nz + ez : 0];
        const T e = in ? ev : T(0);
        T t0, t1, t2;
        elem_rows<T>(Ke, un, a, b, c, t0, t1, t2);
        P[b][0] += (A)(e * t0);
        P[b][1] += (A)(e * t1);
        P[b][2] += (A)(e * t2);
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) Q[a][r] = P[0][r] + P[1][r];
  }
  const int o[3] = {o0, o1, o2};
  const T ax[3] = {(T)(Q[0][0] + Q[1][0]), (T)(Q[0][1] + Q[1][1]), (T)(Q[0][2] + Q[1][2])};
  epilogue3<T, EPI>(o, ax, u, y, ep);
}

// SPLIT = 4: a block is 64 nodes x 4 waves; wave w computes the element pair
// (a,b) = (w>>1, w&1) (c = 0,1) for its node, so the element-local node index,
// and with it every Ke address, is wave-uniform (scalar loads, SGPR operands),
// and each lane loads only the 2x2x3 node window of its two elements.  The
// pair sums P_ab meet in LDS and wave 0 forms (P00+P01)+(P10+P11), the same
// tree as the one-lane-per-node kernel.
template <typename T, typename KP>
__device__ __forceinline__ void pair_rows(KP Ke, const T (&w)[2][2][3][3], int a, int b, int c,
                                          T& t0, T& t1, T& t2) {
  const int L = (1 - c) * 4 + ((1 - b) ? ((1 - a) ? 2 : 3) : ((1 - a) ? 1 : 0));
  const KP K0 = Ke + (3 * L + 0) * 24;
  const KP K1 = Ke + (3 * L + 1) * 24;
  const KP K2 = Ke + (3 * L + 2) * 24;
  t0 = 0;
  t1 = 0;
  t2 = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int px = lpos(j, 0), py = lpos(j, 1), pz = lpos(j, 2);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T uu = w[px][py][c + pz][q];
      t0 += K0[3 * j + q] * uu;
      t1 += K1[3 * j + q] * uu;
      t2 += K2[3 * j + q] * uu;
    }
  }
}

template <typename T, int EPI, bool OFF>
__global__ __launch_bounds__(kBlock) void k_fine_matvec_wave4(Grid g, DofMap map,
                                                              const T* __restrict__ Ke,
                                                              const T* __restrict__ E,
                                                              const T* __restrict__ u,
                                                              T* __restrict__ y, EpiArgs<T> ep,
                                                              const int* gate) {
  using A = typename Acc<T>::type;
  if (gate && *gate == 0) return;
  __shared__ A part[3][3][64];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int a = wv >> 1, b = wv & 1;
  const int n = xcd_tile(blockIdx.x, gridDim.x) * 64 + lane;
  const bool valid = n < g.n_nodes;
  const int nn = valid ? n : 0;
  const int iz = nn % g.NZ, t = nn / g.NZ, iy = t % g.NY, ix = t / g.NY;
  // node window: x in {ix-1+a, ix+a}, y in {iy-1+b, iy+b}, z in {iz-1, iz, iz+1}
  T w[2][2][3][3];
#pragma unroll
  for (int px = 0; px < 2; ++px)
#pragma unroll
    for (int py = 0; py < 2; ++py)
#pragma unroll
      for (int pz = 0; pz < 3; ++pz) {
        const int jx = ix - 1 + a + px, jy = iy - 1 + b + py, jz = iz - 1 + pz;
        const bool in = valid & (jx >= 0) & (jx < g.NX) & (jy >= 0) & (jy < g.NY) & (jz >= 0) &
                        (jz < g.NZ);
        const int m = (jx * g.NY + jy) * g.NZ + jz;
#pragma unroll
        for (int q = 0; q < 3; ++q) w[px][py][pz][q] = map_load<OFF>(map, u, 3 * m + q, in);
      }
  A P[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int ex = ix - 1 + a, ey = iy - 1 + b, ez = iz - 1 + c;
    const bool in = valid & (ex >= 0) & (ex < g.nx) & (ey >= 0) & (ey < g.ny) & (ez >= 0) &
                    (ez < g.nz);
    const T ev = E[in ? (ex * g.ny + ey) * g.nz + ez : 0];
    const T e = in ? ev : T(0);
    T t0, t1, t2;
    pair_rows<T>(Ke, w, a, b, c, t0, t1, t2);
    P[0] += (A)(e * t0);
    P[1] += (A)(e * t1);
    P[2] += (A)(e * t2);
  }
  if (wv > 0) {
#pragma unroll
    for (int r = 0; r < 3; ++r) part[wv - 1][r][lane] = P[r];
  }
  __syncthreads();
  if (wv != 0 || !valid) return;
  const int o0 = map_idx<OFF>(map, 3 * n), o1 = map_idx<OFF>(map, 3 * n + 1),
            o2 = map_idx<OFF>(map, 3 * n + 2);
  A r[3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    r[q] = (P[q] + part[0][q][lane]) + (part[1][q][lane] + part[2][q][lane]);
  const int o[3] = {o0, o1, o2};
  const T ax[3] = {(T)r[0], (T)r[1], (T)r[2]};
  epilogue3<T, EPI>(o, ax, u, y, ep);
}

// Stage the window of the (ix, y0.., z0..) node patch of an LDS tile kernel,
// with branch-free masked loads: us entry ((xx*WY + yy)*WZ + zz)*3 + q = node
// (ix-1+xx, y0-1+yy, z0-1+zz) component q (0 outside the grid and on fixed
// DOFs), es entry (ex*EY + ey)*EZ + ez = element (ix-1+ex, y0-1+ey, z0-1+ez)
// (0 outside).  The window is taken row by row: a (x, y) row of the node
// window is WZ*3 consecutive DOFs in memory, so a thread keeps one position
// cu in a row (its z and component fixed) and walks rows tu, tu + TPU, ...;
// only the row index changes per load.  Every load of a thread is issued
// before the first LDS write (masked loads read a valid address), so they
// are in flight together.  Ends with a barrier.
template <typename T, bool OFF, int TZ>
__device__ __forceinline__ void stage_window(const Grid& g, const DofMap& map,
                                             const T* __restrict__ E, const T* __restrict__ u,
                                             T* us, T* es, int ix, int y0, int z0) {
  constexpr int TY = 64 / TZ;
  constexpr int WY = TY + 2, WZ = TZ + 2;      // node window (x: 3 planes)
  constexpr int EY = TY + 1, EZ = TZ + 1;      // element window (x: 2 planes)
  constexpr int RU = WZ * 3, NRU = 3 * WY;     // node row length, node rows
  constexpr int TPU = kBlock / RU;             // node rows per pass
  constexpr int NUI = (NRU + TPU - 1) / TPU;
  constexpr int NRE = 2 * EY;                  // element rows
  constexpr int TPE = kBlock / EZ;
  constexpr int NEI = (NRE + TPE - 1) / TPE;
  const int tu = threadIdx.x / RU, cu = threadIdx.x - tu * RU;
  const int jz = z0 - 1 + cu / 3;
  const bool zin = (tu < TPU) & ((unsigned)jz < (unsigned)g.NZ);
  const int dz = 3 * (z0 - 1) + cu;  // DOF offset of this position in a row
  const int te = threadIdx.x / EZ, ce = threadIdx.x - te * EZ;
  const int cz = z0 - 1 + ce;
  const bool ezin = (te < TPE) & ((unsigned)cz < (unsigned)g.nz);
  T uv[NUI], ev[NEI];
#pragma unroll
  for (int k = 0; k < NUI; ++k) {
    const int r = tu + k * TPU;
    const int xx = r / WY, yy = r - xx * WY;
    const int jx = ix - 1 + xx, jy = y0 - 1 + yy;
    const bool in = zin & (r < NRU) & ((unsigned)jx < (unsigned)g.NX) &
                    ((unsigned)jy < (unsigned)g.NY);
    uv[k] = map_load<OFF>(map, u, 3 * g.NZ * (jx * g.NY + jy) + dz, in);
  }
#pragma unroll
  for (int k = 0; k < NEI; ++k) {
    const int r = te + k * TPE;
    const int ex = r / EY, ey = r - ex * EY;
    const int cx = ix - 1 + ex, cy = y0 - 1 + ey;
    const bool in = ezin & (r < NRE) & ((unsigned)cx < (unsigned)g.nx) &
                    ((unsigned)cy < (unsigned)g.ny);
    const T e = E[in ? (cx * g.ny + cy) * g.nz + cz : 0];
    ev[k] = in ? e : T(0);
  }
#pragma unroll
  for (int k = 0; k < NUI; ++k) {
    const int r = tu + k * TPU;
    if ((tu < TPU) & (r < NRU)) us[r * RU + cu] = uv[k];
  }
#pragma unroll
  for (int k = 0; k < NEI; ++k) {
    const int r = te + k * TPE;
    if ((te < TPE) & (r < NRE)) es[r * EZ + ce] = ev[k];
  }
  __syncthreads();
}

// LDS-tiled form of the four-wave kernel.  A block covers a TY x TZ patch of
// nodes in one x-plane (64 nodes) and first stages the u values of the
// 3 x (TY+2) x (TZ+2) node window and the moduli of the 2 x (TY+1) x (TZ+1)
// element window in LDS (stage_window).  Wave w then forms the element pair
// (a,b) = (w>>1, w&1) exactly as k_fine_matvec_wave4 does, reading its window
// from LDS, and the pair sums meet in the same fixed tree, so the result is
// bitwise identical.  One global read per window entry replaces the
// 36-per-lane gathers.
template <typename T, int EPI, bool OFF, int TZ>
__global__ __launch_bounds__(kBlock) void k_fine_matvec_tile(Grid g, DofMap map,
                                                             const T* __restrict__ Ke,
                                                             const T* __restrict__ E,
                                                             const T* __restrict__ u,
                                                             T* __restrict__ y, EpiArgs<T> ep,
                                                             const int* gate) {
  using A = typename Acc<T>::type;
  constexpr int TY = 64 / TZ;
  constexpr int WY = TY + 2, WZ = TZ + 2;
  constexpr int EY = TY + 1, EZ = TZ + 1;
  constexpr int NU = 3 * WY * WZ * 3, NE = 2 * EY * EZ;
  if (gate && *gate == 0) return;
  __shared__ T us[NU];
  __shared__ T es[NE];
  __shared__ A part[3][3][64];
  const int tiles_z = (g.NZ + TZ - 1) / TZ, tiles_y = (g.NY + TY - 1) / TY;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tz = tile % tiles_z, ty = (tile / tiles_z) % tiles_y, ix = tile / (tiles_z * tiles_y);
  const int y0 = ty * TY, z0 = tz * TZ;
  stage_window<T, OFF, TZ>(g, map, E, u, us, es, ix, y0, z0);
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int a = wv >> 1, b = wv & 1;
  const int ly = lane / TZ, lz = lane % TZ;
  const int iy = y0 + ly, iz = z0 + lz;
  const bool valid = (iy < g.NY) & (iz < g.NZ);
  T w[2][2][3][3];
#pragma unroll
  for (int px = 0; px < 2; ++px)
#pragma unroll
    for (int py = 0; py < 2; ++py)
#pragma unroll
      for (int pz = 0; pz < 3; ++pz)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          w[px][py][pz][q] = us[(((a + px) * WY + (ly + b + py)) * WZ + (lz + pz)) * 3 + q];
  A P[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const T e = es[(a * EY + (ly + b)) * EZ + (lz + c)];
    T t0, t1, t2;
    pair_rows<T>(Ke, w, a, b, c, t0, t1, t2);
    P[0] += (A)(e * t0);
    P[1] += (A)(e * t1);
    P[2] += (A)(e * t2);
  }
  if (wv > 0) {
#pragma unroll
    for (int r = 0; r < 3; ++r) part[wv - 1][r][lane] = P[r];
  }
  __syncthreads();
  if (wv != 0 || !valid) return;
  A r[3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    r[q] = (P[q] + part[0][q][lane]) + (part[1][q][lane] + part[2][q][lane]);
  const int n = (ix * g.NY + iy) * g.NZ + iz;
  const int o[3] = {map_idx<OFF>(map, 3 * n), map_idx<OFF>(map, 3 * n + 1),
                    map_idx<OFF>(map, 3 * n + 2)};
  const T ax[3] = {(T)r[0], (T)r[1], (T)r[2]};
  epilogue3<T, EPI>(o, ax, u, y, ep);
}

// FP64 LDS tile with the element-row contraction on the matrix cores
// (v_mfma_f64_4x4x4f64: four independent 4x4x4 blocks per instruction).
//
// Same tiles, window staging, element pair per wave and summation tree as
// k_fine_matvec_tile.  Wave w = (a,b) needs, for each of its 64 nodes and
// c = 0, 1, the three row products t_r = sum_k Ke[3L+r][k] u_e[k] of element
// e = (ix-1+a, iy-1+b, iz-1+c), L the node's local index in e.  One
// instruction takes 16 nodes (4 blocks x 4 nodes) with the rows r on the M
// axis (row 3 a zero pad) and 4 consecutive k on the K axis; six chained
// instructions cover k = 0..23 in order.  Each FP64 4x4x4 block forms
// D = fma(a3,b3, fma(a2,b2, fma(a1,b1, fma(a0,b0,C)))) -- the k-ordered chain
// of fused multiply-adds, bit for bit (checked on gfx950 against an exact
// emulation) -- so the chained instructions reproduce the VALU kernels' row
// products exactly, and the moduli enter through the same explicit
// P = fma(e1, t1, fma(e0, t0, 0)) as the contracted P += e*t there: all launch
// shapes stay bitwise identical.
// Operand lanes of the instruction: A lane 16k+4g+i holds A_g[i][k], B lane
// 16k+4g+j holds B_g[k][j], result lane 16i+4g+j holds D_g[i][j].  Here
// A_g[i][k] = Ke[3L+i][4s+k] (the same for every block g), B_g[k][j] =
// u_e[4s+k] of node 16*grp+4g+j, and D_g[i][j] = t_i of that node.
template <int EPI, bool OFF, int TZ>
__global__ __launch_bounds__(kBlock) void k_fine_matvec_mfma(Grid g, DofMap map,
                                                             const double* __restrict__ Ke,
                                                             const double* __restrict__ E,
                                                             const double* __restrict__ u,
                                                             double* __restrict__ y,
                                                             EpiArgs<double> ep, const int* gate) {
  constexpr int TY = 64 / TZ;
  constexpr int WY = TY + 2, WZ = TZ + 2;
  constexpr int EY = TY + 1, EZ = TZ + 1;
  constexpr int NU = 3 * WY * WZ * 3, NE = 2 * EY * EZ;
  if (gate && *gate == 0) return;
  __shared__ double us[NU];
  __shared__ double es[NE];
  __shared__ double part[3][3][64];
  const int tiles_z = (g.NZ + TZ - 1) / TZ, tiles_y = (g.NY + TY - 1) / TY;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tz = tile % tiles_z, ty = (tile / tiles_z) % tiles_y, ix = tile / (tiles_z * tiles_y);
  const int y0 = ty * TY, z0 = tz * TZ;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int a = wv >> 1, b = wv & 1;
  const int hi = lane >> 4;         // k of the A and B operands, row i of the result
  const int blk = (lane >> 2) & 3;  // 4x4x4 block
  const int lo = lane & 3;          // row i of A; node j of B and of the result
  // A operands (issued before the staging loads): rows 3L+lo of Ke (lo = 3
  // pads with zeros), columns 4s+hi
  double ak[2][6];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int L = lidx(1 - a, 1 - b, 1 - c);
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const double kv = Ke[(3 * L + (lo < 3 ? lo : 0)) * 24 + 4 * s + hi];
      ak[c][s] = lo < 3 ? kv : 0.0;
    }
  }
  stage_window<double, OFF, TZ>(g, map, E, u, us, es, ix, y0, z0);
  // B operand: element-local DOF kk = 4s+hi is component kk%3 of corner
  // kk/3, whose window offset from the node's (ly, lz) entry is fixed per lane
  int boff[2][6];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int kk = 4 * s + hi, jc = kk / 3, q = kk - 3 * jc;
      boff[c][s] = (((a + lpos(jc, 0)) * WY + (b + lpos(jc, 1))) * WZ + (c + lpos(jc, 2))) * 3 + q;
    }
  double P[4];
#pragma unroll
  for (int grp = 0; grp < 4; ++grp) {
    const int nd = 16 * grp + 4 * blk + lo;
    const int ly = nd / TZ, lz = nd % TZ;
    const int nb = (ly * WZ + lz) * 3;
    double d0 = 0.0, d1 = 0.0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      d0 = __builtin_amdgcn_mfma_f64_4x4x4f64(ak[0][s], us[nb + boff[0][s]], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f64_4x4x4f64(ak[1][s], us[nb + boff[1][s]], d1, 0, 0, 0);
    }
    const double e0 = es[(a * EY + (ly + b)) * EZ + lz];
    const double e1 = es[(a * EY + (ly + b)) * EZ + lz + 1];
    P[grp] = fma(e1, d1, fma(e0, d0, 0.0));
  }
  if (wv > 0 && hi < 3) {
#pragma unroll
    for (int grp = 0; grp < 4; ++grp) part[wv - 1][hi][16 * grp + 4 * blk + lo] = P[grp];
  }
  __syncthreads();
  if (wv != 0 || hi >= 3) return;
#pragma unroll
  for (int grp = 0; grp < 4; ++grp) {
    const int nd = 16 * grp + 4 * blk + lo;
    const int ly = nd / TZ, lz = nd % TZ;
    const int iy = y0 + ly, iz = z0 + lz;
    if ((iy < g.NY) & (iz < g.NZ)) {
      const double r =
          (P[grp] + part[0][hi][nd]) + (part[1][hi][nd] + part[2][hi][nd]);
      const int o = map_idx<OFF>(map, 3 * ((ix * g.NY + iy) * g.NZ + iz) + hi);
      if (o >= 0) epilogue1<double, EPI>(o, r, u, y, ep);
    }
  }
}

// Node-per-lane box form.  A block of BX*BY*BZ lanes owns a BX x BY x BZ box
// of nodes (z fastest), one node per lane.  It stages the u values of the
// (BX+2) x (BY+2) x (BZ+2) node window (a z-row's 3(BZ+2) DOFs as in memory,
// rows RP entries apart) and the moduli of the (BX+1) x (BY+1) x (BZ+1) element window
// in LDS; then each lane forms all 8 element contributions of its node with
// the one-lane-per-node kernel's arithmetic: the 24-term row products in
// column order as fused multiply-adds from zero, P_ab = E_ab1 t_ab1 +
// E_ab0 t_ab0 in the contracted association, Q_a = P_a0 + P_a1, y = Q_0 + Q_1
// -- bitwise identical to every other shape.  Against the four-wave tile a
// node costs one lane instead of four: no cross-wave pair sums, one epilogue,
// one window decomposition, and (BX+2)(BY+2)(BZ+2) / (BX BY BZ) staged nodes
// per owned node (2.3 for 4x8x8) instead of 4.7.  A lane walks its 3x3x3
// neighbourhood in x and holds at most two x-planes (54 values) at once.
template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c);
template <>
__device__ __forceinline__ double fmadd<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmadd<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

// P += E t for one element of the tree: FP64 contracted into one fused
// multiply-add (as the other shapes compile `P += e * t`); FP32 the rounded
// FP32 product added in FP64.
__device__ __forceinline__ double tree_acc(double p, double e, double t) {
  return __builtin_fma(e, t, p);
}
__device__ __forceinline__ double tree_acc(double p, float e, float t) {
  return p + (double)__fmul_rn(e, t);
}

// Row products of element (a, b, c) from the x-planes a and a + 1 of the
// node's 3x3x3 neighbourhood, w[px][yy][zz][q] (corner (px, py, pz) of the
// element at yy = b + py, zz = c + pz), K the element's Ke rows 3L..3L+2:
// the 24-term products in column order as fused multiply-adds from zero.
template <typename T>
__device__ __forceinline__ void box_rows(const T* __restrict__ K, const T (&w)[2][3][3][3], int b,
                                         int c, T& t0, T& t1, T& t2) {
  t0 = 0;
  t1 = 0;
  t2 = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int px = lpos(j, 0), py = lpos(j, 1), pz = lpos(j, 2);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T uu = w[px][b + py][c + pz][q];
      t0 = fmadd<T>(K[3 * j + q], uu, t0);
      t1 = fmadd<T>(K[24 + 3 * j + q], uu, t1);
      t2 = fmadd<T>(K[48 + 3 * j + q], uu, t2);
    }
  }
}

// Stage the u window ((BX+2) x (BY+2) x (BZ+2) nodes; entry (xx, yy, zz, q)
// at (xx (BY+2) + yy) RP + 3 zz + q) and the E window ((BX+1) x (BY+1) x
// (BZ+1) elements, z-rows EZP entries apart) of the box at (x0, y0, z0).  A window
// z-row is 3(BZ+2) consecutive DOFs; a thread keeps one position of a row.
// Main pass: thread (yy, cu) loads row yy < TPU of every x-plane, so its
// address moves by one constant plane stride per load and only the plane's
// bounds test (uniform) changes.  The rows TPU..BY+1 of all planes follow as
// a flat list of rows.  E: a thread keeps one (y, z) entry of an element
// plane and takes the planes p, p + PPE, ...  All loads are issued before
// the first LDS write; out-of-grid and fixed entries read 0.  Ends with a
// barrier.
template <typename T, bool OFF, int BX, int BY, int BZ, int RP, int EZP, int NT>
__device__ __forceinline__ void stage_box(const Grid& g, const DofMap& map,
                                          const T* __restrict__ E, const T* __restrict__ u,
                                          T* us, T* es, int x0, int y0, int z0) {
  constexpr int WX = BX + 2, WY = BY + 2, WZ = BZ + 2;
  constexpr int RU = 3 * WZ;
  constexpr int TPU = NT / RU < WY ? NT / RU : WY;  // rows per plane in the main pass
  constexpr int RT = WY - TPU;                      // tail rows per plane
  constexpr int NRT = WX * RT;                      // tail rows in all
  constexpr int NTI = (NRT + TPU - 1) / TPU;
  constexpr int EX = BX + 1, EY = BY + 1, EZ = BZ + 1;
  constexpr int PLE = EY * EZ;                      // entries of an element plane
  constexpr int PLP = EY * EZP;                     // its LDS extent (rows padded)
  constexpr int PPE = NT / PLE;                     // element planes per pass
  constexpr int NEI = (EX + PPE - 1) / PPE;
  static_assert(TPU >= 1 && PPE >= 1, "box too small for its window rows");
  const int plane_u = 3 * g.NZ * g.NY;
  const int tu = threadIdx.x / RU, cu = threadIdx.x - tu * RU;
  const int cz = cu / 3;
  const int jz = z0 - 1 + cz;
  const bool zin = (tu < TPU) & ((unsigned)jz < (unsigned)g.NZ);
  const int dz = 3 * (z0 - 1) + cu;  // DOF offset of this position in a row
  const int jy = y0 - 1 + tu;
  const bool rin = zin & ((unsigned)jy < (unsigned)g.NY);
  const int dmain = 3 * g.NZ * jy + dz + (x0 - 1) * plane_u;
  T uv[WX], ut[NTI > 0 ? NTI : 1];
#pragma unroll
  for (int xx = 0; xx < WX; ++xx) {
    const bool in = rin & ((unsigned)(x0 - 1 + xx) < (unsigned)g.NX);
    uv[xx] = map_load<OFF>(map, u, dmain + xx * plane_u, in);
  }
#pragma unroll
  for (int k = 0; k < NTI; ++k) {
    const int r = tu + k * TPU;
    const int xx = r / (RT > 0 ? RT : 1), yy = TPU + (r - xx * RT);
    const int jx = x0 - 1 + xx, jyt = y0 - 1 + yy;
    const bool in = zin & (r < NRT) & ((unsigned)jx < (unsigned)g.NX) &
                    ((unsigned)jyt < (unsigned)g.NY);
    ut[k] = map_load<OFF>(map, u, 3 * g.NZ * (jx * g.NY + jyt) + dz, in);
  }
  const int pe = threadIdx.x / PLE, re = threadIdx.x - pe * PLE;
  const int ey = re / EZ, ez = re - ey * EZ;
  const int cy = y0 - 1 + ey, czz = z0 - 1 + ez;
  const bool yzin =
      (pe < PPE) & ((unsigned)cy < (unsigned)g.ny) & ((unsigned)czz < (unsigned)g.nz);
  const int emain = ((x0 - 1 + pe) * g.ny + cy) * g.nz + czz;
  T ev[NEI];
#pragma unroll
  for (int k = 0; k < NEI; ++k) {
    const int p = pe + k * PPE;
    const bool in = yzin & (p < EX) & ((unsigned)(x0 - 1 + p) < (unsigned)g.nx);
    const T e = E[in ? emain + k * PPE * g.ny * g.nz : 0];
    ev[k] = in ? e : T(0);
  }
  T* ub = us + cu;
  if (tu < TPU) {
#pragma unroll
    for (int xx = 0; xx < WX; ++xx) ub[(xx * WY + tu) * RP] = uv[xx];
  }
#pragma unroll
  for (int k = 0; k < NTI; ++k) {
    const int r = tu + k * TPU;
    const int xx = r / (RT > 0 ? RT : 1), yy = TPU + (r - xx * RT);
    if ((tu < TPU) & (r < NRT)) ub[(xx * WY + yy) * RP] = ut[k];
  }
#pragma unroll
  for (int k = 0; k < NEI; ++k) {
    const int p = pe + k * PPE;
    if ((pe < PPE) & (p < EX)) es[p * PLP + ey * EZP + ez] = ev[k];
  }
  __syncthreads();
}

// The elements with a given a (the x side of the node) in a plane step: the
// x-planes a and a + 1 of the neighbourhood are read from LDS (p0 and p1 at
// the lane's (y, z) window entry; 54 values) with the element plane a (pe at
// the lane's entry) and the four elements (a, b, c) are formed; returns
// Q_a = (E_a00 t_a00 + E_a01 t_a01) + (E_a10 t_a10 + E_a11 t_a11) as the tree
// has it.  Ke rows stream through the scalar registers (288 values).
template <typename T, int RP, int EZ>
__device__ __forceinline__ void box_plane_step(const T* __restrict__ Ke, const T* p0, const T* p1,
                                               const T* pe, int a,
                                               typename Acc<T>::type (&Qa)[3]) {
  using A = typename Acc<T>::type;
  T w[2][3][3][3];
#pragma unroll
  for (int yy = 0; yy < 3; ++yy)
#pragma unroll
    for (int zz = 0; zz < 3; ++zz)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        w[0][yy][zz][q] = p0[yy * RP + 3 * zz + q];
        w[1][yy][zz][q] = p1[yy * RP + 3 * zz + q];
      }
  A P[2][3];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    P[b][0] = P[b][1] = P[b][2] = A(0);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const T e = pe[b * EZ + c];
      const int L = lidx(1 - a, 1 - b, 1 - c);
      T t0, t1, t2;
      box_rows<T>(Ke + 3 * L * 24, w, b, c, t0, t1, t2);
      P[b][0] = tree_acc(P[b][0], e, t0);
      P[b][1] = tree_acc(P[b][1], e, t1);
      P[b][2] = tree_acc(P[b][2], e, t2);
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) Qa[r] = P[0][r] + P[1][r];
}

// AS lanes per node: AS = 1 runs a = 0 and a = 1 as a loop that is not
// unrolled (fully unrolled, the compiler hoists the loads of all 576 Ke
// values and the whole neighbourhood, spilling scalar registers and halving
// the waves per SIMD); AS = 2 gives a = 1 to a second group of waves, whose
// Q_1 meets the first group's Q_0 in LDS (y = Q_0 + Q_1, the same sum) --
// twice the waves for grids too small to fill the device one lane per node.
template <typename T, int EPI, bool OFF, int BX, int BY, int BZ, int RP, int EZP, int AS>
__global__ __launch_bounds__(AS* BX* BY* BZ) void k_fine_matvec_box(Grid g, DofMap map,
                                                                     const T* __restrict__ Ke,
                                                                     const T* __restrict__ E,
                                                                     const T* __restrict__ u,
                                                                     T* __restrict__ y,
                                                                     EpiArgs<T> ep,
                                                                     const int* gate) {
  using A = typename Acc<T>::type;
  constexpr int NB = BX * BY * BZ;
  constexpr int WY = BY + 2;
  constexpr int EY = BY + 1;
  static_assert(AS == 1 || (AS == 2 && NB % 64 == 0), "lane groups must be whole waves");
  const bool open = gate_open(gate);  // tested once the window is staged
  __shared__ T us[(BX + 2) * WY * RP];
  __shared__ T es[(BX + 1) * EY * EZP];
  __shared__ A q1[AS == 2 ? 3 * NB : 1];
  const int tiles_z = (g.NZ + BZ - 1) / BZ, tiles_y = (g.NY + BY - 1) / BY;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tz = tile % tiles_z, ty = (tile / tiles_z) % tiles_y, tx = tile / (tiles_z * tiles_y);
  const int x0 = tx * BX, y0 = ty * BY, z0 = tz * BZ;
  stage_box<T, OFF, BX, BY, BZ, RP, EZP, AS * NB>(g, map, E, u, us, es, x0, y0, z0);
  if (!open) return;  // uniform over the block, after stage_box's barrier
  const int grp = AS == 2 ? __builtin_amdgcn_readfirstlane((int)threadIdx.x / NB) : 0;
  const int t = threadIdx.x - grp * NB;
  const int lz = t % BZ, ly = (t / BZ) % BY, lx = t / (BZ * BY);
  const int ix = x0 + lx, iy = y0 + ly, iz = z0 + lz;
  const T* ub = us + (lx * WY + ly) * RP + 3 * lz;
  const T* eb = es + (lx * EY + ly) * EZP + lz;
  A ax[3];
  int o0, o1, o2;
  EpiOps3<T> ev;  // epilogue operands, loaded before the arithmetic
  if constexpr (AS == 1) {
    if ((ix >= g.NX) | (iy >= g.NY) | (iz >= g.NZ)) return;
    const int n = (ix * g.NY + iy) * g.NZ + iz;
    o0 = map_idx<OFF>(map, 3 * n);
    o1 = map_idx<OFF>(map, 3 * n + 1);
    o2 = map_idx<OFF>(map, 3 * n + 2);
    if ((o0 & o1 & o2) < 0) return;
    const int o[3] = {o0, o1, o2};
    epi_load3<T, EPI>(o, u, ep, ev);
    A Q0[3];
#pragma unroll 1
    for (int a = 0; a < 2; ++a) {
      A Qa[3];
      box_plane_step<T, RP, EZP>(Ke, ub + a * WY * RP, ub + (a + 1) * WY * RP,
                                 eb + a * EY * EZP, a, Qa);
      if (a == 0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) Q0[r] = Qa[r];
      } else {
#pragma unroll
        for (int r = 0; r < 3; ++r) ax[r] = Q0[r] + Qa[r];
      }
    }
  } else {
    const bool in = (ix < g.NX) & (iy < g.NY) & (iz < g.NZ);
    const int n = in ? (ix * g.NY + iy) * g.NZ + iz : 0;
    o0 = map_idx<OFF>(map, 3 * n);
    o1 = map_idx<OFF>(map, 3 * n + 1);
    o2 = map_idx<OFF>(map, 3 * n + 2);
    const bool live = in & ((o0 & o1 & o2) >= 0);
    if (grp == 0) {
      const int o[3] = {o0, o1, o2};
      epi_load3<T, EPI>(o, u, ep, ev);
    }
    A Qa[3] = {A(0), A(0), A(0)};
    if (live)
      box_plane_step<T, RP, EZP>(Ke, ub + grp * WY * RP, ub + (grp + 1) * WY * RP,
                                 eb + grp * EY * EZP, grp, Qa);
    if (grp == 1) {
#pragma unroll
      for (int r = 0; r < 3; ++r) q1[r * NB + t] = Qa[r];
    }
    __syncthreads();
    if (grp == 1 || !live) return;
#pragma unroll
    for (int r = 0; r < 3; ++r) ax[r] = Qa[r] + q1[r * NB + t];
  }
  const int o[3] = {o0, o1, o2};
  const T axt[3] = {(T)ax[0], (T)ax[1], (T)ax[2]};
  epi_store3<T, EPI>(o, axt, ev, y, ep);
}

// Element strain energy and SIMP sensitivity (solver_v4.py:258-265):
//   ce_e = u_e^T KE u_e  (unit modulus; KUe_j = sum_i u_i KE_ij in i order,
//          then ce = sum_j KUe_j u_j in j order),
//   dc_e = -penal * scale * rho_e^(penal-1) * ce_e,  scale = E0 - Emin.
// One lane per element; u_e is gathered through the DOF map (fixed DOFs
// read 0, as the reference's full-length U with zeros at fixed DOFs).
template <bool OFF>
__global__ __launch_bounds__(kBlock) void k_elem_sens(Grid g, DofMap map,
                                                      const double* __restrict__ Ke,
                                                      const double* __restrict__ u,
                                                      const double* __restrict__ rho, double penal,
                                                      double scale, double* __restrict__ ce,
                                                      double* __restrict__ dc) {
  const long ne = (long)g.nx * g.ny * g.nz;
  const long e = (long)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int ez = (int)(e % g.nz), ey = (int)((e / g.nz) % g.ny), ex = (int)(e / ((long)g.ny * g.nz));
  double ue[24];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int node = ((ex + lpos(a, 0)) * g.NY + (ey + lpos(a, 1))) * g.NZ + (ez + lpos(a, 2));
#pragma unroll
    for (int q = 0; q < 3; ++q) ue[3 * a + q] = map_load<OFF>(map, u, 3 * node + q, true);
  }
  double c = 0.0;
#pragma unroll
  for (int j = 0; j < 24; ++j) {
    double k = 0.0;
#pragma unroll
    for (int i = 0; i < 24; ++i) k += ue[i] * Ke[i * 24 + j];
    c += k * ue[j];
  }
  if (ce) ce[e] = c;
  if (dc) dc[e] = -penal * scale * pow(rho[e], penal - 1.0) * c;
}

int elem_sensitivity(const Grid& g, const DofMap& map, const double* Ke, const double* u,
                     const double* rho, double penal, double scale, double* ce, double* dc,
                     hipStream_t st) {
  const long ne = (long)g.nx * g.ny * g.nz;
  if (map.off >= 0)
    hipLaunchKernelGGL(k_elem_sens<true>, dim3(cdiv(ne, kBlock)), dim3(kBlock), 0, st, g, map, Ke,
                       u, rho, penal, scale, ce, dc);
  else
    hipLaunchKernelGGL(k_elem_sens<false>, dim3(cdiv(ne, kBlock)), dim3(kBlock), 0, st, g, map,
                       Ke, u, rho, penal, scale, ce, dc);
  GF_CHECK_LAUNCH();
  return 0;
}

// launch shape of contexts that were not autotuned: the 2x8x8 box, the
// autotuner's choice at 64k, 512k and 1M elements
static std::atomic<int> g_split_f64{kShapeBox288}, g_split_f32{kShapeBox288};

static bool valid_split(int s) {
  return s == 1 || s == 4 || s == 8 || s == 16 || s == kShapeMfma8 || s == kShapeMfma16 ||
         s == kShapeBox488 || s == kShapeBox2816 || s == kShapeBox288;
}

void set_fine_split(int split_f64, int split_f32) {
  if (valid_split(split_f64)) g_split_f64 = split_f64;
  if (valid_split(split_f32)) g_split_f32 = split_f32;
}
int get_fine_split(int is_f64) { return is_f64 ? g_split_f64 : g_split_f32; }

template <typename T, int EPI, bool OFF, int BX, int BY, int BZ, int AS>
static void launch_box(const Grid& g, const DofMap& map, const T* Ke, const T* E, const T* u,
                       T* y, const EpiArgs<T>& ep, const int* gate, hipStream_t st) {
  // LDS row pitches: with 8 nodes per z-row a 16-lane group spans two rows,
  // which fall on disjoint banks when the pitch is 8 (mod 16) doubles (u
  // rows: 30 -> 40; E rows: 9 -> 24); a 16-node row is conflict-free alone
  constexpr int RP = BZ == 8 ? 40 : 3 * (BZ + 2), EZP = BZ == 8 ? 24 : BZ + 1;
  const int tiles = cdiv(g.NX, BX) * cdiv(g.NY, BY) * cdiv(g.NZ, BZ);
  hipLaunchKernelGGL((k_fine_matvec_box<T, EPI, OFF, BX, BY, BZ, RP, EZP, AS>), dim3(tiles),
                     dim3(AS * BX * BY * BZ), 0, st, g, map, Ke, E, u, y, ep, gate);
}

template <typename T, int EPI, bool OFF>
static void launch_fine_m(int split, const Grid& g, const DofMap& map, const T* Ke, const T* E,
                          const T* u, T* y, const EpiArgs<T>& ep, const int* gate,
                          hipStream_t st) {
  if (split == kShapeBox488)
    return launch_box<T, EPI, OFF, 4, 8, 8, 1>(g, map, Ke, E, u, y, ep, gate, st);
  if (split == kShapeBox2816)
    return launch_box<T, EPI, OFF, 2, 8, 16, 1>(g, map, Ke, E, u, y, ep, gate, st);
  if (split == kShapeBox288)
    return launch_box<T, EPI, OFF, 2, 8, 8, 2>(g, map, Ke, E, u, y, ep, gate, st);
  if (split == kShapeMfma8 || split == kShapeMfma16) {
    // matrix-core contraction for FP64; the FP32 operator takes the VALU
    // tile of the same patch
    const int tz = split == kShapeMfma8 ? 8 : 16;
    if constexpr (std::is_same<T, double>::value) {
      const int tiles = g.NX * cdiv(g.NY, 64 / tz) * cdiv(g.NZ, tz);
      if (tz == 8)
        hipLaunchKernelGGL((k_fine_matvec_mfma<EPI, OFF, 8>), dim3(tiles), dim3(kBlock), 0, st,
                           g, map, Ke, E, u, y, ep, gate);
      else
        hipLaunchKernelGGL((k_fine_matvec_mfma<EPI, OFF, 16>), dim3(tiles), dim3(kBlock), 0, st,
                           g, map, Ke, E, u, y, ep, gate);
      return;
    }
    split = tz;
  }
  if (split == 1) {
    hipLaunchKernelGGL((k_fine_matvec_node<T, EPI, OFF>), dim3(cdiv(g.n_nodes, kBlock)),
                       dim3(kBlock), 0, st, g, map, Ke, E, u, y, ep, gate);
  } else if (split == 8 || split == 16) {
    const int tz = split, ty = 64 / tz;
    const int tiles = g.NX * cdiv(g.NY, ty) * cdiv(g.NZ, tz);
    if (split == 8)
      hipLaunchKernelGGL((k_fine_matvec_tile<T, EPI, OFF, 8>), dim3(tiles), dim3(kBlock), 0, st,
                         g, map, Ke, E, u, y, ep, gate);
    else
      hipLaunchKernelGGL((k_fine_matvec_tile<T, EPI, OFF, 16>), dim3(tiles), dim3(kBlock), 0,
                         st, g, map, Ke, E, u, y, ep, gate);
  } else {
    hipLaunchKernelGGL((k_fine_matvec_wave4<T, EPI, OFF>), dim3(cdiv(g.n_nodes, 64)),
                       dim3(kBlock), 0, st, g, map, Ke, E, u, y, ep, gate);
  }
}

template <typename T, int EPI>
static void launch_fine(int split, const Grid& g, const DofMap& map, const T* Ke, const T* E,
                        const T* u, T* y, const EpiArgs<T>& ep, const int* gate, hipStream_t st) {
  if (map.off >= 0)
    launch_fine_m<T, EPI, true>(split, g, map, Ke, E, u, y, ep, gate, st);
  else
    launch_fine_m<T, EPI, false>(split, g, map, Ke, E, u, y, ep, gate, st);
}

template <typename T>
int fine_matvec(const Grid& g, const DofMap& map, const T* Ke, const T* E, const T* u, T* y,
                int epi, const EpiArgs<T>& ep, const int* gate, hipStream_t st) {
  const int split = (sizeof(T) == 4 && g.shape32) ? g.shape32
                    : g.shape                       ? g.shape
                                                    : get_fine_split(sizeof(T) == 8);
  switch (epi) {
    case EPI_PLAIN: launch_fine<T, EPI_PLAIN>(split, g, map, Ke, E, u, y, ep, gate, st); break;
    case EPI_RESID: launch_fine<T, EPI_RESID>(split, g, map, Ke, E, u, y, ep, gate, st); break;
    case EPI_CHEB0: launch_fine<T, EPI_CHEB0>(split, g, map, Ke, E, u, y, ep, gate, st); break;
    case EPI_CHEB: launch_fine<T, EPI_CHEB>(split, g, map, Ke, E, u, y, ep, gate, st); break;
    case EPI_JACOBI: launch_fine<T, EPI_JACOBI>(split, g, map, Ke, E, u, y, ep, gate, st); break;
    default:
      set_error("fine_matvec: bad epilogue");
      return -2;
  }
  GF_CHECK_LAUNCH();
  return 0;
}
template int fine_matvec<double>(const Grid&, const DofMap&, const double*, const double*,
                                 const double*, double*, int, const EpiArgs<double>&, const int*,
                                 hipStream_t);
template int fine_matvec<float>(const Grid&, const DofMap&, const float*, const float*,
                                const float*, float*, int, const EpiArgs<float>&, const int*,
                                hipStream_t);

template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_start(long n, const T* __restrict__ b,
                                                       const T* __restrict__ dinv, T a,
                                                       T* __restrict__ dv, T* __restrict__ x,
                                                       const int* gate) {
  const bool open = gate_open(gate);  // tested after the first element's loads
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const T di = dinv[i], bi = b[i];
    if (!open) return;
    // explicitly rounded, as in the restriction's fused form (ZeroStart)
    if (dv) {
      const T z = rn_mul(di, bi);
      dv[i] = z;
      x[i] = rn_add(T(0), rn_mul(a, z));
    } else {
      x[i] = rn_add(T(0), rn_mul(rn_mul(a, di), bi));
    }
  }
}

template <typename T>
int zero_start_step(long n, const T* b, const T* dinv, T a, T* dv, T* x, const int* gate,
                    hipStream_t st) {
  int grid = cdiv(n, kBlock);
  if (grid > 4096) grid = 4096;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_zero_start<T>, dim3(grid), dim3(kBlock), 0, st, n, b, dinv, a, dv, x,
                     gate);
  GF_CHECK_LAUNCH();
  return 0;
}
template int zero_start_step<double>(long, const double*, const double*, double, double*, double*,
                                     const int*, hipStream_t);
template int zero_start_step<float>(long, const float*, const float*, float, float*, float*,
                                    const int*, hipStream_t);

// diag(K_ff)_dof = sum_{e adjacent, ascending} E_e * Ke[jj, jj]
__global__ __launch_bounds__(kBlock) void k_fine_diag(Grid g, DofMap map,
                                                      const double* __restrict__ Ke,
                                                      const double* __restrict__ E,
                                                      double* __restrict__ diag) {
  const int n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int iz = n % g.NZ, t = n / g.NZ, iy = t % g.NY, ix = t / g.NY;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int ex = ix - 1 + a, ey = iy - 1 + b, ez = iz - 1 + c;
        if (ex < 0 || ex >= g.nx || ey < 0 || ey >= g.ny || ez < 0 || ez >= g.nz) continue;
        const double e = E[(ex * g.ny + ey) * g.nz + ez];
        const int L = lidx(1 - a, 1 - b, 1 - c);
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] += e * Ke[(3 * L + q) * 25];
      }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int f = map(3 * n + q);
    if (f >= 0) diag[f] = acc[q];
  }
}

int fine_diag(const Grid& g, const DofMap& map, const double* Ke, const double* E, double* diag,
              hipStream_t st) {
  hipLaunchKernelGGL(k_fine_diag, dim3(cdiv(g.n_nodes, kBlock)), dim3(kBlock), 0, st, g, map, Ke,
                     E, diag);
  GF_CHECK_LAUNCH();
  return 0;
}

__global__ __launch_bounds__(kBlock) void k_safe_inverse(long n, const double* __restrict__ d,
                                                         double* __restrict__ dinv,
                                                         float* __restrict__ dinv32) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const double v = d[i];
    const double s = fabs(v) > 1e-12 ? v : 1.0;
    const double r = 1.0 / s;
    dinv[i] = r;
    if (dinv32) dinv32[i] = (float)r;
  }
}

int safe_inverse(long n, const double* d, double* dinv, float* dinv32, hipStream_t st) {
  int grid = cdiv(n, kBlock);
  if (grid > 4096) grid = 4096;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_safe_inverse, dim3(grid), dim3(kBlock), 0, st, n, d, dinv, dinv32);
  GF_CHECK_LAUNCH();
  return 0;
}

// Node-block Jacobi: 3x3 node block sum_e E_e Ke[3L:3L+3, 3L:3L+3] (ascending
// elements), inverted by cofactors for nodes whose 3 DOFs are free; scalar
// safe inverse of the supplied diagonal for partially constrained nodes.
__global__ __launch_bounds__(kBlock) void k_nodeblock_update(Grid g, DofMap map,
                                                             const double* __restrict__ Ke,
                                                             const double* __restrict__ E,
                                                             const double* __restrict__ sdiag,
                                                             double* __restrict__ inv9,
                                                             double* __restrict__ pinv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int o[3] = {map(3 * n), map(3 * n + 1), map(3 * n + 2)};
  const bool all = (o[0] >= 0) & (o[1] >= 0) & (o[2] >= 0);
  const bool any = (o[0] >= 0) | (o[1] >= 0) | (o[2] >= 0);
  if (!any) return;
  if (!all) {
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (o[q] >= 0 && sdiag) {
        const double v = sdiag[o[q]];
        pinv[o[q]] = 1.0 / (fabs(v) > 1e-12 ? v : 1.0);
      }
    return;
  }
  const int iz = n % g.NZ, t = n / g.NZ, iy = t % g.NY, ix = t / g.NY;
  double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int ex = ix - 1 + a, ey = iy - 1 + b, ez = iz - 1 + c;
        if (ex < 0 || ex >= g.nx || ey < 0 || ey >= g.ny || ez < 0 || ez >= g.nz) continue;
        const double e = E[(ex * g.ny + ey) * g.nz + ez];
        const int L = lidx(1 - a, 1 - b, 1 - c);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int s = 0; s < 3; ++s) B[r * 3 + s] += e * Ke[(3 * L + r) * 24 + 3 * L + s];
      }
  const double c00 = B[4] * B[8] - B[5] * B[7];
  const double c01 = B[5] * B[6] - B[3] * B[8];
  const double c02 = B[3] * B[7] - B[4] * B[6];
  const double det = B[0] * c00 + B[1] * c01 + B[2] * c02;
  const double id = 1.0 / det;
  double* I = inv9 + 9L * n;
  I[0] = c00 * id;
  I[1] = (B[2] * B[7] - B[1] * B[8]) * id;
  I[2] = (B[1] * B[5] - B[2] * B[4]) * id;
  I[3] = c01 * id;
  I[4] = (B[0] * B[8] - B[2] * B[6]) * id;
  I[5] = (B[2] * B[3] - B[0] * B[5]) * id;
  I[6] = c02 * id;
  I[7] = (B[1] * B[6] - B[0] * B[7]) * id;
  I[8] = (B[0] * B[4] - B[1] * B[3]) * id;
}

int nodeblock_update(const Grid& g, const DofMap& map, const double* Ke, const double* E,
                     const double* scalar_diag, double* inv9, double* pinv, hipStream_t st) {
  hipLaunchKernelGGL(k_nodeblock_update, dim3(cdiv(g.n_nodes, kBlock)), dim3(kBlock), 0, st, g,
                     map, Ke, E, scalar_diag, inv9, pinv);
  GF_CHECK_LAUNCH();
  return 0;
}

__global__ __launch_bounds__(kBlock) void k_nodeblock_apply(Grid g, DofMap map,
                                                            const double* __restrict__ inv9,
                                                            const double* __restrict__ pinv,
                                                            int has_pinv,
                                                            const double* __restrict__ r,
                                                            double* __restrict__ z,
                                                            const int* gate) {
  const bool open = gate_open(gate);  // tested after the node's loads
  const int n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int o0 = map(3 * n), o1 = map(3 * n + 1), o2 = map(3 * n + 2);
  if ((o0 >= 0) & (o1 >= 0) & (o2 >= 0)) {
    const double* I = inv9 + 9L * n;
    const double r0 = r[o0], r1 = r[o1], r2 = r[o2];
    const double z0 = I[0] * r0 + I[1] * r1 + I[2] * r2;
    const double z1 = I[3] * r0 + I[4] * r1 + I[5] * r2;
    const double z2 = I[6] * r0 + I[7] * r1 + I[8] * r2;
    if (!open) return;
    z[o0] = z0;
    z[o1] = z1;
    z[o2] = z2;
    return;
  }
  if (!open) return;
  const int o[3] = {o0, o1, o2};
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (o[q] >= 0) z[o[q]] = has_pinv ? pinv[o[q]] * r[o[q]] : 0.0;
}

int nodeblock_apply(const Grid& g, const DofMap& map, const double* inv9, const double* pinv,
                    int has_pinv, const double* r, double* z, const int* gate, hipStream_t st) {
  hipLaunchKernelGGL(k_nodeblock_apply, dim3(cdiv(g.n_nodes, kBlock)), dim3(kBlock), 0, st, g,
                     map, inv9, pinv, has_pinv, r, z, gate);
  GF_CHECK_LAUNCH();
  return 0;
}

}  // namespace gf
