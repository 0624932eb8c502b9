// BF16 matrix-core fine operator: the fused matvec's bf16 form
// (FusedMatvec dtype="bf16", cuda_fused_matvec.py:138-246, 393-416) and the
// operator of the bf16 fine smoother (multigrid_v4.py:1765-1780, 1831-1848):
//
//   y = sum_e fl32(E_e * (KE_bf16 @ bf16(u_e))),  products accumulated in FP32,
//
// with KE and u rounded to bf16 (round to nearest even), E_e applied in FP32
// to the FP32 contraction and the element results summed per node in FP32,
// as the reference's kernel does with FP32 atomics.  Here the node sums run
// in one fixed order, so the result is deterministic.
//
// One kernel, owner-computes over node tiles; element results never leave
// LDS.  A workgroup (four waves) covers kTX = 4 x-planes of a TY x TZ node
// patch (256 nodes) and
//   1. stages the bf16 node window (kTX+2) x (TY+2) x (TZ+2) x 3 and the FP32
//      moduli of the (kTX+1) x (TY+1) x (TZ+1) element window in LDS;
//   2. contracts every element of the window on the matrix cores: a wave
//      takes 16 elements at a time and forms F = KE(32x32, zero padded) x
//      U(32x16) with two v_mfma_f32_16x16x32_bf16 (KE rows 0-15 and 16-31),
//      scales each column by its element's modulus and keeps the 24 results
//      of each element in LDS;
//   3. one lane per node (wave w = x-plane w of the tile) sums its three rows
//      over its 8 elements in corner order -- corner a = 0..7 of the node in
//      element (ix - px_a, iy - py_a, iz - pz_a) -- and applies the operator
//      epilogue.
// Elements on a tile's faces are contracted by every tile that touches them
// (425 contractions per 256 nodes at TZ = 16): the contraction runs on the
// matrix cores and costs little next to the window traffic.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l holds
// A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16]; the result lane l
// holds C[4*(l/16) .. +3][l%16].
#include "gf_internal.h"
#include "gf_solver.h"

namespace gf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTX = 4;  // x-planes per tile: one wave per plane in the node stage

template <int TZ>
struct Bf16Tile {
  static constexpr int TY = 64 / TZ;
  static constexpr int WX = kTX + 2, WY = TY + 2, WZ = TZ + 2;  // node window
  static constexpr int EX = kTX + 1, EY = TY + 1, EZ = TZ + 1;  // element window
  static constexpr int NU = WX * WY * WZ * 3;
  static constexpr int NEL = EX * EY * EZ;
  static constexpr int NT = (NEL + 15) / 16;  // 16-element contraction tiles
  // element results are stored row-major by DOF row: F[r * NELP + s].  With
  // NELP = 16*NT + 4 the four 16-lane groups of one store land 16 banks
  // apart, and the node stage reads consecutive s on consecutive lanes.
  static constexpr int NELP = 16 * NT + 4;
};

template <int EPI, bool OFF, int TZ>
__global__ __launch_bounds__(kBlock) void k_fine_matvec_bf16(Grid g, DofMap map,
                                                             const unsigned short* __restrict__ KEb,
                                                             const float* __restrict__ E,
                                                             const float* __restrict__ u,
                                                             float* __restrict__ y,
                                                             EpiArgs<float> ep, const int* gate) {
  using TT = Bf16Tile<TZ>;
  constexpr int TY = TT::TY, WY = TT::WY, WZ = TT::WZ, EY = TT::EY, EZ = TT::EZ;
  constexpr int NU = TT::NU, NEL = TT::NEL, NT = TT::NT, NELP = TT::NELP;
  constexpr int NUI = (NU + kBlock - 1) / kBlock;
  constexpr int NEI = (16 * NT + kBlock - 1) / kBlock;
  if (gate && *gate == 0) return;
  __shared__ unsigned short us[NU];
  __shared__ float es[16 * NT];
  __shared__ float F[24 * NELP];
  const int tiles_z = (g.NZ + TZ - 1) / TZ, tiles_y = (g.NY + TY - 1) / TY;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tz = tile % tiles_z, ty = (tile / tiles_z) % tiles_y, tx = tile / (tiles_z * tiles_y);
  const int x0 = tx * kTX, y0 = ty * TY, z0 = tz * TZ;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15;  // element of a contraction tile (B / C column)
  const int kg = lane >> 4;   // K slice 8*kg .. 8*kg+7 (A, B); C rows 4*kg .. 4*kg+3
  // A operands: KE rows col and 16+col, columns 8*kg .. 8*kg+7
  const u16x8 a0 = *reinterpret_cast<const u16x8*>(KEb + col * 32 + 8 * kg);
  const u16x8 a1 = *reinterpret_cast<const u16x8*>(KEb + (16 + col) * 32 + 8 * kg);

  // 1. window staging: every global load is issued before the first LDS write
  {
    float uv[NUI], ev[NEI];
#pragma unroll
    for (int k = 0; k < NUI; ++k) {
      const int t = threadIdx.x + k * kBlock;
      const int q = t % 3, r = t / 3;
      const int zz = r % WZ, r2 = r / WZ, yy = r2 % WY, xx = r2 / WY;
      const int jx = x0 - 1 + xx, jy = y0 - 1 + yy, jz = z0 - 1 + zz;
      const bool in = (t < NU) & (jx >= 0) & (jx < g.NX) & (jy >= 0) & (jy < g.NY) & (jz >= 0) &
                      (jz < g.NZ);
      uv[k] = map_load<OFF>(map, u, 3 * ((jx * g.NY + jy) * g.NZ + jz) + q, in);
    }
#pragma unroll
    for (int k = 0; k < NEI; ++k) {
      const int t = threadIdx.x + k * kBlock;
      const int ez = t % EZ, r = t / EZ, ey = r % EY, ex = r / EY;
      const int cx = x0 - 1 + ex, cy = y0 - 1 + ey, cz = z0 - 1 + ez;
      const bool in = (t < NEL) & (cx >= 0) & (cx < g.nx) & (cy >= 0) & (cy < g.ny) & (cz >= 0) &
                      (cz < g.nz);
      const float e = E[in ? (cx * g.ny + cy) * g.nz + cz : 0];
      ev[k] = in ? e : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < NUI; ++k) {
      const int t = threadIdx.x + k * kBlock;
      if (t < NU) us[t] = bf16_bits_rne(uv[k]);
    }
#pragma unroll
    for (int k = 0; k < NEI; ++k) {
      const int t = threadIdx.x + k * kBlock;
      if (t < 16 * NT) es[t] = ev[k];
    }
  }
  __syncthreads();

  // 2. element contraction.  B entry k = 3a + c (a corner, c component) of
  // element (ex, ey, ez) is window node (ex + px_a, ey + py_a, ez + pz_a),
  // component c: a fixed offset from the element's base entry per (lane, k).
  int boff[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int k = 8 * kg + t;
    const int a = k / 3, c = k - 3 * (k / 3);
    boff[t] = k < 24 ? ((lpos(a, 0) * WY + lpos(a, 1)) * WZ + lpos(a, 2)) * 3 + c : -1;
  }
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int t = wv; t < NT; t += 4) {
    const int s = 16 * t + col;
    const int ez = s % EZ, r = s / EZ, ey = r % EY, ex = r / EY;
    const int base = ((ex * WY + ey) * WZ + ez) * 3;
    const bool sv = s < NEL;
    u16x8 bv;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      // padding columns (s >= NEL) and K rows (k >= 24) read entry 0, then 0
      const bool ok = sv && boff[q] >= 0;
      const unsigned short v = us[ok ? base + boff[q] : 0];
      bv[q] = ok ? v : (unsigned short)0;
    }
    const bf16x8 b = __builtin_bit_cast(bf16x8, bv);
    const f32x4 c0 =
        __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a0), b, zero, 0, 0, 0);
    const f32x4 c1 =
        __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a1), b, zero, 0, 0, 0);
    const float Ee = es[s];
#pragma unroll
    for (int i = 0; i < 4; ++i) F[(4 * kg + i) * NELP + s] = Ee * c0[i];
    if (kg < 2) {
#pragma unroll
      for (int i = 0; i < 4; ++i) F[(16 + 4 * kg + i) * NELP + s] = Ee * c1[i];
    }
  }
  __syncthreads();

  // 3. node sums and epilogue: wave wv = x-plane wv of the tile
  const int ly = lane / TZ, lz = lane % TZ;
  const int ix = x0 + wv, iy = y0 + ly, iz = z0 + lz;
  if ((ix >= g.NX) | (iy >= g.NY) | (iz >= g.NZ)) return;
  const int n = (ix * g.NY + iy) * g.NZ + iz;
  const int o[3] = {map_idx<OFF>(map, 3 * n), map_idx<OFF>(map, 3 * n + 1),
                    map_idx<OFF>(map, 3 * n + 2)};
  if ((o[0] & o[1] & o[2]) < 0) return;
  float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    // element of which this node is corner a (outside the grid: modulus 0,
    // its entries are +-0 and leave the sums unchanged)
    const int s = ((wv + 1 - lpos(a, 0)) * EY + (ly + 1 - lpos(a, 1))) * EZ + (lz + 1 - lpos(a, 2));
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] += F[(3 * a + q) * NELP + s];
  }
  epilogue3<float, EPI>(o, acc, u, y, ep);
}

int Ctx::bf16_alloc() {
  GF_DEVICE_LOCK();
  if (KEb) return 0;
  unsigned short h[32 * 32];
  for (int i = 0; i < 32 * 32; ++i) h[i] = 0;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) h[i * 32 + j] = bf16_bits_rne((float)h_Ke[i * 24 + j]);
  GF_HIP(hipMalloc(&KEb, sizeof(h)));
  GF_TRY(upload_now(KEb, h, sizeof(h)));
  return 0;
}

// TZ = 16 or 8, whichever pads the node grid's (y, z) extent less.
static int bf16_tz(const Grid& g) {
  const long p16 = (long)cdiv(g.NY, 4) * 4 * cdiv(g.NZ, 16) * 16;
  const long p8 = (long)cdiv(g.NY, 8) * 8 * cdiv(g.NZ, 8) * 8;
  return p16 <= p8 ? 16 : 8;
}

template <int EPI, bool OFF>
static void launch_bf16(const Grid& g, const DofMap& map, const unsigned short* KEb,
                        const float* E, const float* u, float* y, const EpiArgs<float>& ep,
                        const int* gate, hipStream_t st) {
  const int tz = bf16_tz(g);
  const int tiles = cdiv(g.NX, kTX) * cdiv(g.NY, 64 / tz) * cdiv(g.NZ, tz);
  if (tz == 16)
    hipLaunchKernelGGL((k_fine_matvec_bf16<EPI, OFF, 16>), dim3(tiles), dim3(kBlock), 0, st, g,
                       map, KEb, E, u, y, ep, gate);
  else
    hipLaunchKernelGGL((k_fine_matvec_bf16<EPI, OFF, 8>), dim3(tiles), dim3(kBlock), 0, st, g,
                       map, KEb, E, u, y, ep, gate);
}

template <int EPI>
static void launch_bf16_e(const Grid& g, const DofMap& map, const unsigned short* KEb,
                          const float* E, const float* u, float* y, const EpiArgs<float>& ep,
                          const int* gate, hipStream_t st) {
  if (map.off >= 0)
    launch_bf16<EPI, true>(g, map, KEb, E, u, y, ep, gate, st);
  else
    launch_bf16<EPI, false>(g, map, KEb, E, u, y, ep, gate, st);
}

int fine_matvec_bf16(const Grid& g, const DofMap& map, const unsigned short* KEb, const float* E,
                     const float* u, float* y, int epi, const EpiArgs<float>& ep, const int* gate,
                     hipStream_t st) {
  switch (epi) {
    case EPI_PLAIN: launch_bf16_e<EPI_PLAIN>(g, map, KEb, E, u, y, ep, gate, st); break;
    case EPI_RESID: launch_bf16_e<EPI_RESID>(g, map, KEb, E, u, y, ep, gate, st); break;
    case EPI_CHEB0: launch_bf16_e<EPI_CHEB0>(g, map, KEb, E, u, y, ep, gate, st); break;
    case EPI_CHEB: launch_bf16_e<EPI_CHEB>(g, map, KEb, E, u, y, ep, gate, st); break;
    case EPI_JACOBI: launch_bf16_e<EPI_JACOBI>(g, map, KEb, E, u, y, ep, gate, st); break;
    default:
      set_error("fine_matvec_bf16: bad epilogue");
      return -2;
  }
  GF_CHECK_LAUNCH();
  return 0;
}

int matvec_full_bf16(Ctx& c, const float* E, const float* u_full, float* y_full, hipStream_t st) {
  GF_TRY(c.bf16_alloc());
  // full layout = identity DOF map (offset 0): y = K u over every DOF
  const DofMap ident{0, nullptr};
  return fine_matvec_bf16(c.g, ident, c.KEb, E, u_full, y_full, EPI_PLAIN, EpiArgs<float>{},
                          nullptr, st);
}

}  // namespace gf
