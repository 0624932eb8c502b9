// BF16 matrix-core element contraction for the fused matvec's bf16 form
// (FusedMatvec dtype="bf16", cuda_fused_matvec.py:112-246, 393-416):
//
//   y_full = sum_e E_e * (KE_bf16 @ bf16(u_e)),  products accumulated in FP32,
//
// with KE and u rounded to bf16 (round to nearest even) and E_e applied in
// FP32 to the FP32 contraction, as the reference's kernel does.  Two passes,
// no atomics, so the result is deterministic:
//   1. k_elem_bf16: one wave per 16-element tile computes the 24x16 tile
//      F = KE(32x32, zero padded) x U(32x16) with two
//      v_mfma_f32_16x16x32_bf16 (M tiles KE rows 0-15 and 16-31), scales each
//      column by its element's E and stores 24 floats per element;
//   2. k_node_gather_f32: one lane per node sums its three rows over the up
//      to 8 adjacent elements in a fixed order.
// Operand layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l holds
// A[l%16][8*(l/16) .. +7], B[8*(l/16) .. +7][l%16]; the result lane l holds
// C[4*(l/16) .. +3][l%16].
#include "gf_internal.h"
#include "gf_solver.h"

namespace gf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_elem_bf16(Grid g, const unsigned short* __restrict__ KEb,
                                                   const float* __restrict__ E,
                                                   const float* __restrict__ u,
                                                   float* __restrict__ F) {
  const int lane = threadIdx.x & 63;
  const long n_e = (long)g.nx * g.ny * g.nz;
  const long e0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (e0 >= n_e) return;  // uniform over the wave
  const int col = lane & 15;  // element of the tile (B / C column)
  const int kg = lane >> 4;   // K slice 8*kg .. 8*kg+7 (A, B); C rows 4*kg .. 4*kg+3
  const u16x8 a0 = *reinterpret_cast<const u16x8*>(KEb + col * 32 + 8 * kg);
  const u16x8 a1 = *reinterpret_cast<const u16x8*>(KEb + (16 + col) * 32 + 8 * kg);
  const long e = e0 + col;
  u16x8 bv = {0, 0, 0, 0, 0, 0, 0, 0};
  float Ee = 0.0f;
  if (e < n_e) {
    const int ex = (int)(e / ((long)g.ny * g.nz));
    const int ey = (int)((e / g.nz) % g.ny);
    const int ez = (int)(e % g.nz);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = 8 * kg + t;
      if (k < 24) {
        const int a = k / 3, c = k - 3 * (k / 3);
        const long node = ((long)(ex + lpos(a, 0)) * g.NY + (ey + lpos(a, 1))) * g.NZ +
                          (ez + lpos(a, 2));
        bv[t] = bf16_bits_rne(u[3 * node + c]);
      }
    }
    Ee = E[e];
  }
  const bf16x8 b = __builtin_bit_cast(bf16x8, bv);
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  const f32x4 c0 =
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a0), b, zero, 0, 0, 0);
  const f32x4 c1 =
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a1), b, zero, 0, 0, 0);
  if (e < n_e) {
    float* Fe = F + e * 24;
    const f32x4 s0 = {Ee * c0[0], Ee * c0[1], Ee * c0[2], Ee * c0[3]};
    *reinterpret_cast<f32x4*>(Fe + 4 * kg) = s0;
    if (kg < 2) {
      const f32x4 s1 = {Ee * c1[0], Ee * c1[1], Ee * c1[2], Ee * c1[3]};
      *reinterpret_cast<f32x4*>(Fe + 16 + 4 * kg) = s1;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_node_gather_f32(Grid g, const float* __restrict__ F,
                                                            float* __restrict__ y) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int iz = (int)(n % g.NZ);
  const int iy = (int)((n / g.NZ) % g.NY);
  const int ix = (int)(n / ((long)g.NY * g.NZ));
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ex = ix - lpos(a, 0), ey = iy - lpos(a, 1), ez = iz - lpos(a, 2);
    if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
    const float* Fe = F + (((long)ex * g.ny + ey) * g.nz + ez) * 24 + 3 * a;
    s0 += Fe[0];
    s1 += Fe[1];
    s2 += Fe[2];
  }
  y[3 * n] = s0;
  y[3 * n + 1] = s1;
  y[3 * n + 2] = s2;
}

int Ctx::bf16_alloc() {
  if (KEb && Fel) return 0;
  GF_DEVICE_LOCK();
  unsigned short h[32 * 32];
  for (int i = 0; i < 32 * 32; ++i) h[i] = 0;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) h[i * 32 + j] = bf16_bits_rne((float)h_Ke[i * 24 + j]);
  const long n_e = (long)g.nx * g.ny * g.nz;
  GF_HIP(hipMalloc(&KEb, sizeof(h)));
  GF_TRY(upload_now(KEb, h, sizeof(h)));
  GF_HIP(hipMalloc(&Fel, sizeof(float) * 24 * (size_t)n_e));
  return 0;
}

int matvec_full_bf16(Ctx& c, const float* E, const float* u_full, float* y_full, hipStream_t st) {
  GF_TRY(c.bf16_alloc());
  const long n_e = (long)c.g.nx * c.g.ny * c.g.nz;
  const long tiles = (n_e + 15) / 16;
  hipLaunchKernelGGL(k_elem_bf16, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, st, c.g,
                     c.KEb, E, u_full, c.Fel);
  GF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_node_gather_f32, dim3(cdiv(c.g.n_nodes, kBlock)), dim3(kBlock), 0, st, c.g,
                     c.Fel, y_full);
  GF_CHECK_LAUNCH();
  return 0;
}

}  // namespace gf
