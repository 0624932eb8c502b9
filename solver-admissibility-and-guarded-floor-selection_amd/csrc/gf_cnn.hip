// Learned floor predictor: a small 3D CNN trained and evaluated by these
// kernels (no reference counterpart; the reference's policy always probes).
//
//   input  density average-pooled to 16x8x8 + three grid features
//   conv1  3x3x3 pad 1, 1 -> 16, ReLU            (VALU, K = 27)
//   conv2  3x3x3 pad 1, 16 -> 32, ReLU           (bf16 MFMA, K = 432)
//   pool   mean over the 1024 positions -> 32
//   fc3    35 -> 32, ReLU;  fc4 32 -> 4 (log10 r50 regression + 3 logits)
//
// BF16 training rule: every product of two tensors (convolutions and dense
// layers, forward and backward) takes bf16 operands rounded to nearest even
// from FP32 and accumulates in FP32; biases, ReLU, pooling, loss and Adam are
// FP32.  One workgroup per sample; every batch sum is a fixed-order loop over
// per-sample partials, so results are deterministic.
//
// conv2 as implicit GEMMs on v_mfma_f32_16x16x32_bf16 (lane l holds
// A[l%16][8*(l/16)..+7], B[8*(l/16)..+7][l%16], result C[4*(l/16)..+3][l%16]):
//   forward   Z2[p][co]   = sum_(t,ci) A1[p+d_t][ci] W2[co][t][ci]
//             M = 16 positions, N = 16 out channels, K = (tap, ci) in steps
//             of 2 taps x 16 ci; the zero-padded 18x10x10x16 bf16 input
//             window sits in LDS.
//   dW2       dW2[co][t][ci] = sum_p dZ2[p][co] A1[p+d_t][ci]
//             M = 16 out channels, N = 16 in channels of one tap, K = 32
//             positions per step.
//   dA1       dA1[q][ci] = sum_(t,co) dZ2[q-d_t][co] W2[co][t][ci]
//             M = 16 positions, N = 16 in channels, K = 32 out channels of
//             one tap per step; W2 transposed to [t][ci][co] bf16 in LDS.
#include "gf_internal.h"

#include <cmath>

namespace gf {
namespace {

constexpr int NP = 1024;                      // 16 x 8 x 8 positions
constexpr int NPP = 18 * 10 * 10;             // zero-padded window
constexpr int C1 = 16, C2 = 32, NT = 27, NF = 3, NI = C2 + NF, NH = 32, NO = 4;
constexpr int OW1 = 0, OB1 = OW1 + C1 * NT, OW2 = OB1 + C1, OB2 = OW2 + C2 * NT * C1,
              OW3 = OB2 + C2, OB3 = OW3 + NH * NI, OW4 = OB3 + NH, OB4 = OW4 + NO * NH,
              NPAR = OB4 + NO;
static_assert(NPAR == 15588, "parameter layout");
// per-sample work layout (floats)
constexpr int WA1 = 0;                   // A1 = relu(Z1)   [1024][16]
constexpr int WZ2 = WA1 + NP * C1;       // Z2              [1024][32]
constexpr int WHD = WZ2 + NP * C2;       // head block (128 floats)
constexpr int HG = 0, HZ3 = 32, HOUT = 64, HDG = 68, HLOSS = 100;
constexpr int WDZ1 = WHD + 128;          // dZ1             [1024][16]
constexpr int WGR = WDZ1 + NP * C1;      // gradient partial [NPAR]
constexpr int WSTRIDE = WGR + 15616;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16r(float f) {
  return __uint_as_float((unsigned int)bf16_bits_rne(f) << 16);
}
// padded-window index of position p = (d*8 + h)*8 + w
__device__ __forceinline__ int padpos(int p) {
  return (((p >> 6) + 1) * 10 + ((p >> 3) & 7) + 1) * 10 + (p & 7) + 1;
}
// window offset of tap t = (kd*3 + kh)*3 + kw (displacement kd-1, kh-1, kw-1)
__device__ __forceinline__ int tapoff(int t) {
  return ((t / 9 - 1) * 10 + ((t / 3) % 3 - 1)) * 10 + (t % 3 - 1);
}
__device__ __forceinline__ bool interior(int pp, int& p) {
  const int pd = pp / 100, ph = (pp / 10) % 10, pw = pp % 10;
  p = ((pd - 1) * 8 + (ph - 1)) * 8 + (pw - 1);
  return (pd >= 1) & (pd <= 16) & (ph >= 1) & (ph <= 8) & (pw >= 1) & (pw <= 8);
}
__device__ __forceinline__ f32x4 mfma(const u16x8& a, const u16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
// A1 (fp32, [1024][16]) of one sample -> bf16 zero-padded window in LDS
__device__ __forceinline__ void stage_a1(const float* __restrict__ A1, unsigned short* a1s) {
  for (int i = threadIdx.x; i < NPP * C1; i += blockDim.x) {
    int p;
    const bool in = interior(i >> 4, p);
    const float v = A1[in ? p * C1 + (i & 15) : 0];
    a1s[i] = in ? bf16_bits_rne(v) : (unsigned short)0;
  }
}

// ---- pooling: one lane per output cell, adaptive bins, fixed loop order ----
__global__ __launch_bounds__(256) void k_cnn_pool(const double* __restrict__ rho, int nx, int ny,
                                                  int nz, float* __restrict__ X) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= NP) return;
  const int i = o >> 6, j = (o >> 3) & 7, k = o & 7;
  const int x0 = i * nx / 16, x1 = ((i + 1) * nx + 15) / 16;
  const int y0 = j * ny / 8, y1 = ((j + 1) * ny + 7) / 8;
  const int z0 = k * nz / 8, z1 = ((k + 1) * nz + 7) / 8;
  double s = 0.0;
  for (int x = x0; x < x1; ++x)
    for (int y = y0; y < y1; ++y)
      for (int z = z0; z < z1; ++z) s += rho[((long)x * ny + y) * nz + z];
  X[o] = (float)(s / (double)((x1 - x0) * (y1 - y0) * (z1 - z0)));
}

// ---- conv1 forward: VALU, one lane per position, 16 channels ----
__global__ __launch_bounds__(256) void k_conv1_fwd(const float* __restrict__ th,
                                                   const float* __restrict__ X,
                                                   float* __restrict__ work) {
  __shared__ float xs[NPP];
  __shared__ float w1[C1 * NT + C1];
  const int b = blockIdx.x;
  const float* Xb = X + (long)b * NP;
  for (int i = threadIdx.x; i < NPP; i += blockDim.x) {
    int p;
    const bool in = interior(i, p);
    const float v = Xb[in ? p : 0];
    xs[i] = in ? bf16r(v) : 0.0f;
  }
  for (int i = threadIdx.x; i < C1 * NT; i += blockDim.x) w1[i] = bf16r(th[OW1 + i]);
  if (threadIdx.x < C1) w1[C1 * NT + threadIdx.x] = th[OB1 + threadIdx.x];
  __syncthreads();
  float* A1 = work + (long)b * WSTRIDE + WA1;
  for (int p = threadIdx.x; p < NP; p += blockDim.x) {
    const int pp = padpos(p);
    float xv[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) xv[t] = xs[pp + tapoff(t)];
#pragma unroll 2
    for (int co = 0; co < C1; ++co) {
      float z = 0.0f;
#pragma unroll
      for (int t = 0; t < NT; ++t) z = fmaf(w1[co * NT + t], xv[t], z);
      z += w1[C1 * NT + co];
      A1[p * C1 + co] = fmaxf(z, 0.0f);
    }
  }
}

// ---- conv2 forward + ReLU + global average pool (bf16 MFMA) ----
__global__ __launch_bounds__(256) void k_conv2_fwd(const float* __restrict__ th,
                                                   float* __restrict__ work) {
  __shared__ __attribute__((aligned(16))) unsigned short a1s[NPP * C1];
  __shared__ float red[4][4][C2];
  const int b = blockIdx.x;
  float* base = work + (long)b * WSTRIDE;
  stage_a1(base + WA1, a1s);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 15, kg = lane >> 4;
  const float bias0 = th[OB2 + col], bias1 = th[OB2 + 16 + col];
  const float* W2 = th + OW2;
  float* Z2 = base + WZ2;
  float g0 = 0.0f, g1 = 0.0f;
  for (int mt = wave; mt < NP / 16; mt += 4) {
    const int pp = padpos(mt * 16 + col);
    f32x4 c0 = {0.f, 0.f, 0.f, 0.f}, c1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int s = 0; s < 14; ++s) {
      const int tap = 2 * s + (kg >> 1), ci0 = 8 * (kg & 1);
      u16x8 a = {0, 0, 0, 0, 0, 0, 0, 0}, w0 = a, w1 = a;
      if (tap < NT) {
        a = *reinterpret_cast<const u16x8*>(&a1s[(pp + tapoff(tap)) * C1 + ci0]);
        const float* r0 = W2 + (col * NT + tap) * C1 + ci0;
        const float* r1 = W2 + ((col + 16) * NT + tap) * C1 + ci0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          w0[j] = bf16_bits_rne(r0[j]);
          w1[j] = bf16_bits_rne(r1[j]);
        }
      }
      c0 = mfma(a, w0, c0);
      c1 = mfma(a, w1, c1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = mt * 16 + 4 * kg + i;
      const float z0 = c0[i] + bias0, z1 = c1[i] + bias1;
      Z2[p * C2 + col] = z0;
      Z2[p * C2 + 16 + col] = z1;
      g0 += fmaxf(z0, 0.0f);
      g1 += fmaxf(z1, 0.0f);
    }
  }
  red[wave][kg][col] = g0;
  red[wave][kg][16 + col] = g1;
  __syncthreads();
  if (threadIdx.x < C2) {
    float s = 0.0f;
    for (int w = 0; w < 4; ++w)
      for (int k = 0; k < 4; ++k) s += red[w][k][threadIdx.x];
    base[WHD + HG + threadIdx.x] = s * (1.0f / NP);
  }
}

// ---- dense head: fc3, fc4, loss and (train) their gradients ----
__global__ __launch_bounds__(64) void k_cnn_head(const float* __restrict__ th,
                                                 const float* __restrict__ feat,
                                                 const float* __restrict__ y_reg,
                                                 const int* __restrict__ y_cls, int B, float w_reg,
                                                 int train, float* __restrict__ work,
                                                 float* __restrict__ out) {
  __shared__ float hb[NI], z3[NH], a3b[NH], os[NO], dout[NO], doutb[NO], dz3b[NH];
  const int b = blockIdx.x, t = threadIdx.x;
  float* hd = work + (long)b * WSTRIDE + WHD;
  if (t < C2) hb[t] = bf16r(hd[HG + t]);
  if (t < NF) hb[C2 + t] = bf16r(feat[b * NF + t]);
  __syncthreads();
  if (t < NH) {
    float z = 0.0f;
    for (int k = 0; k < NI; ++k) z = fmaf(bf16r(th[OW3 + t * NI + k]), hb[k], z);
    z += th[OB3 + t];
    z3[t] = z;
    hd[HZ3 + t] = z;
    a3b[t] = bf16r(fmaxf(z, 0.0f));
  }
  __syncthreads();
  if (t < NO) {
    float o = 0.0f;
    for (int j = 0; j < NH; ++j) o = fmaf(bf16r(th[OW4 + t * NH + j]), a3b[j], o);
    o += th[OB4 + t];
    os[t] = o;
    hd[HOUT + t] = o;
    if (out) out[b * NO + t] = o;
  }
  if (!train) return;
  __syncthreads();
  if (t == 0) {
    const float r = os[0] - y_reg[b];
    const float m = fmaxf(os[1], fmaxf(os[2], os[3]));
    const float se = expf(os[1] - m) + expf(os[2] - m) + expf(os[3] - m);
    const float lse = m + logf(se);
    const int y = y_cls[b];
    hd[HLOSS] = w_reg * r * r + lse - os[1 + y];
    const float inv = 1.0f / (float)B;
    dout[0] = 2.0f * w_reg * r * inv;
    for (int c = 0; c < 3; ++c) dout[1 + c] = expf(os[1 + c] - lse) * inv - (c == y ? inv : 0.0f);
    for (int c = 0; c < NO; ++c) doutb[c] = bf16r(dout[c]);
  }
  __syncthreads();
  float* gr = work + (long)b * WSTRIDE + WGR;
  if (t < NH) {
    for (int o = 0; o < NO; ++o) gr[OW4 + o * NH + t] = doutb[o] * a3b[t];
    float da = 0.0f;
    for (int o = 0; o < NO; ++o) da = fmaf(bf16r(th[OW4 + o * NH + t]), doutb[o], da);
    const float dz = z3[t] > 0.0f ? da : 0.0f;
    dz3b[t] = bf16r(dz);
    gr[OB3 + t] = dz;
  }
  if (t < NO) gr[OB4 + t] = dout[t];
  __syncthreads();
  for (int i = t; i < NH * NI; i += 64) gr[OW3 + i] = dz3b[i / NI] * hb[i % NI];
  if (t < C2) {
    float dg = 0.0f;
    for (int j = 0; j < NH; ++j) dg = fmaf(bf16r(th[OW3 + j * NI + t]), dz3b[j], dg);
    hd[HDG + t] = dg;
  }
}

// ---- dW2 and db2 partials of one sample (bf16 MFMA) ----
__global__ __launch_bounds__(256) void k_conv2_bwd_w(float* __restrict__ work) {
  __shared__ __attribute__((aligned(16))) unsigned short a1s[NPP * C1];
  const int b = blockIdx.x;
  float* base = work + (long)b * WSTRIDE;
  stage_a1(base + WA1, a1s);
  __syncthreads();
  const float* Z2 = base + WZ2;
  const float* hd = base + WHD;
  float* gr = base + WGR;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 15, kg = lane >> 4;
  const int mt = wave & 1, tbeg = (wave >> 1) ? 14 : 0, tcnt = (wave >> 1) ? 13 : 14;
  const int co = mt * 16 + col;
  const float sc = hd[HDG + co] * (1.0f / NP);
  const unsigned short scb = bf16_bits_rne(sc);
  f32x4 acc[14];
#pragma unroll
  for (int k = 0; k < 14; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < NP / 32; ++s) {
    const int p0 = 32 * s + 8 * kg;  // 8 positions with equal (d, h), w = 0..7
    u16x8 a;
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = Z2[(p0 + i) * C2 + co] > 0.0f ? scb : (unsigned short)0;
    const int pp0 = padpos(p0);
#pragma unroll
    for (int k = 0; k < 14; ++k) {
      if (k < tcnt) {
        const int bs = pp0 + tapoff(tbeg + k);
        u16x8 bv;
#pragma unroll
        for (int i = 0; i < 8; ++i) bv[i] = a1s[(bs + i) * C1 + col];
        acc[k] = mfma(a, bv, acc[k]);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 14; ++k) {
    if (k < tcnt) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        gr[OW2 + ((mt * 16 + 4 * kg + i) * NT + tbeg + k) * C1 + col] = acc[k][i];
    }
  }
  if (threadIdx.x < C2) {
    const int c = threadIdx.x;
    int cnt = 0;
    for (int p = 0; p < NP; ++p) cnt += Z2[p * C2 + c] > 0.0f;
    gr[OB2 + c] = hd[HDG + c] * (1.0f / NP) * (float)cnt;
  }
}

// ---- dA1 -> dZ1 of one sample (bf16 MFMA over (tap, out channel)) ----
__global__ __launch_bounds__(256) void k_conv2_bwd_x(const float* __restrict__ th,
                                                     float* __restrict__ work) {
  __shared__ __attribute__((aligned(16))) unsigned short w2t[NT * C1 * C2];
  __shared__ float dgs[C2];
  const int b = blockIdx.x;
  float* base = work + (long)b * WSTRIDE;
  for (int i = threadIdx.x; i < NT * C1 * C2; i += blockDim.x) {
    const int co = i / (NT * C1), r = i % (NT * C1);  // source [co][t][ci], r = t*16 + ci
    w2t[r * C2 + co] = bf16_bits_rne(th[OW2 + i]);
  }
  if (threadIdx.x < C2) dgs[threadIdx.x] = base[WHD + HDG + threadIdx.x] * (1.0f / NP);
  __syncthreads();
  const float* Z2 = base + WZ2;
  const float* A1 = base + WA1;
  float* dZ1 = base + WDZ1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 15, kg = lane >> 4;
  u16x8 dgb;
#pragma unroll
  for (int j = 0; j < 8; ++j) dgb[j] = bf16_bits_rne(dgs[8 * kg + j]);
  for (int mt = wave; mt < NP / 16; mt += 4) {
    const int q = mt * 16 + col;
    const int d = q >> 6, h = (q >> 3) & 7, w = q & 7;
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < NT; ++t) {
      const int sd = d - (t / 9 - 1), sh = h - ((t / 3) % 3 - 1), sw = w - (t % 3 - 1);
      const bool in = (sd >= 0) & (sd < 16) & (sh >= 0) & (sh < 8) & (sw >= 0) & (sw < 8);
      const float* zr = Z2 + (in ? ((sd * 8 + sh) * 8 + sw) * C2 + 8 * kg : 0);
      const f32x4 z0 = *reinterpret_cast<const f32x4*>(zr);
      const f32x4 z1 = *reinterpret_cast<const f32x4*>(zr + 4);
      u16x8 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = (in & (z0[j] > 0.0f)) ? dgb[j] : (unsigned short)0;
        a[4 + j] = (in & (z1[j] > 0.0f)) ? dgb[4 + j] : (unsigned short)0;
      }
      const u16x8 bv = *reinterpret_cast<const u16x8*>(&w2t[(t * C1 + col) * C2 + 8 * kg]);
      c = mfma(a, bv, c);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = mt * 16 + 4 * kg + i;
      dZ1[p * C1 + col] = A1[p * C1 + col] > 0.0f ? c[i] : 0.0f;
    }
  }
}

// ---- dW1 and db1 partials of one sample (VALU) ----
__global__ __launch_bounds__(256) void k_conv1_bwd(const float* __restrict__ X,
                                                   float* __restrict__ work) {
  __shared__ float xs[NPP];
  const int b = blockIdx.x;
  const float* Xb = X + (long)b * NP;
  for (int i = threadIdx.x; i < NPP; i += blockDim.x) {
    int p;
    const bool in = interior(i, p);
    const float v = Xb[in ? p : 0];
    xs[i] = in ? bf16r(v) : 0.0f;
  }
  __syncthreads();
  float* base = work + (long)b * WSTRIDE;
  const float* dZ1 = base + WDZ1;
  float* gr = base + WGR;
  for (int o = threadIdx.x; o < C1 * NT + C1; o += blockDim.x) {
    float s = 0.0f;
    if (o < C1 * NT) {
      const int co = o / NT, off = tapoff(o % NT);
      for (int p = 0; p < NP; ++p) s = fmaf(bf16r(dZ1[p * C1 + co]), xs[padpos(p) + off], s);
      gr[OW1 + o] = s;
    } else {
      const int co = o - C1 * NT;
      for (int p = 0; p < NP; ++p) s += dZ1[p * C1 + co];
      gr[OB1 + co] = s;
    }
  }
}

// ---- batch sums of the per-sample partials (fixed order) ----
__global__ __launch_bounds__(256) void k_cnn_reduce(const float* __restrict__ work, int B,
                                                    float* __restrict__ grad,
                                                    float* __restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NPAR) {
    float s = 0.0f;
    for (int b = 0; b < B; ++b) s += work[(long)b * WSTRIDE + WGR + i];
    grad[i] = s;
  }
  if (i == 0 && loss) {
    float l = 0.0f;
    for (int b = 0; b < B; ++b) l += work[(long)b * WSTRIDE + WHD + HLOSS];
    *loss = l / (float)B;
  }
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ th, float* __restrict__ m,
                                              float* __restrict__ v,
                                              const float* __restrict__ g, int n, float lr,
                                              float b1, float b2, float eps, float bc1,
                                              float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

int forward_part(const float* th, const float* X, const float* feat, const float* y_reg,
                 const int* y_cls, int B, float w_reg, int train, float* out, float* work,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_conv1_fwd, dim3(B), dim3(256), 0, st, th, X, work);
  GF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_conv2_fwd, dim3(B), dim3(256), 0, st, th, work);
  GF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_cnn_head, dim3(B), dim3(64), 0, st, th, feat, y_reg, y_cls, B, w_reg,
                     train, work, out);
  GF_CHECK_LAUNCH();
  return 0;
}

}  // namespace
}  // namespace gf

using namespace gf;

#define CNN_CHECK(cond, msg)   \
  do {                         \
    if (!(cond)) {             \
      set_error(msg);          \
      return -2;               \
    }                          \
  } while (0)

extern "C" {

int32_t gf_cnn_sizes(int32_t* n_params, int64_t* work_floats_per_sample) {
  if (n_params) *n_params = NPAR;
  if (work_floats_per_sample) *work_floats_per_sample = WSTRIDE;
  return 0;
}

int32_t gf_cnn_pool_f64(const double* rho, int32_t nx, int32_t ny, int32_t nz, float* x_out,
                        void* stream) {
  CNN_CHECK(rho && x_out, "gf_cnn_pool_f64: null argument");
  CNN_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "gf_cnn_pool_f64: grid dims must be >= 1");
  hipLaunchKernelGGL(k_cnn_pool, dim3(NP / 256), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), rho, nx, ny, nz, x_out);
  GF_CHECK_LAUNCH();
  return 0;
}

int32_t gf_cnn_forward(const float* theta, const float* x, const float* feat, int32_t batch,
                       float* out, float* work, void* stream) {
  CNN_CHECK(theta && x && feat && out && work, "gf_cnn_forward: null argument");
  CNN_CHECK(batch >= 1, "gf_cnn_forward: batch must be >= 1");
  return forward_part(theta, x, feat, nullptr, nullptr, batch, 0.0f, 0, out, work,
                      reinterpret_cast<hipStream_t>(stream));
}

int32_t gf_cnn_grad(const float* theta, const float* x, const float* feat, const float* y_reg,
                    const int32_t* y_cls, int32_t batch, float w_reg, float* loss_out,
                    float* grad_out, float* out, float* work, void* stream) {
  CNN_CHECK(theta && x && feat && y_reg && y_cls && grad_out && work,
            "gf_cnn_grad: null argument");
  CNN_CHECK(batch >= 1, "gf_cnn_grad: batch must be >= 1");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  GF_TRY(forward_part(theta, x, feat, y_reg, y_cls, batch, w_reg, 1, out, work, st));
  hipLaunchKernelGGL(k_conv2_bwd_w, dim3(batch), dim3(256), 0, st, work);
  GF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_conv2_bwd_x, dim3(batch), dim3(256), 0, st, theta, work);
  GF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_conv1_bwd, dim3(batch), dim3(256), 0, st, x, work);
  GF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_cnn_reduce, dim3(cdiv(NPAR, 256)), dim3(256), 0, st, work, batch, grad_out,
                     loss_out);
  GF_CHECK_LAUNCH();
  return 0;
}

int32_t gf_cnn_adam(float* theta, float* m, float* v, const float* grad, int32_t n, float lr,
                    float beta1, float beta2, float eps, int32_t step, void* stream) {
  CNN_CHECK(theta && m && v && grad, "gf_cnn_adam: null argument");
  CNN_CHECK(n >= 1 && step >= 1, "gf_cnn_adam: n and step must be >= 1");
  const float bc1 = (float)(1.0 - std::pow((double)beta1, step));
  const float bc2 = (float)(1.0 - std::pow((double)beta2, step));
  hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), theta, m, v, grad, n, lr, beta1, beta2,
                     eps, bc1, bc2);
  GF_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
