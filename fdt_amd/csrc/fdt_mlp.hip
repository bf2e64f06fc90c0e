// fdt_mlp.hip -- kernels of the learned initialiser: a BF16 MFMA GEMM with a
// fused bias/GELU epilogue and the element-wise passes of training
// (GELU backward, bias gradients, Adam, loss) and of inference (signal
// normalisation, decoding of the network outputs into chain start values and
// proposal widths, the fallback test).
//
// GEMM: C[M x N] = A[M x K] . B[K x N] with arbitrary element strides for A
// and B (so X.W, dZ.W^T and X^T.dZ are one kernel).  A 256-thread workgroup
// computes a 64x64 tile; each wave a 32x32 quadrant as 2x2 blocks of
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation), K in steps of 32 staged in
// LDS (A row-major, B column-major so that every lane's 8 K values are one
// 16-byte LDS read).
#include "fdt_device.h"
#include "fdt_mlp.h"

namespace fdt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

FDT_DEV float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
// round to nearest even (finite inputs)
FDT_DEV uint16_t f2bf(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
FDT_DEV float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
FDT_DEV float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

constexpr int GT = 64;   // tile edge
constexpr int GK = 32;   // K step
constexpr int GLD = 40;  // LDS row stride in bf16 (80 bytes)

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[GT * GLD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[GT * GLD];  // [n][k]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int qm = (wave >> 1) * 32, qn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // staging: thread t fills row (t / 4) of the tile, K elements (t % 4) * 8 .. + 8
  const int sr = tid >> 2, sk = (tid & 3) * 8;
  for (int k0 = 0; k0 < g.K; k0 += GK) {
    {
      const int m = m0 + sr;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + sk + e;
        As[sr * GLD + sk + e] = (m < g.M && k < g.K) ? g.A[(size_t)m * g.sam + (size_t)k * g.sak] : (uint16_t)0;
      }
      const int n = n0 + sr;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + sk + e;
        Bs[sr * GLD + sk + e] = (n < g.N && k < g.K) ? g.B[(size_t)k * g.sbk + (size_t)n * g.sbn] : (uint16_t)0;
      }
    }
    __syncthreads();
    const int r = lane & 15, kb = (lane >> 4) * 8;
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[(qm + 16 * i + r) * GLD + kb]);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[(qn + 16 * j + r) * GLD + kb]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // lane holds C[(lane / 16) * 4 + v][lane % 16] of each 16x16 block
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + qm + 16 * i + (lane >> 4) * 4 + v;
        const int n = n0 + qn + 16 * j + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float x = acc[i][j][v];
        const size_t ci = (size_t)m * g.ldc + n;
        if (g.accumulate) x += g.C[ci];
        if (g.bias) x += g.bias[n];
        if (g.C) g.C[ci] = x;
        if (g.Cb) g.Cb[ci] = f2bf(g.act ? gelu(x) : x);
      }
}

// dZ (bf16) = dA * gelu'(Z)
__global__ void gelu_bwd_kernel(const float* dA, const float* Z, uint16_t* dZ, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dZ[i] = f2bf(dA[i] * gelu_grad(Z[i]));
}

// db[n] = sum_m dZ[m][n] (dZ bf16, M x N), one thread per column, fixed order
__global__ void colsum_kernel(const uint16_t* dZ, int M, int N, float* db) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += bf2f(dZ[(size_t)m * N + n]);
  db[n] = s;
}

// Adam (Kingma & Ba 2015) on FP32 master weights, with a BF16 copy for the GEMMs
__global__ void adam_kernel(float* w, float* m, float* v, const float* g, uint16_t* wb, size_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void to_bf16_kernel(const float* x, uint16_t* y, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Network input: the voxel's signal over the mean of its b=0 volumes (the
// volumes with b <= b0_max), one thread per voxel
__global__ void normalize_kernel(const float* data, int N, int nvox, const float* bvals, float b0_max, uint16_t* X,
                                 float* S0) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  const float* y = data + (size_t)v * N;
  double s = 0.0;
  int nb = 0;
  for (int i = 0; i < N; ++i)
    if (bvals[i] <= b0_max) { s += y[i]; ++nb; }
  float s0 = nb ? (float)(s / nb) : 1.0f;
  if (!(s0 > 0.0f)) s0 = 1.0f;
  float inv = 1.0f / s0;
  for (int i = 0; i < N; ++i) X[(size_t)v * N + i] = f2bf(y[i] * inv);
  if (S0) S0[v] = s0;
}

// Loss on the network outputs O [M][SG_OUT] against targets T [M][SG_OUT] and
// its gradient dO (bf16, the last layer has no activation).  Targets: f_k (3),
// dyads x_k (3 x 3, sign-free: loss |p|^4 - 2 (p.x)^2 + 1 = ||p p^T - x x^T||^2
// weighted by f_k), 1000 d, 1000 d_std, log proposal widths (weight w_lw,
// zero where the target slot is unused, marked by a target >= 1e29).
__global__ void sg_loss_kernel(const float* O, const float* T, int M, uint16_t* dO, float* loss_part, float w_lw) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (m < M) {
    const float* o = O + (size_t)m * SG_OUT;
    const float* t = T + (size_t)m * SG_OUT;
    float g[SG_OUT];
    for (int i = 0; i < SG_OUT; ++i) g[i] = 0.0f;
    for (int k = 0; k < 3; ++k) {
      float e = o[k] - t[k];
      l += e * e;
      g[k] = 2.0f * e;
    }
    for (int k = 0; k < 3; ++k) {
      const float* p = o + 3 + 3 * k;
      const float* x = t + 3 + 3 * k;
      float w = fmaxf(t[k], 0.0f);
      float pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
      float px = p[0] * x[0] + p[1] * x[1] + p[2] * x[2];
      l += w * (pp * pp - 2.0f * px * px + 1.0f);
      for (int c = 0; c < 3; ++c) g[3 + 3 * k + c] = w * (4.0f * pp * p[c] - 4.0f * px * x[c]);
    }
    for (int i = 12; i < 14; ++i) {
      float e = o[i] - t[i];
      l += e * e;
      g[i] = 2.0f * e;
    }
    for (int i = 14; i < SG_OUT; ++i) {
      if (t[i] >= 1e29f) continue;
      float e = o[i] - t[i];
      l += w_lw * e * e;
      g[i] = 2.0f * w_lw * e;
    }
    for (int i = 0; i < SG_OUT; ++i) dO[(size_t)m * SG_OUT + i] = f2bf(g[i] / (float)M);
  }
  // block partial sums of the loss (fixed order within the block)
  __shared__ float red[256];
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Network outputs -> chain start [S0, d, (d_std), f_k, th_k, ph_k] and the
// proposal widths of the slots (f0, d, d_std, R, S0, th/ph/f of each fibre);
// also the predicted principal direction (for the fallback test).
__global__ void sg_decode_kernel(SgDecodeArgs a) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.nvox) return;
  const float* o = a.O + (size_t)v * SG_OUT;
  float f[3];
  float fs = 0.0f;
  for (int k = 0; k < 3; ++k) { f[k] = fminf(fmaxf(o[k], 0.01f), 0.95f); fs += f[k]; }
  if (fs > 0.95f) for (int k = 0; k < 3; ++k) f[k] *= 0.95f / fs;
  float* P = a.params + (size_t)v * a.np;
  const int fb = a.model >= 2 ? 3 : 2;
  P[0] = a.S0[v];
  P[1] = fminf(fmaxf(o[12] * 1e-3f, 1e-5f), 0.0049f);
  if (a.model >= 2) {
    float up = a.model == 3 ? 0.0039f : 0.0099f;
    P[2] = fminf(fmaxf(o[13] * 1e-3f, 1e-5f), up);
  }
  for (int k = 0; k < a.K; ++k) {
    const float* p = o + 3 + 3 * k;
    float n2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    float x0 = p[0], x1 = p[1], x2 = p[2];
    if (!(n2 > 1e-20f)) { x0 = 0.0f; x1 = 0.0f; x2 = 1.0f; }
    float th = atan2f(sqrtf(x0 * x0 + x1 * x1), x2), ph = atan2f(x1, x0);
    P[fb + 3 * k] = f[k];
    P[fb + 3 * k + 1] = th;
    P[fb + 3 * k + 2] = ph;
    if (k == 0 && a.dir) {
      float inv = n2 > 1e-20f ? rsqrtf(n2) : 1.0f;
      a.dir[3 * v] = x0 * inv; a.dir[3 * v + 1] = x1 * inv; a.dir[3 * v + 2] = x2 * inv;
    }
  }
  if (a.f0) P[a.np - 1] = 0.001f;
  if (a.widths) {
    float* w = a.widths + (size_t)v * 16;
    for (int s = 0; s < 16; ++s) w[s] = -1.0f;  // default width
    const int slot_of[14] = {0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14};
    // slot 0 (f0) keeps the default width: the training targets carry no f0
    for (int i = 1; i < 14; ++i) {
      float lw = fminf(fmaxf(o[14 + i], -20.0f), 5.0f);
      w[slot_of[i]] = expf(lw);
    }
  }
}

// Fallback test after the short chain: the voxel is rerun the plain way when
// its final sum of squares exceeds chi2_max * N * sigma^2 (sigma^2 from the
// spread of its b=0 volumes; flag bit 1) or the predicted direction of fibre 1
// lies more than ang_max from every posterior mean dyad (flag bit 2).
__global__ void sg_fallback_kernel(SgFallbackArgs a) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.nvox) return;
  const float* y = a.data + (size_t)v * a.N;
  double s = 0.0, s2 = 0.0;
  int nb = 0;
  for (int i = 0; i < a.N; ++i)
    if (a.bvals[i] <= a.b0_max) { s += y[i]; s2 += (double)y[i] * y[i]; ++nb; }
  double var = nb > 1 ? (s2 - s * s / nb) / (nb - 1) : 0.0;
  int why = 0;
  if (var > 0.0 && a.energy[v] > a.chi2_max * a.N * var) why |= 1;
  // the predicted direction of fibre 1 against the closest posterior mean dyad
  float c = 0.0f;
  for (int k = 0; k < a.K; ++k) {
    const float* dy = a.dyads + (size_t)k * 3 * a.nvox + v;  // [K][3][nvox]
    c = fmaxf(c, fabsf(dy[0] * a.dir[3 * v] + dy[(size_t)a.nvox] * a.dir[3 * v + 1] +
                       dy[(size_t)2 * a.nvox] * a.dir[3 * v + 2]));
  }
  if (c < a.cos_min) why |= 2;
  if (a.flag) a.flag[v] = (unsigned char)why;
  if (why) a.list[atomicAdd(a.count, 1)] = v;
}

}  // namespace fdt

extern "C" hipError_t fdt_launch_gemm_bf16(const fdt::GemmArgs* g, hipStream_t s) {
  if (g->M <= 0 || g->N <= 0) return hipSuccess;
  dim3 grid((g->N + fdt::GT - 1) / fdt::GT, (g->M + fdt::GT - 1) / fdt::GT), block(256);
  hipLaunchKernelGGL(fdt::gemm_bf16_kernel, grid, block, 0, s, *g);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_mlp_ew(int op, const void* a, const void* b, void* c, size_t n, int M, int N,
                                        hipStream_t s) {
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  switch (op) {
    case 0: hipLaunchKernelGGL(fdt::gelu_bwd_kernel, grid, block, 0, s, (const float*)a, (const float*)b, (uint16_t*)c, n); break;
    case 1: hipLaunchKernelGGL(fdt::colsum_kernel, dim3((N + 255) / 256), block, 0, s, (const uint16_t*)a, M, N, (float*)c); break;
    case 2: hipLaunchKernelGGL(fdt::to_bf16_kernel, grid, block, 0, s, (const float*)a, (uint16_t*)c, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_adam(float* w, float* m, float* v, const float* g, uint16_t* wb, size_t n, float lr,
                                      float b1, float b2, float eps, int step, hipStream_t s) {
  float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(fdt::adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, m, v, g, wb, n, lr, b1, b2,
                     eps, c1, c2);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_normalize(const float* data, int N, int nvox, const float* bvals, float b0_max,
                                           uint16_t* X, float* S0, hipStream_t s) {
  if (nvox <= 0) return hipSuccess;
  hipLaunchKernelGGL(fdt::normalize_kernel, dim3((nvox + 255) / 256), dim3(256), 0, s, data, N, nvox, bvals, b0_max, X, S0);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_sg_loss(const float* O, const float* T, int M, uint16_t* dO, float* loss_part,
                                         float w_lw, hipStream_t s) {
  hipLaunchKernelGGL(fdt::sg_loss_kernel, dim3((M + 255) / 256), dim3(256), 0, s, O, T, M, dO, loss_part, w_lw);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_sg_decode(const fdt::SgDecodeArgs* a, hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  hipLaunchKernelGGL(fdt::sg_decode_kernel, dim3((a->nvox + 255) / 256), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_sg_fallback(const fdt::SgFallbackArgs* a, hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  hipLaunchKernelGGL(fdt::sg_fallback_kernel, dim3((a->nvox + 255) / 256), dim3(256), 0, s, *a);
  return hipGetLastError();
}
