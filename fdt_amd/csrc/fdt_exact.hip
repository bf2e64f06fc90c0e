// fdt_exact.hip -- kernels that restate the reference's arithmetic operation
// by operation; this file is compiled with -ffp-contract=off, so every
// multiply and add rounds as written (no fused multiply-add).
//
// grad_acq_kernel: the per-voxel acquisitions of --gradnonlin (below).
// pred_kernel: the posterior-predictive signal of the ball-and-sticks models
// from the posterior means (xfibres_pred.cc:153-190), one thread per
// (volume, voxel): consecutive threads take consecutive voxels of one volume,
// so the per-voxel means are read and the [N][nvox] prediction written with
// unit stride; the volume's (g, b) is uniform over the workgroup.
//
// Arithmetic follows the reference's promotions (bvals/bvecs are double
// matrices there, the images float): the projection of the dyad on the
// gradient and every exponent in double, each `output += term` rounded to
// float, the model-3 (1-R) a + R factor in float, the sum scaled by S0 in float.
#include "fdt_device.h"
#include "fdt_kernels.h"

namespace fdt {

// --gradnonlin: the acquisition of every voxel through its gradient-deviation
// tensor L (grad [nvox][9], L(i,j) = grad[3j+i]), correct_bvals_bvecs of
// CUDA/xfibres_gpu.cu:151-170: for b > 0, v -> (I+L) v / |(I+L) v| and
// b -> |(I+L) v|^2 b.  Products and left-to-right sums in double, |.| rounded to float as the reference's
// `float mag`, results rounded to float as the reference's float host vectors.
// One thread per (voxel, volume slot); slots from N to stride are zero.
__global__ void grad_acq_kernel(const float4* __restrict__ base, int N, int stride, const float* __restrict__ grad,
                                int nvox, float4* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nvox * stride) return;
  const int vox = (int)(t / stride), i = (int)(t % stride);
  float4 q = i < N ? base[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (i < N && q.w > 0.0f) {
    const float* g = grad + (size_t)vox * 9;
    const double v[3] = {q.x, q.y, q.z};
    double w[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double m0 = (r == 0 ? __dadd_rn(1.0, (double)g[0]) : (double)g[r]);
      double m1 = (r == 1 ? __dadd_rn(1.0, (double)g[4]) : (double)g[3 + r]);
      double m2 = (r == 2 ? __dadd_rn(1.0, (double)g[8]) : (double)g[6 + r]);
      w[r] = __dadd_rn(__dadd_rn(__dmul_rn(m0, v[0]), __dmul_rn(m1, v[1])), __dmul_rn(m2, v[2]));
    }
    double ss = __dadd_rn(__dadd_rn(__dmul_rn(w[0], w[0]), __dmul_rn(w[1], w[1])), __dmul_rn(w[2], w[2]));
    float mag = __double2float_rn(sqrt(ss));
    if (mag != 0.0f)
#pragma unroll
      for (int r = 0; r < 3; ++r) w[r] = __ddiv_rn(w[r], (double)mag);
    float m2 = __fmul_rn(mag, mag);
    q = make_float4(__double2float_rn(w[0]), __double2float_rn(w[1]), __double2float_rn(w[2]),
                    __double2float_rn(__dmul_rn((double)m2, (double)q.w)));
  }
  out[t] = q;
}


FDT_DEV double dm(double x, double y) { return __dmul_rn(x, y); }
FDT_DEV double da(double x, double y) { return __dadd_rn(x, y); }

__global__ __launch_bounds__(256) void pred_kernel(PredArgs a) {
  const int l = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nvox || l >= a.N) return;
  const double b = a.bvals[l];
  const double gx = a.bvecs[l], gy = a.bvecs[a.N + l], gz = a.bvecs[2 * a.N + l];
  const float d = a.d[i];
  // stick form: exp(-b d (g.x)^2) for model 1, and for model 2 with d_std <= 1e-5;
  // ball form: exp(-b d) for model 1 and whenever d_std <= 1e-5
  const bool small = a.model >= 2 && a.dstd[i] <= 1e-5f;
  const bool stick1 = a.model == 1 || (a.model == 2 && small);
  const bool ball1 = a.model == 1 || small;
  float sig2 = 0.0f, dalpha = 0.0f;
  if (!stick1 || !ball1) {
    sig2 = __fmul_rn(a.dstd[i], a.dstd[i]);
    dalpha = __fdiv_rn(__fmul_rn(d, d), sig2);
  }
  const float Rv = a.model == 3 ? a.R[i] : 0.0f;
  float out = 0.0f, sumf = 0.0f;
  for (int k = 0; k < a.K; ++k) {
    const float f = a.f[(size_t)k * a.nvox + i];
    sumf = __fadd_rn(sumf, f);
    const float* dy = a.dyads + (size_t)k * 3 * a.nvox;
    const float angp = __double2float_rn(
        da(da(dm(dy[i], gx), dm(dy[a.nvox + i], gy)), dm(dy[2 * (size_t)a.nvox + i], gz)));
    double e;
    if (stick1) {
      e = exp(dm(dm(dm(-b, d), angp), angp));
    } else if (a.model == 2) {
      e = exp(dm(dalpha, log(__ddiv_rn(d, da(d, dm(dm(dm(b, angp), angp), sig2))))));
    } else {
      const float t = __fadd_rn(__fmul_rn(__fmul_rn(__fsub_rn(1.0f, Rv), angp), angp), Rv);
      e = exp(dm(__ddiv_rn(dm(dm(-b, 3.0), d), da(__fmul_rn(2.0f, Rv), 1.0)), t));
    }
    out = __double2float_rn(da(out, dm(f, e)));
  }
  const double iso = ball1 ? exp(dm(-b, d)) : exp(dm(dalpha, log(__ddiv_rn(d, da(d, dm(b, sig2))))));
  if (a.f0v) {
    const float f0 = a.f0v[i];
    out = __double2float_rn(da(out, da(f0, dm(__fsub_rn(__fsub_rn(1.0f, sumf), f0), iso))));
  } else {
    out = __double2float_rn(da(out, dm(__fsub_rn(1.0f, sumf), iso)));
  }
  a.out[(size_t)l * a.nvox + i] = __fmul_rn(out, a.S0[i]);
}

}  // namespace fdt

extern "C" hipError_t fdt_launch_pred(const fdt::PredArgs* a, hipStream_t s) {
  if (a->nvox <= 0 || a->N <= 0) return hipSuccess;
  dim3 grid((a->nvox + 255) / 256, a->N), block(256);
  hipLaunchKernelGGL(fdt::pred_kernel, grid, block, 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_grad_acq(const float4* base, int N, int stride, const float* grad, int nvox,
                                          float4* out, hipStream_t s) {
  if (nvox <= 0) return hipSuccess;
  if (N > stride) return hipErrorInvalidValue;
  size_t n = (size_t)nvox * stride;
  hipLaunchKernelGGL(fdt::grad_acq_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, base, N, stride, grad,
                     nvox, out);
  return hipGetLastError();
}
