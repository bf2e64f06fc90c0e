// fdt_finish.hip -- per-voxel sample post-processing and data staging.
//
// finish_kernel: the work the reference does on the host after copying the
// samples back (record_finish_voxels, CUDA/xfibres_gpu.cu:659-738, with
// Samples::record/finish_voxel/save, CUDA/samples.cu:134-347): fibres are
// sorted by descending sum over samples of f (ties: higher fibre index first,
// the order of the reference's std::sort over (sum, index) pairs read from
// the end, CUDA/samples.cu:297-330), means are formed, and the dyad of each
// fibre is the principal eigenvector of the mean dyadic tensor with
// dispersion 1 - lambda_max (make_dyadic_vectors.cc:65-88).  One thread per
// voxel; every sample array is [.][s][v] so the sample loop is coalesced
// across the wave.
#include "fdt_device.h"
#include "fdt_kernels.h"

namespace fdt {

FDT_DEV void jacobi3_sym(double A[9], double w[3], double V[9]) {
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    double tot = off * 2 + A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off == 0.0 || off <= 1e-30 * tot) break;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      double apq = A[p * 3 + q];
      if (apq == 0.0) continue;
      double app = A[p * 3 + p], aqq = A[q * 3 + q];
      double theta = (aqq - app) / (2 * apq);
      double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
      double c = 1 / sqrt(t * t + 1), s = t * c;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double akp = A[k * 3 + p], akq = A[k * 3 + q];
        A[k * 3 + p] = c * akp - s * akq;
        A[k * 3 + q] = s * akp + c * akq;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double apk = A[p * 3 + k], aqk = A[q * 3 + k];
        A[p * 3 + k] = c * apk - s * aqk;
        A[q * 3 + k] = s * apk + c * aqk;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double vkp = V[k * 3 + p], vkq = V[k * 3 + q];
        V[k * 3 + p] = c * vkp - s * vkq;
        V[k * 3 + q] = s * vkp + c * vkq;
      }
    }
  }
  w[0] = A[0]; w[1] = A[4]; w[2] = A[8];
}

template <int K>
__global__ __launch_bounds__(256) void finish_kernel(FinishArgs a) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.nvox) return;
  const size_t nv = (size_t)a.nvox;
  const int ns = a.ns;
  double sf[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double acc = 0;
    for (int s = 0; s < ns; ++s) acc += (double)a.f[((size_t)k * ns + s) * nv + v];
    sf[k] = acc;
  }
  // descending by (sum, index)
  int ord[K];
#pragma unroll
  for (int k = 0; k < K; ++k) ord[k] = k;
#pragma unroll
  for (int i = 0; i < K; ++i)
#pragma unroll
    for (int j = 0; j < K - 1 - i; ++j) {
      int x = ord[j], y = ord[j + 1];
      bool swap = (sf[x] < sf[y]) || (sf[x] == sf[y] && x < y);
      if (swap) { ord[j] = y; ord[j + 1] = x; }
    }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    int src = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) if (q == k) src = ord[q];
    double T[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < ns; ++s) {
      size_t is = ((size_t)src * ns + s) * nv + v, id = ((size_t)k * ns + s) * nv + v;
      float th = a.th[is], ph = a.ph[is], ff = a.f[is];
      if (a.oth) a.oth[id] = th;
      if (a.oph) a.oph[id] = ph;
      if (a.of) a.of[id] = ff;
      double st = sin((double)th), ct = cos((double)th), sp = sin((double)ph), cp = cos((double)ph);
      double x = st * cp, y = st * sp, z = ct;
      T[0] += x * x; T[1] += x * y; T[2] += x * z; T[3] += y * y; T[4] += y * z; T[5] += z * z;
    }
    if (a.mean_f) a.mean_f[(size_t)k * nv + v] = (float)(sf[src] / ns);
    if (a.dyads || a.disp) {
      double A[9] = {T[0] / ns, T[1] / ns, T[2] / ns, T[1] / ns, T[3] / ns, T[4] / ns, T[2] / ns, T[4] / ns, T[5] / ns};
      double w[3], V[9];
      jacobi3_sym(A, w, V);
      int m = 0;
      if (w[1] > w[m]) m = 1;
      if (w[2] > w[m]) m = 2;
      double e[3] = {V[0 * 3 + m], V[1 * 3 + m], V[2 * 3 + m]};
      int big = 0;
      for (int i = 1; i < 3; ++i) if (fabs(e[i]) > fabs(e[big])) big = i;
      if (e[big] < 0) for (int i = 0; i < 3; ++i) e[i] = -e[i];
      if (a.dyads) for (int i = 0; i < 3; ++i) a.dyads[((size_t)k * 3 + i) * nv + v] = (float)e[i];
      if (a.disp) a.disp[(size_t)k * nv + v] = (float)(1.0 - fmax(fabs(w[0]), fmax(fabs(w[1]), fabs(w[2]))));
    }
  }
  auto mean_of = [&](const float* q) {
    double acc = 0;
    for (int s = 0; s < ns; ++s) acc += (double)q[(size_t)s * nv + v];
    return (float)(acc / ns);
  };
  if (a.mean_S0) a.mean_S0[v] = mean_of(a.S0);
  if (a.mean_d) a.mean_d[v] = mean_of(a.d);
  if (a.mean_dstd && a.dstd) a.mean_dstd[v] = mean_of(a.dstd);
  if (a.mean_R && a.R) a.mean_R[v] = mean_of(a.R);
  if (a.mean_f0 && a.f0s) a.mean_f0[v] = mean_of(a.f0s);
  if (a.mean_tau && a.taus) a.mean_tau[v] = mean_of(a.taus);
}

// What the CPU xfibres keeps of each finished voxel for the starts of its
// neighbours (Samples::record / finish_voxel, xfibres.cc:215-303), in chain
// fibre order (the CPU path does not sort fibres): float sums over the
// samples divided by the sample count, the dyadic tensor summed in double
// from float direction components, its principal eigenvector (sign: largest
// component positive), and the count of fibres whose mean f exceeds 0.05.
// One thread per listed voxel.
__global__ __launch_bounds__(256) void spatial_stats_kernel(SpatialStatsArgs a) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.count) return;
  const int v = a.list ? a.list[q] : q;
  const size_t nv = (size_t)a.vstride;
  const int ns = a.ns;
  float sd = 0.0f, sS0 = 0.0f;
  for (int s = 0; s < ns; ++s) {
    sd += a.d[(size_t)s * nv + v];
    sS0 += a.S0[(size_t)s * nv + v];
  }
  a.mean_d[v] = sd / (float)ns;
  a.mean_S0[v] = sS0 / (float)ns;
  int nfib = 0;
  for (int k = 0; k < a.K; ++k) {
    float sf = 0.0f;
    double T[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < ns; ++s) {
      const size_t is = ((size_t)k * ns + s) * nv + v;
      const float th = a.th[is], ph = a.ph[is];
      sf += a.f[is];
      const double x = (double)(sinf(th) * cosf(ph)), y = (double)(sinf(th) * sinf(ph)), z = (double)cosf(th);
      T[0] += x * x; T[1] += x * y; T[2] += x * z; T[3] += y * y; T[4] += y * z; T[5] += z * z;
    }
    const float mf = sf / (float)ns;
    a.mean_f[(size_t)k * nv + v] = mf;
    if ((double)mf > 0.05) ++nfib;
    double A[9] = {T[0], T[1], T[2], T[1], T[3], T[4], T[2], T[4], T[5]};
    double w[3], V[9];
    jacobi3_sym(A, w, V);
    int m = 0;
    if (w[1] > w[m]) m = 1;
    if (w[2] > w[m]) m = 2;
    double e[3] = {V[0 * 3 + m], V[1 * 3 + m], V[2 * 3 + m]};
    int big = 0;
    for (int i = 1; i < 3; ++i) if (fabs(e[i]) > fabs(e[big])) big = i;
    if (e[big] < 0) for (int i = 0; i < 3; ++i) e[i] = -e[i];
    for (int i = 0; i < 3; ++i) a.dyad[((size_t)k * 3 + i) * nv + v] = e[i];
  }
  a.been[v] = nfib;
}

// [N][nvox] -> [nvox][N] through 32x33 LDS tiles (the reference transposes on
// the host, prepare_data_gpu_FIT, CUDA/xfibres_gpu.cu:388-399)
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in, float* __restrict__ out, int N, int nvox) {
  __shared__ float tile[32][33];
  int v0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    int i = i0 + r, v = v0 + tx;
    if (i < N && v < nvox) tile[r][tx] = in[(size_t)i * nvox + v];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    int v = v0 + r, i = i0 + tx;
    if (i < N && v < nvox) out[(size_t)v * N + i] = tile[tx][r];
  }
}

// Implausible-f0 test of CUDA/xfibres_gpu.cu:216-246, 305-339 on the fitted
// parameters: (K > 0 && f_1 < 0.05) || d > 0.007 || f0 > 0.4 puts the voxel on
// the refit list (list order is irrelevant: voxels are fitted independently).
__global__ void f0_flag_kernel(F0FlagArgs a) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.nvox) return;
  const float* q = a.params + (size_t)v * a.np;
  float md = q[1], mf = q[a.fb], f0v = q[a.np - 1];
  bool bad = (a.K > 0 && mf < 0.05f) || md > 0.007f || f0v > 0.4f;
  a.refit[v] = bad ? 1 : 0;
  if (bad) a.list[atomicAdd(a.count, 1)] = v;
}

// work-queue counters to zero with a kernel (capturable in a graph)
__global__ void zero_ints_kernel(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace fdt

extern "C" hipError_t fdt_launch_f0_flag(const fdt::F0FlagArgs* a, hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  hipLaunchKernelGGL(fdt::f0_flag_kernel, dim3((a->nvox + 255) / 256), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_zero_ints(int* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fdt::zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_finish(const fdt::FinishArgs* a, hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  dim3 grid((a->nvox + 255) / 256), block(256);
  if (a->K == 1) hipLaunchKernelGGL(fdt::finish_kernel<1>, grid, block, 0, s, *a);
  else if (a->K == 2) hipLaunchKernelGGL(fdt::finish_kernel<2>, grid, block, 0, s, *a);
  else if (a->K == 3) hipLaunchKernelGGL(fdt::finish_kernel<3>, grid, block, 0, s, *a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_spatial_stats(const fdt::SpatialStatsArgs* a, hipStream_t s) {
  if (a->count <= 0) return hipSuccess;
  if (a->ns <= 0 || a->K < 1 || a->K > 3) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fdt::spatial_stats_kernel, dim3((a->count + 255) / 256), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t fdt_launch_transpose(const float* in, float* out, int N, int nvox, hipStream_t s) {
  if (nvox <= 0 || N <= 0) return hipSuccess;
  dim3 grid((nvox + 31) / 32, (N + 31) / 32), block(256);
  hipLaunchKernelGGL(fdt::transpose_kernel, grid, block, 0, s, in, out, N, nvox);
  return hipGetLastError();
}
