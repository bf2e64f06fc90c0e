// fdt_fit.hip -- per-voxel Levenberg-Marquardt initialisation on gfx950.
//
// One wave64 per voxel, lanes striding over the diffusion volumes.  The whole
// per-voxel pipeline of the reference's fit() (CUDA/xfibres_gpu.cu:173-348)
// runs inside the wave without host round trips:
//   tensor fit (DTI::linfit, diffmodels.cc:17-46, done per voxel on the host
//   by the reference, CUDA/diffmodels.cu:57-94) -> start values
//   (CUDA/diffmodels.cu:72-93, 149-183 with PVM_single_c::fit_pvf,
//   diffmodels.cc:561-603) -> LM stage(s) -> finalisation -> the --f0
//   implausibility refit -> Rician tau (CUDA/diffmodels.cu:259-347).
// The residual, Jacobian, gradient (2 J^T r) and Gauss-Newton Hessian
// (2 J^T J) are accumulated per lane and summed across the wave with DPP row
// reductions and lane reads; the damped normal equations are solved in FP64 by LU with
// partial pivoting (the algorithm of CUDA/solver_mult_inverse.cu:14-166) in a
// per-wave LDS workspace.  The model and Jacobian run in T = float
// (production) or T = double (parity mode).
#include "fdt_device.h"
#include "fdt_kernels.h"
#include <type_traits>

namespace fdt {

FDT_DEV int lane_id() { return __lane_id(); }

// The voxel's volumes, streamed: the acquisition (gx, gy, gz, b) is staged in
// LDS for the workgroup, the signal row is read from the voxel-major data.
// With --rician, non-positive samples are replaced on the fly by the smallest
// positive one (remove_NonPositive_entries, CUDA/xfibres_gpu.cu:350-367).
struct Vols {
  const float* y;
  const float4* acq;
  int N;
  bool san;
  float ysub;
  FDT_DEV float yv(int i) const {
    float v = y[i];
    return (san && v <= 0.0f) ? ysub : v;
  }
};

template <class T>
FDT_DEV T wsum(T v);
template <>
FDT_DEV float wsum<float>(float v) { return uni(wave_sum(v)); }
template <>
FDT_DEV double wsum<double>(double v) { return uni(wave_sum_d(v)); }

// ------------------------------------------------------------------ models
// Forward model, Jacobian row and cost of the three LM models with the
// reference's exact formulas (including its sign conventions):
//   PVM_single   CUDA/PVM_single.cu:107-373
//   PVM_single_c CUDA/PVM_single_c.cu:104-514
//   PVM_multi    CUDA/PVM_multi.cu:19-466
template <class T, int KIND, int K, bool F0>
struct Mdl {
  static constexpr int FB = (KIND == KIND_MULTI) ? 3 : 2;
  static constexpr int NP = FB + 3 * K + (F0 ? 1 : 0);
  T fs[K], x0[K], x1[K], x2[K], sth[K], cth[K], sph[K], cph[K];
  T fder[K][K];
  T sumf, d, lam, a, bb, tf0, part;
  int G;
  T Rv, invR;

  FDT_DEV void prep(const T* p) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      int kk = FB + 3 * k;
      sth[k] = sin(p[kk + 1]); cth[k] = cos(p[kk + 1]);
      sph[k] = sin(p[kk + 2]); cph[k] = cos(p[kk + 2]);
      x0[k] = sth[k] * cph[k]; x1[k] = sth[k] * sph[k]; x2[k] = cth[k];
    }
    sumf = T(0);
    if constexpr (KIND == KIND_SINGLE_C) {
      lam = p[1];
      d = p[1] * p[1];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        T pt = T(1);
#pragma unroll
        for (int j = 0; j < k; ++j) pt -= fs[j];
        fs[k] = beta2f<T>(p[2 + 3 * k]) * pt;
        sumf += fs[k];
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int j = 0; j < K; ++j) fder[j][k] = T(0);
        T s2b = sin(T(2) * p[2 + 3 * k]);
#pragma unroll
        for (int j = 0; j < K; ++j) {
          if (j == k) {
            T fsum = T(1);
#pragma unroll
            for (int n = 0; n < j; ++n) fsum -= fs[n];
            fder[j][k] = s2b * fsum;
          } else if (j > k) {
            T sp = sin(p[2 + 3 * j]);
            T fsum = T(0);
#pragma unroll
            for (int n = 0; n < j; ++n) fsum += fder[n][k];
            fder[j][k] = -(sp * sp) * fsum;
          }
        }
      }
      part = T(1);
#pragma unroll
      for (int j = 0; j < K; ++j) part -= fs[j];
      if constexpr (F0) tf0 = beta2f<T>(p[NP - 1]) * part;
    } else if constexpr (KIND == KIND_SINGLE) {
      d = fabs(p[1]);
#pragma unroll
      for (int k = 0; k < K; ++k) { fs[k] = x2f<T>(p[2 + 3 * k]); sumf += fs[k]; }
      if constexpr (F0) tf0 = x2f<T>(p[NP - 1]);
    } else {
      a = fabs(p[1]);
      bb = fabs(p[2]);
#pragma unroll
      for (int k = 0; k < K; ++k) { fs[k] = x2f<T>(p[3 + 3 * k]); sumf += fs[k]; }
      if constexpr (F0) tf0 = x2f<T>(p[NP - 1]);
    }
  }

  FDT_DEV T aniso(T b, T dp) const {
    if constexpr (KIND != KIND_MULTI) return exp(-b * d * dp * dp);
    else {
      if (G == 1) return exp(-b * a * bb * dp * dp);
      if (G == 2) return exp(-b * T(3) * a * bb * invR * ((T(1) - Rv) * dp * dp + Rv));
      return exp(-a * log(T(1) + b * bb * (dp * dp)));
    }
  }

  FDT_DEV T pred(const T* p, T b, T gx, T gy, T gz) const {
    T err = T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      T dp = gx * x0[k] + gy * x1[k] + gz * x2[k];
      err += fs[k] * aniso(b, dp);
    }
    if constexpr (KIND == KIND_MULTI) {
      T iso = exp(-a * log(T(1) + b * bb));
      T S0 = fabs(p[0]);
      if constexpr (F0) return S0 * (tf0 + ((T(1) - sumf - tf0) * iso + err));
      else return S0 * ((T(1) - sumf) * iso + err);
    } else {
      T iso = exp(-b * d);
      if constexpr (F0) return p[0] * ((tf0 + (T(1) - sumf - tf0) * iso) + err);
      else return p[0] * ((T(1) - sumf) * iso + err);
    }
  }

  // Jacobian row for the gradient (Jg); the Hessian row equals it except in
  // column 0 for PVM_multi (CUDA/PVM_multi.cu:317 vs :451), returned in J0h.
  FDT_DEV T jrow(const T* p, T b, T gx, T gy, T gz, T* Jg, T& J0h) const {
#pragma unroll
    for (int j = 0; j < NP; ++j) Jg[j] = T(0);
    T sig = T(0);
    if constexpr (KIND == KIND_SINGLE) {
      T S0 = p[0], sg1 = p[1] > T(0) ? T(1) : T(-1);
      T iso = exp(-b * d), iso_d = -b * iso;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int kk = 2 + 3 * k;
        T dp = gx * x0[k] + gy * x1[k] + gz * x2[k];
        T e = exp(-b * d * dp * dp);
        T dp1t = cth[k] * (gx * cph[k] + gy * sph[k]) - gz * sth[k];
        T dp1p = sth[k] * (-gx * sph[k] + gy * cph[k]);
        sig += fs[k] * e;
        Jg[1] += sg1 * S0 * fs[k] * (-b * dp * dp * e);
        Jg[kk] = S0 * (e - iso) * T(TWO_PI_D) * sgn<T>(p[kk]) * T(1) / (T(1) + p[kk] * p[kk]);
        Jg[kk + 1] = S0 * fs[k] * (T(-2) * b * d * dp * dp1t * e);
        Jg[kk + 2] = S0 * fs[k] * (T(-2) * b * d * dp * dp1p * e);
      }
      if constexpr (F0) {
        Jg[NP - 1] = S0 * (T(1) - iso) * T(TWO_PI_D) * sgn<T>(p[NP - 1]) * T(1) / (T(1) + p[NP - 1] * p[NP - 1]);
        sig = S0 * ((tf0 + (T(1) - sumf - tf0) * iso) + sig);
        Jg[1] += sg1 * S0 * (T(1) - sumf - tf0) * iso_d;
      } else {
        sig = S0 * ((T(1) - sumf) * iso + sig);
        Jg[1] += sg1 * S0 * (T(1) - sumf) * iso_d;
      }
      Jg[0] = sig / S0;
      J0h = Jg[0];
    } else if constexpr (KIND == KIND_SINGLE_C) {
      T S0 = p[0];
      T iso = exp(-b * d);
      T A[K], dps[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        dps[k] = gx * x0[k] + gy * x1[k] + gz * x2[k];
        A[k] = exp(-b * d * dps[k] * dps[k]);
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int kk = 2 + 3 * k;
        T dp = dps[k];
        sig += fs[k] * A[k];
        Jg[1] += S0 * fs[k] * (T(-2) * b * lam * dp * dp * exp(-b * lam * lam * dp * dp));
        T jb = T(0);
#pragma unroll
        for (int j = 0; j < K; ++j)
          if (fder[j][k] != T(0)) jb += S0 * (A[j] - iso) * fder[j][k];
        Jg[kk] = jb;
        T dp1t = cth[k] * (gx * cph[k] + gy * sph[k]) - gz * sth[k];
        T dp1p = sth[k] * (-gx * sph[k] + gy * cph[k]);
        Jg[kk + 1] = S0 * fs[k] * (T(-2) * b * d * dp * dp1t * A[k]);
        Jg[kk + 2] = S0 * fs[k] * (T(-2) * b * d * dp * dp1p * A[k]);
      }
      T isol = T(-2) * b * lam * exp(-b * lam * lam);
      if constexpr (F0) {
        Jg[NP - 1] = S0 * (T(1) - iso) * sin(T(2) * p[NP - 1]) * part;
        sig = S0 * ((tf0 + (T(1) - sumf - tf0) * iso) + sig);
        Jg[1] += S0 * (T(1) - sumf - tf0) * isol;
      } else {
        sig = S0 * ((T(1) - sumf) * iso + sig);
        Jg[1] += S0 * (T(1) - sumf) * isol;
      }
      Jg[0] = sig / S0;
      J0h = Jg[0];
    } else {
      T S0a = fabs(p[0]);
      T sg1 = p[1] > T(0) ? T(1) : T(-1), sg2 = p[2] > T(0) ? T(1) : T(-1);
      T la = log(T(1) + b * bb);
      T iso = exp(-a * la);
      T iso_a = -la * iso;
      T iso_b = -a * b / (T(1) + b * bb) * iso;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int kk = 3 + 3 * k;
        T dp = gx * x0[k] + gy * x1[k] + gz * x2[k];
        T dp1t = cth[k] * (gx * cph[k] + gy * sph[k]) - gz * sth[k];
        T dp1p = sth[k] * (-gx * sph[k] + gy * cph[k]);
        T an, an_a, an_b, an_th, an_ph;
        if (G == 1) {
          an = exp(-b * a * bb * dp * dp);
          an_a = -b * bb * dp * dp * an;
          an_b = -b * a * dp * dp * an;
          an_th = T(-2) * b * a * bb * dp * dp1t * an;
          an_ph = T(-2) * b * a * bb * dp * dp1p * an;
        } else if (G == 2) {
          T t = (T(1) - Rv) * dp * dp + Rv;
          an = exp(-b * T(3) * a * bb * invR * t);
          T qa = b * T(3) * bb * invR * t;
          an_a = -qa * exp(-qa * a);
          T qb = b * T(3) * a * invR * t;
          an_b = -qb * exp(-qb * bb);
          // the reference's G=2 angle derivative carries no factor (g.x)
          // (CUDA/PVM_multi.cu:77-78, :95-96); kept, it decides where LM stops
          an_th = -(T(2) * b * T(3) * a * bb * invR * (T(1) - Rv) * dp1t) * an;
          an_ph = -(T(2) * b * T(3) * a * bb * invR * (T(1) - Rv) * dp1p) * an;
        } else {
          T den = T(1) + b * (dp * dp) * bb;
          T l1 = log(den);
          an = exp(-a * l1);
          an_a = -l1 * an;
          an_b = -a * b * (dp * dp) / den * an;
          T c = -a * bb * b / den * an;
          an_th = c * T(2) * dp * dp1t;
          an_ph = c * T(2) * dp * dp1p;
        }
        sig += fs[k] * an;
        Jg[1] += sg1 * S0a * fs[k] * an_a;
        Jg[2] += sg2 * S0a * fs[k] * an_b;
        Jg[kk] = S0a * (an - iso) * T(TWO_PI_D) * sgn<T>(p[kk]) * T(1) / (T(1) + p[kk] * p[kk]);
        Jg[kk + 1] = S0a * fs[k] * an_th;
        Jg[kk + 2] = S0a * fs[k] * an_ph;
      }
      if constexpr (F0) {
        Jg[NP - 1] = S0a * (T(1) - iso) * T(TWO_PI_D) * sgn<T>(p[NP - 1]) * T(1) / (T(1) + p[NP - 1] * p[NP - 1]);
        sig = S0a * ((tf0 + (T(1) - sumf - tf0) * iso) + sig);
        Jg[1] += sg1 * S0a * (T(1) - sumf - tf0) * iso_a;
        Jg[2] += sg2 * S0a * (T(1) - sumf - tf0) * iso_b;
      } else {
        sig = S0a * ((T(1) - sumf) * iso + sig);
        Jg[1] += sg1 * S0a * (T(1) - sumf) * iso_a;
        Jg[2] += sg2 * S0a * (T(1) - sumf) * iso_b;
      }
      T sg0 = p[0] > T(0) ? T(1) : T(-1);
      Jg[0] = sg0 * sig / p[0];
      J0h = sig / p[0];
    }
    return sig;
  }
};

// cost = sum_i (pred_i - y_i)^2 with FP64 accumulation
template <class T, int KIND, int K, bool F0>
FDT_DEV double lm_cost(const Vols& V, const T* p, int G, T Rv, T invR) {
  Mdl<T, KIND, K, F0> m;
  m.G = G; m.Rv = Rv; m.invR = invR;
  m.prep(p);
  double c = 0.0;
  for (int i = lane_id(); i < V.N; i += 64) {
    float4 q = V.acq[i];
    T e = m.pred(p, (T)q.w, (T)q.x, (T)q.y, (T)q.z) - (T)V.yv(i);
    c += (double)e * (double)e;
  }
  return uni(wave_sum_d(c));
}

// LU with partial pivoting and substitution, distributed over the wave: lane
// i (< NP) holds row i of the matrix and entry i of the right-hand side.  The
// elimination order, the pivot rule (first maximum of the updated column
// below the diagonal; a zero pivot skips elimination) and the interchange of
// the right-hand side follow CUDA/solver_mult_inverse.cu:14-166.  Pivot
// search, row swaps and the pivot-row broadcast use v_readlane, so the solve
// stays in registers.
template <int NP>
FDT_DEV void lu_rows(const double (&A)[NP], double b, double (&x)[NP]) {
  const int lane = lane_id();
  double r[NP];
#pragma unroll
  for (int l = 0; l < NP; ++l) r[l] = A[l];
  double rb = b;
  int mu = 0;
  double big = fabs(rld(r[0], 0));
#pragma unroll
  for (int i = 1; i < NP; ++i) {
    double t = fabs(rld(r[0], i));
    if (big < t) { big = t; mu = i; }
  }
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    if (mu != k) {
#pragma unroll
      for (int l = 0; l < NP; ++l) {
        double vk = rld(r[l], k), vm = rld(r[l], mu);
        r[l] = lane == k ? vm : (lane == mu ? vk : r[l]);
      }
      double bk = rld(rb, k), bm = rld(rb, mu);
      rb = lane == k ? bm : (lane == mu ? bk : rb);
    }
    double pr[NP];
#pragma unroll
    for (int l = k; l < NP; ++l) pr[l] = rld(r[l], k);
    const double diag = pr[k];
    mu = k + 1;
    big = 0.0;
    if (diag != 0.0) {
      if (lane > k && lane < NP) {
        double mult = r[k] / diag;
        r[k] = mult;
#pragma unroll
        for (int l = k + 1; l < NP; ++l) r[l] = r[l] - mult * pr[l];
      }
      const int kn = k + 1 < NP ? k + 1 : k;
#pragma unroll
      for (int i = k + 1; i < NP; ++i) {
        double t = fabs(rld(r[kn], i));
        if (big < t) { big = t; mu = i; }
      }
    }
  }
  // forward substitution with the unit lower factor (lane i: L[i][j] = r[j], j < i)
  double acc = rb;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    double yj = rld(acc, j);
    if (lane > j && lane < NP) acc = acc - r[j] * yj;
  }
  // back substitution with the upper factor
#pragma unroll
  for (int i = NP - 1; i >= 0; --i) {
    double xi = rld(acc, i) / rld(r[i], i);
    x[i] = xi;
    if (lane < i) acc = acc - r[i] * xi;
  }
}

// The Levenberg loop of CUDA/levenberg_marquardt.cu:24-115: lambda0 = 0.1,
// /10 after a success, x10 after a failure (the Hessian diagonal carries
// lambda - lambda_old), relative cost test 1e-8, ltol 1e20, 200 successes.
template <class T, int KIND, int K, bool F0>
FDT_DEV int lm_stage(const Vols& V, T* p, int G, T Rv, T invR) {
  using M = Mdl<T, KIND, K, F0>;
  constexpr int NP = M::NP;
  constexpr int NH = NP * (NP + 1) / 2;
  const int lane = lane_id();
  double lambda = 0.1, olambda = 0.0;
  const double cftol = 1.0e-8, ltol = 1.0e20;
  bool success = true, end = false;
  int niter = 0, total = 0;
  double pcf = lm_cost<T, KIND, K, F0>(V, p, G, Rv, invR), ncf;
  T step[NP];
  double Hrow[NP], grow = 0.0;  // lane i < NP: row i of the damped Hessian, gradient entry i
#pragma unroll
  for (int i = 0; i < NP; ++i) Hrow[i] = 0.0;
  while (!(success && niter++ >= 200)) {
    ++total;
    if (success) {
      M m;
      m.G = G; m.Rv = Rv; m.invR = invR;
      m.prep(p);
      T gacc[NP], hacc[NH];
#pragma unroll
      for (int j = 0; j < NP; ++j) gacc[j] = T(0);
#pragma unroll
      for (int j = 0; j < NH; ++j) hacc[j] = T(0);
      for (int i = lane; i < V.N; i += 64) {
        {
          float4 q = V.acq[i];
          T Jg[NP], J0h;
          T sig = m.jrow(p, (T)q.w, (T)q.x, (T)q.y, (T)q.z, Jg, J0h);
          T diff = sig - (T)V.yv(i);
#pragma unroll
          for (int a = 0; a < NP; ++a) gacc[a] += T(2) * Jg[a] * diff;
          int h = 0;
#pragma unroll
          for (int a = 0; a < NP; ++a) {
            T ja = a == 0 ? J0h : Jg[a];
#pragma unroll
            for (int c = a; c < NP; ++c) {
              T jc = c == 0 ? J0h : Jg[c];
              hacc[h++] += T(2) * (ja * jc);
            }
          }
        }
      }
#pragma unroll
      for (int a = 0; a < NP; ++a) {
        double ga = (double)wsum<T>(gacc[a]);
        grow = lane == a ? ga : grow;
      }
      int h = 0;
#pragma unroll
      for (int a = 0; a < NP; ++a)
#pragma unroll
        for (int c = a; c < NP; ++c) {
          double hv = (double)wsum<T>(hacc[h++]);
          if (lane == a) Hrow[c] = hv;
          if (lane == c) Hrow[a] = hv;
        }
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) if (lane == i) Hrow[i] += (lambda - olambda);
    double sol[NP];
    lu_rows<NP>(Hrow, grow, sol);
#pragma unroll
    for (int i = 0; i < NP; ++i) step[i] = p[i] - (T)sol[i];
    ncf = lm_cost<T, KIND, K, F0>(V, step, G, Rv, invR);
    if ((success = (ncf < pcf))) {
      olambda = 0.0;
#pragma unroll
      for (int i = 0; i < NP; ++i) p[i] = step[i];
      lambda = lambda / 10.0;
      if (2.0 * fabs(pcf - ncf) <= cftol * (fabs(pcf) + fabs(ncf) + 2.0e-16)) end = true;
      pcf = ncf;
    } else {
      olambda = lambda;
      lambda = lambda * 10.0;
      if (lambda > ltol) end = true;
    }
    if (end) break;
  }
  return total;
}

// ------------------------------------------------------- finalisation
template <class T, int K>
FDT_DEV void sort_fibres(int fb, T* p) {  // bubble sort by f descending
#pragma unroll
  for (int i = 1; i < K; ++i)
#pragma unroll
    for (int j = 0; j < K - i; ++j)
      if (p[fb + j * 3] < p[fb + (j + 1) * 3]) {
#pragma unroll
        for (int t = 0; t < 3; ++t) { T x = p[fb + j * 3 + t]; p[fb + j * 3 + t] = p[fb + (j + 1) * 3 + t]; p[fb + (j + 1) * 3 + t] = x; }
      }
}

// CUDA/PVM_single.cu:441-468, PVM_single_c.cu:581-611, PVM_multi.cu:548-571
template <class T, int KIND, int K, bool F0>
FDT_DEV void finalise(T* p) {
  constexpr int NP = Mdl<T, KIND, K, F0>::NP;
  if constexpr (KIND == KIND_SINGLE) {
    p[1] = fabs(p[1]);
#pragma unroll
    for (int k = 0; k < K; ++k) p[2 + 3 * k] = x2f<T>(p[2 + 3 * k]);
    if constexpr (F0) p[NP - 1] = x2f<T>(p[NP - 1]);
    sort_fibres<T, K>(2, p);
    T sum = F0 ? p[NP - 1] : T(0);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      sum += p[2 + 3 * i];
      if (sum >= T(1)) {
#pragma unroll
        for (int j = 0; j < K; ++j) if (j >= i) p[2 + 3 * j] = T(FSMALL_F);
        break;
      }
    }
  } else if constexpr (KIND == KIND_SINGLE_C) {
    p[1] = p[1] * p[1];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      T pt = T(1);
#pragma unroll
      for (int j = 0; j < k; ++j) pt -= p[2 + 3 * j];
      p[2 + 3 * k] = beta2f<T>(p[2 + 3 * k]) * pt;
    }
    if constexpr (F0) {
      T pt = T(1);
#pragma unroll
      for (int j = 0; j < K; ++j) pt -= p[2 + 3 * j];
      p[NP - 1] = beta2f<T>(p[NP - 1]) * pt;
    }
    sort_fibres<T, K>(2, p);
  } else {
    T aux = p[1];
    p[1] = fabs(aux * p[2]);
    p[2] = sqrt(fabs(aux * p[2] * p[2]));
#pragma unroll
    for (int k = 0; k < K; ++k) p[3 + 3 * k] = x2f<T>(p[3 + 3 * k]);
    if constexpr (F0) p[NP - 1] = x2f<T>(p[NP - 1]);
    sort_fibres<T, K>(3, p);
    T sum = F0 ? p[NP - 1] : T(0);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (p[3 + 3 * i] == T(0)) p[3 + 3 * i] = T(FSMALL_F);
      sum += p[3 + 3 * i];
      if (sum >= T(1)) {
#pragma unroll
        for (int j = 0; j < K; ++j) if (j >= i) p[3 + 3 * j] = T(FSMALL_F);
        break;
      }
    }
  }
}

// ------------------------------------------------------------ tensor fit
FDT_DEV void jacobi3(double A[9], double w[3], double V[9]) {
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

// eigenvalues descending with vectors; sign: largest-magnitude component
// positive (the reference's newmat sign is unspecified; LM is invariant).
FDT_DEV void eig3_desc(const double T_[9], double l[3], double v[3][3]) {
  double A[9], w[3], V[9];
  for (int i = 0; i < 9; ++i) A[i] = T_[i];
  jacobi3(A, w, V);
  int o[3] = {0, 1, 2};
  // sort descending (stable on ties by index)
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2 - i; ++j)
      if (w[o[j]] < w[o[j + 1]]) { int t = o[j]; o[j] = o[j + 1]; o[j + 1] = t; }
  for (int k = 0; k < 3; ++k) {
    l[k] = w[o[k]];
    for (int i = 0; i < 3; ++i) v[k][i] = V[i * 3 + o[k]];
    int m = 0;
    for (int i = 1; i < 3; ++i) if (fabs(v[k][i]) > fabs(v[k][m])) m = i;
    if (v[k][m] < 0) for (int i = 0; i < 3; ++i) v[k][i] = -v[k][i];
  }
}

struct Dti { double s0, md, fa, l[3], v[3][3]; };

FDT_DEV void cart2sph(const double v[3], double& th, double& ph) {
  double mag = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (mag == 0.0) { th = M_PI / 2; ph = M_PI / 2; return; }
  ph = atan2(v[1], v[0]);
  th = atan2(sqrt(v[0] * v[0] + v[1] * v[1]), v[2]);
}

// DTI::linfit (diffmodels.cc:17-46) + calc_tensor_parameters (:239-265)
FDT_DEV Dti dti_fit(const Vols& V, const double* pinvA) {
  const int lane = lane_id();
  const int N = V.N;
  double ysum = 0;
  float ymx = 0.0f;
  for (int i = lane; i < N; i += 64) { float y = V.yv(i); ysum += (double)y; ymx = fmaxf(ymx, fabsf(y)); }
  ysum = uni(wave_sum_d(ysum));
  double ymax = (double)uni(-wave_min(-ymx));
  double mean = ysum / N;
  double D[7];
  for (int j = 0; j < 7; ++j) D[j] = 0;
  for (int i = lane; i < N; i += 64) {
    float y = V.yv(i);
    double l = y > 0.0f ? log((double)y) : 0.0;
#pragma unroll
    for (int j = 0; j < 7; ++j) D[j] += pinvA[j * N + i] * l;
  }
#pragma unroll
  for (int j = 0; j < 7; ++j) D[j] = -uni(wave_sum_d(D[j]));
  double s0 = (D[6] > -23) ? exp(-D[6]) : ymax;
  if (s0 < mean) s0 = ymax;
  for (int j = 0; j < 7; ++j) D[j] = 0;
  for (int i = lane; i < N; i += 64) {
    double y = (double)V.yv(i);
    double l = (y / s0) > 0.01 ? log(y) : log(0.01 * s0);
#pragma unroll
    for (int j = 0; j < 7; ++j) D[j] += pinvA[j * N + i] * l;
  }
#pragma unroll
  for (int j = 0; j < 7; ++j) D[j] = -uni(wave_sum_d(D[j]));
  s0 = exp(-D[6]);
  if (s0 < mean) s0 = mean;
  Dti o;
  o.s0 = s0;
  double Tn[9] = {D[0], D[1], D[2], D[1], D[3], D[4], D[2], D[4], D[5]};
  eig3_desc(Tn, o.l, o.v);
  o.md = (o.l[0] + o.l[1] + o.l[2]) / 3.0;
  double numer = 1.5 * ((o.l[0] - o.md) * (o.l[0] - o.md) + (o.l[1] - o.md) * (o.l[1] - o.md) + (o.l[2] - o.md) * (o.l[2] - o.md));
  double denom = o.l[0] * o.l[0] + o.l[1] * o.l[1] + o.l[2] * o.l[2];
  double fa = denom > 0 ? numer / denom : 0.0;
  o.fa = fa > 0 ? sqrt(fa) : 0.0;
  return o;
}

// PVM_single start values (CUDA/diffmodels.cu:72-93)
template <class T, int K, bool F0>
FDT_DEV void start_single(const Dti& dti, T* p) {
  constexpr int NP = 2 + 3 * K + (F0 ? 1 : 0);
  double th, ph;
  cart2sph(dti.v[0], th, ph);
  p[0] = (T)dti.s0;
  p[1] = (T)(dti.md > 0 ? dti.md * 2 : 0.001);
  p[2] = dti.fa < 1 ? f2x<T>((T)dti.fa) : f2x<T>((T)0.95);
  p[3] = (T)th; p[4] = (T)ph;
  T sumf = x2f<T>(p[2]);
#pragma unroll
  for (int ii = 2; ii <= K; ++ii) {
    const int i = 5 + 3 * (ii - 2);
    T denom = T(2), tmpsumf;
    do {
      p[i] = f2x<T>(x2f<T>(p[i - 3]) / denom);
      denom *= T(2);
      tmpsumf = sumf + x2f<T>(p[i]);
    } while (tmpsumf >= T(1));
    sumf += x2f<T>(p[i]);
    cart2sph(dti.v[ii - 1], th, ph);
    p[i + 1] = (T)th; p[i + 2] = (T)ph;
  }
  if constexpr (F0) p[NP - 1] = f2x<T>((T)FSMALL_F);
}

// PVM_single_c start values with fit_pvf (CUDA/diffmodels.cu:149-183,
// diffmodels.cc:561-603): LLS for the fractions through the pseudo-inverse
// of the K-column design, formed from wave-reduced M^T M and M^T (y - s0 iso).
template <class T, int K, bool F0>
FDT_DEV void start_single_c(const Dti& dti, const Vols& V, T* p) {
  const int N = V.N;
  constexpr int NP = 2 + 3 * K + (F0 ? 1 : 0);
  double th, ph;
  p[0] = (T)dti.s0;
  p[1] = sqrt((T)(dti.l[0] > 0 ? dti.l[0] : 0.002));
#pragma unroll
  for (int k = 0; k < K; ++k) {
    cart2sph(dti.v[k], th, ph);
    p[3 + 3 * k] = (T)th; p[4 + 3 * k] = (T)ph;
  }
  T s0 = p[0], d = p[1] * p[1];
  T dir[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    dir[k][0] = sin(p[3 + 3 * k]) * cos(p[4 + 3 * k]);
    dir[k][1] = sin(p[3 + 3 * k]) * sin(p[4 + 3 * k]);
    dir[k][2] = cos(p[3 + 3 * k]);
  }
  double MtM[K * K], MtY[K];
#pragma unroll
  for (int a = 0; a < K * K; ++a) MtM[a] = 0;
#pragma unroll
  for (int a = 0; a < K; ++a) MtY[a] = 0;
  for (int i = lane_id(); i < N; i += 64) {
    {
      float4 q = V.acq[i];
      T b = (T)q.w;
      T iso = exp(-b * d);
      double yi = (double)((T)V.yv(i) - s0 * iso);
      double m[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        T dp = (T)q.x * dir[k][0] + (T)q.y * dir[k][1] + (T)q.z * dir[k][2];
        m[k] = (double)(s0 * (exp(-b * d * dp * dp) - iso));
      }
#pragma unroll
      for (int a = 0; a < K; ++a) {
        MtY[a] += m[a] * yi;
#pragma unroll
        for (int c = 0; c < K; ++c) MtM[a * K + c] += m[a] * m[c];
      }
    }
  }
#pragma unroll
  for (int a = 0; a < K * K; ++a) MtM[a] = uni(wave_sum_d(MtM[a]));
#pragma unroll
  for (int a = 0; a < K; ++a) MtY[a] = uni(wave_sum_d(MtY[a]));
  // pinv(M) y = V diag(1/s^2) V^T M^T y, singular values s <= N*s_max*1e-15 dropped
  double fsd[K];
  {
    double A[9] = {0}, w[3], Vv[9];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int c = 0; c < K; ++c) A[a * 3 + c] = MtM[a * K + c];
    for (int a = K; a < 3; ++a) A[a * 3 + a] = 0.0;
    jacobi3(A, w, Vv);
    double smax = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) smax = fmax(smax, sqrt(fmax(w[j], 0.0)));
    double tol = (double)(N > K ? N : K) * smax * 1e-15;
#pragma unroll
    for (int a = 0; a < K; ++a) fsd[a] = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      double s = sqrt(fmax(w[j], 0.0));
      if (s > tol) {
        double proj = 0;
#pragma unroll
        for (int c = 0; c < K; ++c) proj += Vv[c * 3 + j] * MtY[c];
        proj /= (s * s);
#pragma unroll
        for (int a = 0; a < K; ++a) fsd[a] += Vv[a * 3 + j] * proj;
      }
    }
  }
  T fs[K];
#pragma unroll
  for (int k = 0; k < K; ++k) fs[k] = (T)fsd[k];
  T f0v = T(0);
  if constexpr (F0) { f0v = T(FSMALL_F); fs[0] -= f0v; }
#pragma unroll
  for (int k = 0; k < K; ++k) fs[k] = fabs(fs[k]);
  {
    T sumf = T(0);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      sumf += fs[i];
      if (sumf >= T(1)) {
#pragma unroll
        for (int j = 0; j < K; ++j) if (j >= i) fs[j] = T(FSMALL_F);
        break;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    T pt = T(1);
#pragma unroll
    for (int j = 0; j < k; ++j) pt -= fs[j];
    T r = fs[k] / pt;
    if (r > T(1)) r = T(1);
    p[2 + 3 * k] = f2beta<T>(r);
  }
  if constexpr (F0) {
    T pt = T(1);
#pragma unroll
    for (int j = 0; j < K; ++j) pt -= fs[j];
    T r = f0v / pt;
    if (r > T(1)) r = T(1);
    p[NP - 1] = f2beta<T>(r);
  }
}

// PVM_multi start from the finalised stage-1 vector (CUDA/PVM_multi.cu:516-534)
template <class T, int K, bool F0>
FDT_DEV void start_multi(const T* s1, T* p) {
  constexpr int NP1 = 2 + 3 * K + (F0 ? 1 : 0), NP = NP1 + 1;
  p[0] = s1[0];
  p[1] = T(1);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int ii = 3 + 3 * i;
    p[ii] = f2x<T>(s1[ii - 1]);
    p[ii + 1] = s1[ii];
    p[ii + 2] = s1[ii + 1];
  }
  p[2] = s1[1];
  if constexpr (F0) p[NP - 1] = f2x<T>(s1[NP1 - 1]);
}

// ---------------------------------------------------------- pipelines
// model 1: PVM_single (--nonlinear) or PVM_single_c; model >= 2: PVM_single
// then PVM_multi (refit after an implausible --f0 fit: PVM_single_c then
// PVM_multi, CUDA/xfibres_gpu.cu:331-333).  Output [S0,d,(dstd),f,th,ph..,(f0)].
template <class T, int K, bool F0>
FDT_DEV int run_model1(const Vols& V, const Dti& dti, bool nonlin, T* out) {
  constexpr int NP = 2 + 3 * K + (F0 ? 1 : 0);
  T p[NP];
  int it;
  if (nonlin) {
    start_single<T, K, F0>(dti, p);
    it = lm_stage<T, KIND_SINGLE, K, F0>(V, p, 0, T(0), T(1));
    finalise<T, KIND_SINGLE, K, F0>(p);
  } else {
    start_single_c<T, K, F0>(dti, V, p);
    it = lm_stage<T, KIND_SINGLE_C, K, F0>(V, p, 0, T(0), T(1));
    finalise<T, KIND_SINGLE_C, K, F0>(p);
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) out[i] = p[i];
  return it;
}

template <class T, int K, bool F0>
FDT_DEV int run_multi(const Vols& V, const Dti& dti, bool refit, int G, T Rv, T* out) {
  constexpr int NP1 = 2 + 3 * K + (F0 ? 1 : 0), NP = NP1 + 1;
  T s1[NP1], p[NP];
  int it;
  if (!refit) {
    start_single<T, K, F0>(dti, s1);
    it = lm_stage<T, KIND_SINGLE, K, F0>(V, s1, 0, T(0), T(1));
    finalise<T, KIND_SINGLE, K, F0>(s1);
  } else {
    start_single_c<T, K, F0>(dti, V, s1);
    it = lm_stage<T, KIND_SINGLE_C, K, F0>(V, s1, 0, T(0), T(1));
    finalise<T, KIND_SINGLE_C, K, F0>(s1);
  }
  start_multi<T, K, F0>(s1, p);
  T invR = T(1) / (T(2) * Rv + T(1));
  it += lm_stage<T, KIND_MULTI, K, F0>(V, p, G, Rv, invR);
  finalise<T, KIND_MULTI, K, F0>(p);
#pragma unroll
  for (int i = 0; i < NP; ++i) out[i] = p[i];
  return it;
}

// 1/var of the residuals of the final fit (N-1 normalisation)
template <class T, int K>
FDT_DEV T tau_of(const Vols& V, int model, bool f0, const T* q, T Rv) {
  const int fb = model >= 2 ? 3 : 2;
  const int N = V.N;
  int npo = fb + 3 * K + (f0 ? 1 : 0);
  T S0 = model >= 2 ? fabs(q[0]) : q[0];
  T dd = q[1], a_ = T(0), b_ = T(0);
  if (model >= 2) { a_ = q[1] * q[1] / q[2] / q[2]; b_ = q[2] * q[2] / q[1]; }
  T iR = T(1) / (T(2) * Rv + T(1));
  T f0v = f0 ? q[npo - 1] : T(0);
  T sumf = T(0), x[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    sumf += q[fb + 3 * k];
    T th = q[fb + 3 * k + 1], ph = q[fb + 3 * k + 2];
    x[k][0] = sin(th) * cos(ph); x[k][1] = sin(th) * sin(ph); x[k][2] = cos(th);
  }
  // two passes: mean, then centred sum of squares
  double sums[2] = {0.0, 0.0}, mean = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    double acc = 0.0;
    for (int i = lane_id(); i < N; i += 64) {
      float4 qq = V.acq[i];
      T b = (T)qq.w;
      T iso = model >= 2 ? exp(-a_ * log(T(1) + b * b_)) : exp(-b * dd);
      T an = T(0);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        T dp = (T)qq.x * x[k][0] + (T)qq.y * x[k][1] + (T)qq.z * x[k][2];
        T e;
        if (model == 1) e = exp(-b * dd * dp * dp);
        else if (model == 3) e = exp(-b * T(3) * a_ * b_ * iR * ((T(1) - Rv) * dp * dp + Rv));
        else e = exp(-a_ * log(T(1) + b * b_ * (dp * dp)));
        an += q[fb + 3 * k] * e;
      }
      T pred = S0 * (f0v + (T(1) - sumf - f0v) * iso + an);
      double r = (double)V.yv(i) - (double)pred;
      acc += pass == 0 ? r : (r - mean) * (r - mean);
    }
    sums[pass] = uni(wave_sum_d(acc));
    if (pass == 0) mean = sums[0] / N;
  }
  return (T)(1.0 / (sums[1] / (N - 1)));
}

template <class T, int K>
FDT_DEV void fit_pipeline(const FitArgs& a, const Vols& V, int vox) {
  const int lane = lane_id();
  Dti dti = dti_fit(V, a.pinvA);
  T out[NPMAX];
  int npo = a.out_np;
  int it = 0;
  bool used_f0 = a.f0 != 0;
  T Rv = (T)a.Rmean;
  if (a.model == 1) {
    if (a.f0) {
      it = run_model1<T, K, true>(V, dti, a.nonlin, out);
      T md = out[1], mf = out[2], f0v = out[npo - 1];
      if ((K > 0 && mf < T(0.05)) || md > T(0.007) || f0v > T(0.4)) {
        it += run_model1<T, K, false>(V, dti, a.nonlin, out);
        out[npo - 1] = T(0.001);
        used_f0 = false;
      }
    } else {
      it = run_model1<T, K, false>(V, dti, a.nonlin, out);
    }
  } else {
    int G = a.model == 3 ? 2 : 0;
    if (a.f0) {
      it = run_multi<T, K, true>(V, dti, false, G, Rv, out);
      T md = out[1], mf = out[3], f0v = out[npo - 1];
      if ((K > 0 && mf < T(0.05)) || md > T(0.007) || f0v > T(0.4)) {
        it += run_multi<T, K, false>(V, dti, true, G, Rv, out);
        out[npo - 1] = T(0.001);
        used_f0 = false;
      }
    } else {
      it = run_multi<T, K, false>(V, dti, false, G, Rv, out);
    }
  }
  T tau = T(0);
  if (a.rician) tau = tau_of<T, K>(V, a.model, used_f0, out, Rv);
  if (lane < npo) {
    T v = out[0];
#pragma unroll
    for (int i = 1; i < NPMAX; ++i) if (i == lane) v = out[i];
    a.out[(size_t)vox * npo + lane] = (float)v;
  }
  if (lane == 0) {
    if (a.tau) a.tau[vox] = (float)tau;
    if (a.iters) a.iters[vox] = it;
  }
}

template <class T, int K, int KIND, bool F0>
FDT_DEV void fit_stage(const FitArgs& a, const Vols& V, int vox) {
  constexpr int NP = Mdl<T, KIND, K, F0>::NP;
  const int lane = lane_id();
  T p[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) p[i] = (T)a.start[(size_t)vox * NP + i];
  T Rv = (T)a.R;
  int it = lm_stage<T, KIND, K, F0>(V, p, a.G, Rv, T(1) / (T(2) * Rv + T(1)));
  finalise<T, KIND, K, F0>(p);
  if (lane < NP) {
    T v = p[0];
#pragma unroll
    for (int i = 1; i < NP; ++i) if (i == lane) v = p[i];
    a.out[(size_t)vox * NP + lane] = (float)v;
  }
  if (lane == 0 && a.iters) a.iters[vox] = it;
}

template <class T, int K, bool F0>
FDT_DEV void fit_start(const FitArgs& a, const Vols& V, int vox) {
  constexpr int NP = 2 + 3 * K + (F0 ? 1 : 0);
  const int lane = lane_id();
  Dti dti = dti_fit(V, a.pinvA);
  T p[NP];
  if (a.kind == KIND_SINGLE) start_single<T, K, F0>(dti, p);
  else start_single_c<T, K, F0>(dti, V, p);
  if (lane < NP) {
    T v = p[0];
#pragma unroll
    for (int i = 1; i < NP; ++i) if (i == lane) v = p[i];
    a.out[(size_t)vox * NP + lane] = (float)v;
  }
}

template <int K, bool FP64>
__global__ __launch_bounds__(64 * FIT_WAVES) void fit_kernel(FitArgs a) {
  using T = typename std::conditional<FP64, double, float>::type;
  __shared__ float4 sacq[MAX_N];
  for (int i = threadIdx.x; i < a.N; i += blockDim.x)
    sacq[i] = make_float4(a.bvecs[i], a.bvecs[a.N + i], a.bvecs[2 * a.N + i], a.bvals[i]);
  __syncthreads();
  const int wave = threadIdx.x / 64;
  const int lane = lane_id();
  const int vox = blockIdx.x * FIT_WAVES + wave;
  if (vox >= a.nvox) return;
  Vols V;
  V.y = a.data + (size_t)vox * a.N;
  V.acq = sacq;
  V.N = a.N;
  V.san = false;
  V.ysub = 0.0f;
  if (a.rician && a.mode == FIT_PIPELINE) {
    float mn = 3.0e38f, mx = -3.0e38f, mnp = 3.0e38f;
    for (int i = lane; i < a.N; i += 64) {
      float v = V.y[i];
      mn = fminf(mn, v); mx = fmaxf(mx, v);
      if (v > 0.0f) mnp = fminf(mnp, v);
    }
    mn = uni(wave_min(mn)); mx = uni(-wave_min(-mx)); mnp = uni(wave_min(mnp));
    V.san = (mn <= 0.0f && mx > 0.0f);
    V.ysub = mnp;
  }
  if (a.mode == FIT_PIPELINE) {
    fit_pipeline<T, K>(a, V, vox);
  } else if (a.mode == FIT_STAGE) {
    if (a.kind == KIND_SINGLE) {
      if (a.stage_f0) fit_stage<T, K, KIND_SINGLE, true>(a, V, vox);
      else fit_stage<T, K, KIND_SINGLE, false>(a, V, vox);
    } else if (a.kind == KIND_SINGLE_C) {
      if (a.stage_f0) fit_stage<T, K, KIND_SINGLE_C, true>(a, V, vox);
      else fit_stage<T, K, KIND_SINGLE_C, false>(a, V, vox);
    } else {
      if (a.stage_f0) fit_stage<T, K, KIND_MULTI, true>(a, V, vox);
      else fit_stage<T, K, KIND_MULTI, false>(a, V, vox);
    }
  } else {
    if (a.stage_f0) fit_start<T, K, true>(a, V, vox);
    else fit_start<T, K, false>(a, V, vox);
  }
}

}  // namespace fdt

// One instantiation per translation unit (K x precision), compiled in
// parallel: -DFDT_FIT_K=1..3 -DFDT_FIT_FP64=0/1.
#ifndef FDT_FIT_K
#define FDT_FIT_K 3
#endif
#ifndef FDT_FIT_FP64
#define FDT_FIT_FP64 0
#endif
#define FDT_CAT3_(a, b, c) a##b##_##c
#define FDT_CAT3(a, b, c) FDT_CAT3_(a, b, c)
extern "C" hipError_t FDT_CAT3(fdt_launch_fit_k, FDT_FIT_K, FDT_FIT_FP64)(const fdt::FitArgs* a, hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  dim3 grid((a->nvox + fdt::FIT_WAVES - 1) / fdt::FIT_WAVES), block(64 * fdt::FIT_WAVES);
  hipLaunchKernelGGL((fdt::fit_kernel<FDT_FIT_K, FDT_FIT_FP64 != 0>), grid, block, 0, s, *a);
  return hipGetLastError();
}
