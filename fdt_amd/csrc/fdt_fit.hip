// fdt_fit.hip -- per-voxel Levenberg-Marquardt initialisation on gfx950.
//
// One voxel per 16-lane DPP row, four voxels per wave64, 16 per 256-thread
// workgroup.  Launches are persistent: each row takes a voxel from a device
// work queue, runs one LM stage on it to convergence and takes the next, so a
// slow voxel never holds back the rest of its wave.  The stages of the
// reference's fit() (CUDA/xfibres_gpu.cu:173-348) are separate launches:
//   tensor fit (DTI::linfit, diffmodels.cc:17-46, done per voxel on the host
//   by the reference, CUDA/diffmodels.cu:57-94) -> start values
//   (CUDA/diffmodels.cu:72-93, 149-183 with PVM_single_c::fit_pvf,
//   diffmodels.cc:561-603) -> PVM_single / PVM_single_c LM -> finalisation;
//   PVM_multi LM from the stage-1 result; the --f0 implausibility refit over a
//   device list; Rician tau (CUDA/diffmodels.cu:259-347).
// Lane l of a row holds volumes l, l+16, ...; the voxel's signal row is
// staged in LDS next to the workgroup's acquisition table.  The residual,
// gradient (2 J^T r) and Gauss-Newton Hessian (2 J^T J) are accumulated per
// lane and summed over the row with DPP; the damped normal equations are
// solved in FP64 by LU with partial pivoting (the algorithm of
// CUDA/solver_mult_inverse.cu:14-166) with row i of the matrix in lane i.  The
// model and Jacobian run in T = float (production) or T = double (parity
// mode).
#include "fdt_device.h"
#include "fdt_kernels.h"
#include <type_traits>

namespace fdt {

// The voxel as seen by its row: lane rl, the voxel's volumes in LDS.
struct Row {
  int rl;
  int N;
  const float4* acq;  // [N]: the workgroup's table in LDS, or the voxel's own (--gradnonlin)
  const float* y;     // LDS, the voxel's signal (sanitised for --rician)
  // b shells (acquisitions with <= 16 distinct b values, not --gradnonlin):
  // sh[i] the shell of volume i (LDS), shb[l] its b value, isot the row's
  // table of the cost model's isotropic signal per shell (LDS); nshell 0 = none
  int nshell = 0;
  const int* sh = nullptr;
  const float* shb = nullptr;
  double* isot = nullptr;
};

template <class T>
FDT_DEV T rsum(T v) {
  if constexpr (std::is_same<T, double>::value) return row_sum16_d(v);
  else return row_sum16(v);
}

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
  // per-stage constants of the Jacobian rows, hoisted out of the volume loop:
  // d f / d x of the f = x2f(x) fractions (and of f0), and 1 / p[0]
  T dfx[K], dfx0, ip0;
  int G;
  T Rv, invR;
  // exp / log of the model and Jacobian rows: the hardware exp2 / log2 with one
  // scaling in FP32 (a few ulp, below the FP32 noise of the fit); exp_d /
  // log_d (1-3 ulp) in FP64
  static FDT_DEV T ex(T x) {
    if constexpr (std::is_same<T, float>::value) return __expf(x);
    else return exp_d(x);
  }
  static FDT_DEV T rc(T x) {  // 1/x: hardware reciprocal (1 ulp) in FP32
    if constexpr (std::is_same<T, float>::value) return __builtin_amdgcn_rcpf(x);
    else return T(1) / x;
  }
  static FDT_DEV T lg(T x) {  // x >= 1 in every model here
    if constexpr (std::is_same<T, float>::value) return __builtin_amdgcn_logf(x) * 0.6931471805599453f;
    else return log_d(x);
  }

  template <int L>
  static FDT_DEV T bc(T v) {
    if constexpr (std::is_same<T, double>::value) return row_bcast_d<L>(v);
    else return row_bcast<L>(v);
  }
  // The parameters are the same in every lane of the row, so the per-row
  // transcendentals are evaluated once: lane l takes sin and cos of p[l] (and,
  // for the x2f fractions, x2f(p[l]) and its derivative), and the row
  // broadcasts each from its lane, instead of all 16 lanes evaluating all of
  // them (a double sin or cos is ~150 instructions with its reduction).
  FDT_DEV void prep(const T* p) {
    const int rl = __lane_id() & 15;
    // lane l's parameter where l is an angle (theta, phi) or an x2f fraction
    // slot, 0 elsewhere (no lane takes the slow argument reduction of sin/cos
    // for S0 or a diffusivity)
    T pl = T(0);
#pragma unroll
    for (int i = FB; i < NP; ++i) if (rl == i) pl = p[i];
    const T sl = sin(pl), cl = cos(pl);
    sfor<K>([&](auto kc) {
      constexpr int k = kc, kk = FB + 3 * k;
      sth[k] = bc<kk + 1>(sl); cth[k] = bc<kk + 1>(cl);
      sph[k] = bc<kk + 2>(sl); cph[k] = bc<kk + 2>(cl);
      x0[k] = sth[k] * cph[k]; x1[k] = sth[k] * sph[k]; x2[k] = cth[k];
    });
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
    } else {
      if constexpr (KIND == KIND_SINGLE) d = fabs(p[1]);
      else { a = fabs(p[1]); bb = fabs(p[2]); }
      // the fractions f = x2f(x) and d f / d x from the lane of each x
      const T fl = x2f<T>(pl), dl = T(TWO_PI_D) * sgn<T>(pl) * T(1) / (T(1) + pl * pl);
      sfor<K>([&](auto kc) {
        constexpr int k = kc;
        fs[k] = bc<FB + 3 * k>(fl);
        dfx[k] = bc<FB + 3 * k>(dl);
        sumf += fs[k];
      });
      if constexpr (F0) { tf0 = bc<NP - 1>(fl); dfx0 = bc<NP - 1>(dl); }
    }
    ip0 = T(1) / p[0];
  }

  FDT_DEV T aniso(T b, T dp) const {
    if constexpr (KIND != KIND_MULTI) return ex(-b * d * dp * dp);
    else {
      if (G == 1) return ex(-b * a * bb * dp * dp);
      if (G == 2) return ex(-b * T(3) * a * bb * invR * ((T(1) - Rv) * dp * dp + Rv));
      return ex(-a * lg(T(1) + b * bb * (dp * dp)));
    }
  }

  // isotropic signal at b (a function of b alone: evaluated once per b shell
  // where the acquisition has shells)
  FDT_DEV T iso_val(T b) const {
    if constexpr (KIND == KIND_MULTI) return ex(-a * lg(T(1) + b * bb));
    else return ex(-b * d);
  }
  FDT_DEV T pred_iso(const T* p, T b, T gx, T gy, T gz, T iso) const {
    T err = T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      T dp = gx * x0[k] + gy * x1[k] + gz * x2[k];
      err += fs[k] * aniso(b, dp);
    }
    if constexpr (KIND == KIND_MULTI) {
      T S0 = fabs(p[0]);
      if constexpr (F0) return S0 * (tf0 + ((T(1) - sumf - tf0) * iso + err));
      else return S0 * ((T(1) - sumf) * iso + err);
    } else {
      if constexpr (F0) return p[0] * ((tf0 + (T(1) - sumf - tf0) * iso) + err);
      else return p[0] * ((T(1) - sumf) * iso + err);
    }
  }
  FDT_DEV T pred(const T* p, T b, T gx, T gy, T gz) const { return pred_iso(p, b, gx, gy, gz, iso_val(b)); }

  // Jacobian row for the gradient (Jg); the Hessian row equals it except in
  // column 0 for PVM_multi (CUDA/PVM_multi.cu:317 vs :451), returned in J0h.
  FDT_DEV T jrow(const T* p, T b, T gx, T gy, T gz, T* Jg, T& J0h) const {
#pragma unroll
    for (int j = 0; j < NP; ++j) Jg[j] = T(0);
    T sig = T(0);
    if constexpr (KIND == KIND_SINGLE) {
      T S0 = p[0], sg1 = p[1] > T(0) ? T(1) : T(-1);
      T iso = ex(-b * d), iso_d = -b * iso;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int kk = 2 + 3 * k;
        T dp = gx * x0[k] + gy * x1[k] + gz * x2[k];
        T e = ex(-b * d * dp * dp);
        T dp1t = cth[k] * (gx * cph[k] + gy * sph[k]) - gz * sth[k];
        T dp1p = sth[k] * (-gx * sph[k] + gy * cph[k]);
        sig += fs[k] * e;
        Jg[1] += sg1 * S0 * fs[k] * (-b * dp * dp * e);
        Jg[kk] = S0 * (e - iso) * dfx[k];
        Jg[kk + 1] = S0 * fs[k] * (T(-2) * b * d * dp * dp1t * e);
        Jg[kk + 2] = S0 * fs[k] * (T(-2) * b * d * dp * dp1p * e);
      }
      if constexpr (F0) {
        Jg[NP - 1] = S0 * (T(1) - iso) * dfx0;
        sig = S0 * ((tf0 + (T(1) - sumf - tf0) * iso) + sig);
        Jg[1] += sg1 * S0 * (T(1) - sumf - tf0) * iso_d;
      } else {
        sig = S0 * ((T(1) - sumf) * iso + sig);
        Jg[1] += sg1 * S0 * (T(1) - sumf) * iso_d;
      }
      Jg[0] = sig * ip0;
      J0h = Jg[0];
    } else if constexpr (KIND == KIND_SINGLE_C) {
      T S0 = p[0];
      T iso = ex(-b * d);
      T A[K], dps[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        dps[k] = gx * x0[k] + gy * x1[k] + gz * x2[k];
        A[k] = ex(-b * d * dps[k] * dps[k]);
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int kk = 2 + 3 * k;
        T dp = dps[k];
        sig += fs[k] * A[k];
        Jg[1] += S0 * fs[k] * (T(-2) * b * lam * dp * dp * A[k]);  // exp(-b lam^2 (g.x)^2) = A[k]
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
      T isol = T(-2) * b * lam * ex(-b * lam * lam);
      if constexpr (F0) {
        Jg[NP - 1] = S0 * (T(1) - iso) * sin(T(2) * p[NP - 1]) * part;
        sig = S0 * ((tf0 + (T(1) - sumf - tf0) * iso) + sig);
        Jg[1] += S0 * (T(1) - sumf - tf0) * isol;
      } else {
        sig = S0 * ((T(1) - sumf) * iso + sig);
        Jg[1] += S0 * (T(1) - sumf) * isol;
      }
      Jg[0] = sig * ip0;
      J0h = Jg[0];
    } else {
      T S0a = fabs(p[0]);
      T sg1 = p[1] > T(0) ? T(1) : T(-1), sg2 = p[2] > T(0) ? T(1) : T(-1);
      T la = lg(T(1) + b * bb);
      T iso = ex(-a * la);
      T iso_a = -la * iso;
      T iso_b = -a * b * rc(T(1) + b * bb) * iso;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int kk = 3 + 3 * k;
        T dp = gx * x0[k] + gy * x1[k] + gz * x2[k];
        T dp1t = cth[k] * (gx * cph[k] + gy * sph[k]) - gz * sth[k];
        T dp1p = sth[k] * (-gx * sph[k] + gy * cph[k]);
        T an, an_a, an_b, an_th, an_ph;
        if (G == 1) {
          an = ex(-b * a * bb * dp * dp);
          an_a = -b * bb * dp * dp * an;
          an_b = -b * a * dp * dp * an;
          an_th = T(-2) * b * a * bb * dp * dp1t * an;
          an_ph = T(-2) * b * a * bb * dp * dp1p * an;
        } else if (G == 2) {
          T t = (T(1) - Rv) * dp * dp + Rv;
          an = ex(-b * T(3) * a * bb * invR * t);
          T qa = b * T(3) * bb * invR * t;
          an_a = -qa * ex(-qa * a);
          T qb = b * T(3) * a * invR * t;
          an_b = -qb * ex(-qb * bb);
          // the reference's G=2 angle derivative carries no factor (g.x)
          // (CUDA/PVM_multi.cu:77-78, :95-96); kept, it decides where LM stops
          an_th = -(T(2) * b * T(3) * a * bb * invR * (T(1) - Rv) * dp1t) * an;
          an_ph = -(T(2) * b * T(3) * a * bb * invR * (T(1) - Rv) * dp1p) * an;
        } else {
          T den = T(1) + b * (dp * dp) * bb;
          T rden = rc(den);
          T l1 = lg(den);
          an = ex(-a * l1);
          an_a = -l1 * an;
          an_b = -a * b * (dp * dp) * rden * an;
          T c = -a * bb * b * rden * an;
          an_th = c * T(2) * dp * dp1t;
          an_ph = c * T(2) * dp * dp1p;
        }
        sig += fs[k] * an;
        Jg[1] += sg1 * S0a * fs[k] * an_a;
        Jg[2] += sg2 * S0a * fs[k] * an_b;
        Jg[kk] = S0a * (an - iso) * dfx[k];
        Jg[kk + 1] = S0a * fs[k] * an_th;
        Jg[kk + 2] = S0a * fs[k] * an_ph;
      }
      if constexpr (F0) {
        Jg[NP - 1] = S0a * (T(1) - iso) * dfx0;
        sig = S0a * ((tf0 + (T(1) - sumf - tf0) * iso) + sig);
        Jg[1] += sg1 * S0a * (T(1) - sumf - tf0) * iso_a;
        Jg[2] += sg2 * S0a * (T(1) - sumf - tf0) * iso_b;
      } else {
        sig = S0a * ((T(1) - sumf) * iso + sig);
        Jg[1] += sg1 * S0a * (T(1) - sumf) * iso_a;
        Jg[2] += sg2 * S0a * (T(1) - sumf) * iso_b;
      }
      T sg0 = p[0] > T(0) ? T(1) : T(-1);
      J0h = sig * ip0;
      Jg[0] = sg0 * J0h;
    }
    return sig;
  }
};

// cost = sum_i (pred_i - y_i)^2 with FP64 accumulation.  With b shells the
// isotropic signal is evaluated once per shell (lane l: shell l) into the
// row's table and looked up per volume (the same value as per volume).
template <class T, class M>
FDT_DEV double row_cost(const Row& R, const M& m, const T* p) {
  double c = 0.0;
  if (R.nshell) {
    if (R.rl < R.nshell) R.isot[R.rl] = (double)m.iso_val((T)R.shb[R.rl]);
    for (int i = R.rl; i < R.N; i += 16) {
      float4 q = R.acq[i];
      T e = m.pred_iso(p, (T)q.w, (T)q.x, (T)q.y, (T)q.z, (T)R.isot[R.sh[i]]) - (T)R.y[i];
      c += (double)e * (double)e;
    }
  } else {
    for (int i = R.rl; i < R.N; i += 16) {
      float4 q = R.acq[i];
      T e = m.pred(p, (T)q.w, (T)q.x, (T)q.y, (T)q.z) - (T)R.y[i];
      c += (double)e * (double)e;
    }
  }
  return row_sum16_d(c);
}

// LU with partial pivoting and substitution, row i of the matrix and entry i
// of the right-hand side in lane i of the row.  The elimination order, the
// pivot rule (first maximum of the column at and below the diagonal; after a
// zero pivot no search and no interchange) and the interchange of the
// right-hand side follow CUDA/solver_mult_inverse.cu:14-166.  The pivot is
// found with a DPP row maximum and a ballot, rows are interchanged with
// ds_bpermute and the pivot row is broadcast with DPP row_newbcast.
template <int NP>
FDT_DEV void lu_row(const double (&A)[NP], double b, double (&x)[NP], int rl) {
  const int base = __lane_id() & ~15;
  double r[NP];
#pragma unroll
  for (int l = 0; l < NP; ++l) r[l] = A[l];
  double rb = b;
  bool search = true;
  [&]<int... ks>(std::integer_sequence<int, ks...>) {
    ([&] {
      constexpr int k = ks;
      int mu = k;
      if (search) {
        double v = (rl >= k && rl < NP) ? fabs(r[k]) : -1.0;
        double m = row_max16_d(v);
        uint64_t bal = __builtin_amdgcn_ballot_w64(v == m);
        mu = __builtin_ctz((uint32_t)(bal >> base) & 0xFFFFu);
      }
      if (mu != k) {
        int src = base + (rl == k ? mu : (rl == mu ? k : rl));
#pragma unroll
        for (int l = 0; l < NP; ++l) r[l] = bperm_d(src, r[l]);
        rb = bperm_d(src, rb);
      }
      double pr[NP];
#pragma unroll
      for (int l = k; l < NP; ++l) pr[l] = row_bcast_d<k>(r[l]);
      const double diag = pr[k];
      search = diag != 0.0;
      if (diag != 0.0 && rl > k && rl < NP) {
        double mult = r[k] / diag;
        r[k] = mult;
#pragma unroll
        for (int l = k + 1; l < NP; ++l) r[l] = r[l] - mult * pr[l];
      }
    }(), ...);
  }(std::make_integer_sequence<int, NP>{});
  // forward substitution with the unit lower factor (lane i: L[i][j] = r[j], j < i)
  double acc = rb;
  [&]<int... js>(std::integer_sequence<int, js...>) {
    ([&] {
      constexpr int j = js;
      double yj = row_bcast_d<j>(acc);
      if (rl > j && rl < NP) acc = acc - r[j] * yj;
    }(), ...);
  }(std::make_integer_sequence<int, NP>{});
  // back substitution with the upper factor
  [&]<int... is>(std::integer_sequence<int, is...>) {
    ([&] {
      constexpr int i = NP - 1 - is;
      double xi = row_bcast_d<i>(acc) / row_bcast_d<i>(r[i]);
      x[i] = xi;
      if (rl < i) acc = acc - r[i] * xi;
    }(), ...);
  }(std::make_integer_sequence<int, NP>{});
}

// Damped Gauss-Newton system of the row's voxel at parameters p: lane a < NP
// receives row a of 2 J^T J (Hessian Jacobian) and entry a of 2 J^T r.
template <class T, int KIND, int K, bool F0>
FDT_DEV void row_normal_eq(const Row& R, const Mdl<T, KIND, K, F0>& m, const T* p,
                           double (&Hrow)[Mdl<T, KIND, K, F0>::NP], double& grow) {
  constexpr int NP = Mdl<T, KIND, K, F0>::NP;
  constexpr int NH = NP * (NP + 1) / 2;
  T gacc[NP], hacc[NH];
#pragma unroll
  for (int j = 0; j < NP; ++j) gacc[j] = T(0);
#pragma unroll
  for (int j = 0; j < NH; ++j) hacc[j] = T(0);
  for (int i = R.rl; i < R.N; i += 16) {
    float4 q = R.acq[i];
    T Jg[NP], J0h;
    T sig = m.jrow(p, (T)q.w, (T)q.x, (T)q.y, (T)q.z, Jg, J0h);
    T diff = sig - (T)R.y[i];
    // J^T r and J^T J; the factor 2 of the gradient and Hessian is applied
    // after the row sums
#pragma unroll
    for (int a = 0; a < NP; ++a) gacc[a] += Jg[a] * diff;
    int h = 0;
#pragma unroll
    for (int a = 0; a < NP; ++a) {
      T ja = a == 0 ? J0h : Jg[a];
#pragma unroll
      for (int c = a; c < NP; ++c) {
        T jc = c == 0 ? J0h : Jg[c];
        hacc[h++] += ja * jc;
      }
    }
  }
  T hsel[NP], gsel = T(0);
#pragma unroll
  for (int a = 0; a < NP; ++a) {
    T ga = T(2) * rsum<T>(gacc[a]);
    gsel = R.rl == a ? ga : gsel;
  }
  int h = 0;
#pragma unroll
  for (int a = 0; a < NP; ++a)
#pragma unroll
    for (int c = a; c < NP; ++c) {
      T hv = T(2) * rsum<T>(hacc[h++]);
      if (R.rl == a) hsel[c] = hv;
      if (R.rl == c) hsel[a] = hv;
    }
  grow = (double)gsel;
#pragma unroll
  for (int c = 0; c < NP; ++c) Hrow[c] = (double)hsel[c];
}

// The Levenberg loop of CUDA/levenberg_marquardt.cu:24-115 as a resumable
// per-row state: lambda0 = 0.1, /10 after a success, x10 after a failure (the
// Hessian diagonal carries lambda - lambda_old), relative cost test 1e-8,
// ltol 1e20, stop at the 201st success test.
// T is the type of the Jacobian rows and the normal equations, TC that of the
// parameters and of the cost.  TC = double with T = float (the mixed mode)
// keeps every decision of the loop -- accept/reject and the 1e-8 relative
// stop test -- on an FP64 cost (an FP32 model's cost carries ~1e-7 relative
// noise, above the stop tolerance) while the O(N np^2) normal equations stay
// in FP32: they set the step direction only.  (Residuals of the FP64 model in
// the gradient, measured, changed nothing: the FP32 Jacobian sets where the
// steps lead along the flat directions of the fit.)
template <class T, class TC, int KIND, int K, bool F0>
struct LmState {
  static constexpr int NP = Mdl<T, KIND, K, F0>::NP;
  TC p[NP];
  double Hrow[NP], grow;
  double pcf, lambda, olambda;
  int niter, total;
  bool success;

  FDT_DEV void begin(const Row& R, const Mdl<TC, KIND, K, F0>& mc0) {
    Mdl<TC, KIND, K, F0> m = mc0;
    m.prep(p);
    pcf = row_cost<TC>(R, m, p);
    lambda = 0.1; olambda = 0.0;
    niter = 0; total = 0;
    success = true;
  }
  // one trip of the loop; returns true when the stage has ended
  FDT_DEV bool step(const Row& R, const Mdl<T, KIND, K, F0>& m0, const Mdl<TC, KIND, K, F0>& mc0) {
    if (success && niter++ >= 200) return true;
    ++total;
    if (success) {
      Mdl<T, KIND, K, F0> m = m0;
      T pj[NP];
#pragma unroll
      for (int i = 0; i < NP; ++i) pj[i] = (T)p[i];
      m.prep(pj);
      row_normal_eq<T, KIND, K, F0>(R, m, pj, Hrow, grow);
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) if (R.rl == i) Hrow[i] += (lambda - olambda);
    double sol[NP];
    lu_row<NP>(Hrow, grow, sol, R.rl);
    TC st[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) st[i] = p[i] - (TC)sol[i];
    Mdl<TC, KIND, K, F0> mc = mc0;
    mc.prep(st);
    double ncf = row_cost<TC>(R, mc, st);
    bool end = false;
    if ((success = (ncf < pcf))) {
      olambda = 0.0;
#pragma unroll
      for (int i = 0; i < NP; ++i) p[i] = st[i];
      lambda = lambda / 10.0;
      if (2.0 * fabs(pcf - ncf) <= 1.0e-8 * (fabs(pcf) + fabs(ncf) + 2.0e-16)) end = true;
      pcf = ncf;
    } else {
      olambda = lambda;
      lambda = lambda * 10.0;
      if (lambda > 1.0e20) end = true;
    }
    return end;
  }
};

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
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i] = T_[i];
  jacobi3(A, w, V);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    l[k] = w[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[k][i] = V[i * 3 + k];
  }
  // sort descending, stable on ties (compare-exchange, no register indexing)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2 - i; ++j) {
      bool sw = l[j] < l[j + 1];
      double t = l[j];
      l[j] = sw ? l[j + 1] : t;
      l[j + 1] = sw ? t : l[j + 1];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double u = v[j][c];
        v[j][c] = sw ? v[j + 1][c] : u;
        v[j + 1][c] = sw ? u : v[j + 1][c];
      }
    }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double big = v[k][0];
#pragma unroll
    for (int i = 1; i < 3; ++i) big = fabs(v[k][i]) > fabs(big) ? v[k][i] : big;
    if (big < 0) {
#pragma unroll
      for (int i = 0; i < 3; ++i) v[k][i] = -v[k][i];
    }
  }
}

struct Dti { double s0, md, fa, l[3], v[3][3]; };

FDT_DEV void cart2sph(const double v[3], double& th, double& ph) {
  double mag = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (mag == 0.0) { th = M_PI / 2; ph = M_PI / 2; return; }
  ph = atan2(v[1], v[0]);
  th = atan2(sqrt(v[0] * v[0] + v[1] * v[1]), v[2]);
}

FDT_DEV Dti tensor_of(double s0, const double (&D)[7]);

// DTI::linfit (diffmodels.cc:17-46) + calc_tensor_parameters (:239-265)
FDT_DEV Dti dti_row(const Row& R, const double* pinvA) {
  const int N = R.N;
  double ysum = 0;
  float ymx = 0.0f;
  for (int i = R.rl; i < N; i += 16) { float y = R.y[i]; ysum += (double)y; ymx = fmaxf(ymx, fabsf(y)); }
  ysum = row_sum16_d(ysum);
  double ymax = (double)row_max16(ymx);
  double mean = ysum / N;
  double D[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) D[j] = 0;
  for (int i = R.rl; i < N; i += 16) {
    float y = R.y[i];
    double l = y > 0.0f ? log((double)y) : 0.0;
#pragma unroll
    for (int j = 0; j < 7; ++j) D[j] += pinvA[j * N + i] * l;
  }
#pragma unroll
  for (int j = 0; j < 7; ++j) D[j] = -row_sum16_d(D[j]);
  double s0 = (D[6] > -23) ? exp(-D[6]) : ymax;
  if (s0 < mean) s0 = ymax;
#pragma unroll
  for (int j = 0; j < 7; ++j) D[j] = 0;
  for (int i = R.rl; i < N; i += 16) {
    double y = (double)R.y[i];
    double l = (y / s0) > 0.01 ? log(y) : log(0.01 * s0);
#pragma unroll
    for (int j = 0; j < 7; ++j) D[j] += pinvA[j * N + i] * l;
  }
#pragma unroll
  for (int j = 0; j < 7; ++j) D[j] = -row_sum16_d(D[j]);
  s0 = exp(-D[6]);
  if (s0 < mean) s0 = mean;
  return tensor_of(s0, D);
}

// calc_tensor_parameters (diffmodels.cc:239-265) of the tensor D
FDT_DEV Dti tensor_of(double s0, const double (&D)[7]) {
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
// DTI::linfit of a voxel with its own acquisition (--gradnonlin: the
// reference forms and pseudo-inverts the design matrix per voxel,
// CUDA/diffmodels.cu:57-63): the same two log-linear passes as dti_row, each
// solved through the column-equilibrated normal equations (G = As^T As, row a
// of G in lane a, LU in FP64); equal to the pseudo-inverse for a full-rank design.
FDT_DEV void design_row(float4 q, double (&r)[7]) {
  double x = q.x, y = q.y, z = q.z, b = q.w;
  r[0] = x * x * b; r[1] = 2 * x * y * b; r[2] = 2 * x * z * b;
  r[3] = y * y * b; r[4] = 2 * y * z * b; r[5] = z * z * b; r[6] = 1.0;
}

// D = -pinv(A) l for l_i = lval(y_i)
template <class F>
FDT_DEV void lls_row(const Row& R, const double (&Grow)[7], const double (&sc)[7], F&& lval, double (&D)[7]) {
  double h[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) h[j] = 0.0;
  for (int i = R.rl; i < R.N; i += 16) {
    double r[7];
    design_row(R.acq[i], r);
    const double l = lval(R.y[i]);
#pragma unroll
    for (int j = 0; j < 7; ++j) h[j] += r[j] / sc[j] * l;
  }
  double hr = 0.0;
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    h[j] = row_sum16_d(h[j]);
    if (R.rl == j) hr = h[j];
  }
  double x[7];
  lu_row<7>(Grow, hr, x, R.rl);
#pragma unroll
  for (int j = 0; j < 7; ++j) D[j] = -x[j] / sc[j];
}

FDT_DEV Dti dti_row_gn(const Row& R) {
  const int N = R.N;
  double ysum = 0;
  float ymx = 0.0f;
  for (int i = R.rl; i < N; i += 16) { float y = R.y[i]; ysum += (double)y; ymx = fmaxf(ymx, fabsf(y)); }
  ysum = row_sum16_d(ysum);
  double ymax = (double)row_max16(ymx);
  double mean = ysum / N;
  // column scales and the equilibrated normal matrix
  double sc[7], G[28];
#pragma unroll
  for (int j = 0; j < 7; ++j) sc[j] = 0.0;
  for (int i = R.rl; i < N; i += 16) {
    double r[7];
    design_row(R.acq[i], r);
#pragma unroll
    for (int j = 0; j < 7; ++j) sc[j] += r[j] * r[j];
  }
#pragma unroll
  for (int j = 0; j < 7; ++j) { double v = row_sum16_d(sc[j]); sc[j] = v > 0 ? sqrt(v) : 1.0; }
#pragma unroll
  for (int e = 0; e < 28; ++e) G[e] = 0.0;
  for (int i = R.rl; i < N; i += 16) {
    double r[7];
    design_row(R.acq[i], r);
#pragma unroll
    for (int j = 0; j < 7; ++j) r[j] /= sc[j];
    int e = 0;
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
      for (int c = a; c < 7; ++c) G[e++] += r[a] * r[c];
  }
  double Grow[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) Grow[c] = 0.0;
  {
    int e = 0;
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
      for (int c = a; c < 7; ++c) {
        double v = row_sum16_d(G[e++]);
        if (R.rl == a) Grow[c] = v;
        if (R.rl == c) Grow[a] = v;
      }
  }
  double D[7];
  lls_row(R, Grow, sc, [](float y) { return y > 0.0f ? log((double)y) : 0.0; }, D);
  double s0 = (D[6] > -23) ? exp(-D[6]) : ymax;
  if (s0 < mean) s0 = ymax;
  lls_row(R, Grow, sc, [s0](float yf) { double y = (double)yf; return (y / s0) > 0.01 ? log(y) : log(0.01 * s0); }, D);
  s0 = exp(-D[6]);
  if (s0 < mean) s0 = mean;
  return tensor_of(s0, D);
}

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
// of the K-column design, formed from row-reduced M^T M and M^T (y - s0 iso).
template <class T, int K, bool F0>
FDT_DEV void start_single_c(const Dti& dti, const Row& R, T* p) {
  const int N = R.N;
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
  for (int i = R.rl; i < N; i += 16) {
    float4 q = R.acq[i];
    T b = (T)q.w;
    T iso = exp(-b * d);
    double yi = (double)((T)R.y[i] - s0 * iso);
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
#pragma unroll
  for (int a = 0; a < K * K; ++a) MtM[a] = row_sum16_d(MtM[a]);
#pragma unroll
  for (int a = 0; a < K; ++a) MtY[a] = row_sum16_d(MtY[a]);
  // pinv(M) y = V diag(1/s^2) V^T M^T y, singular values s <= N*s_max*1e-15 dropped
  double fsd[K];
  {
    double A[9] = {0}, w[3], Vv[9];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int c = 0; c < K; ++c) A[a * 3 + c] = MtM[a * K + c];
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

// ------------------------------------------------------------ kernels
// Signal row stride in LDS: N rounded up to whole 16-lane groups, an odd
// number of them, so that the four rows of a wave fall on different banks.
FDT_DEV int ystride(int N) { return 16 * (((N + 15) / 16) | 1); }

// lane j < NP of the row stores v[j] (no dynamic register indexing)
template <int NP, class T>
FDT_DEV void store_vec(float* dst, const T (&v)[NP], int rl) {
#pragma unroll
  for (int j = 0; j < NP; ++j)
    if (rl == j) dst[j] = (float)v[j];
}

// GN: --gradnonlin, every voxel reads its own corrected acquisition from
// a.acqv instead of the workgroup's table in LDS.
template <class T, class TC, int KIND, int K, bool F0, bool GN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(std::is_same<T, TC>::value && std::is_same<T, double>::value ? 1 : 2, 2))) void lm_kernel(LmArgs a) {
  extern __shared__ float4 lds4[];
  float4* sacq = lds4;
  // without shells the table regions are empty (the launcher sizes them)
  const bool shells = !GN && a.nshell > 0;
  double* sisot = reinterpret_cast<double*>(lds4 + a.N);
  float* sy = reinterpret_cast<float*>(sisot + (shells ? LM_ROWS * 16 : 0));
  int* ssh = reinterpret_cast<int*>(sy + LM_ROWS * ystride(a.N));
  if constexpr (!GN)
    for (int i = threadIdx.x; i < a.N; i += blockDim.x) {
      sacq[i] = a.acq[i];
      if (shells) ssh[i] = a.shell[i];
    }
  __syncthreads();
  using M = Mdl<T, KIND, K, F0>;
  constexpr int NP = M::NP;
  const int rl = __lane_id() & 15;
  float* ym = sy + (threadIdx.x >> 4) * ystride(a.N);
  Row R;
  R.rl = rl; R.N = a.N; R.acq = sacq; R.y = ym;
  if (shells) {
    R.nshell = a.nshell; R.sh = ssh; R.shb = a.shell_b;
    R.isot = sisot + (threadIdx.x >> 4) * 16;
  }
  M m0;
  m0.G = a.G;
  m0.Rv = (T)a.R;
  m0.invR = T(1) / (T(2) * (T)a.R + T(1));
  Mdl<TC, KIND, K, F0> mc0;
  mc0.G = a.G;
  mc0.Rv = (TC)a.R;
  mc0.invR = TC(1) / (TC(2) * (TC)a.R + TC(1));
  const int count = a.count ? *a.count : a.nvox;
  LmState<T, TC, KIND, K, F0> S;
  int vox = 0;
  bool fresh = true;
  for (;;) {
    if (fresh) {
      int q = 0;
      if (rl == 0) q = atomicAdd(a.queue, 1);
      q = row_bcast_i<0>(q);
      if (q >= count) break;
      vox = a.list ? a.list[q] : q;
      if constexpr (GN) R.acq = a.acqv + (size_t)vox * a.acqv_stride;
      // the voxel's signal row, non-positive samples replaced by the smallest
      // positive one for --rician (remove_NonPositive_entries, CUDA/xfibres_gpu.cu:350-367)
      const float* src = a.data + (size_t)vox * a.N;
      float mn = 3.0e38f, mx = -3.0e38f, mnp = 3.0e38f;
      for (int i = rl; i < a.N; i += 16) {
        float v = src[i];
        ym[i] = v;
        mn = fminf(mn, v); mx = fmaxf(mx, v);
        if (v > 0.0f) mnp = fminf(mnp, v);
      }
      if (a.rician) {
        mn = row_min16(mn); mx = row_max16(mx); mnp = row_min16(mnp);
        if (mn <= 0.0f && mx > 0.0f)
          for (int i = rl; i < a.N; i += 16) if (ym[i] <= 0.0f) ym[i] = mnp;
      }
      if (a.init == INIT_BUF) {
#pragma unroll
        for (int i = 0; i < NP; ++i) S.p[i] = (TC)a.in[(size_t)vox * a.in_np + i];
      } else if constexpr (KIND == KIND_MULTI) {
        TC s1[NP - 1];
#pragma unroll
        for (int i = 0; i < NP - 1; ++i) s1[i] = (TC)a.in[(size_t)vox * a.in_np + i];
        start_multi<TC, K, F0>(s1, S.p);
      } else {
        Dti dti;
        if constexpr (GN) dti = dti_row_gn(R);
        else dti = dti_row(R, a.pinvA);
        if constexpr (KIND == KIND_SINGLE) start_single<TC, K, F0>(dti, S.p);
        else start_single_c<TC, K, F0>(dti, R, S.p);
      }
      if (!a.lm) {
        store_vec<NP>(a.out + (size_t)vox * a.out_np, S.p, rl);
        continue;
      }
      S.begin(R, mc0);
      fresh = false;
    }
    if (S.step(R, m0, mc0)) {
      finalise<TC, KIND, K, F0>(S.p);
      store_vec<NP>(a.out + (size_t)vox * a.out_np, S.p, rl);
      if (a.set_f0_slot && rl == 15) a.out[(size_t)vox * a.out_np + a.out_np - 1] = 0.001f;
      if (a.iters && rl == 0) a.iters[vox] = (a.iters_add ? a.iters[vox] : 0) + S.total;
      fresh = true;
    }
  }
}

// 1/var of the residuals of the final fit (N-1 normalisation), one row per voxel
template <class T, int K>
__global__ __launch_bounds__(256) void tau_kernel(TauArgs a) {
  __shared__ float4 sacq[MAX_N];
  for (int i = threadIdx.x; i < a.N; i += blockDim.x) sacq[i] = a.acq[i];
  __syncthreads();
  const int rl = __lane_id() & 15;
  const int vox = blockIdx.x * 16 + (threadIdx.x >> 4);
  if (vox >= a.nvox) return;
  const float* y = a.data + (size_t)vox * a.N;
  const float4* acq = a.acqv ? a.acqv + (size_t)vox * a.acqv_stride : sacq;
  float mn = 3.0e38f, mx = -3.0e38f, mnp = 3.0e38f;
  for (int i = rl; i < a.N; i += 16) {
    float v = y[i];
    mn = fminf(mn, v); mx = fmaxf(mx, v);
    if (v > 0.0f) mnp = fminf(mnp, v);
  }
  mn = row_min16(mn); mx = row_max16(mx); mnp = row_min16(mnp);
  const bool san = (mn <= 0.0f && mx > 0.0f);
  const float* q = a.params + (size_t)vox * a.np;
  const int model = a.model;
  const bool f0 = a.f0 && !(a.refit && a.refit[vox]);
  const int fb = model >= 2 ? 3 : 2;
  const T Rv = (T)a.R;
  T S0 = model >= 2 ? fabs((T)q[0]) : (T)q[0];
  T dd = (T)q[1], a_ = T(0), b_ = T(0);
  if (model >= 2) { a_ = (T)q[1] * (T)q[1] / (T)q[2] / (T)q[2]; b_ = (T)q[2] * (T)q[2] / (T)q[1]; }
  T iR = T(1) / (T(2) * Rv + T(1));
  T f0v = f0 ? (T)q[a.np - 1] : T(0);
  T sumf = T(0), x[K][3], fk[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    fk[k] = (T)q[fb + 3 * k];
    sumf += fk[k];
    T th = (T)q[fb + 3 * k + 1], ph = (T)q[fb + 3 * k + 2];
    x[k][0] = sin(th) * cos(ph); x[k][1] = sin(th) * sin(ph); x[k][2] = cos(th);
  }
  // two passes: mean, then centred sum of squares
  double mean = 0.0, ss = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    double acc = 0.0;
    for (int i = rl; i < a.N; i += 16) {
      float4 g = acq[i];
      T b = (T)g.w;
      T iso = model >= 2 ? exp(-a_ * log(T(1) + b * b_)) : exp(-b * dd);
      T an = T(0);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        T dp = (T)g.x * x[k][0] + (T)g.y * x[k][1] + (T)g.z * x[k][2];
        T e;
        if (model == 1) e = exp(-b * dd * dp * dp);
        else if (model == 3) e = exp(-b * T(3) * a_ * b_ * iR * ((T(1) - Rv) * dp * dp + Rv));
        else e = exp(-a_ * log(T(1) + b * b_ * (dp * dp)));
        an += fk[k] * e;
      }
      T pred = S0 * (f0v + (T(1) - sumf - f0v) * iso + an);
      float yv = y[i];
      if (san && yv <= 0.0f) yv = mnp;
      double r = (double)yv - (double)pred;
      acc += pass == 0 ? r : (r - mean) * (r - mean);
    }
    acc = row_sum16_d(acc);
    if (pass == 0) mean = acc / a.N; else ss = acc;
  }
  if (rl == 0) a.tau[vox] = (float)(1.0 / (ss / (a.N - 1)));
}

// The start of a chain as the CPU xfibres sets it for Gaussian noise without
// --nonlinear / --cnonlinear (xfibresVoxelManager::initialise, xfibres.cc:517-541):
// the means of the best finished neighbour (Samples::neighbour_initialise,
// :306-357) when the grid is given and one qualifies, else the tensor
// (initialise_tensor, :546-572).  Values are rounded to float where the
// reference holds them in float.  One 16-lane row per listed voxel; every
// lane of the row scans the neighbourhood, lane 0 stores.
template <bool GN>
__global__ __launch_bounds__(256) void cpu_start_kernel(CpuStartArgs a) {
  const int rl = __lane_id() & 15;
  const int q = blockIdx.x * 16 + (threadIdx.x >> 4);
  if (q >= a.count) return;
  const int vox = a.list ? a.list[q] : q;
  const size_t nv = (size_t)a.nvox;
  const int K = a.K, add = a.model >= 2 ? 1 : 0;
  // neighbour_initialise: scan x, then y, then z (outer to inner) over the
  // 3x3x3 block; a finished voxel with at least maxfib supported fibres and a
  // larger sum of mean f becomes the best and raises maxfib to its count.
  // Outside the grid nothing is finished (newimage's zero extrapolation).
  int best = -1;
  if (a.vol2mat) {
    const long long g = a.gidx[vox];
    const int xx = (int)(g % a.nx), yy = (int)((g / a.nx) % a.ny), zz = (int)(g / ((long long)a.nx * a.ny));
    int maxfib = 1;
    float maxsf = 0.0f;
    for (int x = xx - 1; x <= xx + 1; ++x)
      for (int y = yy - 1; y <= yy + 1; ++y)
        for (int z = zz - 1; z <= zz + 1; ++z) {
          if (x < 0 || y < 0 || z < 0 || x >= a.nx || y >= a.ny || z >= a.nz) continue;
          const int vn = a.vol2mat[((size_t)z * a.ny + y) * a.nx + x];
          const int been = vn > 0 ? a.been[vn - 1] : 0;
          if (been < maxfib) continue;
          float sumf = 0.0f;
          for (int k = 0; k < K; ++k) sumf += a.mean_f[(size_t)k * nv + (vn - 1)];
          if (sumf > maxsf) { maxsf = sumf; maxfib = been; best = vn - 1; }
        }
    if (maxfib <= 1) best = -1;
  }
  float S0, d, dstd = 0.0f, f[3], th[3], ph[3];
  if (best >= 0) {
    d = a.mean_d[best];
    S0 = a.mean_S0[best];
    for (int k = 0; k < K; ++k) {
      double v[3] = {a.dyad[((size_t)k * 3 + 0) * nv + best], a.dyad[((size_t)k * 3 + 1) * nv + best],
                     a.dyad[((size_t)k * 3 + 2) * nv + best]};
      double t, p;
      cart2sph(v, t, p);
      th[k] = (float)t; ph[k] = (float)p;
      f[k] = a.mean_f[(size_t)k * nv + best];
    }
  } else {
    Row R;
    R.rl = rl; R.N = a.N; R.y = a.data + (size_t)vox * a.N;
    R.acq = GN ? a.acqv + (size_t)vox * a.acqv_stride : nullptr;
    Dti dti;
    if constexpr (GN) dti = dti_row_gn(R);
    else dti = dti_row(R, a.pinvA);
    float D = (float)dti.md;
    if (a.model >= 2) D = D * 2.0f;  // the mono-exponential tensor underestimates d (:554-556)
    if (D <= 0.0f) D = 2e-3f;
    d = D;
    if (a.model >= 2) dstd = D;
    S0 = (float)dti.s0;
    double t, p;
    cart2sph(dti.v[0], t, p);
    th[0] = (float)t; ph[0] = (float)p;
    f[0] = (float)dti.fa;
    for (int k = 1; k < K; ++k) { th[k] = (float)(M_PI / 2); ph[k] = 0.0f; f[k] = 0.01f; }  // Fibre() defaults, fibre.h:103-110
  }
  if (rl != 0) return;
  float* P = a.params + (size_t)vox * a.np;
  P[0] = S0;
  P[1] = d;
  if (add) P[2] = dstd;
  for (int k = 0; k < K; ++k) {
    P[2 + add + 3 * k] = f[k];
    P[3 + add + 3 * k] = th[k];
    P[4 + add + 3 * k] = ph[k];
  }
  if (a.f0) P[a.np - 1] = 0.0f;  // Multifibre starts f0 at 0 (fibre.h:666)
}

}  // namespace fdt

// One translation unit per (K, precision), compiled in parallel:
// -DFDT_FIT_K=1..3 -DFDT_FIT_FP64=0 (FP32), 1 (FP64) or 2 (FP32 normal
// equations with FP64 parameters and cost).
#ifndef FDT_FIT_K
#define FDT_FIT_K 3
#endif
#ifndef FDT_FIT_FP64
#define FDT_FIT_FP64 0
#endif
#define FDT_CAT3_(a, b, c) a##b##_##c
#define FDT_CAT3(a, b, c) FDT_CAT3_(a, b, c)
using FitReal = std::conditional<FDT_FIT_FP64 == 1, double, float>::type;
using FitCost = std::conditional<FDT_FIT_FP64 != 0, double, float>::type;

namespace {
// persistent grid: as many workgroups as stay resident, at most one row per voxel
template <int KIND, bool F0, bool GN>
hipError_t launch_lm(const fdt::LmArgs* a, hipStream_t s) {
  auto kern = fdt::lm_kernel<FitReal, FitCost, KIND, FDT_FIT_K, F0, GN>;
  const bool shells = !GN && a->nshell > 0;
  size_t lds = (size_t)a->N * sizeof(float4) + (size_t)fdt::LM_ROWS * (16 * (((a->N + 15) / 16) | 1)) * sizeof(float) +
               (shells ? (size_t)fdt::LM_ROWS * 16 * sizeof(double) + (size_t)a->N * sizeof(int) : 0);
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, lds)) != hipSuccess) return e;
  long want = ((long)a->nvox + fdt::LM_ROWS - 1) / fdt::LM_ROWS;
  long cap = (long)(per_cu > 0 ? per_cu : 1) * cus;
  int grid = (int)(want < cap ? want : cap);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, *a);
  return hipGetLastError();
}

template <bool GN>
hipError_t launch_kind(const fdt::LmArgs* a, int kind, int f0, hipStream_t s) {
  using namespace fdt;
  if (kind == KIND_SINGLE) return f0 ? launch_lm<KIND_SINGLE, true, GN>(a, s) : launch_lm<KIND_SINGLE, false, GN>(a, s);
  if (kind == KIND_SINGLE_C) return f0 ? launch_lm<KIND_SINGLE_C, true, GN>(a, s) : launch_lm<KIND_SINGLE_C, false, GN>(a, s);
  if (kind == KIND_MULTI) return f0 ? launch_lm<KIND_MULTI, true, GN>(a, s) : launch_lm<KIND_MULTI, false, GN>(a, s);
  return hipErrorInvalidValue;
}
}  // namespace

extern "C" hipError_t FDT_CAT3(fdt_launch_lm_k, FDT_FIT_K, FDT_FIT_FP64)(const fdt::LmArgs* a, int kind, int f0,
                                                                       hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  if (a->N > fdt::MAX_N) return hipErrorInvalidValue;
  if (a->acqv) return launch_kind<true>(a, kind, f0, s);
  return launch_kind<false>(a, kind, f0, s);
}

extern "C" hipError_t FDT_CAT3(fdt_launch_tau_k, FDT_FIT_K, FDT_FIT_FP64)(const fdt::TauArgs* a, hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  if (a->N > fdt::MAX_N) return hipErrorInvalidValue;
  dim3 grid((a->nvox + 15) / 16), block(256);
  hipLaunchKernelGGL((fdt::tau_kernel<FitCost, FDT_FIT_K>), grid, block, 0, s, *a);
  return hipGetLastError();
}

#if FDT_FIT_K == 1 && FDT_FIT_FP64 == 1
// the CPU xfibres starts (one translation unit: K is a run-time value there)
extern "C" hipError_t fdt_launch_cpu_start(const fdt::CpuStartArgs* a, hipStream_t s) {
  if (a->count <= 0) return hipSuccess;
  if (a->N > fdt::MAX_N || a->K < 1 || a->K > 3) return hipErrorInvalidValue;
  dim3 grid((a->count + 15) / 16), block(256);
  if (a->acqv) hipLaunchKernelGGL(fdt::cpu_start_kernel<true>, grid, block, 0, s, *a);
  else hipLaunchKernelGGL(fdt::cpu_start_kernel<false>, grid, block, 0, s, *a);
  return hipGetLastError();
}
#endif
