// xfibres_oracle.cpp -- CPU restatement of the reference xfibres_gpu hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU product is
// checked against; it is compiled into oracle/liborc.so and may only be loaded
// by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The
// product (fdt_amd/) never includes, links or calls it.
//
// Every routine is a from-scratch restatement of the algorithm the reference
// implements; each cites the reference file:line it follows.  It is templated
// on the arithmetic type: Real = double is the FP64 oracle, Real = float is the
// FP32 "faithful" mode used as the CPU throughput baseline (the reference
// fibre.h / CUDA code compute in float).
//
// Parity pins for each function live in tests/test_oracle_*.py (closed forms,
// central finite differences, numpy/scipy cross-checks, an independent Python
// Philox with the Random123 known-answer vectors, brute force on tiny inputs).
//
// RNG: the reference draws from FSL normrnd/unifrnd (CPU) or cuRAND XORWOW
// (GPU); neither is reproduced.  Both this oracle and the GPU product draw from
// Philox4x32-10 (Salmon et al., SC'11) with the counter layout documented at
// draw_pair(), so the two are comparable draw for draw.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <limits>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- constants
// diffmodels.h:30-31 (two_pi is 2/pi as the reference defines it), fibre.h:19,24
static const double TWO_PI_C = 0.636619772;
static const double FSMALL_C = 0.001;
static const double UPPERDIFF_C = 0.005;
static const double MAXFLOAT_C = 1e10;
static const int NPMAX = 20;

template <class R> static inline R f2x(R f) { return std::tan(f / R(TWO_PI_C)); }      // diffmodels.h:33
template <class R> static inline R x2f(R x) { return std::fabs(R(TWO_PI_C) * std::atan(x)); } // diffmodels.h:34
template <class R> static inline R f2beta(R f) { return std::asin(std::sqrt(f)); }     // diffmodels.h:36
template <class R> static inline R beta2f(R b) { R s = std::sin(b); return s * s; }    // diffmodels.h:37
template <class R> static inline int sgn(R x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); } // CUDA/diffmodels_utils.h:45-47

// ------------------------------------------------------------------ Philox
// Philox4x32-10 round function as published (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11, section 4.1).
static inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Draw layout shared with the GPU product: one Philox block per
// (voxel, jump, proposal slot).  key = (lo32 seed, hi32 seed);
// counter = (lo32 voxel, hi32 voxel, jump, slot).  Normal from Box-Muller
// (cosine branch) on the first two words, acceptance uniform from the third.
struct Draw { double n; double logu; };
static inline Draw draw_pair(uint64_t seed, uint64_t vox, uint32_t jump, uint32_t slot) {
  uint32_t c[4] = {(uint32_t)vox, (uint32_t)(vox >> 32), jump, slot};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t x[4];
  philox4x32_10(c, k, x);
  const double s24 = 1.0 / 16777216.0;
  double u1 = ((double)(x[0] >> 8) + 1.0) * s24;   // (0,1]
  double u2 = ((double)(x[1] >> 8)) * s24;         // [0,1)
  double u3 = ((double)(x[2] >> 8) + 1.0) * s24;   // (0,1]
  Draw d;
  d.n = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  d.logu = std::log(u3);
  return d;
}

// ---------------------------------------------------------------- linear alg
// Symmetric eigen-decomposition by cyclic Jacobi rotations (n <= 8).
// Eigenvalues ascending in w, eigenvectors as columns of V (row-major n x n).
static void jacobi_eig(int n, const double* Ain, double* w, double* V) {
  double A[64];
  std::memcpy(A, Ain, sizeof(double) * n * n);
  for (int i = 0; i < n * n; ++i) V[i] = 0;
  for (int i = 0; i < n; ++i) V[i * n + i] = 1;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += A[i * n + j] * A[i * n + j];
        if (i != j) off += A[i * n + j] * A[i * n + j];
      }
    if (off <= 1e-30 * tot || off == 0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (apq == 0) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double theta = (aqq - app) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {  // A = A * G
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A = G^T * A
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  int idx[8];
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx, idx + n, [&](int a, int b) { return A[a * n + a] < A[b * n + b]; });
  double Vs[64];
  for (int j = 0; j < n; ++j) {
    w[j] = A[idx[j] * n + idx[j]];
    for (int i = 0; i < n; ++i) Vs[i * n + j] = V[i * n + idx[j]];
  }
  std::memcpy(V, Vs, sizeof(double) * n * n);
}

// Moore-Penrose pseudo-inverse of an m x n matrix (n <= 8) by one-sided
// Jacobi SVD (Hestenes): columns are orthogonalised by plane rotations, the
// column norms are the singular values.  Singular values at or below
// max(m,n) * s_max * 1e-15 are treated as zero (the FSL pinv tolerance is
// un-vendored; this matters only for rank-deficient input).
static void pinv(int m, int n, const double* A, double* P /* n x m */) {
  std::vector<double> U((size_t)m * n);
  std::memcpy(U.data(), A, sizeof(double) * m * n);
  double V[64];
  for (int i = 0; i < n * n; ++i) V[i] = 0;
  for (int i = 0; i < n; ++i) V[i * n + i] = 1;
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double al = 0, be = 0, ga = 0;
        for (int k = 0; k < m; ++k) {
          double up = U[(size_t)k * n + p], uq = U[(size_t)k * n + q];
          al += up * up; be += uq * uq; ga += up * uq;
        }
        if (ga == 0 || std::fabs(ga) <= 1e-15 * std::sqrt(al * be)) continue;
        rotated = true;
        double zeta = (be - al) / (2 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
        double c = 1 / std::sqrt(1 + t * t), sn = c * t;
        for (int k = 0; k < m; ++k) {
          double up = U[(size_t)k * n + p], uq = U[(size_t)k * n + q];
          U[(size_t)k * n + p] = c * up - sn * uq;
          U[(size_t)k * n + q] = sn * up + c * uq;
        }
        for (int k = 0; k < n; ++k) {
          double vp = V[k * n + p], vq = V[k * n + q];
          V[k * n + p] = c * vp - sn * vq;
          V[k * n + q] = sn * vp + c * vq;
        }
      }
    if (!rotated) break;
  }
  double sv[8], smax = 0;
  for (int j = 0; j < n; ++j) {
    double s = 0;
    for (int k = 0; k < m; ++k) s += U[(size_t)k * n + j] * U[(size_t)k * n + j];
    sv[j] = std::sqrt(s);
    smax = std::max(smax, sv[j]);
  }
  double tol = std::max(m, n) * smax * 1e-15;
  // P = V diag(1/s) U_normalised^T = V diag(1/s^2) U^T (U columns carry s)
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < m; ++k) {
      double acc = 0;
      for (int j = 0; j < n; ++j)
        if (sv[j] > tol) acc += V[i * n + j] * U[(size_t)k * n + j] / (sv[j] * sv[j]);
      P[i * m + k] = acc;
    }
}

// In-place LU with partial pivoting followed by forward/back substitution:
// the algorithm of CUDA/solver_mult_inverse.cu:14-166 ("as NEWMAT
// GeneralSolvI").  Pivot for the next column is the first maximum of the
// updated column below the diagonal; a zero pivot skips elimination.
template <class R>
static void lu_solve(int n, const R* A, const R* b, R* x) {
  R C[NPMAX * NPMAX], el[NPMAX];
  int indx[NPMAX];
  for (int i = 0; i < n * n; ++i) C[i] = A[i];
  int mu = 0;
  R big = std::fabs(C[0]);
  for (int k = 1; k < n; ++k)
    if (big < std::fabs(C[k * n])) { big = std::fabs(C[k * n]); mu = k; }
  for (int k = 0; k < n; ++k) {
    indx[k] = mu;
    if (mu != k)
      for (int j = 0; j < n; ++j) std::swap(C[k * n + j], C[mu * n + j]);
    R diag = C[k * n + k];
    big = 0;
    mu = k + 1;
    if (diag != 0) {
      for (int i = k + 1; i < n; ++i) {
        R mult = C[i * n + k] / diag;
        C[i * n + k] = mult;
        for (int l = k + 1; l < n; ++l) {
          C[i * n + l] = C[i * n + l] - mult * C[k * n + l];
          if (l == k + 1 && big < std::fabs(C[i * n + l])) { big = std::fabs(C[i * n + l]); mu = i; }
        }
      }
    }
  }
  for (int i = 0; i < n; ++i) el[i] = b[i];
  int ii = n;
  for (int i = 0; i < n; ++i) {
    int ip = indx[i];
    R t = el[ip]; el[ip] = el[i]; el[i] = t;
    if (t != 0) { ii = i; break; }
  }
  for (int i = ii + 1; i < n; ++i) {
    int ip = indx[i];
    R sum = el[ip]; el[ip] = el[i];
    for (int j = ii; j < i; ++j) sum -= C[i * n + j] * el[j];
    el[i] = sum;
  }
  for (int i = n - 1; i >= 0; --i) {
    R sum = el[i];
    for (int j = i + 1; j < n; ++j) sum -= C[i * n + j] * el[j];
    el[i] = sum / C[i * n + i];
  }
  for (int i = 0; i < n; ++i) x[i] = el[i];
}

// --------------------------------------------------------------- acquisition
template <class R>
struct Acq {
  int N;
  std::vector<R> b, gx, gy, gz;
  Acq(int n, const double* bvals, const double* bvecs /* 3 x N row-major */) : N(n), b(n), gx(n), gy(n), gz(n) {
    for (int i = 0; i < n; ++i) {
      b[i] = (R)bvals[i];
      gx[i] = (R)bvecs[i];
      gy[i] = (R)bvecs[n + i];
      gz[i] = (R)bvecs[2 * n + i];
    }
  }
};

// correct_bvals_bvecs (CUDA/xfibres_gpu.cu:151-170; xfibres.cc:797-817): the
// voxel's gradient-deviation tensor L (9 values g, L(i,j) = g[3j+i]) maps each
// diffusion-weighted direction v to (I+L) v; the direction is renormalised and
// b scaled by |(I+L) v|^2; b <= 0 volumes are left alone.  The products are in
// double, |.| is rounded to float (the reference's `float mag`), and the
// results are rounded to float as the reference hands them to its kernels in
// float host vectors (CUDA/xfibres_gpu.cu:75, 423-427).
static void correct_acq(int N, const double* bvals, const double* bvecs, const double* g,
                        double* bvals_c, double* bvecs_c) {
  double M[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = (i == j ? 1.0 : 0.0) + g[3 * j + i];
  for (int l = 0; l < N; ++l) {
    double v[3] = {bvecs[l], bvecs[N + l], bvecs[2 * N + l]};
    double b = bvals[l];
    if (b > 0) {
      // left-to-right sums, no fused multiply-add (the build has no -mfma)
      double w[3];
      for (int i = 0; i < 3; ++i) w[i] = M[i][0] * v[0] + M[i][1] * v[1] + M[i][2] * v[2];
      float mag = (float)std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      if (mag != 0.0f)
        for (int i = 0; i < 3; ++i) w[i] = w[i] / (double)mag;
      float m2 = mag * mag;
      b = (double)m2 * b;
      for (int i = 0; i < 3; ++i) v[i] = w[i];
    }
    bvals_c[l] = (double)(float)b;
    for (int i = 0; i < 3; ++i) bvecs_c[i * N + l] = (double)(float)v[i];
  }
}

// MISCMATHS::cart2sph (un-vendored): polar angle from +z, azimuth atan2(y,x);
// pinned by the forward map (sin th cos ph, sin th sin ph, cos th) round trip.
static inline void cart2sph(const double v[3], double& th, double& ph) {
  double mag = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (mag == 0) { th = M_PI / 2; ph = M_PI / 2; return; }
  ph = std::atan2(v[1], v[0]);
  th = std::atan2(std::sqrt(v[0] * v[0] + v[1] * v[1]), v[2]);
}

// ----------------------------------------------------------------- DTI fit
// DTI::form_Amat (diffmodels.h:148-163) and DTI::linfit (diffmodels.cc:17-46)
// with calc_tensor_parameters (diffmodels.cc:239-265).  Always double.
struct DtiOut { double s0, md, fa, l[3]; double v[3][3]; };  // v[0] = principal

static void dti_amat_pinv(int N, const double* bvals, const double* bvecs, std::vector<double>& A, std::vector<double>& iA) {
  A.assign((size_t)N * 7, 0.0);
  iA.assign((size_t)N * 7, 0.0);
  for (int i = 0; i < N; ++i) {
    double x = bvecs[i], y = bvecs[N + i], z = bvecs[2 * N + i], b = bvals[i];
    double* r = &A[(size_t)i * 7];
    r[0] = x * x * b; r[1] = 2 * x * y * b; r[2] = 2 * x * z * b;
    r[3] = y * y * b; r[4] = 2 * y * z * b; r[5] = z * z * b; r[6] = 1;
  }
  pinv(N, 7, A.data(), iA.data());
}

// The acquisition one voxel is fitted with: the shared one, or under
// --gradnonlin its own corrected copy (per-voxel bvecs/bvals and tensor
// design, CUDA/xfibres_gpu.cu:403-436, CUDA/diffmodels.cu:57-63;
// xfibres.cc:909-919).
template <class R>
struct VoxAcq {
  const Acq<R>* acq;
  const std::vector<double>* iA;
  std::vector<double> bv, bc, A, iAv;
  std::unique_ptr<Acq<R>> own;
  VoxAcq(int N, const double* bvals, const double* bvecs, const double* g, const Acq<R>& acq0,
         const std::vector<double>& iA0, bool need_pinv = true)
      : acq(&acq0), iA(&iA0) {
    if (!g) return;
    bv.resize(N);
    bc.resize(3 * (size_t)N);
    correct_acq(N, bvals, bvecs, g, bv.data(), bc.data());
    own.reset(new Acq<R>(N, bv.data(), bc.data()));
    acq = own.get();
    if (need_pinv) {
      dti_amat_pinv(N, bv.data(), bc.data(), A, iAv);
      iA = &iAv;
    }
  }
};

// eigenvector sign convention shared with the GPU product: the component of
// largest magnitude is made positive (the reference's newmat EigenValues sign
// is unspecified; the LM cost is invariant to x -> -x).
static inline void canon_sign(double* v) {
  int m = 0;
  for (int i = 1; i < 3; ++i) if (std::fabs(v[i]) > std::fabs(v[m])) m = i;
  if (v[m] < 0) for (int i = 0; i < 3; ++i) v[i] = -v[i];
}

static void dti_linfit(int N, const double* iA, const double* Y, DtiOut& o) {
  std::vector<double> logS(N);
  double maxabs = 0, sum = 0;
  for (int i = 0; i < N; ++i) { maxabs = std::max(maxabs, std::fabs(Y[i])); sum += Y[i]; }
  double mean = sum / N;
  for (int i = 0; i < N; ++i) logS[i] = Y[i] > 0 ? std::log(Y[i]) : 0.0;
  double D[7];
  for (int j = 0; j < 7; ++j) { double s = 0; for (int i = 0; i < N; ++i) s += iA[j * N + i] * logS[i]; D[j] = -s; }
  double s0 = (D[6] > -23) ? std::exp(-D[6]) : maxabs;
  if (s0 < mean) s0 = maxabs;
  for (int i = 0; i < N; ++i) logS[i] = (Y[i] / s0) > 0.01 ? std::log(Y[i]) : std::log(0.01 * s0);
  for (int j = 0; j < 7; ++j) { double s = 0; for (int i = 0; i < N; ++i) s += iA[j * N + i] * logS[i]; D[j] = -s; }
  s0 = std::exp(-D[6]);
  if (s0 < mean) s0 = mean;
  double T[9] = {D[0], D[1], D[2], D[1], D[3], D[4], D[2], D[4], D[5]};
  double w[3], V[9];
  jacobi_eig(3, T, w, V);
  o.s0 = s0;
  o.l[0] = w[2]; o.l[1] = w[1]; o.l[2] = w[0];
  for (int k = 0; k < 3; ++k) {
    int col = 2 - k;
    for (int i = 0; i < 3; ++i) o.v[k][i] = V[i * 3 + col];
    canon_sign(o.v[k]);
  }
  o.md = (w[0] + w[1] + w[2]) / 3.0;
  double numer = 1.5 * ((o.l[0] - o.md) * (o.l[0] - o.md) + (o.l[1] - o.md) * (o.l[1] - o.md) + (o.l[2] - o.md) * (o.l[2] - o.md));
  double denom = o.l[0] * o.l[0] + o.l[1] * o.l[1] + o.l[2] * o.l[2];
  double fa = denom > 0 ? numer / denom : 0;
  o.fa = fa > 0 ? std::sqrt(fa) : 0;
}

// ------------------------------------------------------------- LM models
// kind 0 = PVM_single  (CUDA/PVM_single.cu:107-373)
// kind 1 = PVM_single_c (CUDA/PVM_single_c.cu:104-514)
// kind 2 = PVM_multi   (CUDA/PVM_multi.cu:19-466); G = Gamma_for_ball_only
enum { K_SINGLE = 0, K_SINGLE_C = 1, K_MULTI = 2 };

template <class R>
struct LMModel {
  int kind, K, G;
  bool f0;
  R Rv, invR;
  int np;
  const Acq<R>* acq;
  LMModel(int kind_, int K_, bool f0_, int G_, R Rv_, const Acq<R>* a)
      : kind(kind_), K(K_), G(G_), f0(f0_), Rv(Rv_), invR(R(1) / (R(2) * Rv_ + R(1))), acq(a) {
    np = (kind == K_MULTI ? 3 : 2) + 3 * K + (f0 ? 1 : 0);
  }
  int fbase() const { return kind == K_MULTI ? 3 : 2; }

  // per-parameter-vector state
  struct St { R fs[8], x[8][3], sth[8], cth[8], sph[8], cph[8], sumf, d, lam, a, b, fderiv[8][8]; };

  void prep(const R* p, St& s) const {
    int fb = fbase();
    for (int k = 0; k < K; ++k) {
      int kk = fb + 3 * k;
      s.sth[k] = std::sin(p[kk + 1]); s.cth[k] = std::cos(p[kk + 1]);
      s.sph[k] = std::sin(p[kk + 2]); s.cph[k] = std::cos(p[kk + 2]);
      s.x[k][0] = s.sth[k] * s.cph[k]; s.x[k][1] = s.sth[k] * s.sph[k]; s.x[k][2] = s.cth[k];
    }
    s.sumf = 0;
    if (kind == K_SINGLE_C) {
      // fs_k = sin^2(beta_k) * (1 - sum_{j<k} fs_j)   (CUDA/PVM_single_c.cu:170-182)
      s.lam = p[1];
      s.d = p[1] * p[1];
      for (int k = 0; k < K; ++k) {
        R part = 1;
        for (int j = 0; j < k; ++j) part -= s.fs[j];
        s.fs[k] = beta2f(p[2 + 3 * k]) * part;
        s.sumf += s.fs[k];
      }
      // fractions_deriv (CUDA/PVM_single_c.cu:104-139, diffmodels.cc:3032-3053)
      for (int k = 0; k < K; ++k) {
        for (int j = 0; j < K; ++j) s.fderiv[j][k] = 0;
        R s2b = std::sin(R(2) * p[2 + 3 * k]);
        for (int j = 0; j < K; ++j) {
          if (j == k) {
            R fsum = 1;
            for (int n = 0; n <= j - 1; ++n) fsum -= s.fs[n];
            s.fderiv[j][k] = s2b * fsum;
          } else if (j > k) {
            R sp = std::sin(p[2 + 3 * j]);
            R fsum = 0;
            for (int n = 0; n <= j - 1; ++n) fsum += s.fderiv[n][k];
            s.fderiv[j][k] = -(sp * sp) * fsum;
          }
        }
      }
    } else if (kind == K_SINGLE) {
      s.d = std::fabs(p[1]);
      for (int k = 0; k < K; ++k) { s.fs[k] = x2f(p[2 + 3 * k]); s.sumf += s.fs[k]; }
    } else {
      s.a = std::fabs(p[1]);
      s.b = std::fabs(p[2]);
      for (int k = 0; k < K; ++k) { s.fs[k] = x2f(p[3 + 3 * k]); s.sumf += s.fs[k]; }
    }
  }

  // predicted signal of volume i (cf_* in the CUDA files)
  R pred(const R* p, const St& s, int i) const {
    const Acq<R>& q = *acq;
    R b = q.b[i];
    R err = 0, iso;
    for (int k = 0; k < K; ++k) {
      R dp = q.gx[i] * s.x[k][0] + q.gy[i] * s.x[k][1] + q.gz[i] * s.x[k][2];
      err += s.fs[k] * aniso(b, dp, s);
    }
    if (kind == K_MULTI) {
      iso = std::exp(-s.a * std::log(1 + b * s.b));
      R S0 = std::fabs(p[0]);
      if (f0) { R tf0 = x2f(p[np - 1]); return S0 * (tf0 + ((1 - s.sumf - tf0) * iso + err)); }
      return S0 * ((1 - s.sumf) * iso + err);
    }
    iso = std::exp(-b * s.d);
    if (f0) {
      R tf0;
      if (kind == K_SINGLE_C) { R part = 1; for (int j = 0; j < K; ++j) part -= s.fs[j]; tf0 = beta2f(p[np - 1]) * part; }
      else tf0 = x2f(p[np - 1]);
      return p[0] * ((tf0 + (1 - s.sumf - tf0) * iso) + err);
    }
    return p[0] * ((1 - s.sumf) * iso + err);
  }

  R aniso(R b, R dp, const St& s) const {
    if (kind != K_MULTI) return std::exp(-b * s.d * dp * dp);
    if (G == 1) return std::exp(-b * s.a * s.b * dp * dp);
    if (G == 2) return std::exp(-b * 3 * s.a * s.b * invR * ((1 - Rv) * dp * dp + Rv));
    return std::exp(-s.a * std::log(1 + b * s.b * (dp * dp)));
  }

  // J row of volume i for the gradient (Jg) and the Hessian (Jh); they
  // differ only in column 0 for PVM_multi (grad uses sign(S0)*sig/S0,
  // CUDA/PVM_multi.cu:317; hess uses sig/S0, CUDA/PVM_multi.cu:451).
  void jrow(const R* p, const St& s, int i, R& sig, R* Jg, R* Jh) const {
    const Acq<R>& q = *acq;
    R b = q.b[i], gx = q.gx[i], gy = q.gy[i], gz = q.gz[i];
    for (int j = 0; j < np; ++j) Jg[j] = 0;
    if (kind == K_SINGLE) {
      // CUDA/PVM_single.cu:217-255 with the *_PVM_single helpers :23-61
      R S0 = p[0], sg1 = p[1] > 0 ? R(1) : R(-1);
      R iso = std::exp(-b * s.d), iso_d = -b * std::exp(-b * s.d);
      sig = 0;
      for (int k = 0; k < K; ++k) {
        int kk = 2 + 3 * k;
        R dp = gx * s.x[k][0] + gy * s.x[k][1] + gz * s.x[k][2];
        R e = std::exp(-b * s.d * dp * dp);
        R an_d = -b * dp * dp * e;
        R dp1t = s.cth[k] * (gx * s.cph[k] + gy * s.sph[k]) - gz * s.sth[k];
        R dp1p = s.sth[k] * (-gx * s.sph[k] + gy * s.cph[k]);
        R an_th = -2 * b * s.d * dp * dp1t * e;
        R an_ph = -2 * b * s.d * dp * dp1p * e;
        sig += s.fs[k] * e;
        Jg[1] += sg1 * S0 * s.fs[k] * an_d;
        Jg[kk] = S0 * (e - iso) * R(TWO_PI_C) * R(sgn(p[kk])) * R(1) / (1 + p[kk] * p[kk]);
        Jg[kk + 1] = S0 * s.fs[k] * an_th;
        Jg[kk + 2] = S0 * s.fs[k] * an_ph;
      }
      if (f0) {
        R tf0 = x2f(p[np - 1]);
        Jg[np - 1] = S0 * (1 - iso) * R(TWO_PI_C) * R(sgn(p[np - 1])) * R(1) / (1 + p[np - 1] * p[np - 1]);
        sig = S0 * ((tf0 + (1 - s.sumf - tf0) * iso) + sig);
        Jg[1] += sg1 * S0 * (1 - s.sumf - tf0) * iso_d;
      } else {
        sig = S0 * ((1 - s.sumf) * iso + sig);
        Jg[1] += sg1 * S0 * (1 - s.sumf) * iso_d;
      }
      Jg[0] = sig / S0;
      for (int j = 0; j < np; ++j) Jh[j] = Jg[j];
    } else if (kind == K_SINGLE_C) {
      // CUDA/PVM_single_c.cu:264-333
      R S0 = p[0], lam = s.lam;
      R iso = std::exp(-b * s.d);
      R A[8];
      for (int k = 0; k < K; ++k) {
        R dp = gx * s.x[k][0] + gy * s.x[k][1] + gz * s.x[k][2];
        A[k] = std::exp(-b * s.d * dp * dp);
      }
      sig = 0;
      for (int k = 0; k < K; ++k) {
        int kk = 2 + 3 * k;
        R dp = gx * s.x[k][0] + gy * s.x[k][1] + gz * s.x[k][2];
        sig += s.fs[k] * A[k];
        Jg[1] += S0 * s.fs[k] * (-2 * b * lam * dp * dp * std::exp(-b * lam * lam * dp * dp));
        R jb = 0;
        for (int j = 0; j < K; ++j)
          if (s.fderiv[j][k] != 0) jb += S0 * (A[j] - iso) * s.fderiv[j][k];
        Jg[kk] = jb;
      }
      for (int k = 0; k < K; ++k) {
        int kk = 2 + 3 * k;
        R dp = gx * s.x[k][0] + gy * s.x[k][1] + gz * s.x[k][2];
        R dp1t = s.cth[k] * (gx * s.cph[k] + gy * s.sph[k]) - gz * s.sth[k];
        R dp1p = s.sth[k] * (-gx * s.sph[k] + gy * s.cph[k]);
        R e = std::exp(-b * s.d * dp * dp);
        Jg[kk + 1] = S0 * s.fs[k] * (-2 * b * s.d * dp * dp1t * e);
        Jg[kk + 2] = S0 * s.fs[k] * (-2 * b * s.d * dp * dp1p * e);
      }
      R isol = -2 * b * lam * std::exp(-b * lam * lam);
      if (f0) {
        R part = 1;
        for (int j = 0; j < K; ++j) part -= s.fs[j];
        R tf0 = beta2f(p[np - 1]) * part;
        Jg[np - 1] = S0 * (1 - iso) * std::sin(R(2) * p[np - 1]) * part;
        sig = S0 * ((tf0 + (1 - s.sumf - tf0) * iso) + sig);
        Jg[1] += S0 * (1 - s.sumf - tf0) * isol;
      } else {
        sig = S0 * ((1 - s.sumf) * iso + sig);
        Jg[1] += S0 * (1 - s.sumf) * isol;
      }
      Jg[0] = sig / S0;
      for (int j = 0; j < np; ++j) Jh[j] = Jg[j];
    } else {
      // CUDA/PVM_multi.cu:279-318 (grad) and :413-452 (hess)
      R S0a = std::fabs(p[0]);
      R sg1 = p[1] > 0 ? R(1) : R(-1), sg2 = p[2] > 0 ? R(1) : R(-1);
      R la = std::log(1 + b * s.b);
      R iso = std::exp(-s.a * la);
      R iso_a = -la * iso;
      R iso_b = -s.a * b / (1 + b * s.b) * iso;
      sig = 0;
      for (int k = 0; k < K; ++k) {
        int kk = 3 + 3 * k;
        R dp = gx * s.x[k][0] + gy * s.x[k][1] + gz * s.x[k][2];
        R dp1t = s.cth[k] * (gx * s.cph[k] + gy * s.sph[k]) - gz * s.sth[k];
        R dp1p = s.sth[k] * (-gx * s.sph[k] + gy * s.cph[k]);
        R an, an_a, an_b, an_th, an_ph;
        if (G == 1) {
          an = std::exp(-b * s.a * s.b * dp * dp);
          an_a = -b * s.b * dp * dp * an;
          an_b = -b * s.a * dp * dp * an;
          an_th = -2 * b * s.a * s.b * dp * dp1t * an;
          an_ph = -2 * b * s.a * s.b * dp * dp1p * an;
        } else if (G == 2) {
          R t = (1 - Rv) * dp * dp + Rv;
          an = std::exp(-b * 3 * s.a * s.b * invR * t);
          R q2a = b * 3 * s.b * invR * t;
          an_a = -q2a * std::exp(-q2a * s.a);
          R q2b = b * 3 * s.a * invR * t;
          an_b = -q2b * std::exp(-q2b * s.b);
          // the reference's G=2 angle derivative has no factor (g.x)
          // (CUDA/PVM_multi.cu:77-78, :95-96; diffmodels.cc:3189, :3202);
          // kept because it decides where the reference LM stops.
          an_th = -(2 * b * 3 * s.a * s.b * invR * (1 - Rv) * dp1t) * an;
          an_ph = -(2 * b * 3 * s.a * s.b * invR * (1 - Rv) * dp1p) * an;
        } else {
          R l1 = std::log(1 + b * (dp * dp) * s.b);
          an = std::exp(-s.a * l1);
          an_a = -l1 * an;
          an_b = -s.a * b * (dp * dp) / (1 + b * (dp * dp) * s.b) * an;
          R c = -s.a * s.b * b / (1 + b * (dp * dp) * s.b) * an;
          an_th = c * 2 * dp * dp1t;
          an_ph = c * 2 * dp * dp1p;
        }
        sig += s.fs[k] * an;
        Jg[1] += sg1 * S0a * s.fs[k] * an_a;
        Jg[2] += sg2 * S0a * s.fs[k] * an_b;
        Jg[kk] = S0a * (an - iso) * R(TWO_PI_C) * R(sgn(p[kk])) * R(1) / (1 + p[kk] * p[kk]);
        Jg[kk + 1] = S0a * s.fs[k] * an_th;
        Jg[kk + 2] = S0a * s.fs[k] * an_ph;
      }
      if (f0) {
        R tf0 = x2f(p[np - 1]);
        Jg[np - 1] = S0a * (1 - iso) * R(TWO_PI_C) * R(sgn(p[np - 1])) * R(1) / (1 + p[np - 1] * p[np - 1]);
        sig = S0a * ((tf0 + (1 - s.sumf - tf0) * iso) + sig);
        Jg[1] += sg1 * S0a * (1 - s.sumf - tf0) * iso_a;
        Jg[2] += sg2 * S0a * (1 - s.sumf - tf0) * iso_b;
      } else {
        sig = S0a * ((1 - s.sumf) * iso + sig);
        Jg[1] += sg1 * S0a * (1 - s.sumf) * iso_a;
        Jg[2] += sg2 * S0a * (1 - s.sumf) * iso_b;
      }
      for (int j = 0; j < np; ++j) Jh[j] = Jg[j];
      R sg0 = p[0] > 0 ? R(1) : R(-1);
      Jg[0] = sg0 * sig / p[0];
      Jh[0] = sig / p[0];
    }
  }

  double cost(const R* p, const R* y) const {
    St s;
    prep(p, s);
    double c = 0;
    for (int i = 0; i < acq->N; ++i) {
      R e = pred(p, s, i) - y[i];
      c += (double)e * (double)e;
    }
    return c;
  }
};

struct LMStat { int iters; int success_iters; double cost; };

// Levenberg loop of CUDA/levenberg_marquardt.cu:24-115 (identical for the
// three models): lambda0 = 0.1, x10 on failure, /10 on success, the
// relative cost test zero_cf_diff_conv (:17-22), ltol 1e20, 200 successes.
template <class R>
static LMStat lm_fit(const LMModel<R>& m, const R* y, R* p) {
  const int np = m.np, N = m.acq->N;
  double lambda = 0.1, cftol = 1.0e-8, ltol = 1.0e20, olambda = 0.0;
  bool success = true, end = false;
  int niter = 0, total = 0;
  const int maxiter = 200;
  double pcf = m.cost(p, y), ncf = 0;
  R g[NPMAX], H[NPMAX * NPMAX], step[NPMAX], sol[NPMAX];
  R Jg[NPMAX], Jh[NPMAX];
  typename LMModel<R>::St s;
  while (!(success && niter++ >= maxiter)) {
    ++total;
    if (success) {
      m.prep(p, s);
      for (int j = 0; j < np; ++j) g[j] = 0;
      for (int j = 0; j < np * np; ++j) H[j] = 0;
      for (int i = 0; i < N; ++i) {
        R sig;
        m.jrow(p, s, i, sig, Jg, Jh);
        R diff = sig - y[i];
        for (int a = 0; a < np; ++a) {
          g[a] += 2 * Jg[a] * diff;
          for (int c = a; c < np; ++c) H[a * np + c] += 2 * (Jh[a] * Jh[c]);
        }
      }
      for (int a = 0; a < np; ++a)
        for (int c = a + 1; c < np; ++c) H[c * np + a] = H[a * np + c];
    }
    for (int i = 0; i < np; ++i) H[i * np + i] += (R)(lambda - olambda);
    lu_solve<R>(np, H, g, sol);
    for (int i = 0; i < np; ++i) step[i] = p[i] - sol[i];
    ncf = m.cost(step, y);
    if ((success = (ncf < pcf))) {
      olambda = 0.0;
      for (int i = 0; i < np; ++i) p[i] = step[i];
      lambda = lambda / 10.0;
      if (2.0 * std::fabs(pcf - ncf) <= cftol * (std::fabs(pcf) + std::fabs(ncf) + 2.0e-16)) end = true;
      pcf = ncf;
    } else {
      olambda = lambda;
      lambda = lambda * 10.0;
      if (lambda > ltol) end = true;
    }
    if (end) break;
  }
  LMStat st;
  st.iters = total;
  st.success_iters = niter;
  st.cost = pcf;
  return st;
}

// ----------------------------------------------------- start values & finals
// PVM_single start from the tensor (CUDA/diffmodels.cu:72-93)
template <class R>
static void start_single(const DtiOut& dti, int K, bool f0, R* p) {
  int np = 2 + 3 * K + (f0 ? 1 : 0);
  double th, ph;
  cart2sph(dti.v[0], th, ph);
  p[0] = (R)dti.s0;
  p[1] = (R)(dti.md > 0 ? dti.md * 2 : 0.001);
  p[2] = dti.fa < 1 ? f2x<R>((R)dti.fa) : f2x<R>((R)0.95);
  p[3] = (R)th; p[4] = (R)ph;
  R sumf = x2f<R>(p[2]);
  R tmpsumf = sumf;
  for (int ii = 2, i = 5; ii <= K; ++ii, i += 3) {
    R denom = 2;
    do {
      p[i] = f2x<R>(x2f<R>(p[i - 3]) / denom);
      denom *= 2;
      tmpsumf = sumf + x2f<R>(p[i]);
    } while (tmpsumf >= 1);
    sumf += x2f<R>(p[i]);
    const double* v = dti.v[std::min(ii, 3) - 1];
    cart2sph(v, th, ph);
    p[i + 1] = (R)th; p[i + 2] = (R)ph;
  }
  if (f0) p[np - 1] = f2x<R>((R)FSMALL_C);
}

// PVM_single_c start: tensor + PVM_single_c::fit_pvf
// (CUDA/diffmodels.cu:149-183, diffmodels.cc:561-603)
template <class R>
static void start_single_c(const DtiOut& dti, const Acq<R>& acq, const R* Y, int K, bool f0, R* p) {
  int np = 2 + 3 * K + (f0 ? 1 : 0), N = acq.N;
  double th, ph;
  p[0] = (R)dti.s0;
  p[1] = std::sqrt((R)(dti.l[0] > 0 ? dti.l[0] : 0.002));
  for (int k = 0; k < K; ++k) {
    cart2sph(dti.v[std::min(k, 2)], th, ph);
    p[2 + 3 * k + 1] = (R)th;
    p[2 + 3 * k + 2] = (R)ph;
  }
  // fit_pvf
  R s0 = p[0], d = p[1] * p[1];
  std::vector<double> M((size_t)N * K), YI(N), P((size_t)K * N);
  R dir[8][3];
  for (int k = 0; k < K; ++k) {
    int kk = 2 + 3 * k;
    dir[k][0] = std::sin(p[kk + 1]) * std::cos(p[kk + 2]);
    dir[k][1] = std::sin(p[kk + 1]) * std::sin(p[kk + 2]);
    dir[k][2] = std::cos(p[kk + 1]);
  }
  for (int i = 0; i < N; ++i) {
    R iso = std::exp(-acq.b[i] * d);
    YI[i] = Y[i] - s0 * iso;
    for (int k = 0; k < K; ++k) {
      R dp = acq.gx[i] * dir[k][0] + acq.gy[i] * dir[k][1] + acq.gz[i] * dir[k][2];
      M[(size_t)i * K + k] = s0 * (std::exp(-acq.b[i] * d * dp * dp) - iso);
    }
  }
  pinv(N, K, M.data(), P.data());
  R fs[8];
  for (int k = 0; k < K; ++k) { double s = 0; for (int i = 0; i < N; ++i) s += P[(size_t)k * N + i] * YI[i]; fs[k] = (R)s; }
  R f0v = 0;
  if (f0) { f0v = (R)FSMALL_C; fs[0] -= f0v; }
  for (int k = 0; k < K; ++k) fs[k] = std::fabs(fs[k]);
  R sumf = 0;  // fix_fsum (diffmodels.cc:546-556)
  for (int i = 0; i < K; ++i) {
    sumf += fs[i];
    if (sumf >= 1) { for (int j = i; j < K; ++j) fs[j] = (R)FSMALL_C; break; }
  }
  for (int k = 0; k < K; ++k) {
    R part = 1;
    for (int j = 0; j < k; ++j) part -= fs[j];
    R t = fs[k] / part;
    if (t > 1) t = 1;
    p[2 + 3 * k] = f2beta<R>(t);
  }
  if (f0) {
    R part = 1;
    for (int j = 0; j < K; ++j) part -= fs[j];
    R t = f0v / part;
    if (t > 1) t = 1;
    p[np - 1] = f2beta<R>(t);
  }
}

// finalisations: CUDA/PVM_single.cu:441-468, PVM_single_c.cu:581-611,
// PVM_multi.cu:548-571 (sort by f descending, bubble order; fix_fsum).
template <class R>
static void sort_fibres(int K, int fb, R* p) {
  for (int i = 1; i < K; ++i)
    for (int j = 0; j < K - i; ++j)
      if (p[fb + j * 3] < p[fb + (j + 1) * 3])
        for (int t = 0; t < 3; ++t) std::swap(p[fb + j * 3 + t], p[fb + (j + 1) * 3 + t]);
}

template <class R>
static void finalise(int kind, int K, bool f0, R* p) {
  int np = (kind == K_MULTI ? 3 : 2) + 3 * K + (f0 ? 1 : 0);
  if (kind == K_SINGLE) {
    p[1] = std::fabs(p[1]);
    for (int k = 0; k < K; ++k) p[2 + 3 * k] = x2f<R>(p[2 + 3 * k]);
    if (f0) p[np - 1] = x2f<R>(p[np - 1]);
    sort_fibres<R>(K, 2, p);
    R sum = f0 ? p[np - 1] : R(0);
    for (int i = 0; i < K; ++i) {
      sum += p[2 + 3 * i];
      if (sum >= 1) { for (int j = i; j < K; ++j) p[2 + 3 * j] = (R)FSMALL_C; break; }
    }
  } else if (kind == K_SINGLE_C) {
    p[1] = p[1] * p[1];
    for (int k = 0; k < K; ++k) {
      R part = 1;
      for (int j = 0; j < k; ++j) part -= p[2 + 3 * j];
      p[2 + 3 * k] = beta2f<R>(p[2 + 3 * k]) * part;
    }
    if (f0) {
      R part = 1;
      for (int j = 0; j < K; ++j) part -= p[2 + 3 * j];
      p[np - 1] = beta2f<R>(p[np - 1]) * part;
    }
    sort_fibres<R>(K, 2, p);
  } else {
    R aux = p[1];
    p[1] = std::fabs(aux * p[2]);
    p[2] = std::sqrt(std::fabs(aux * p[2] * p[2]));
    for (int k = 0; k < K; ++k) p[3 + 3 * k] = x2f<R>(p[3 + 3 * k]);
    if (f0) p[np - 1] = x2f<R>(p[np - 1]);
    sort_fibres<R>(K, 3, p);
    R sum = f0 ? p[np - 1] : R(0);
    for (int i = 0; i < K; ++i) {
      if (p[3 + 3 * i] == 0) p[3 + 3 * i] = (R)FSMALL_C;
      sum += p[3 + 3 * i];
      if (sum >= 1) { for (int j = i; j < K; ++j) p[3 + 3 * j] = (R)FSMALL_C; break; }
    }
  }
}

// PVM_multi start from a finalised stage-1 vector (CUDA/PVM_multi.cu:516-534)
template <class R>
static void start_multi_from_stage1(const R* s1, int K, bool f0, R* p) {
  int np1 = 2 + 3 * K + (f0 ? 1 : 0), np = np1 + 1;
  p[0] = s1[0];
  p[1] = 1.0;
  for (int i = 0, ii = 3; i < K; ++i, ii += 3) {
    p[ii] = f2x<R>(s1[ii - 1]);
    p[ii + 1] = s1[ii];
    p[ii + 2] = s1[ii + 1];
  }
  p[2] = s1[1];
  if (f0) p[np - 1] = f2x<R>(s1[np1 - 1]);
}

// ------------------------------------------------------------------ options
}  // namespace orc

extern "C" {
typedef struct orc_opts {
  int nfibres;
  int model;
  double fudge;
  int njumps;
  int nburn;
  int sampleevery;
  int updateproposalevery;
  uint64_t seed;
  int no_ard;
  int all_ard;
  int nonlin;
  int cnonlin;
  int rician;
  int f0;
  int ardf0;
  double Rmean;
  double Rstd;
  double Rfudge;
} orc_opts;
}

namespace orc {

// remove_NonPositive_entries (CUDA/xfibres_gpu.cu:350-367; xfibres.cc:820-835)
template <class R>
static void remove_nonpositive(std::vector<R>& y) {
  R mn = y[0], mx = y[0];
  for (R v : y) { mn = std::min(mn, v); mx = std::max(mx, v); }
  if (mn <= 0 && mx > 0) {
    R minpos = mx;
    for (R v : y) if (v > 0) minpos = std::min(minpos, v);
    for (R& v : y) if (v <= 0) v = minpos;
  }
}

// The fit() dispatch of CUDA/xfibres_gpu.cu:173-348 for one voxel:
// model 1 + --nonlinear -> PVM_single; model 1 otherwise -> PVM_single_c;
// model >= 2 -> PVM_single then PVM_multi (G = 0 model 2, G = 2 model 3).
// With --f0 an implausible fit (f1 < 0.05, d > 0.007 or f0 > 0.4) is refit
// without f0 and given f0 = 0.001 (:216-246, :305-339, mix_params :552-576;
// the model>=2 refit uses PVM_single_c as stage 1, :331-333).
// Output layout: [S0, d, (d_std), f_k, th_k, ph_k ..., (f0)]; returns tau.
template <class R>
struct FitOut { std::vector<R> p; bool used_f0; R tau; int iters; };

template <class R>
static void fit_one(const orc_opts& o, const Acq<R>& acq, const std::vector<double>& iA, const std::vector<R>& y,
                    FitOut<R>& out) {
  const int K = o.nfibres, N = acq.N;
  std::vector<double> yd(y.begin(), y.end());
  DtiOut dti;
  dti_linfit(N, iA.data(), yd.data(), dti);
  auto run = [&](bool f0, bool refit, std::vector<R>& res) {
    R p[NPMAX];
    int iters = 0;
    if (o.model == 1) {
      int kind = o.nonlin ? K_SINGLE : K_SINGLE_C;  // :207-208 and the refits :236, :283
      if (kind == K_SINGLE) start_single<R>(dti, K, f0, p);
      else start_single_c<R>(dti, acq, y.data(), K, f0, p);
      LMModel<R> m(kind, K, f0, 0, R(0), &acq);
      LMStat st = lm_fit<R>(m, y.data(), p);
      iters += st.iters;
      finalise<R>(kind, K, f0, p);
      res.assign(p, p + m.np);
    } else {
      int G = o.model == 3 ? 2 : 0;
      R s1[NPMAX];
      int kind1 = refit ? K_SINGLE_C : K_SINGLE;
      if (kind1 == K_SINGLE) start_single<R>(dti, K, f0, s1);
      else start_single_c<R>(dti, acq, y.data(), K, f0, s1);
      LMModel<R> m1(kind1, K, f0, 0, R(0), &acq);
      LMStat st1 = lm_fit<R>(m1, y.data(), s1);
      finalise<R>(kind1, K, f0, s1);
      start_multi_from_stage1<R>(s1, K, f0, p);
      LMModel<R> m2(K_MULTI, K, f0, G, (R)o.Rmean, &acq);
      LMStat st2 = lm_fit<R>(m2, y.data(), p);
      iters += st1.iters + st2.iters;
      finalise<R>(K_MULTI, K, f0, p);
      res.assign(p, p + m2.np);
    }
    return iters;
  };
  std::vector<R> res;
  out.iters = run(o.f0 != 0, false, res);
  out.used_f0 = o.f0 != 0;
  if (o.f0) {
    int np = (int)res.size();
    R md = res[1], mf = res[o.model >= 2 ? 3 : 2], f0v = res[np - 1];
    if ((K > 0 && mf < 0.05) || md > 0.007 || f0v > 0.4) {
      std::vector<R> r2;
      out.iters += run(false, true, r2);
      r2.push_back((R)0.001);
      res = r2;
      out.used_f0 = false;
    }
  }
  out.p = res;
  out.tau = 0;
  if (o.rician) {
    // calculate_tau (CUDA/diffmodels.cu:259-347): tau = 1/var(residuals),
    // residuals of the final fit; var with the N-1 normalisation
    // (MISCMATHS::var is un-vendored).
    // residual of the final (untransformed) fit; the reference's
    // get_residuals_* kernels re-transform the parameters and evaluate the
    // same forward model (CUDA/PVM_single.cu:476-585, PVM_single_c.cu:619-763,
    // PVM_multi.cu:579-697).
    bool f0 = out.used_f0;
    int np = (int)res.size();
    int fb = o.model >= 2 ? 3 : 2;
    R S0 = o.model >= 2 ? std::fabs(res[0]) : res[0];
    R dd = res[1], a_ = 0, b_ = 0;
    if (o.model >= 2) { a_ = res[1] * res[1] / res[2] / res[2]; b_ = res[2] * res[2] / res[1]; }
    R Rv = (R)o.Rmean, iR = R(1) / (R(2) * Rv + R(1));
    R f0v = f0 ? res[np - 1] : R(0);
    R sumf = 0;
    for (int k = 0; k < K; ++k) sumf += res[fb + 3 * k];
    double mean = 0;
    std::vector<double> r(N);
    for (int i = 0; i < N; ++i) {
      R b = acq.b[i];
      R iso = o.model >= 2 ? std::exp(-a_ * std::log(1 + b * b_)) : std::exp(-b * dd);
      R an = 0;
      for (int k = 0; k < K; ++k) {
        R th = res[fb + 3 * k + 1], ph = res[fb + 3 * k + 2];
        R dp = acq.gx[i] * std::sin(th) * std::cos(ph) + acq.gy[i] * std::sin(th) * std::sin(ph) + acq.gz[i] * std::cos(th);
        R e;
        if (o.model == 1) e = std::exp(-b * dd * dp * dp);
        else if (o.model == 3) e = std::exp(-b * 3 * a_ * b_ * iR * ((1 - Rv) * dp * dp + Rv));
        else e = std::exp(-a_ * std::log(1 + b * b_ * (dp * dp)));
        an += res[fb + 3 * k] * e;
      }
      R pred = S0 * (f0v + (1 - sumf - f0v) * iso + an);
      r[i] = (double)y[i] - (double)pred;
      mean += r[i];
    }
    mean /= N;
    double var = 0;
    for (int i = 0; i < N; ++i) var += (r[i] - mean) * (r[i] - mean);
    var /= (N - 1);
    out.tau = (R)(1.0 / var);
  }
}

// ------------------------------------------------------------------ MCMC
// Semantics of CUDA/runmcmc_kernels.cu: init_Fibres_Multifibres_kernel
// (:193-505) and runmcmc_kernel (:507-1474), burn-in and record as one
// sequence of nburn + njumps sweeps (the record launch continues the
// proposal-update counter from nburn, :707).
// Draw slots (fixed per parameter): f0 0, tau 1, d 2, d_std 3, R 4, S0 5,
// fibre k: th 6+3k, ph 7+3k, f 8+3k.
template <class R>
struct Chain {
  const orc_opts& o;
  const Acq<R>& acq;
  const R* y;
  int K, N, model;
  bool incf0, rician;
  // state
  R S0, d, dstd, Rr, f0, tau;
  R th[8], ph[8], f[8];
  R S0_prop, d_prop, dstd_prop, R_prop, f0_prop, tau_prop, th_prop[8], ph_prop[8], f_prop[8];
  int S0_acc, S0_rej, d_acc, d_rej, dstd_acc, dstd_rej, R_acc, R_rej, f0_acc, f0_rej, tau_acc, tau_rej;
  int th_acc[8], th_rej[8], ph_acc[8], ph_rej[8], f_acc[8], f_rej[8];
  R d_prior, dstd_prior, R_prior, f0_prior, th_prior[8], f_prior[8];
  bool lam_jump[8];
  R energy;  // prior + likelihood
  std::vector<R> ang, sig, iso;  // ang[k*N+i] = (g.x)^2, sig[k*N+i], iso[i]
  std::vector<R> ns, niso, na, nsk;  // candidate signals

  Chain(const orc_opts& oo, const Acq<R>& a, const R* yy) : o(oo), acq(a), y(yy) {
    K = o.nfibres; N = a.N; model = o.model; incf0 = o.f0 != 0; rician = o.rician != 0;
    ang.assign((size_t)K * N, 0); sig.assign((size_t)K * N, 0); iso.assign(N, 0);
    ns.assign((size_t)K * N, 0); niso.assign(N, 0); na.assign(N, 0); nsk.assign(N, 0);
  }

  R aniso_sig(R a, R b) const {  // compute_signal, CUDA/runmcmc_kernels.cu:67-82
    if (model == 1 || (dstd < R(1e-5) && model == 2)) return std::exp(-d * b * a);
    if (model == 2) {
      R sig2 = dstd * dstd, al = d * d / sig2;
      return std::exp(std::log(d / (d + b * a * sig2)) * al);
    }
    R iR = R(1) / (R(2) * Rr + R(1));
    return std::exp(-b * 3 * d * iR * ((1 - Rr) * a + Rr));
  }
  R iso_sig(R b) const {  // compute_iso_signal, :83-95
    if (model == 1 || dstd < R(1e-5)) return std::exp(-d * b);
    R sig2 = dstd * dstd, al = d * d / sig2;
    return std::exp(std::log(d / (d + b * sig2)) * al);
  }
  R angle(int k, int i) const {
    R st = std::sin(th[k]), ct = std::cos(th[k]), sp = std::sin(ph[k]), cp = std::cos(ph[k]);
    R dp = acq.gx[i] * st * cp + acq.gy[i] * st * sp + acq.gz[i] * ct;
    return dp * dp;
  }
  R fsum() const { R s = f0; for (int k = 0; k < K; ++k) s += f[k]; return s; }
  R prior_en() const {  // compute_prior, :118-124
    R p = d_prior + dstd_prior + R_prior + f0_prior;  // S0 and tau priors are 0
    for (int k = 0; k < K; ++k) p += th_prior[k] + f_prior[k];
    return p;
  }
  R loglik(const R* s_, const R* iso_, R f0v, R S0v, R tauv, const R* fv, int kover = -1, const R* sover = nullptr) const {  // compute_likelihood, :147-191
    double acc = 0;
    R fs = f0v;
    for (int k = 0; k < K; ++k) fs += fv[k];
    for (int i = 0; i < N; ++i) {
      R pred = 0;
      for (int k = 0; k < K; ++k) pred += fv[k] * (k == kover ? sover[i] : s_[(size_t)k * N + i]);
      pred = S0v * (pred + (1 - fs) * iso_[i] + f0v);
      if (!rician) { R df = y[i] - pred; acc += (double)df * (double)df; }
      else acc += (double)(std::log(y[i]) + (R(-0.5) * tauv * (y[i] * y[i] + pred * pred) + logI0(tauv * pred * y[i])));
    }
    if (!rician) return (R)((N / 2.0) * std::log(acc / 2.0));
    return (R)(-N * std::log((double)tauv) - acc);
  }
  static R logI0(R x) {  // logIo, fibre.h:29-48 (Abramowitz-Stegun 9.8.1/9.8.2)
    R b = std::fabs(x), yv;
    if (b < R(3.75)) {
      R a = x / R(3.75);
      a *= a;
      yv = R(1.0) + a * (R(3.5156229) + a * (R(3.0899424) + a * (R(1.2067492) + a * (R(0.2659732) + a * (R(0.0360768) + a * R(0.0045813))))));
      return std::log(yv);
    }
    R a = R(3.75) / b;
    return b + std::log((R(0.39894228) + a * (R(0.01328592) + a * (R(0.00225319) + a * (R(-0.00157565) + a * (R(0.00916281) + a * (R(-0.02057706) + a * (R(0.02635537) + a * (R(-0.01647633) + a * R(0.00392377))))))))) / std::sqrt(b));
  }
  R lik_cur() const { return loglik(sig.data(), iso.data(), f0, S0, tau, f); }

  void init(const R* params, R tau_in) {  // init_Fibres_Multifibres_kernel, :265-504
    int add = model >= 2 ? 1 : 0;
    int npf = 2 + 3 * K + add + (incf0 ? 1 : 0);
    S0 = params[0];
    d = params[1];
    if (d < 0 || d > R(UPPERDIFF_C)) d = R(2e-3);
    if (model >= 2) {
      dstd = params[2];
      R up = model == 3 ? R(0.004) : R(0.01);
      if (dstd < 0 || dstd > up) dstd = d / 10;
      Rr = model == 3 ? (R)o.Rmean : R(0);
    } else { dstd = 0; Rr = 0; }
    f0 = incf0 ? params[npf - 1] : R(0);
    tau = tau_in;
    S0_acc = S0_rej = d_acc = d_rej = dstd_acc = dstd_rej = R_acc = R_rej = f0_acc = f0_rej = tau_acc = tau_rej = 0;
    d_prior = dstd_prior = R_prior = f0_prior = 0;
    for (int k = 0; k < K; ++k) {
      th[k] = params[2 + 3 * k + 1 + add];
      ph[k] = params[2 + 3 * k + 2 + add];
      f[k] = params[2 + 3 * k + add];
      th_prop[k] = ph_prop[k] = f_prop[k] = R(0.2);
      th_acc[k] = th_rej[k] = ph_acc[k] = ph_rej[k] = f_acc[k] = f_rej[k] = 0;
      th_prior[k] = th[k] == 0 ? R(0) : (R)(-std::log(std::fabs(std::sin((double)th[k])) / 2));
      lam_jump[k] = (k == 0) ? (o.all_ard != 0) : (o.no_ard == 0);
      f_prior[k] = 0;
      if (!(f[k] <= 0 || f[k] >= 1)) f_prior[k] = (R)o.fudge * (lam_jump[k] ? (R)std::log((double)f[k]) : R(0));
    }
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < N; ++i) {
        ang[(size_t)k * N + i] = angle(k, i);
        sig[(size_t)k * N + i] = aniso_sig(ang[(size_t)k * N + i], acq.b[i]);
      }
    if (d >= 0 && d <= R(UPPERDIFF_C) && model == 3) d_prior = (R(1) - R(3)) * std::log(d) + R(4000) * d;
    if (model >= 2) {
      R up = model == 3 ? R(0.004) : R(0.01);
      if (dstd > 0 && dstd <= up) dstd_prior = std::log(dstd);
      if (model == 3) R_prior = r_prior_value(Rr, nullptr);
    }
    if (incf0 && !(f0 <= 0 || f0 >= 1) && o.ardf0) f0_prior = (R)std::log((double)f0);
    for (int i = 0; i < N; ++i) iso[i] = iso_sig(acq.b[i]);
    energy = prior_en() + lik_cur();
    S0_prop = S0 / 10; d_prop = d / 10; dstd_prop = dstd / 10; tau_prop = tau / 2; f0_prop = R(0.2); R_prop = Rr / 10;
  }
  // R prior (fibre.h:816-843; CUDA/runmcmc_kernels.cu:1058-1085); returns
  // the prior and sets *rej when the value is outside its support.
  R r_prior_value(R Rv, bool* rej) const {
    R upper = R(2) * (R)o.Rmean, lower = (R)o.Rmean - R(2) * (R)o.Rstd;
    if ((R)o.Rmean > R(0.5)) upper = 1;
    if (lower < 0) lower = R(1e-8);
    if ((R)o.Rfudge > 0 && d > R(UPPERDIFF_C) / R(2)) {
      bool r = (Rv < R(1e-8) || Rv > upper);
      if (rej) *rej = r;
      return r ? R_prior : (R)o.Rfudge * std::log(Rv);
    }
    bool r = (Rv <= lower || Rv > upper);
    if (rej) *rej = r;
    R s2 = (R)o.Rstd * (R)o.Rstd;
    return r ? R_prior : (Rv - (R)o.Rmean) * (Rv - (R)o.Rmean) / s2;
  }

  // one Metropolis test; E_new is the candidate energy
  bool accept(R E_new, double logu) {
    double dE = (double)energy - (double)E_new;
    if (dE > logu) { energy = E_new; return true; }
    return false;
  }

  void jump(uint64_t seed, uint64_t vox, uint32_t j) {
    // f0 (:842-888)
    if (incf0) {
      Draw dr = draw_pair(seed, vox, j, 0);
      R old = f0;
      f0 = f0 + (R)dr.n * f0_prop;
      bool rej = (f0 <= 0 || f0 >= 1) || fsum() > 1;
      if (!rej) {
        R old_prior = f0_prior;
        f0_prior = o.ardf0 ? (R)std::log((double)f0) : R(0);
        R E = prior_en() + lik_cur();
        if (accept(E, dr.logu)) f0_acc++;
        else { f0 = old; f0_prior = old_prior; f0_rej++; }
      } else { f0 = old; f0_rej++; }
    }
    // tau (:890-922)
    if (rician) {
      Draw dr = draw_pair(seed, vox, j, 1);
      R old = tau;
      tau = tau + (R)dr.n * tau_prop;
      if (tau > 0) {
        R E = prior_en() + lik_cur();
        if (accept(E, dr.logu)) tau_acc++;
        else { tau = old; tau_rej++; }
      } else { tau = old; tau_rej++; }
    }
    // d (:924-985)
    {
      Draw dr = draw_pair(seed, vox, j, 2);
      R old = d;
      d = d + (R)dr.n * d_prop;
      if (!(d < 0 || d > R(UPPERDIFF_C))) {
        R old_prior = d_prior;
        d_prior = model == 3 ? (R(1) - R(3)) * std::log(d) + R(4000) * d : R(0);
        for (int k = 0; k < K; ++k)
          for (int i = 0; i < N; ++i) ns[(size_t)k * N + i] = aniso_sig(ang[(size_t)k * N + i], acq.b[i]);
        for (int i = 0; i < N; ++i) niso[i] = iso_sig(acq.b[i]);
        R E = prior_en() + loglik(ns.data(), niso.data(), f0, S0, tau, f);
        if (accept(E, dr.logu)) { d_acc++; sig.swap(ns); iso.swap(niso); }
        else { d = old; d_prior = old_prior; d_rej++; }
      } else { d = old; d_rej++; }
    }
    // d_std (:987-1051)
    if (model >= 2) {
      Draw dr = draw_pair(seed, vox, j, 3);
      R old = dstd;
      dstd = dstd + (R)dr.n * dstd_prop;
      R up = model == 3 ? R(0.004) : R(0.01);
      if (!(dstd <= 0 || dstd > up)) {
        R old_prior = dstd_prior;
        dstd_prior = std::log(dstd);
        if (model == 2)
          for (int k = 0; k < K; ++k)
            for (int i = 0; i < N; ++i) ns[(size_t)k * N + i] = aniso_sig(ang[(size_t)k * N + i], acq.b[i]);
        for (int i = 0; i < N; ++i) niso[i] = iso_sig(acq.b[i]);
        R E = prior_en() + loglik(model == 2 ? ns.data() : sig.data(), niso.data(), f0, S0, tau, f);
        if (accept(E, dr.logu)) { dstd_acc++; if (model == 2) sig.swap(ns); iso.swap(niso); }
        else { dstd = old; dstd_prior = old_prior; dstd_rej++; }
      } else { dstd = old; dstd_rej++; }
      // R (:1053-1122)
      if (model == 3) {
        Draw dr2 = draw_pair(seed, vox, j, 4);
        R oldR = Rr;
        Rr = Rr + (R)dr2.n * R_prop;
        bool rej = false;
        R old_prior = R_prior;
        R np_ = r_prior_value(Rr, &rej);
        if (!rej) {
          R_prior = np_;
          for (int k = 0; k < K; ++k)
            for (int i = 0; i < N; ++i) ns[(size_t)k * N + i] = aniso_sig(ang[(size_t)k * N + i], acq.b[i]);
          R E = prior_en() + loglik(ns.data(), iso.data(), f0, S0, tau, f);
          if (accept(E, dr2.logu)) { R_acc++; sig.swap(ns); }
          else { Rr = oldR; R_prior = old_prior; R_rej++; }
        } else { Rr = oldR; R_rej++; }
      }
    }
    // S0 (:1125-1155)
    {
      Draw dr = draw_pair(seed, vox, j, 5);
      R old = S0;
      S0 = S0 + (R)dr.n * S0_prop;
      if (!(S0 < 0)) {
        R E = prior_en() + lik_cur();
        if (accept(E, dr.logu)) S0_acc++;
        else { S0 = old; S0_rej++; }
      } else { S0 = old; S0_rej++; }
    }
    // fibres: th (:1157-1211), ph (:1213-1265), f (:1267-1322)
    for (int k = 0; k < K; ++k) {
      for (int which = 0; which < 2; ++which) {
        Draw dr = draw_pair(seed, vox, j, 6 + 3 * k + which);
        R& v = which == 0 ? th[k] : ph[k];
        R prop = which == 0 ? th_prop[k] : ph_prop[k];
        R old = v;
        v = v + (R)dr.n * prop;
        R old_prior = th_prior[k];
        if (which == 0) th_prior[k] = th[k] == 0 ? R(0) : (R)(-std::log(std::fabs(std::sin((double)th[k])) / 2));
        for (int i = 0; i < N; ++i) { na[i] = angle(k, i); nsk[i] = aniso_sig(na[i], acq.b[i]); }
        R E = prior_en() + loglik(sig.data(), iso.data(), f0, S0, tau, f, k, nsk.data());
        if (accept(E, dr.logu)) {
          (which == 0 ? th_acc[k] : ph_acc[k])++;
          for (int i = 0; i < N; ++i) { ang[(size_t)k * N + i] = na[i]; sig[(size_t)k * N + i] = nsk[i]; }
        } else {
          v = old;
          th_prior[k] = old_prior;
          (which == 0 ? th_rej[k] : ph_rej[k])++;
        }
      }
      {
        Draw dr = draw_pair(seed, vox, j, 8 + 3 * k);
        R old = f[k];
        f[k] = f[k] + (R)dr.n * f_prop[k];
        bool rej = (f[k] <= 0 || f[k] >= 1) || fsum() > 1;
        if (!rej) {
          R old_prior = f_prior[k];
          R fp = 0;
          if (!o.no_ard && lam_jump[k]) fp = (R)std::log((double)f[k]);
          f_prior[k] = (R)o.fudge * fp;
          R E = prior_en() + lik_cur();
          if (accept(E, dr.logu)) f_acc[k]++;
          else { f[k] = old; f_prior[k] = old_prior; f_rej[k]++; }
        } else { f[k] = old; f_rej[k]++; }
      }
    }
  }

  // initial proposal widths per draw slot (<= 0: keep the default); the
  // product's fdt_set_widths, used by the learned start
  void set_widths(const double* w) {
    auto put = [&](R& prop, int slot) { if (w[slot] > 0) prop = (R)w[slot]; };
    put(f0_prop, 0); put(tau_prop, 1); put(d_prop, 2); put(dstd_prop, 3); put(R_prop, 4); put(S0_prop, 5);
    for (int k = 0; k < K; ++k) { put(th_prop[k], 6 + 3 * k); put(ph_prop[k], 7 + 3 * k); put(f_prop[k], 8 + 3 * k); }
  }

  static void upd(R& prop, int& acc, int& rej) {  // :1345-1393
    prop *= std::sqrt(R(acc + 1) / R(rej + 1));
    prop = std::min(prop, (R)MAXFLOAT_C);
    acc = 0; rej = 0;
  }
  void update_proposals() {
    upd(d_prop, d_acc, d_rej);
    if (rician) upd(tau_prop, tau_acc, tau_rej);
    if (incf0) upd(f0_prop, f0_acc, f0_rej);
    if (model >= 2) { upd(dstd_prop, dstd_acc, dstd_rej); if (model == 3) upd(R_prop, R_acc, R_rej); }
    upd(S0_prop, S0_acc, S0_rej);
    for (int k = 0; k < K; ++k) { upd(th_prop[k], th_acc[k], th_rej[k]); upd(ph_prop[k], ph_acc[k], ph_rej[k]); upd(f_prop[k], f_acc[k], f_rej[k]); }
  }
};

}  // namespace orc

// ------------------------------------------------------------------ C API
// Sample layout (shared with the product):
//   per-fibre quantity q in {th, ph, f}: q[k][s][v]  (K x ns x nvox)
//   scalars S0, d, d_std, R, f0, tau:   q[s][v]      (ns x nvox)
extern "C" {

typedef struct orc_samples {
  float* th;   // K*ns*nvox
  float* ph;
  float* f;
  float* S0;   // ns*nvox
  float* d;
  float* dstd; // may be NULL
  float* R;    // may be NULL
  float* f0;   // may be NULL
  float* tau;  // may be NULL
} orc_samples;

#define ORC_INSTANTIATE(RT, PFX)                                                                              \
  int PFX##_lm(int kind, int K, int f0, int G, double Rv, int N, const double* bvals, const double* bvecs,    \
               int nvox, const double* data, const double* start, double* out, int* iters, double* cost) {   \
    orc::Acq<RT> acq(N, bvals, bvecs);                                                                         \
    orc::LMModel<RT> m(kind, K, f0 != 0, G, (RT)Rv, &acq);                                                      \
    int np = m.np;                                                                                            \
    _Pragma("omp parallel for schedule(dynamic, 4)") for (int v = 0; v < nvox; ++v) {                          \
      std::vector<RT> y(N);                                                                                    \
      RT p[orc::NPMAX];                                                                                        \
      for (int i = 0; i < N; ++i) y[i] = (RT)data[(size_t)v * N + i];                                          \
      for (int i = 0; i < np; ++i) p[i] = (RT)start[(size_t)v * np + i];                                       \
      orc::LMStat st = orc::lm_fit<RT>(m, y.data(), p);                                                        \
      if (iters) iters[v] = st.iters;                                                                         \
      if (cost) cost[v] = st.cost;                                                                            \
      orc::finalise<RT>(kind, K, f0 != 0, p);                                                                  \
      for (int i = 0; i < np; ++i) out[(size_t)v * np + i] = (double)p[i];                                    \
    }                                                                                                         \
    return 0;                                                                                                 \
  }                                                                                                           \
  int PFX##_model_eval(int kind, int K, int f0, int G, double Rv, int N, const double* bvals,                  \
                       const double* bvecs, const double* p_in, double* pred, double* Jg, double* Jh,         \
                       double* cost_of_y, const double* y_in) {                                               \
    orc::Acq<RT> acq(N, bvals, bvecs);                                                                         \
    orc::LMModel<RT> m(kind, K, f0 != 0, G, (RT)Rv, &acq);                                                      \
    int np = m.np;                                                                                            \
    RT p[orc::NPMAX];                                                                                          \
    for (int i = 0; i < np; ++i) p[i] = (RT)p_in[i];                                                           \
    typename orc::LMModel<RT>::St s;                                                                           \
    m.prep(p, s);                                                                                             \
    RT jg[orc::NPMAX], jh[orc::NPMAX];                                                                         \
    for (int i = 0; i < N; ++i) {                                                                             \
      RT sig;                                                                                                  \
      if (pred) pred[i] = (double)m.pred(p, s, i);                                                            \
      m.jrow(p, s, i, sig, jg, jh);                                                                           \
      for (int j = 0; j < np; ++j) {                                                                          \
        if (Jg) Jg[(size_t)i * np + j] = jg[j];                                                               \
        if (Jh) Jh[(size_t)i * np + j] = jh[j];                                                               \
      }                                                                                                       \
    }                                                                                                         \
    if (cost_of_y && y_in) {                                                                                  \
      std::vector<RT> y(N);                                                                                    \
      for (int i = 0; i < N; ++i) y[i] = (RT)y_in[i];                                                          \
      *cost_of_y = m.cost(p, y.data());                                                                       \
    }                                                                                                         \
    return np;                                                                                                \
  }                                                                                                           \
  int PFX##_lu_solve(int n, const double* A, const double* b, double* x) {                                    \
    RT a[orc::NPMAX * orc::NPMAX] = {}, bb[orc::NPMAX] = {}, xx[orc::NPMAX] = {};                                           \
    for (int i = 0; i < n * n; ++i) a[i] = (RT)A[i];                                                           \
    for (int i = 0; i < n; ++i) bb[i] = (RT)b[i];                                                              \
    orc::lu_solve<RT>(n, a, bb, xx);                                                                           \
    for (int i = 0; i < n; ++i) x[i] = (double)xx[i];                                                         \
    return 0;                                                                                                 \
  }                                                                                                           \
  int PFX##_start(int kind, int K, int f0, int N, const double* bvals, const double* bvecs, int nvox,         \
                  const double* data, double* start_out, const double* grad) {                                \
    orc::Acq<RT> acq0(N, bvals, bvecs);                                                                        \
    std::vector<double> A0, iA0;                                                                              \
    orc::dti_amat_pinv(N, bvals, bvecs, A0, iA0);                                                             \
    int np = 2 + 3 * K + (f0 ? 1 : 0);                                                                        \
    _Pragma("omp parallel for schedule(static)") for (int v = 0; v < nvox; ++v) {                             \
      orc::VoxAcq<RT> va(N, bvals, bvecs, grad ? grad + (size_t)v * 9 : nullptr, acq0, iA0);                   \
      orc::DtiOut dti;                                                                                        \
      orc::dti_linfit(N, va.iA->data(), data + (size_t)v * N, dti);                                           \
      RT p[orc::NPMAX];                                                                                        \
      std::vector<RT> y(N);                                                                                    \
      for (int i = 0; i < N; ++i) y[i] = (RT)data[(size_t)v * N + i];                                          \
      if (kind == orc::K_SINGLE) orc::start_single<RT>(dti, K, f0 != 0, p);                                    \
      else orc::start_single_c<RT>(dti, *va.acq, y.data(), K, f0 != 0, p);                                     \
      for (int i = 0; i < np; ++i) start_out[(size_t)v * np + i] = (double)p[i];                              \
    }                                                                                                         \
    return np;                                                                                                \
  }                                                                                                           \
  int PFX##_fit(const orc_opts* o, int N, const double* bvals, const double* bvecs, int nvox,                 \
                const double* data, double* params, double* tau, int* iters, const double* grad) {            \
    orc::Acq<RT> acq0(N, bvals, bvecs);                                                                        \
    std::vector<double> A0, iA0;                                                                              \
    orc::dti_amat_pinv(N, bvals, bvecs, A0, iA0);                                                             \
    int npo = 2 + 3 * o->nfibres + (o->model >= 2 ? 1 : 0) + (o->f0 ? 1 : 0);                                 \
    _Pragma("omp parallel for schedule(dynamic, 4)") for (int v = 0; v < nvox; ++v) {                          \
      orc::VoxAcq<RT> va(N, bvals, bvecs, grad ? grad + (size_t)v * 9 : nullptr, acq0, iA0);                   \
      std::vector<RT> y(N);                                                                                    \
      for (int i = 0; i < N; ++i) y[i] = (RT)data[(size_t)v * N + i];                                          \
      if (o->rician) orc::remove_nonpositive<RT>(y);                                                           \
      orc::FitOut<RT> fo;                                                                                      \
      orc::fit_one<RT>(*o, *va.acq, *va.iA, y, fo);                                                            \
      for (int i = 0; i < npo; ++i) params[(size_t)v * npo + i] = (double)fo.p[i];                            \
      if (tau) tau[v] = (double)fo.tau;                                                                       \
      if (iters) iters[v] = fo.iters;                                                                         \
    }                                                                                                         \
    return npo;                                                                                               \
  }                                                                                                           \
  int PFX##_mcmc(const orc_opts* o, int N, const double* bvals, const double* bvecs, int nvox,                \
                 const double* data, const double* params, const double* tau, uint64_t vox_offset,            \
                 orc_samples* out, double* final_energy, const double* grad, const double* widths) {          \
    orc::Acq<RT> acq0(N, bvals, bvecs);                                                                        \
    int K = o->nfibres;                                                                                       \
    int npo = 2 + 3 * K + (o->model >= 2 ? 1 : 0) + (o->f0 ? 1 : 0);                                          \
    int ns = o->njumps / o->sampleevery;                                                                      \
    size_t nv = (size_t)nvox;                                                                                 \
    _Pragma("omp parallel for schedule(dynamic, 1)") for (int v = 0; v < nvox; ++v) {                          \
      std::vector<RT> y(N);                                                                                    \
      for (int i = 0; i < N; ++i) y[i] = (RT)data[(size_t)v * N + i];                                          \
      if (o->rician) orc::remove_nonpositive<RT>(y);                                                           \
      std::vector<RT> p(npo);                                                                                  \
      for (int i = 0; i < npo; ++i) p[i] = (RT)params[(size_t)v * npo + i];                                    \
      std::vector<double> none;                                                                                \
      orc::VoxAcq<RT> va(N, bvals, bvecs, grad ? grad + (size_t)v * 9 : nullptr, acq0, none, false);            \
      orc::Chain<RT> ch(*o, *va.acq, y.data());                                                                \
      ch.init(p.data(), tau ? (RT)tau[v] : RT(0));                                                              \
      if (widths) ch.set_widths(widths + (size_t)v * 16);                                                      \
      uint64_t gv = vox_offset + (uint64_t)v;                                                                 \
      int total = o->nburn + o->njumps;                                                                       \
      for (int j = 0; j < total; ++j) {                                                                       \
        ch.jump(o->seed, gv, (uint32_t)j);                                                                    \
        if (j >= o->nburn) {                                                                                  \
          int r = j - o->nburn + 1;                                                                           \
          if (o->sampleevery > 0 && r % o->sampleevery == 0) {                                                \
            int s = r / o->sampleevery - 1;                                                                   \
            if (s < ns) {                                                                                     \
              size_t sv = (size_t)s * nv + v;                                                                 \
              out->S0[sv] = (float)ch.S0; out->d[sv] = (float)ch.d;                                           \
              if (out->dstd) out->dstd[sv] = (float)ch.dstd;                                                  \
              if (out->R) out->R[sv] = (float)ch.Rr;                                                          \
              if (out->f0) out->f0[sv] = (float)ch.f0;                                                        \
              if (out->tau) out->tau[sv] = (float)ch.tau;                                                     \
              for (int k = 0; k < K; ++k) {                                                                   \
                size_t kv = ((size_t)k * ns + s) * nv + v;                                                    \
                out->th[kv] = (float)ch.th[k]; out->ph[kv] = (float)ch.ph[k]; out->f[kv] = (float)ch.f[k];    \
              }                                                                                               \
            }                                                                                                 \
          }                                                                                                   \
        }                                                                                                     \
        if ((j + 1) % o->updateproposalevery == 0) ch.update_proposals();                                     \
      }                                                                                                       \
      if (final_energy) final_energy[v] = (double)ch.energy;                                                  \
    }                                                                                                         \
    return ns;                                                                                                \
  }

ORC_INSTANTIATE(double, orc64)
ORC_INSTANTIATE(float, orc32)

// DTI of one voxel: out = [s0, md, fa, l1, l2, l3, v1(3), v2(3), v3(3)]
int orc_dti(int N, const double* bvals, const double* bvecs, int nvox, const double* data, double* out) {
  std::vector<double> A, iA;
  orc::dti_amat_pinv(N, bvals, bvecs, A, iA);
  for (int v = 0; v < nvox; ++v) {
    orc::DtiOut o;
    orc::dti_linfit(N, iA.data(), data + (size_t)v * N, o);
    double* r = out + (size_t)v * 15;
    r[0] = o.s0; r[1] = o.md; r[2] = o.fa; r[3] = o.l[0]; r[4] = o.l[1]; r[5] = o.l[2];
    for (int k = 0; k < 3; ++k) for (int i = 0; i < 3; ++i) r[6 + 3 * k + i] = o.v[k][i];
  }
  return 0;
}

int orc_correct_acq(int N, const double* bvals, const double* bvecs, const double* g, double* bvals_c,
                    double* bvecs_c) {
  orc::correct_acq(N, bvals, bvecs, g, bvals_c, bvecs_c);
  return 0;
}
int orc_pinv(int m, int n, const double* A, double* P) { orc::pinv(m, n, A, P); return 0; }
int orc_eig(int n, const double* A, double* w, double* V) { orc::jacobi_eig(n, A, w, V); return 0; }
int orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { orc::philox4x32_10(ctr, key, out); return 0; }
int orc_draw(uint64_t seed, uint64_t vox, uint32_t jump, uint32_t slot, double* n, double* logu) {
  orc::Draw d = orc::draw_pair(seed, vox, jump, slot);
  *n = d.n; *logu = d.logu;
  return 0;
}
double orc_logI0(double x) { return (double)orc::Chain<double>::logI0(x); }

// MCMC forward model of one voxel (for closed-form tests): returns pred[N].
int orc_mcmc_predict(const orc_opts* o, int N, const double* bvals, const double* bvecs, const double* state,
                     double* pred) {
  // state = [S0, d, d_std, R, f0, th_k, ph_k, f_k ...]
  orc::Acq<double> acq(N, bvals, bvecs);
  std::vector<double> y(N, 1.0);
  orc::Chain<double> ch(*o, acq, y.data());
  int K = o->nfibres;
  ch.S0 = state[0]; ch.d = state[1]; ch.dstd = state[2]; ch.Rr = state[3]; ch.f0 = state[4];
  for (int k = 0; k < K; ++k) { ch.th[k] = state[5 + 3 * k]; ch.ph[k] = state[6 + 3 * k]; ch.f[k] = state[7 + 3 * k]; }
  double fs = ch.fsum();
  for (int i = 0; i < N; ++i) {
    double a = 0;
    for (int k = 0; k < K; ++k) a += ch.f[k] * ch.aniso_sig(ch.angle(k, i), acq.b[i]);
    pred[i] = ch.S0 * (a + (1 - fs) * ch.iso_sig(acq.b[i]) + ch.f0);
  }
  return 0;
}

// Finish: per voxel sort fibres by sum_s f (descending; ties keep the
// reference's std::sort-of-pairs order, i.e. the higher index first,
// CUDA/samples.cu:297-330), means (CUDA/samples.cu:173-229), dyads and
// dispersion (principal eigenvector of the mean dyadic tensor,
// make_dyadic_vectors.cc:65-88).  Samples are permuted in place.
int orc_finish(int K, int ns, int nvox, orc_samples* s, double* mean_S0, double* mean_d, double* mean_dstd,
               double* mean_R, double* mean_f0, double* mean_tau, double* mean_f /*K*nvox*/,
               double* dyads /*K*3*nvox*/, double* disp /*K*nvox*/) {
  size_t nv = (size_t)nvox;
  for (int v = 0; v < nvox; ++v) {
    auto mean_of = [&](float* q) { double a = 0; for (int j = 0; j < ns; ++j) a += q[(size_t)j * nv + v]; return a / ns; };
    if (mean_S0) mean_S0[v] = mean_of(s->S0);
    if (mean_d) mean_d[v] = mean_of(s->d);
    if (mean_dstd && s->dstd) mean_dstd[v] = mean_of(s->dstd);
    if (mean_R && s->R) mean_R[v] = mean_of(s->R);
    if (mean_f0 && s->f0) mean_f0[v] = mean_of(s->f0);
    if (mean_tau && s->tau) mean_tau[v] = mean_of(s->tau);
    std::vector<std::pair<double, int>> sf(K);
    for (int k = 0; k < K; ++k) {
      double a = 0;
      for (int j = 0; j < ns; ++j) a += s->f[((size_t)k * ns + j) * nv + v];
      sf[k] = {a, k};
    }
    std::sort(sf.begin(), sf.end());
    std::vector<float> tmp((size_t)K * ns * 3);
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < ns; ++j) {
        size_t src = ((size_t)k * ns + j) * nv + v;
        tmp[((size_t)k * ns + j) * 3 + 0] = s->th[src];
        tmp[((size_t)k * ns + j) * 3 + 1] = s->ph[src];
        tmp[((size_t)k * ns + j) * 3 + 2] = s->f[src];
      }
    for (int k = 0; k < K; ++k) {
      int from = sf[K - 1 - k].second;
      double T[9] = {0};
      double fsum = 0;
      for (int j = 0; j < ns; ++j) {
        size_t dst = ((size_t)k * ns + j) * nv + v;
        float th = tmp[((size_t)from * ns + j) * 3 + 0], ph = tmp[((size_t)from * ns + j) * 3 + 1], ff = tmp[((size_t)from * ns + j) * 3 + 2];
        s->th[dst] = th; s->ph[dst] = ph; s->f[dst] = ff;
        double x[3] = {std::sin((double)th) * std::cos((double)ph), std::sin((double)th) * std::sin((double)ph), std::cos((double)th)};
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) T[a * 3 + b] += x[a] * x[b];
        fsum += ff;
      }
      for (int a = 0; a < 9; ++a) T[a] /= ns;
      double w[3], V[9];
      orc::jacobi_eig(3, T, w, V);
      if (mean_f) mean_f[(size_t)k * nv + v] = fsum / ns;
      if (dyads) for (int a = 0; a < 3; ++a) dyads[((size_t)k * 3 + a) * nv + v] = V[a * 3 + 2];
      if (disp) disp[(size_t)k * nv + v] = 1 - std::max(std::fabs(w[0]), std::max(std::fabs(w[1]), std::fabs(w[2])));
    }
  }
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

}  // extern "C"
