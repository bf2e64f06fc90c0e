// fdt_chain.h -- one Metropolis-Hastings chain of the ball-and-sticks models
// per 16-lane row on gfx950, shared by the per-volume kernel (fdt_mcmc.hip)
// and the batched queue kernel (fdt_queue.hip).
//
// Semantics follow the reference GPU MCMC: init_Fibres_Multifibres_kernel
// (CUDA/runmcmc_kernels.cu:193-505) and runmcmc_kernel (:507-1474), with the
// burn-in and record launches (CUDA/runmcmc.cu:108-316) fused into one loop of
// nburn + njumps sweeps.  Layout: one voxel per 16-lane DPP row, four voxels
// per wave64.  Lane l of a row holds volumes l, l+16, l+32, ... (VPL volumes
// per lane) and keeps, in VGPRs,
//   the residual r = y - S0*M (Gaussian) or the data y (Rician),
//   the mixture M = sum_k f_k s_k + (1 - f0 - sum_k f_k) iso + f0,
//   the isotropic signal iso and the stick signals s_k;
// The candidate stick signals of the diffusivity proposals are parked in LDS
// (one block per wave, lane-interleaved) so that an accepted move copies them;
// K=3 at 18 volumes per lane fits two waves per SIMD.
// A proposal evaluates only what it changes; its likelihood change is
// summed over the row with DPP adds.  For Gaussian noise a move of every
// residual by -c*q gives the change of the sum of squares c*(c*sum q^2 -
// 2*sum r*q), and the energy difference (N/2)*log1p(dSSR/SSR) keeps full
// relative precision (the reference differences two FP32 energies).  The
// acquisition (gx,gy,gz,b) is staged in LDS.  Draws: lane l of the row
// computes the Philox block of proposal slot l for the current jump; DPP
// row_newbcast hands it to the row.
#pragma once
#include "fdt_device.h"
#include "fdt_kernels.h"
#include <utility>

namespace fdt {

constexpr float L2E = 1.4426950408889634f;
constexpr float LN2 = 0.6931471805599453f;
FDT_DEV float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
FDT_DEV float lg2(float x) { return __builtin_amdgcn_logf(x); }
FDT_DEV float flog(float x) { return lg2(x) * LN2; }

// sin and cos of a float angle of any magnitude.  Width adaptation drives the
// angles of unsupported fibres to 1e8..1e11 rad, and the reference evaluates
// them as sin/cos(double(angle)) with a full-range reduction
// (CUDA/runmcmc_kernels.cu:340, :1165, :1177-1182).  Here the angle is turned
// into revolutions x/(2 pi) and reduced to its fraction in double, with
// 1/(2 pi) split into two doubles: the first FMA forms x * HI exactly before
// subtracting the nearest integer, so the fraction is accurate to ~2^-53 for
// |x| < 2^60.  The hardware sin/cos take revolutions (valid in [-256, 256]).
FDT_DEV void sincos_full(float x, float& s, float& c) {
  constexpr double R_HI = 0x1.45f306dc9c883p-3, R_LO = -0x1.6b01ec5417056p-57;
  const double xd = (double)x;
  const double n = __builtin_rint(xd * R_HI);
  double fr = __builtin_fma(xd, R_HI, -n);
  fr = __builtin_fma(xd, R_LO, fr);
  const float f = (float)fr;
  s = __builtin_amdgcn_sinf(f);
  c = __builtin_amdgcn_cosf(f);
}

// draw slots (fixed per parameter)
enum { SL_F0 = 0, SL_TAU = 1, SL_D = 2, SL_DSTD = 3, SL_R = 4, SL_S0 = 5, SL_FIB = 6 };

// Tiles whose chains stage b = 0 volumes first and skip their slot (see
// stage_acq): the 18-volume tile, where it saves 3-4% of the chain time.  In
// the smaller tiles the uniform branch around slot 0 costs more than it saves
// (it splits the few volumes a lane interleaves); at 24 and 32 volumes per
// lane (one wave per SIMD) it saved nothing (N = 384 with 24 b = 0 volumes)
// and cost 1.4% (N = 432).
template <int VPL> constexpr bool ZERO_SLOT = VPL == 18;

// Signal forms with row-uniform coefficients (compute_signal /
// compute_iso_signal, CUDA/runmcmc_kernels.cu:67-95):
//   kind 0: exp(-d*u)                        model 1
//   kind 1: (1 + w)^(-d^2/dstd^2), w = (dstd^2/d) u   model 2 (gamma diffusivities);
//           below d_std 1e-5 the reference switches to exp(-d*u), which the
//           same expression evaluates with c0 = 2^-64, c1 = -d 2^64: then
//           1 + w rounds to 1, lg2 of it is 0, and the recovered rounding
//           error is w = c0 u exactly (a power-of-two scaling), so the
//           exponent is c1 w / ln2 = -d u / ln2 with the roundings of the
//           exponential form.  Rows on either side of the switch share one
//           code path (a wave whose rows took different forms would run both).
//   kind 2: exp(-3 d invR ((1-R) b a + R b))  model 3 zeppelin, u = b a
// Kind 1 in FP32: u1 = fl(1 + w) keeps w only to 2^-24 absolute, and the
// exponent d^2/dstd^2 multiplies that error (1e-3 relative signal error at
// dstd/d ~ 0.01; the reference evaluates these signals in double).  The
// rounding error e = (1 + w) - u1 is recovered with one FMA and
// lg2(1 + w) = lg2(u1) + lg2(1 + e/u1) ~ lg2(u1) + e/ln2; the dropped 1/u1
// factor leaves an exponent error below d u 2^-24 for every w (e <= half an
// ulp of u1, and (d^2/dstd^2) w = d u).
struct SigC {
  int kind;
  float c0, c1;
};
FDT_DEV SigC gamma_coef(float d, float dstd) {
  SigC c;
  c.kind = 1;
  if (dstd < 1e-5f) { c.c0 = 0x1p-64f; c.c1 = -d * 0x1p64f; }
  else { const float s2 = dstd * dstd; c.c0 = s2 / d; c.c1 = -(d * d / s2); }
  return c;
}
FDT_DEV SigC aniso_coef(int model, float d, float dstd, float R) {
  SigC c;
  if (model == 1) { c.kind = 0; c.c0 = -d * L2E; c.c1 = 0.0f; }
  else if (model == 2) c = gamma_coef(d, dstd);
  else {
    float iR = 1.0f / (2.0f * R + 1.0f);
    float k = -3.0f * d * iR * L2E;
    c.kind = 2; c.c0 = k * (1.0f - R); c.c1 = k * R;
  }
  return c;
}
FDT_DEV SigC iso_coef(int model, float d, float dstd) {
  SigC c;
  if (model == 1) { c.kind = 0; c.c0 = -d * L2E; c.c1 = 0.0f; }
  else c = gamma_coef(d, dstd);
  return c;
}
template <int SK>
FDT_DEV float sig(const SigC& c, float u, float b) {
  if constexpr (SK == 0) return ex2(c.c0 * u);
  else if constexpr (SK == 1) {
    const float u1 = fmaf(c.c0, u, 1.0f);
    const float e = fmaf(c.c0, u, 1.0f - u1);  // (1 + w) - u1 (1 - u1 is exact)
    return ex2(c.c1 * fmaf(e, L2E, lg2(u1)));
  }
  else return ex2(fmaf(c.c0, u, c.c1 * b));
}
template <class F>
FDT_DEV void by_kind(int kind, F&& f) {
  if (kind == 0) f(std::integral_constant<int, 0>{});
  else if (kind == 1) f(std::integral_constant<int, 1>{});
  else f(std::integral_constant<int, 2>{});
}

// log I0 by Abramowitz-Stegun 9.8.1/9.8.2 (fibre.h:29-48, CUDA/runmcmc_kernels.cu:126-145)
FDT_DEV float logI0f(float x) {
  float b = fabsf(x);
  float a = x * (1.0f / 3.75f);
  a *= a;
  float ys = 1.0f + a * (3.5156229f + a * (3.0899424f + a * (1.2067492f + a * (0.2659732f + a * (0.0360768f + a * 0.0045813f)))));
  float Ls = flog(ys);
  float c = 3.75f * __builtin_amdgcn_rcpf(b);
  float pl = 0.39894228f + c * (0.01328592f + c * (0.00225319f + c * (-0.00157565f + c * (0.00916281f + c * (-0.02057706f + c * (0.02635537f + c * (-0.01647633f + c * 0.00392377f)))))));
  // one logarithm of the selected branch: log ys, or b + log(pl / sqrt b)
  const bool sm = b < 3.75f;
  return flog(sm ? ys : pl * __builtin_amdgcn_rsqf(b)) + (sm ? 0.0f : b);
}
// Rician log-likelihood term of one volume (fibre.h:949-953),
//   log y - tau (y^2 + p^2) / 2 + log I0(tau p y),
// without the part that does not depend on the model signal p: log y and
// -tau y^2 / 2 cancel from every energy difference except the tau move's,
// which adds -(tau' - tau) sum y^2 / 2 itself; the energy of a state adds
// both back (sum log y - tau sum y^2 / 2)
FDT_DEV float ell_of(float y, float p, float tau) {
  const float tp = tau * p;
  return fmaf(-0.5f * tp, p, logI0f(tp * y));
}

// Gaussian likelihood change when every residual moves by -c*q_t:
// sum (r - c q)^2 - r^2 = c (c sum q^2 - 2 sum r q); two FMAs per volume and
// two row sums
struct Quad {
  float A = 0.0f, B = 0.0f;
  FDT_DEV void add(float r, float q) { A = fmaf(r, q, A); B = fmaf(q, q, B); }
  FDT_DEV float delta(float c) {
    A = row_sum16(A);
    B = row_sum16(B);
    return c * fmaf(c, B, -2.0f * A);
  }
};
// Compiler-only memory barrier: keeps the acquisition table in LDS (re-read
// per proposal) instead of letting the loop-invariant loads be hoisted into
// ~4*VPL VGPRs for the whole chain.
FDT_DEV void reload_acq() { asm volatile("" ::: "memory"); }
FDT_DEV int acc1(int c) { return c + 1; }
FDT_DEV int rej1(int c) { return c + (1 << 16); }


// One chain per 16-lane row.  Volume state lives in registers per lane;
// everything that belongs to a proposal slot -- the parameter value, its
// proposal width, its acceptance counters (acc + rej << 16) and its current
// prior energy -- lives in the lane of that slot (sval/sprop/scnt/sprior) and
// is read by a row broadcast.  Slots: f0 0, tau 1, d 2, d_std 3, R 4, S0 5,
// theta_k 6+3k, phi_k 7+3k, f_k 8+3k (the theta prior sits in the theta slot).
// GN (--gradnonlin): the voxel sees gradient g through its deviation M = I+L
// as direction Mg/|Mg| and b value |Mg|^2 b, so b'(g'.x)^2 = b (g.M^T x)^2:
// the stick signals use the shared acquisition with the fibre directions
// mapped by M^T, and the isotropic signal (and the model-3 R term) the
// voxel's own b values, kept per row in LDS.
template <int K, int VPL, bool RIC, bool GN = false>
struct Chain {
  // lane-resident volume state, volume i = rl + 16 t.  One per-volume
  // array lives in the wave's LDS block `vol` instead of registers (18
  // VGPRs at 18 volumes per lane, the difference between spilling and not):
  // the isotropic signal of the Gaussian chain, the data of the Rician one.
  // The --gradnonlin kernels keep both in registers (their LDS holds the
  // rows' b values).
  static constexpr bool ISO_LDS = !GN && !RIC, Y_LDS = !GN && RIC;
  float r[Y_LDS ? 1 : VPL];  // Gaussian: y - S0*M; Rician: y
  float M[VPL];
  float iso_r[ISO_LDS ? 1 : VPL];
  float s[K][VPL];
  float* vol;   // wave's LDS block: vol[t*64]
  float* cand;  // candidate stick signals of the diffusivity proposals, wave's LDS block: cand[(k*VPL+t)*64]
  float ell[RIC ? VPL : 1];
  // isotropic signal per b-shell (acquisitions with <= 16 distinct b values):
  // sh[i] is the shell of volume i, shb[l] its b value, isot the row's table
  const int* sh;
  const float* shb;
  float* isot;
  int nshell;
  float sval, sprop, sprior;
  int scnt;
  // sin and cos of the slot's angle (theta_k / phi_k slots), kept with the
  // value so that fibre directions are row broadcasts, not re-evaluated.
  // The d slot holds the stick-signal coefficients of the current state
  // there instead (c0, c1), set whenever d, d_std or R moves, so that the
  // angle proposals broadcast them instead of dividing again.
  float ssin, scos;
  double like;  // Gaussian: SSR; Rician: sum of ell
  int nv;       // slots of this lane that hold a real volume (rl + 16 t < N)
  int rl;
  // slot 0 holds b = 0 volumes only (see stage_acq): the moves that leave
  // their signals at 1 skip it (18-volume tiles only)
  bool z0;
  FDT_DEV bool skip(int t) const { return ZERO_SLOT<VPL> && t == 0 && z0; }
  // options
  int model;
  bool incf0, ardf0, can_ard;
  bool lam[K];
  float fudge, halfN, Nf, Rmean, Rstd, Rfudge;
  const float4* acq;
  const float* bp;  // GN: the voxel's b values, bp[rl + 16 t] (LDS)
  const float* gd;  // GN: the voxel's 9 deviation values L(i,j) = gd[3j+i] (LDS)

  // b of slot t for the isotropic signal and the model-3 R term
  FDT_DEV float bv(int t, const float4& g) const {
    if constexpr (GN) return bp[rl + 16 * t];
    else return g.w;
  }
  // fibre direction as the voxel's gradients see it: x -> M^T x
  FDT_DEV void xform(float& x0, float& x1, float& x2) const {
    if constexpr (GN) {
      float y0 = x0 + fmaf(gd[0], x0, fmaf(gd[1], x1, gd[2] * x2));
      float y1 = x1 + fmaf(gd[3], x0, fmaf(gd[4], x1, gd[5] * x2));
      float y2 = x2 + fmaf(gd[6], x0, fmaf(gd[7], x1, gd[8] * x2));
      x0 = y0; x1 = y1; x2 = y2;
    }
  }

  static constexpr int TH(int k) { return SL_FIB + 3 * k; }
  static constexpr int PH(int k) { return SL_FIB + 3 * k + 1; }
  static constexpr int FK(int k) { return SL_FIB + 3 * k + 2; }
  template <int SL> FDT_DEV float v() const { return row_bcast<SL>(sval); }
  template <int SL> FDT_DEV float width() const { return row_bcast<SL>(sprop); }
  template <int SL> FDT_DEV float prior() const { return row_bcast<SL>(sprior); }
  template <int SL> FDT_DEV void reject() { if (rl == SL) scnt += 1 << 16; }
  template <int SL> FDT_DEV void accept(float val, float prn) {
    if (rl == SL) { sval = val; scnt += 1; sprior = prn; }
  }

  FDT_DEV float& cnd(int k, int t) const { return cand[(k * VPL + t) * 64]; }
  FDT_DEV float& isov(int t) {
    if constexpr (ISO_LDS) return vol[t * 64];
    else return iso_r[t];
  }
  FDT_DEV float isov(int t) const {
    if constexpr (ISO_LDS) return vol[t * 64];
    else return iso_r[t];
  }
  // the Rician chain's data of slot t
  FDT_DEV float yv(int t) const {
    if constexpr (Y_LDS) return vol[t * 64];
    else return r[t];
  }
  // stick-signal coefficients of the current state (the kind follows the model)
  FDT_DEV SigC aniso_cached() const {
    SigC c;
    c.kind = model == 1 ? 0 : (model == 2 ? 1 : 2);
    c.c0 = row_bcast<SL_D>(ssin);
    c.c1 = row_bcast<SL_D>(scos);
    return c;
  }
  FDT_DEV void set_aniso_cache(const SigC& c) {
    if (rl == SL_D) { ssin = c.c0; scos = c.c1; }
  }
  // unit directions of the row's fibres (from the theta / phi slots)
  struct Dirs {
    float x[K][3];
  };
  FDT_DEV Dirs dirs() const {
    Dirs D;
    sfor<K>([&](auto kc) {
      constexpr int k = kc;
      const float st = row_bcast<TH(k)>(ssin), ct = row_bcast<TH(k)>(scos);
      const float sp = row_bcast<PH(k)>(ssin), cp = row_bcast<PH(k)>(scos);
      D.x[k][0] = st * cp; D.x[k][1] = st * sp; D.x[k][2] = ct;
      xform(D.x[k][0], D.x[k][1], D.x[k][2]);
    });
    return D;
  }
  // Contribution of slot t: zero where the slot is padding (volume >= N).
  // Padding slots see g = 0 and b = 0 in every acquisition table (shared,
  // per-voxel, b-shell), so their isotropic and stick signals are exactly 1
  // for every parameter value.  The Gaussian chain keeps their residual at
  // exactly 0 (masked where it is formed and where S0 or the mixture moves),
  // so the f0, f and angle proposals, whose per-volume change is then exactly
  // 0 there, need no mask; the S0 and diffusivity proposals, where the
  // mixture at a padding slot moves by rounding, and every Rician term do.
  FDT_DEV float mask(int t, float x) const { return t < nv ? x : 0.0f; }
  // the fractions of the row: fk[k] and f0, their sum
  struct Fr {
    float fk[K];
    float f0, sum;
  };
  FDT_DEV Fr fractions() const {
    Fr q;
    q.f0 = v<SL_F0>();
    q.sum = q.f0;
    sfor<K>([&](auto kc) { q.fk[kc] = v<FK(kc)>(); q.sum += q.fk[kc]; });
    return q;
  }
  // energy change of the likelihood from a row-summed change `dl`
  FDT_DEV float dlik(float dl) const {
    if constexpr (RIC) return -dl;
    else {
      // log1p(x): log of u = fl(1 + x) plus the rounding error of u,
      // x - (u - 1) (exact for x in [-1/2, 1]; outside, its error is below an
      // ulp of u against |log u| > 0.4); the dropped factor 1/u leaves an
      // error below 2^-24 |x| (the same device as the gamma-form signal)
      const float x = dl * __builtin_amdgcn_rcpf((float)like);
      const float u = 1.0f + x;
      return halfN * fmaf(lg2(u), LN2, x - (u - 1.0f));
    }
  }
  // theta prior -log(|sin t| / 2) from st = sin t (fibre.h:283-290)
  FDT_DEV static float theta_prior(float t, float st) {
    return t == 0.0f ? 0.0f : -flog(fabsf(st) * 0.5f);
  }
  // R prior (fibre.h:816-843) at diffusivity d; false outside its support
  FDT_DEV bool r_prior(float Rv, float d, float& pr) const {
    float upper = 2.0f * Rmean, lower = Rmean - 2.0f * Rstd;
    if (Rmean > 0.5f) upper = 1.0f;
    if (lower < 0.0f) lower = 1e-8f;
    if (Rfudge > 0.0f && d > UPPERDIFF_F / 2.0f) {
      if (Rv < 1e-8f || Rv > upper) return false;
      pr = Rfudge * flog(Rv);
      return true;
    }
    if (Rv <= lower || Rv > upper) return false;
    pr = (Rv - Rmean) * (Rv - Rmean) / (Rstd * Rstd);
    return true;
  }
  // mixture of slot t from the stored signals
  FDT_DEV float mix(int t, const Fr& q) const {
    float m = fmaf(1.0f - q.sum, isov(t), q.f0);
#pragma unroll
    for (int k = 0; k < K; ++k) m = fmaf(q.fk[k], s[k][t], m);
    return m;
  }

  // rebuild M and the residual (or the Rician terms) from the stored signals
  FDT_DEV void rebuild(const float* y) {
    const Fr q = fractions();
    const float S0 = v<SL_S0>(), tau = v<SL_TAU>();
    double ss = 0.0;
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      float m = mix(t, q);
      M[t] = m;
      if constexpr (RIC) {
        ell[t] = ell_of(yv(t), S0 * m, tau);
        ss += (double)mask(t, ell[t]);
      } else {
        r[t] = mask(t, y[t] - S0 * m);
        ss += (double)(r[t] * r[t]);
      }
    }
    like = row_sum16_d(ss);
  }

  // ---------------------------------------------------------- proposals
  // n: standard normal draw, lu: log of the uniform draw of the slot.
  // f0 (CUDA/runmcmc_kernels.cu:842-888)
  FDT_DEV void prop_f0(float n, float lu) {
    const float f0 = v<SL_F0>();
    const float f0n = f0 + n * width<SL_F0>();
    const Fr q = fractions();
    const float fsn = q.sum - f0 + f0n;
    if (f0n <= 0.0f || f0n >= 1.0f || fsn > 1.0f) { reject<SL_F0>(); return; }
    const float prn = ardf0 ? flog(f0n) : 0.0f;
    const float df = f0n - f0, S0 = v<SL_S0>(), tau = v<SL_TAU>();
    float a = 0.0f, en[RIC ? VPL : 1];
    Quad qd;
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      if (skip(t)) continue;
      if constexpr (RIC) {
        en[t] = ell_of(yv(t), S0 * fmaf(df, 1.0f - isov(t), M[t]), tau);
        a += mask(t, en[t] - ell[t]);
      } else qd.add(r[t], 1.0f - isov(t));
    }
    a = RIC ? row_sum16(a) : qd.delta(S0 * df);
    if (dlik(a) + (prn - prior<SL_F0>()) < -lu) {
      const float c = S0 * df;
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        if (skip(t)) continue;
        const float q = 1.0f - isov(t);
        M[t] = fmaf(df, q, M[t]);
        if constexpr (RIC) ell[t] = en[t]; else r[t] = fmaf(-c, q, r[t]);
      }
      like += (double)a;
      accept<SL_F0>(f0n, prn);
    } else reject<SL_F0>();
  }
  // tau (:890-922), Rician only
  FDT_DEV void prop_tau(float n, float lu) {
    if constexpr (RIC) {
      const float tau = v<SL_TAU>();
      const float tn = tau + n * width<SL_TAU>();
      if (tn <= 0.0f) { reject<SL_TAU>(); return; }
      const float S0 = v<SL_S0>();
      float a = 0.0f, yy = 0.0f, en[VPL];
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        en[t] = ell_of(yv(t), S0 * M[t], tn);
        const float y = yv(t);
        a += mask(t, en[t] - ell[t]);
        yy += mask(t, y * y);
      }
      a = row_sum16(a);  // change of the stored terms (the running sum's)
      // the -tau y^2 / 2 part of the full terms moves with tau too
      float dE = -Nf * (flog(tn) - flog(tau)) - fmaf(-0.5f * (tn - tau), row_sum16(yy), a);
      if (dE < -lu) {
#pragma unroll
        for (int t = 0; t < VPL; ++t) ell[t] = en[t];
        like += (double)a;
        accept<SL_TAU>(tn, 0.0f);
      } else reject<SL_TAU>();
    }
  }
  // S0 (:1125-1155)
  FDT_DEV void prop_S0(float n, float lu) {
    const float S0 = v<SL_S0>();
    const float Sn = S0 + n * width<SL_S0>();
    if (Sn < 0.0f) { reject<SL_S0>(); return; }
    const float dS = Sn - S0, tau = v<SL_TAU>();
    float a = 0.0f, en[RIC ? VPL : 1];
    Quad qd;
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      if constexpr (RIC) { en[t] = ell_of(yv(t), Sn * M[t], tau); a += mask(t, en[t] - ell[t]); }
      else qd.add(r[t], mask(t, M[t]));
    }
    a = RIC ? row_sum16(a) : qd.delta(dS);
    if (dlik(a) < -lu) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        if constexpr (RIC) ell[t] = en[t]; else r[t] -= dS * mask(t, M[t]);
      }
      like += (double)a;
      accept<SL_S0>(Sn, 0.0f);
    } else reject<SL_S0>();
  }
  // candidate isotropic signal of every shell into the row's table (lane l: shell l)
  template <int SI>
  FDT_DEV void fill_iso(const SigC& ci) const {
    float b = shb[rl];
    isot[rl] = sig<SI>(ci, b, b);
  }
  // candidate isotropic signal of slot t: from the shell table (SH) or evaluated
  template <int SI, bool SH>
  FDT_DEV float iso_new(int t, const SigC& ci, float b) const {
    if constexpr (SH) return isot[sh[rl + 16 * t]];
    else return sig<SI>(ci, b, b);
  }
  // candidate mixture of slot t when the stick signals (NA) and/or the
  // isotropic signal (NI) are re-evaluated with coefficients ca / ci
  // The candidate stick signals are evaluated from b*(g.x_k)^2 and parked in
  // LDS, so that an accepted move copies them instead of re-evaluating.
  template <int SA, int SI, bool NA, bool NI, bool SH>
  FDT_DEV float mix_new(int t, const Fr& q, const SigC& ca, const SigC& ci, const Dirs& D) const {
    const float4 g = acq[rl + 16 * t];
    float isn = NI ? iso_new<SI, SH>(t, ci, bv(t, g)) : isov(t);
    float m = fmaf(1.0f - q.sum, isn, q.f0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      float sk = s[k][t];
      if constexpr (NA) {
        float dp = fmaf(g.x, D.x[k][0], fmaf(g.y, D.x[k][1], g.z * D.x[k][2]));
        sk = sig<SA>(ca, g.w * dp * dp, bv(t, g));
        cnd(k, t) = sk;
      }
      m = fmaf(q.fk[k], sk, m);
    }
    return m;
  }
  template <int SA, int SI, bool NA, bool NI, bool SH>
  FDT_DEV void store_new(int t, const SigC& ca, const SigC& ci) {
    float b = bv(t, acq[rl + 16 * t]);
    if (NI) isov(t) = iso_new<SI, SH>(t, ci, b);
    if (NA) {
#pragma unroll
      for (int k = 0; k < K; ++k) s[k][t] = cnd(k, t);
    }
  }
  // dispatch over the coefficient forms that occur: model 1 (0,0); model 2
  // (1,1); model 3 sticks 2 with iso 1
  // run f(SH) with SH = true when the isotropic signal changes and the
  // acquisition has a shell table (the loop over slots is then the table
  // variant only; no per-slot branch)
  template <bool NI, class F>
  FDT_DEV void by_shells(F&& f) const {
    if constexpr (NI && !GN) {
      if (nshell) f(std::true_type{});
      else f(std::false_type{});
    } else f(std::false_type{});
  }
  template <class F>
  FDT_DEV void by_forms(const SigC& ca, const SigC& ci, bool na, bool ni, F&& f) const {
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using T_ = std::true_type;
    using F_ = std::false_type;
    // the isotropic form is kind 0 for model 1 and kind 1 otherwise
    if (na && ni) {
      if (ca.kind == 0) f(I0{}, I0{}, T_{}, T_{});
      else if (ca.kind == 1) f(I1{}, I1{}, T_{}, T_{});
      else f(I2{}, I1{}, T_{}, T_{});
    } else if (ni) {
      f(I0{}, I1{}, F_{}, T_{});
    } else {
      f(I2{}, I0{}, T_{}, F_{});
    }
  }
  // all stick signals and/or iso change: d (:924-985), d_std (:987-1051), R
  // (:1053-1122).  The row sum of the likelihood change is accumulated on the
  // fly; the candidate stick signals go to the wave's LDS block (mix_new) and
  // an accepted move copies them back (store_new), re-evaluating only the
  // isotropic signal, so the live registers stay the chain state.
  template <int SL>
  FDT_DEV void prop_global(float n, float lu) {
    const float d = v<SL_D>(), dstd = v<SL_DSTD>(), Rr = v<SL_R>();
    float dn = d, dsn = dstd, Rn = Rr, prn = 0.0f, xn;
    bool new_aniso = true, new_iso = true;
    if constexpr (SL == SL_D) {
      dn = xn = d + n * width<SL>();
      if (dn < 0.0f || dn > UPPERDIFF_F) { reject<SL>(); return; }
      prn = model == 3 ? (1.0f - 3.0f) * flog(dn) + 4000.0f * dn : 0.0f;
    } else if constexpr (SL == SL_DSTD) {
      dsn = xn = dstd + n * width<SL>();
      float up = model == 3 ? 0.004f : 0.01f;
      if (dsn <= 0.0f || dsn > up) { reject<SL>(); return; }
      prn = flog(dsn);
      new_aniso = (model == 2);
    } else {
      Rn = xn = Rr + n * width<SL>();
      if (!r_prior(Rn, d, prn)) { reject<SL>(); return; }
      new_iso = false;
    }
    const SigC ca = aniso_coef(model, dn, dsn, Rn), ci = iso_coef(model, dn, dsn);
    const Fr q = fractions();
    const Dirs D = new_aniso ? dirs() : Dirs{};
    const float S0 = v<SL_S0>(), tau = v<SL_TAU>();
    reload_acq();
    float a = 0.0f;
    Quad qd;
    by_forms(ca, ci, new_aniso, new_iso, [&](auto SA, auto SI, auto NA, auto NI) {
      by_shells<decltype(NI)::value>([&](auto SH) {
        if constexpr (decltype(SH)::value) fill_iso<decltype(SI)::value>(ci);
#pragma unroll
        for (int t = 0; t < VPL; ++t) {
          if (skip(t)) continue;
          float mn = mix_new<decltype(SA)::value, decltype(SI)::value, decltype(NA)::value, decltype(NI)::value,
                             decltype(SH)::value>(t, q, ca, ci, D);
          if constexpr (RIC) a += mask(t, ell_of(yv(t), S0 * mn, tau) - ell[t]);
          else qd.add(r[t], mask(t, mn - M[t]));
          // the gamma form has several temporaries per volume and signal: keep
          // the scheduler from overlapping many volumes of it (register spills)
          if constexpr (decltype(SA)::value == 1 || decltype(SI)::value == 1) reload_acq();
        }
      });
    });
    a = RIC ? row_sum16(a) : qd.delta(S0);
    if (dlik(a) + (prn - prior<SL>()) < -lu) {
      // store the new signals, then form the mixture from them with the
      // same expression as the candidate (bit-identical to it)
      reload_acq();
      by_forms(ca, ci, new_aniso, new_iso, [&](auto SA, auto SI, auto NA, auto NI) {
        by_shells<decltype(NI)::value>([&](auto SH) {
#pragma unroll
          for (int t = 0; t < VPL; ++t) {
            if (skip(t)) continue;
            store_new<decltype(SA)::value, decltype(SI)::value, decltype(NA)::value, decltype(NI)::value,
                      decltype(SH)::value>(t, ca, ci);
            float mn = mix(t, q);
            if constexpr (RIC) ell[t] = ell_of(yv(t), S0 * mn, tau); else r[t] -= S0 * mask(t, mn - M[t]);
            M[t] = mn;
          }
        });
      });
      like += (double)a;
      accept<SL>(xn, prn);
      set_aniso_cache(ca);
    } else reject<SL>();
  }
  // theta_k / phi_k (:1157-1265); the theta prior lives in the theta slot
  template <int k, bool IS_TH>
  FDT_DEV void prop_angle(float n, float lu) {
    constexpr int SL = IS_TH ? TH(k) : PH(k);
    constexpr int SO = IS_TH ? PH(k) : TH(k);  // the other angle's slot: its sin/cos are cached there
    const float xn = v<SL>() + n * width<SL>();
    float sn_, cn_;
    sincos_full(xn, sn_, cn_);
    const float so = row_bcast<SO>(ssin), co = row_bcast<SO>(scos);
    float prn = 0.0f, prold = 0.0f;
    if constexpr (IS_TH) { prn = theta_prior(xn, sn_); prold = prior<SL>(); }
    const float st = IS_TH ? sn_ : so, ct = IS_TH ? cn_ : co;
    const float sp = IS_TH ? so : sn_, cp = IS_TH ? co : cn_;
    float x0 = st * cp, x1 = st * sp, x2 = ct;
    xform(x0, x1, x2);
    const SigC ca = aniso_cached();
    const float fk = v<FK(k)>(), S0 = v<SL_S0>(), tau = v<SL_TAU>();
    reload_acq();
    float sn[VPL], en[RIC ? VPL : 1];
    float a = 0.0f;
    by_kind(ca.kind, [&](auto SK) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        if (skip(t)) continue;
        float4 g = acq[rl + 16 * t];
        float dp = fmaf(g.x, x0, fmaf(g.y, x1, g.z * x2));
        sn[t] = sig<decltype(SK)::value>(ca, g.w * dp * dp, bv(t, g));
      }
    });
    Quad qd;
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      if (skip(t)) continue;
      if constexpr (RIC) {
        en[t] = ell_of(yv(t), S0 * fmaf(fk, sn[t] - s[k][t], M[t]), tau);
        a += mask(t, en[t] - ell[t]);
      } else qd.add(r[t], sn[t] - s[k][t]);
    }
    a = RIC ? row_sum16(a) : qd.delta(S0 * fk);
    if (dlik(a) + (prn - prold) < -lu) {
      reload_acq();
      const float c = S0 * fk;
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        if (skip(t)) continue;
        const float q = sn[t] - s[k][t];
        M[t] = fmaf(fk, q, M[t]);
        if constexpr (RIC) ell[t] = en[t]; else r[t] = fmaf(-c, q, r[t]);
        s[k][t] = sn[t];
      }
      like += (double)a;
      if (rl == SL) { ssin = sn_; scos = cn_; }
      if constexpr (IS_TH) accept<SL>(xn, prn);
      else if (rl == SL) { sval = xn; scnt += 1; }
    } else reject<SL>();
  }
  // f_k (:1267-1322)
  template <int k>
  FDT_DEV void prop_f(float n, float lu) {
    constexpr int SL = FK(k);
    const Fr q = fractions();
    const float fn = q.fk[k] + n * width<SL>();
    const float fsn = q.sum - q.fk[k] + fn;
    if (fn <= 0.0f || fn >= 1.0f || fsn > 1.0f) { reject<SL>(); return; }
    const float prn = (can_ard && lam[k]) ? fudge * flog(fn) : 0.0f;
    const float df = fn - q.fk[k], S0 = v<SL_S0>(), tau = v<SL_TAU>();
    float a = 0.0f, en[RIC ? VPL : 1];
    Quad qd;
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      if (skip(t)) continue;
      if constexpr (RIC) {
        en[t] = ell_of(yv(t), S0 * fmaf(df, s[k][t] - isov(t), M[t]), tau);
        a += mask(t, en[t] - ell[t]);
      } else qd.add(r[t], s[k][t] - isov(t));
    }
    a = RIC ? row_sum16(a) : qd.delta(S0 * df);
    if (dlik(a) + (prn - prior<SL>()) < -lu) {
      const float c = S0 * df;
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        if (skip(t)) continue;
        const float q = s[k][t] - isov(t);
        M[t] = fmaf(df, q, M[t]);
        if constexpr (RIC) ell[t] = en[t]; else r[t] = fmaf(-c, q, r[t]);
      }
      like += (double)a;
      accept<SL>(fn, prn);
    } else reject<SL>();
  }

  // proposal-width adaptation every upe jumps (:1341-1395), all slots of the
  // row in parallel; slots of parameters that are not sampled keep counters
  // 0/0 and are never read
  FDT_DEV void update_proposals() {
    float acc = (float)(scnt & 0xFFFF), rej = (float)(scnt >> 16);
    sprop *= sqrtf((acc + 1.0f) / (rej + 1.0f));
    sprop = fminf(sprop, MAXPROP_F);
    scnt = 0;
  }
};

// LDS of a chain workgroup: the acquisition table, candidate stick signals per wave, the
// b-shell index of each volume and the shells' b values, the rows' isotropic
// shell tables
struct ChainLds {
  float4* acq;
  float* cand;
  int* sh;     // null without a shell table (--gradnonlin)
  float* shb;
  float* iso;
  float* bp;   // --gradnonlin: [16 rows][16 VPL] b values of the rows' voxels
  float* gd;   // --gradnonlin: [16 rows][9] deviation values
  float* vol;  // [4 waves][VPL][64] per-volume array of the chains (null for --gradnonlin)
  int* perm;   // [16 VPL] volume held by each slot position (null: identity)
  int* zc;     // [ZC_CHUNKS] b = 0 volumes per 64-volume chunk of the acquisition
};

// Volumes with b = 0 have stick and isotropic signals of exactly 1 in every
// model, so an angle, fraction or f0 move leaves their mixture, residual and
// likelihood term unchanged, and a diffusivity move changes their mixture by
// rounding only.  When the acquisition has at least 16 of them, the staging
// puts them first (a stable partition: b = 0 volumes in order, then the others
// in order, padding last), so that slot 0 of every row holds b = 0 volumes
// only, and those proposals skip slot 0.  The likelihood is the same sum over
// the same volumes; with fewer than 16 the order is the acquisition's.  Only
// the 18-volume tile does this (ZERO_SLOT).  With --gradnonlin the per-voxel
// b values are b |(I+L) g|^2, zero exactly where the shared table's are, and
// the rows read them in the staged order.
constexpr int ZC_CHUNKS = 8;  // 64-volume chunks of the largest table (16 x 32)

// stage the acquisition of `a` (16*VPL volumes, zero padded) into LDS; the
// caller synchronises the workgroup afterwards.  The block size is a multiple
// of 64.
template <int VPL>
FDT_DEV void stage_acq(const McmcArgs& a, const ChainLds& L) {
  constexpr int NV = 16 * VPL, NCH = (NV + 63) / 64;
  static_assert(NCH <= ZC_CHUNKS, "acquisition table larger than the chunk counts");
  if (!ZERO_SLOT<VPL> || !L.perm) {
    for (int i = threadIdx.x; i < NV; i += blockDim.x) {
      L.acq[i] = a.acq[i];
      if (L.sh) L.sh[i] = a.nshell ? a.shell[i] : 0;
    }
  } else {
    // b = 0 count per 64-volume chunk (one wave per chunk), then each volume's
    // position from the counts of the chunks before it and a ballot
    const int lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int ch = w; ch < NCH; ch += nw) {
      const int i = 64 * ch + lane;
      bool z = false;
      if (i < a.N) z = a.acq[i].w == 0.0f;
      const uint64_t m = __builtin_amdgcn_ballot_w64(z);
      if (lane == 0) L.zc[ch] = __builtin_popcountll(m);
    }
    __syncthreads();
    int nz = 0;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) nz += L.zc[ch];
    const bool part = nz >= 16;
    for (int ch = w; ch < NCH; ch += nw) {
      const int i = 64 * ch + lane;
      const float4 g = i < NV ? a.acq[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const bool z = i < a.N && g.w == 0.0f;
      const uint64_t m = __builtin_amdgcn_ballot_w64(z);
      int before = __builtin_popcountll(m & ((1ull << lane) - 1ull));  // b = 0 volumes below i
#pragma unroll
      for (int c = 0; c < NCH; ++c) before += c < ch ? L.zc[c] : 0;
      const int pos = (!part || i >= a.N) ? i : (z ? before : nz + i - before);
      if (i < NV) {
        L.perm[pos] = i;
        L.acq[pos] = g;
        if (L.sh) L.sh[pos] = a.nshell ? a.shell[i] : 0;
      }
    }
  }
  if (threadIdx.x < 16) L.shb[threadIdx.x] = a.nshell ? a.shell_b[threadIdx.x] : 0.0f;
}
// slot 0 of every row holds b = 0 volumes only (after the caller's
// synchronisation); uniform over the workgroup
template <int VPL>
FDT_DEV bool zero_slot0(const ChainLds& L) {
  constexpr int NCH = (16 * VPL + 63) / 64;
  if (!ZERO_SLOT<VPL> || !L.perm) return false;
  int nz = 0;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) nz += L.zc[ch];
  return __builtin_amdgcn_readfirstlane(nz) >= 16;
}

// The whole chain of voxel `vox` on this thread's 16-lane row: initialisation
// (init_Fibres_Multifibres_kernel, CUDA/runmcmc_kernels.cu:193-505), nburn +
// njumps sweeps with recording and proposal adaptation (runmcmc_kernel,
// :507-1474, run twice by CUDA/runmcmc.cu:108-316).  GN (--gradnonlin): the
// chain reads its voxel's own corrected acquisition from a.acqv (the reference
// indexes per-voxel bvals/alpha/beta with posBV, :266, :711).
template <int K, int VPL, bool RIC, bool GN = false>
FDT_DEV void run_chain(const McmcArgs& a, int vox, const ChainLds& L) {
  using C = Chain<K, VPL, RIC, GN>;
  const int lane = __lane_id();
  const int rl = lane & 15;
  const float4* sacq = L.acq;
  C c;
  c.rl = rl;
  c.acq = sacq;
  c.cand = L.cand + (threadIdx.x >> 6) * (K * VPL * 64) + lane;
  c.vol = L.vol ? L.vol + (threadIdx.x >> 6) * (VPL * 64) + lane : nullptr;
  c.sh = L.sh;
  c.shb = L.shb;
  c.isot = GN ? L.iso : L.iso + (threadIdx.x >> 4) * 16;
  c.nshell = GN ? 0 : a.nshell;
  if constexpr (GN) {
    // the row's b values (from the corrected acquisition table) and deviations
    float* bp = L.bp + (threadIdx.x >> 4) * (16 * VPL);
    float* gd = L.gd + (threadIdx.x >> 4) * 9;
    const float4* av = a.acqv + (size_t)vox * a.acqv_stride;
    const int* pm = ZERO_SLOT<VPL> ? L.perm : nullptr;  // the staged order of the shared table
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      const int i = rl + 16 * t;
      bp[i] = av[pm ? pm[i] : i].w;
    }
    if (rl < 9) gd[rl] = a.gdev[(size_t)vox * 9 + rl];
    c.bp = bp;
    c.gd = gd;
  }
  c.model = a.model;
  c.incf0 = a.f0 != 0;
  c.ardf0 = a.ardf0 != 0;
  c.can_ard = a.no_ard == 0;
  c.fudge = a.fudge;
  c.halfN = 0.5f * (float)a.N;
  c.Nf = (float)a.N;
  c.Rmean = a.Rmean; c.Rstd = a.Rstd; c.Rfudge = a.Rfudge;
  c.nv = a.N > rl ? min(VPL, (a.N - rl + 15) >> 4) : 0;
#pragma unroll
  for (int k = 0; k < K; ++k) c.lam[k] = (k == 0) ? (a.all_ard != 0) : (a.no_ard == 0);

  // ---- data (with the Rician non-positive replacement, CUDA/xfibres_gpu.cu:350-367)
  float y[VPL];
  const float* yrow = a.data + (size_t)vox * a.N;
  const int* perm = ZERO_SLOT<VPL> ? L.perm : nullptr;  // volume of each slot position (null: identity)
  c.z0 = zero_slot0<VPL>(L);
#pragma unroll
  for (int t = 0; t < VPL; ++t) {
    int i = rl + 16 * t;
    y[t] = i < a.N ? yrow[perm ? perm[i] : i] : 0.0f;
  }
  if (RIC) {
    float mn = 3.0e38f, mx = -3.0e38f, mnp = 3.0e38f;
#pragma unroll
    for (int t = 0; t < VPL; ++t)
      if (rl + 16 * t < a.N) { mn = fminf(mn, y[t]); mx = fmaxf(mx, y[t]); if (y[t] > 0.0f) mnp = fminf(mnp, y[t]); }
    mn = row_min16(mn); mx = row_max16(mx); mnp = row_min16(mnp);
    bool san = (mn <= 0.0f && mx > 0.0f);
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      if (san && y[t] <= 0.0f) y[t] = mnp;
      if (rl + 16 * t >= a.N) y[t] = 1.0f;
    }
  }

  // ---- chain initialisation (init_Fibres_Multifibres_kernel, :265-504):
  // every lane loads the parameter of its slot and evaluates its prior
  const float* P = a.params + (size_t)vox * a.np;
  const int add = a.model >= 2 ? 1 : 0;
  float d = P[1];
  if (d < 0.0f || d > UPPERDIFF_F) d = 2e-3f;
  float dstd = 0.0f, Rr = 0.0f;
  if (a.model >= 2) {
    dstd = P[2];
    float up = a.model == 3 ? 0.004f : 0.01f;
    if (dstd < 0.0f || dstd > up) dstd = d / 10.0f;
    if (a.model == 3) Rr = a.Rmean;
  }
  float val = 0.0f, pr = 0.0f, wd = 0.2f;  // f0 and the fibre slots start at width 0.2 (:497-503)
  float vs = 0.0f, vc = 1.0f;              // sin / cos of an angle slot's value
  if (rl == SL_F0) {
    val = c.incf0 ? P[a.np - 1] : 0.0f;
    if (c.incf0 && c.ardf0 && val > 0.0f && val < 1.0f) pr = flog(val);
  } else if (rl == SL_TAU) {
    val = (RIC && a.tau) ? a.tau[vox] : 0.0f;
    wd = val / 2.0f;
  } else if (rl == SL_D) {
    val = d;
    if (a.model == 3) pr = (1.0f - 3.0f) * flog(d) + 4000.0f * d;
    wd = d / 10.0f;
  } else if (rl == SL_DSTD) {
    val = dstd;
    float up = a.model == 3 ? 0.004f : 0.01f;
    if (a.model >= 2 && dstd > 0.0f && dstd <= up) pr = flog(dstd);
    wd = dstd / 10.0f;
  } else if (rl == SL_R) {
    val = Rr;
    float q;
    if (a.model == 3 && c.r_prior(Rr, d, q)) pr = q;
    wd = Rr / 10.0f;
  } else if (rl == SL_S0) {
    val = P[0];
    wd = val / 10.0f;
  } else if (rl < SL_FIB + 3 * K) {
    int k = (rl - SL_FIB) / 3, w = (rl - SL_FIB) % 3;
    if (w == 0) { val = P[3 + 3 * k + add]; sincos_full(val, vs, vc); pr = C::theta_prior(val, vs); }
    else if (w == 1) { val = P[4 + 3 * k + add]; sincos_full(val, vs, vc); }
    else {
      val = P[2 + 3 * k + add];
      bool lk = (k == 0) ? (a.all_ard != 0) : (a.no_ard == 0);
      pr = (val <= 0.0f || val >= 1.0f) ? 0.0f : a.fudge * (lk ? flog(val) : 0.0f);
    }
  }
  if (a.widths_in) {
    float w = a.widths_in[(size_t)vox * 16 + rl];
    if (w > 0.0f) wd = w;
  }
  c.sval = val;
  c.sprior = pr;
  c.sprop = wd;
  c.ssin = vs;
  c.scos = vc;
  c.scnt = 0;
  {
    SigC ca = aniso_coef(a.model, d, dstd, Rr), ci = iso_coef(a.model, d, dstd);
    c.set_aniso_cache(ca);
    const typename C::Dirs D = c.dirs();
    by_kind(ca.kind, [&](auto SK) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        float4 g = sacq[rl + 16 * t];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          float dp = fmaf(g.x, D.x[k][0], fmaf(g.y, D.x[k][1], g.z * D.x[k][2]));
          c.s[k][t] = sig<decltype(SK)::value>(ca, g.w * dp * dp, c.bv(t, g));
        }
      }
    });
    by_kind(ci.kind, [&](auto SK) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) { float b = c.bv(t, sacq[rl + 16 * t]); c.isov(t) = sig<decltype(SK)::value>(ci, b, b); }
    });
  }
  if constexpr (RIC) {
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      if constexpr (C::Y_LDS) c.vol[t * 64] = y[t];
      else c.r[t] = y[t];
    }
  }
  c.rebuild(y);

  // record destination of the lane's slot (:1326-1339)
  float* rec_base = nullptr;
  size_t rec_stride = (size_t)a.vstride;  // per sample
  if (rl == SL_F0) rec_base = a.f0s;
  else if (rl == SL_TAU) rec_base = a.taus;
  else if (rl == SL_D) rec_base = a.d;
  else if (rl == SL_DSTD) rec_base = a.dstd;
  else if (rl == SL_R) rec_base = a.R;
  else if (rl == SL_S0) rec_base = a.S0;
  else if (rl < SL_FIB + 3 * K) {
    int k = (rl - SL_FIB) / 3, w = (rl - SL_FIB) % 3;
    float* arr = w == 0 ? a.th : (w == 1 ? a.ph : a.f);
    if (arr) rec_base = arr + (size_t)k * a.nsamples * a.vstride;
  }
  if (rec_base) rec_base += vox;

  const uint64_t gvox = a.vox_offset + (uint64_t)vox;
  const int total = a.nburn + a.njumps;
  for (int j = 0; j < total; ++j) {
    float nl, lul;
    draw(a.seed, gvox, (uint32_t)j, (uint32_t)rl, nl, lul);
    if (c.incf0) c.prop_f0(row_bcast<SL_F0>(nl), row_bcast<SL_F0>(lul));
    if (RIC) c.prop_tau(row_bcast<SL_TAU>(nl), row_bcast<SL_TAU>(lul));
    c.template prop_global<SL_D>(row_bcast<SL_D>(nl), row_bcast<SL_D>(lul));
    if (a.model >= 2) {
      c.template prop_global<SL_DSTD>(row_bcast<SL_DSTD>(nl), row_bcast<SL_DSTD>(lul));
      if (a.model == 3) c.template prop_global<SL_R>(row_bcast<SL_R>(nl), row_bcast<SL_R>(lul));
    }
    c.prop_S0(row_bcast<SL_S0>(nl), row_bcast<SL_S0>(lul));
    sfor<K>([&](auto kc) {
      constexpr int k = kc;
      c.template prop_angle<k, true>(row_bcast<C::TH(k)>(nl), row_bcast<C::TH(k)>(lul));
      c.template prop_angle<k, false>(row_bcast<C::PH(k)>(nl), row_bcast<C::PH(k)>(lul));
      c.template prop_f<k>(row_bcast<C::FK(k)>(nl), row_bcast<C::FK(k)>(lul));
    });

    // record: the lane of each sampled slot writes its value
    if (j >= a.nburn) {
      int rec = j - a.nburn + 1;
      if (a.se > 0 && rec % a.se == 0) {
        int smp = rec / a.se - 1;
        if (smp < a.nsamples && rec_base) rec_base[(size_t)smp * rec_stride] = c.sval;
      }
    }
    if ((j + 1) % a.upe == 0) {
      c.update_proposals();
      // Re-anchor the incremental state with the widths: the mixtures, the
      // residuals (Gaussian: from the data, re-read) and the running
      // likelihood recomputed from the stored signals, so rounding in the
      // per-move updates cannot accumulate over the whole chain (the
      // reference recomputes its likelihood at every proposal).
      if constexpr (RIC) c.rebuild(nullptr);
      else {
        float yr[VPL];
#pragma unroll
        for (int t = 0; t < VPL; ++t) {
          const int i = rl + 16 * t;
          yr[t] = i < a.N ? yrow[perm ? perm[i] : i] : 0.0f;
        }
        c.rebuild(yr);
      }
    }
  }
  if (a.energy) {
    const float tau = c.template v<SL_TAU>();
    double e = c.like;
    if constexpr (RIC) {
      double sl = 0.0, sy = 0.0;  // the model-independent part of the Rician terms
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        const double yd = (double)c.yv(t);
        sl += (double)c.mask(t, flog(c.yv(t)));
        sy += t < c.nv ? yd * yd : 0.0;
      }
      e = -(double)a.N * (double)flog(tau) - (c.like + row_sum16_d(sl) - 0.5 * (double)tau * row_sum16_d(sy));
    }
    if (rl == 0) a.energy[vox] = (float)e;
  }
  if (a.ssr) {  // sum of squares of the final state's residuals (the Rician chain keeps the data in r)
    double ss;
    if constexpr (RIC) {
      const float S0 = c.template v<SL_S0>();
      float acc = 0.0f;
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        const float e = c.yv(t) - S0 * c.M[t];
        acc += c.mask(t, e * e);
      }
      ss = (double)row_sum16(acc);
    } else {
      ss = c.like;
    }
    if (rl == 0) a.ssr[vox] = (float)ss;
  }
  if (a.widths_out) a.widths_out[(size_t)vox * 16 + rl] = c.sprop;
}


}  // namespace fdt
