// fdt_mcmc.hip -- persistent single-site Metropolis-Hastings for the
// ball-and-sticks models on gfx950.
//
// Semantics follow the reference GPU MCMC: init_Fibres_Multifibres_kernel
// (CUDA/runmcmc_kernels.cu:193-505) and runmcmc_kernel (:507-1474), with the
// burn-in and record launches (CUDA/runmcmc.cu:108-316) fused into one loop of
// nburn + njumps sweeps.  Layout: one voxel per 16-lane DPP row, four voxels
// per wave64, 16 voxels per 256-thread workgroup.  Lane l of a row holds
// volumes l, l+16, l+32, ... (VPL volumes per lane) and keeps, in VGPRs,
//   the residual r = y - S0*M (Gaussian) or the data y (Rician),
//   the mixture M = sum_k f_k s_k + (1 - f0 - sum_k f_k) iso + f0,
//   the isotropic signal iso, the stick signals s_k and b*(g.x_k)^2.
// A proposal evaluates only what it changes; its likelihood change is
// summed over the row with four DPP adds.  For Gaussian noise the change of
// the sum of squares is accumulated as sum dr*(2r+dr) so that the energy
// difference (N/2)*log1p(dSSR/SSR) keeps full relative precision (the
// reference differences two FP32 energies).  The acquisition (gx,gy,gz,b) is
// staged in LDS.  Draws: lane l of the row computes the Philox block of
// proposal slot l for the current jump; DPP row_newbcast hands it to the row.
#include "fdt_device.h"
#include "fdt_kernels.h"

namespace fdt {

constexpr float L2E = 1.4426950408889634f;
constexpr float LN2 = 0.6931471805599453f;
FDT_DEV float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
FDT_DEV float lg2(float x) { return __builtin_amdgcn_logf(x); }
FDT_DEV float flog(float x) { return lg2(x) * LN2; }

// draw slots (fixed per parameter)
enum { SL_F0 = 0, SL_TAU = 1, SL_D = 2, SL_DSTD = 3, SL_R = 4, SL_S0 = 5, SL_FIB = 6 };

// Signal forms with row-uniform coefficients (compute_signal /
// compute_iso_signal, CUDA/runmcmc_kernels.cu:67-95):
//   kind 0: exp(-d*u)                        model 1, or d_std < 1e-5
//   kind 1: (1 + (dstd^2/d) u)^(-d^2/dstd^2)  model 2 (gamma diffusivities)
//   kind 2: exp(-3 d invR ((1-R) b a + R b))  model 3 zeppelin, u = b a
struct SigC {
  int kind;
  float c0, c1;
};
FDT_DEV SigC aniso_coef(int model, float d, float dstd, float R) {
  SigC c;
  if (model == 1 || (model == 2 && dstd < 1e-5f)) { c.kind = 0; c.c0 = -d * L2E; c.c1 = 0.0f; }
  else if (model == 2) { float s2 = dstd * dstd; c.kind = 1; c.c0 = s2 / d; c.c1 = -(d * d / s2); }
  else {
    float iR = 1.0f / (2.0f * R + 1.0f);
    float k = -3.0f * d * iR * L2E;
    c.kind = 2; c.c0 = k * (1.0f - R); c.c1 = k * R;
  }
  return c;
}
FDT_DEV SigC iso_coef(int model, float d, float dstd) {
  SigC c;
  if (model == 1 || dstd < 1e-5f) { c.kind = 0; c.c0 = -d * L2E; c.c1 = 0.0f; }
  else { float s2 = dstd * dstd; c.kind = 1; c.c0 = s2 / d; c.c1 = -(d * d / s2); }
  return c;
}
template <int SK>
FDT_DEV float sig(const SigC& c, float u, float b) {
  if constexpr (SK == 0) return ex2(c.c0 * u);
  else if constexpr (SK == 1) return ex2(c.c1 * lg2(fmaf(c.c0, u, 1.0f)));
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
  float Ll = b + flog(pl * __builtin_amdgcn_rsqf(b));
  return b < 3.75f ? Ls : Ll;
}
// Rician log-likelihood term of one volume (fibre.h:949-953)
FDT_DEV float ell_of(float y, float p, float tau) {
  return flog(y) - 0.5f * tau * (y * y + p * p) + logI0f(tau * p * y);
}

FDT_DEV float g_term(float r, float dr) { return dr * fmaf(2.0f, r, dr); }
FDT_DEV int acc1(int c) { return c + 1; }
FDT_DEV int rej1(int c) { return c + (1 << 16); }

template <int K, int VPL, bool RIC>
struct Chain {
  // lane-resident volume state, volume i = rl + 16 t
  float r[VPL];  // Gaussian: y - S0*M; Rician: y
  float M[VPL];
  float iso[VPL];
  float s[K][VPL];
  float ba[K][VPL];
  float ell[RIC ? VPL : 1];
  // row-uniform state
  float S0, d, dstd, Rr, f0, tau;
  float th[K], ph[K], f[K];
  float pS0, pd, pdstd, pR, pf0, ptau, pth[K], pph[K], pf[K];
  int cS0, cd, cdstd, cR, cf0, ctau, cth[K], cph[K], cf[K];  // acc + (rej << 16)
  float prd, prdstd, prR, prf0, prth[K], prf[K];
  double like;  // Gaussian: SSR; Rician: sum of ell
  bool last_ok;  // the lane's last slot holds a real volume
  int rl;
  // options
  int model;
  bool incf0, ardf0, can_ard;
  bool lam[K];
  float fudge, halfN, Nf, Rmean, Rstd, Rfudge;
  const float4* acq;

  FDT_DEV float mask_last(int t, float v) const { return (t == VPL - 1 && !last_ok) ? 0.0f : v; }
  FDT_DEV float fsum() const {
    float a = f0;
#pragma unroll
    for (int k = 0; k < K; ++k) a += f[k];
    return a;
  }
  // energy change of the likelihood from a row-summed change `dl`
  FDT_DEV float dlik(float dl) const {
    if constexpr (RIC) return -dl;
    else {
      float x = dl / (float)like;
      float u = 1.0f + x;
      float lp = (u == 1.0f) ? x : flog(u) * (x / (u - 1.0f));
      return halfN * lp;
    }
  }
  FDT_DEV float theta_prior(float t) const {  // fibre.h:283-290
    return t == 0.0f ? 0.0f : -flog(fabsf(__sinf(t)) * 0.5f);
  }
  // R prior (fibre.h:816-843); returns false when outside its support
  FDT_DEV bool r_prior(float Rv, float& pr) const {
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

  // rebuild M and the residual from the stored signals
  FDT_DEV void rebuild(const float* y) {
    float fs = fsum();
    double ss = 0.0;
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      float m = fmaf(1.0f - fs, iso[t], f0);
#pragma unroll
      for (int k = 0; k < K; ++k) m = fmaf(f[k], s[k][t], m);
      M[t] = m;
      if constexpr (RIC) {
        ell[t] = ell_of(r[t], S0 * m, tau);
        ss += (double)mask_last(t, ell[t]);
      } else {
        r[t] = y[t] - S0 * m;
        ss += (double)mask_last(t, r[t] * r[t]);
      }
    }
    like = (double)row_sum16((float)ss);
  }

  // ---------------------------------------------------------- proposals
  // Each returns through `acc` whether the move was accepted.
  // f0 (CUDA/runmcmc_kernels.cu:842-888)
  FDT_DEV void prop_f0(float n, float lu) {
    float f0n = f0 + n * pf0;
    float fsn = fsum() - f0 + f0n;
    if (f0n <= 0.0f || f0n >= 1.0f || fsn > 1.0f) { cf0 = rej1(cf0); return; }
    float prn = ardf0 ? flog(f0n) : 0.0f;
    float df = f0n - f0;
    float a = 0.0f, en[RIC ? VPL : 1];
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      float dM = df * (1.0f - iso[t]);
      if constexpr (RIC) { en[t] = ell_of(r[t], S0 * (M[t] + dM), tau); a += mask_last(t, en[t] - ell[t]); }
      else a += mask_last(t, g_term(r[t], -S0 * dM));
    }
    a = row_sum16(a);
    if (dlik(a) + (prn - prf0) < -lu) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        float dM = df * (1.0f - iso[t]);
        M[t] += dM;
        if constexpr (RIC) ell[t] = en[t]; else r[t] -= S0 * dM;
      }
      like += (double)a;
      f0 = f0n; prf0 = prn; cf0 = acc1(cf0);
    } else cf0 = rej1(cf0);
  }
  // tau (:890-922), Rician only
  FDT_DEV void prop_tau(float n, float lu) {
    if constexpr (RIC) {
      float tn = tau + n * ptau;
      if (tn <= 0.0f) { ctau = rej1(ctau); return; }
      float a = 0.0f, en[VPL];
#pragma unroll
      for (int t = 0; t < VPL; ++t) { en[t] = ell_of(r[t], S0 * M[t], tn); a += mask_last(t, en[t] - ell[t]); }
      a = row_sum16(a);
      float dE = -Nf * (flog(tn) - flog(tau)) - a;
      if (dE < -lu) {
#pragma unroll
        for (int t = 0; t < VPL; ++t) ell[t] = en[t];
        like += (double)a;
        tau = tn; ctau = acc1(ctau);
      } else ctau = rej1(ctau);
    }
  }
  // S0 (:1125-1155)
  FDT_DEV void prop_S0(float n, float lu) {
    float Sn = S0 + n * pS0;
    if (Sn < 0.0f) { cS0 = rej1(cS0); return; }
    float dS = Sn - S0;
    float a = 0.0f, en[RIC ? VPL : 1];
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      if constexpr (RIC) { en[t] = ell_of(r[t], Sn * M[t], tau); a += mask_last(t, en[t] - ell[t]); }
      else a += mask_last(t, g_term(r[t], -dS * M[t]));
    }
    a = row_sum16(a);
    if (dlik(a) < -lu) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        if constexpr (RIC) ell[t] = en[t]; else r[t] -= dS * M[t];
      }
      like += (double)a;
      S0 = Sn; cS0 = acc1(cS0);
    } else cS0 = rej1(cS0);
  }
  // all stick signals and/or iso change: d (:924-985), d_std (:987-1051), R (:1053-1122)
  FDT_DEV void prop_global(int which, float n, float lu) {
    float dn = d, dsn = dstd, Rn = Rr, prn = 0.0f, prold = 0.0f;
    bool new_aniso = true, new_iso = true;
    if (which == SL_D) {
      dn = d + n * pd;
      if (dn < 0.0f || dn > UPPERDIFF_F) { cd = rej1(cd); return; }
      prn = model == 3 ? (1.0f - 3.0f) * flog(dn) + 4000.0f * dn : 0.0f;
      prold = prd;
    } else if (which == SL_DSTD) {
      dsn = dstd + n * pdstd;
      float up = model == 3 ? 0.004f : 0.01f;
      if (dsn <= 0.0f || dsn > up) { cdstd = rej1(cdstd); return; }
      prn = flog(dsn);
      prold = prdstd;
      new_aniso = (model == 2);
    } else {
      Rn = Rr + n * pR;
      if (!r_prior(Rn, prn)) { cR = rej1(cR); return; }
      prold = prR;
      new_iso = false;
    }
    SigC ca = aniso_coef(model, dn, dsn, Rn), ci = iso_coef(model, dn, dsn);
    float sn[K][VPL], isn[VPL], Mn[VPL], en[RIC ? VPL : 1];
    float fs = fsum();
    float a = 0.0f;
    if (new_aniso) {
      by_kind(ca.kind, [&](auto SK) {
#pragma unroll
        for (int t = 0; t < VPL; ++t) {
          float b = acq[rl + 16 * t].w;
#pragma unroll
          for (int k = 0; k < K; ++k) sn[k][t] = sig<decltype(SK)::value>(ca, ba[k][t], b);
        }
      });
    } else {
#pragma unroll
      for (int t = 0; t < VPL; ++t)
#pragma unroll
        for (int k = 0; k < K; ++k) sn[k][t] = s[k][t];
    }
    if (new_iso) {
      by_kind(ci.kind, [&](auto SK) {
#pragma unroll
        for (int t = 0; t < VPL; ++t) {
          float b = acq[rl + 16 * t].w;
          isn[t] = sig<decltype(SK)::value>(ci, b, b);
        }
      });
    } else {
#pragma unroll
      for (int t = 0; t < VPL; ++t) isn[t] = iso[t];
    }
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      float m = fmaf(1.0f - fs, isn[t], f0);
#pragma unroll
      for (int k = 0; k < K; ++k) m = fmaf(f[k], sn[k][t], m);
      Mn[t] = m;
      if constexpr (RIC) { en[t] = ell_of(r[t], S0 * m, tau); a += mask_last(t, en[t] - ell[t]); }
      else a += mask_last(t, g_term(r[t], -S0 * (m - M[t])));
    }
    a = row_sum16(a);
    bool ok = dlik(a) + (prn - prold) < -lu;
    if (ok) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
#pragma unroll
        for (int k = 0; k < K; ++k) s[k][t] = sn[k][t];
        iso[t] = isn[t];
        if constexpr (RIC) ell[t] = en[t]; else r[t] -= S0 * (Mn[t] - M[t]);
        M[t] = Mn[t];
      }
      like += (double)a;
    }
    if (which == SL_D) { if (ok) { d = dn; prd = prn; cd = acc1(cd); } else cd = rej1(cd); }
    else if (which == SL_DSTD) { if (ok) { dstd = dsn; prdstd = prn; cdstd = acc1(cdstd); } else cdstd = rej1(cdstd); }
    else { if (ok) { Rr = Rn; prR = prn; cR = acc1(cR); } else cR = rej1(cR); }
  }
  // theta_k / phi_k (:1157-1265)
  template <int k>
  FDT_DEV void prop_angle(bool is_th, float n, float lu) {
    float tn = th[k], pn = ph[k], prn = prth[k];
    if (is_th) { tn = th[k] + n * pth[k]; prn = theta_prior(tn); }
    else pn = ph[k] + n * pph[k];
    float st, ct, sp, cp;
    __sincosf(tn, &st, &ct);
    __sincosf(pn, &sp, &cp);
    float x0 = st * cp, x1 = st * sp, x2 = ct;
    SigC ca = aniso_coef(model, d, dstd, Rr);
    float ban[VPL], sn[VPL], en[RIC ? VPL : 1];
    float a = 0.0f;
    by_kind(ca.kind, [&](auto SK) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        float4 q = acq[rl + 16 * t];
        float dp = fmaf(q.x, x0, fmaf(q.y, x1, q.z * x2));
        ban[t] = q.w * dp * dp;
        sn[t] = sig<decltype(SK)::value>(ca, ban[t], q.w);
      }
    });
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      float dM = f[k] * (sn[t] - s[k][t]);
      if constexpr (RIC) { en[t] = ell_of(r[t], S0 * (M[t] + dM), tau); a += mask_last(t, en[t] - ell[t]); }
      else a += mask_last(t, g_term(r[t], -S0 * dM));
    }
    a = row_sum16(a);
    if (dlik(a) + (prn - prth[k]) < -lu) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        float dM = f[k] * (sn[t] - s[k][t]);
        M[t] += dM;
        if constexpr (RIC) ell[t] = en[t]; else r[t] -= S0 * dM;
        s[k][t] = sn[t];
        ba[k][t] = ban[t];
      }
      like += (double)a;
      th[k] = tn; ph[k] = pn; prth[k] = prn;
      if (is_th) cth[k] = acc1(cth[k]); else cph[k] = acc1(cph[k]);
    } else {
      if (is_th) cth[k] = rej1(cth[k]); else cph[k] = rej1(cph[k]);
    }
  }
  // f_k (:1267-1322)
  template <int k>
  FDT_DEV void prop_f(float n, float lu) {
    float fn = f[k] + n * pf[k];
    float fsn = fsum() - f[k] + fn;
    if (fn <= 0.0f || fn >= 1.0f || fsn > 1.0f) { cf[k] = rej1(cf[k]); return; }
    float prn = (can_ard && lam[k]) ? fudge * flog(fn) : 0.0f;
    float df = fn - f[k];
    float a = 0.0f, en[RIC ? VPL : 1];
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
      float dM = df * (s[k][t] - iso[t]);
      if constexpr (RIC) { en[t] = ell_of(r[t], S0 * (M[t] + dM), tau); a += mask_last(t, en[t] - ell[t]); }
      else a += mask_last(t, g_term(r[t], -S0 * dM));
    }
    a = row_sum16(a);
    if (dlik(a) + (prn - prf[k]) < -lu) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        float dM = df * (s[k][t] - iso[t]);
        M[t] += dM;
        if constexpr (RIC) ell[t] = en[t]; else r[t] -= S0 * dM;
      }
      like += (double)a;
      f[k] = fn; prf[k] = prn; cf[k] = acc1(cf[k]);
    } else cf[k] = rej1(cf[k]);
  }

  // proposal-width adaptation every upe jumps (:1341-1395)
  FDT_DEV static void upd(float& p, int& c) {
    float acc = (float)(c & 0xFFFF), rej = (float)(c >> 16);
    p *= sqrtf((acc + 1.0f) / (rej + 1.0f));
    p = fminf(p, MAXPROP_F);
    c = 0;
  }
  FDT_DEV void update_proposals(bool ric) {
    upd(pd, cd);
    if (ric) upd(ptau, ctau);
    if (incf0) upd(pf0, cf0);
    if (model >= 2) { upd(pdstd, cdstd); if (model == 3) upd(pR, cR); }
    upd(pS0, cS0);
#pragma unroll
    for (int k = 0; k < K; ++k) { upd(pth[k], cth[k]); upd(pph[k], cph[k]); upd(pf[k], cf[k]); }
  }
};

template <int K, int VPL, bool RIC>
__global__ __launch_bounds__(256) void mcmc_kernel(McmcArgs a) {
  __shared__ float4 sacq[16 * VPL];
  for (int i = threadIdx.x; i < 16 * VPL; i += blockDim.x) sacq[i] = a.acq[i];
  __syncthreads();
  const int lane = __lane_id();
  const int rl = lane & 15;
  const int vox = blockIdx.x * 16 + (threadIdx.x >> 4);
  if (vox >= a.nvox) return;

  Chain<K, VPL, RIC> c;
  c.rl = rl;
  c.acq = sacq;
  c.model = a.model;
  c.incf0 = a.f0 != 0;
  c.ardf0 = a.ardf0 != 0;
  c.can_ard = a.no_ard == 0;
  c.fudge = a.fudge;
  c.halfN = 0.5f * (float)a.N;
  c.Nf = (float)a.N;
  c.Rmean = a.Rmean; c.Rstd = a.Rstd; c.Rfudge = a.Rfudge;
  c.last_ok = (rl + 16 * (VPL - 1)) < a.N;

  // ---- data (with the Rician non-positive replacement, CUDA/xfibres_gpu.cu:350-367)
  float y[VPL];
  const float* yrow = a.data + (size_t)vox * a.N;
#pragma unroll
  for (int t = 0; t < VPL; ++t) {
    int i = rl + 16 * t;
    y[t] = i < a.N ? yrow[i] : 0.0f;
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

  // ---- chain initialisation (init_Fibres_Multifibres_kernel, :265-504)
  const float* P = a.params + (size_t)vox * a.np;
  const int add = a.model >= 2 ? 1 : 0;
  c.S0 = P[0];
  c.d = P[1];
  if (c.d < 0.0f || c.d > UPPERDIFF_F) c.d = 2e-3f;
  if (a.model >= 2) {
    c.dstd = P[2];
    float up = a.model == 3 ? 0.004f : 0.01f;
    if (c.dstd < 0.0f || c.dstd > up) c.dstd = c.d / 10.0f;
    c.Rr = a.model == 3 ? a.Rmean : 0.0f;
  } else { c.dstd = 0.0f; c.Rr = 0.0f; }
  c.f0 = c.incf0 ? P[a.np - 1] : 0.0f;
  c.tau = (RIC && a.tau) ? a.tau[vox] : 0.0f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    c.f[k] = P[2 + 3 * k + add];
    c.th[k] = P[3 + 3 * k + add];
    c.ph[k] = P[4 + 3 * k + add];
    c.pth[k] = c.pph[k] = c.pf[k] = 0.2f;
    c.cth[k] = c.cph[k] = c.cf[k] = 0;
    c.prth[k] = c.theta_prior(c.th[k]);
    c.lam[k] = (k == 0) ? (a.all_ard != 0) : (a.no_ard == 0);
    c.prf[k] = (c.f[k] <= 0.0f || c.f[k] >= 1.0f) ? 0.0f : a.fudge * (c.lam[k] ? flog(c.f[k]) : 0.0f);
  }
  c.cS0 = c.cd = c.cdstd = c.cR = c.cf0 = c.ctau = 0;
  c.prd = (a.model == 3 && c.d >= 0.0f && c.d <= UPPERDIFF_F) ? (1.0f - 3.0f) * flog(c.d) + 4000.0f * c.d : 0.0f;
  c.prdstd = 0.0f;
  c.prR = 0.0f;
  if (a.model >= 2) {
    float up = a.model == 3 ? 0.004f : 0.01f;
    if (c.dstd > 0.0f && c.dstd <= up) c.prdstd = flog(c.dstd);
    if (a.model == 3) { float pr; if (c.r_prior(c.Rr, pr)) c.prR = pr; }
  }
  c.prf0 = (c.incf0 && c.ardf0 && c.f0 > 0.0f && c.f0 < 1.0f) ? flog(c.f0) : 0.0f;
  {
    SigC ca = aniso_coef(a.model, c.d, c.dstd, c.Rr), ci = iso_coef(a.model, c.d, c.dstd);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      float st, ct, sp, cp;
      __sincosf(c.th[k], &st, &ct);
      __sincosf(c.ph[k], &sp, &cp);
      float x0 = st * cp, x1 = st * sp, x2 = ct;
#pragma unroll
      for (int t = 0; t < VPL; ++t) {
        float4 q = sacq[rl + 16 * t];
        float dp = fmaf(q.x, x0, fmaf(q.y, x1, q.z * x2));
        c.ba[k][t] = q.w * dp * dp;
      }
    }
    by_kind(ca.kind, [&](auto SK) {
#pragma unroll
      for (int t = 0; t < VPL; ++t)
#pragma unroll
        for (int k = 0; k < K; ++k) c.s[k][t] = sig<decltype(SK)::value>(ca, c.ba[k][t], sacq[rl + 16 * t].w);
    });
    by_kind(ci.kind, [&](auto SK) {
#pragma unroll
      for (int t = 0; t < VPL; ++t) { float b = sacq[rl + 16 * t].w; c.iso[t] = sig<decltype(SK)::value>(ci, b, b); }
    });
  }
  if constexpr (RIC) {
#pragma unroll
    for (int t = 0; t < VPL; ++t) c.r[t] = y[t];
  }
  c.rebuild(y);
  c.pS0 = c.S0 / 10.0f; c.pd = c.d / 10.0f; c.pdstd = c.dstd / 10.0f;
  c.ptau = c.tau / 2.0f; c.pf0 = 0.2f; c.pR = c.Rr / 10.0f;

  const uint64_t gvox = a.vox_offset + (uint64_t)vox;
  const int total = a.nburn + a.njumps;
  for (int j = 0; j < total; ++j) {
    float nl, lul;
    draw(a.seed, gvox, (uint32_t)j, (uint32_t)rl, nl, lul);
    if (c.incf0) c.prop_f0(row_bcast<SL_F0>(nl), row_bcast<SL_F0>(lul));
    if (RIC) c.prop_tau(row_bcast<SL_TAU>(nl), row_bcast<SL_TAU>(lul));
    c.prop_global(SL_D, row_bcast<SL_D>(nl), row_bcast<SL_D>(lul));
    if (a.model >= 2) {
      c.prop_global(SL_DSTD, row_bcast<SL_DSTD>(nl), row_bcast<SL_DSTD>(lul));
      if (a.model == 3) c.prop_global(SL_R, row_bcast<SL_R>(nl), row_bcast<SL_R>(lul));
    }
    c.prop_S0(row_bcast<SL_S0>(nl), row_bcast<SL_S0>(lul));
    [&]<int... ks>(std::integer_sequence<int, ks...>) {
      ((c.template prop_angle<ks>(true, row_bcast<SL_FIB + 3 * ks>(nl), row_bcast<SL_FIB + 3 * ks>(lul)),
        c.template prop_angle<ks>(false, row_bcast<SL_FIB + 3 * ks + 1>(nl), row_bcast<SL_FIB + 3 * ks + 1>(lul)),
        c.template prop_f<ks>(row_bcast<SL_FIB + 3 * ks + 2>(nl), row_bcast<SL_FIB + 3 * ks + 2>(lul))), ...);
    }(std::make_integer_sequence<int, K>{});

    // record (:1326-1339): lane q of the row writes quantity q
    if (j >= a.nburn) {
      int rec = j - a.nburn + 1;
      if (a.se > 0 && rec % a.se == 0) {
        int smp = rec / a.se - 1;
        if (smp < a.nsamples) {
          size_t sv = (size_t)smp * a.nvox + vox;
          float val = 0.0f;
          float* dst = nullptr;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            size_t kv = ((size_t)k * a.nsamples + smp) * a.nvox + vox;
            if (rl == k) { val = c.th[k]; dst = a.th + kv; }
            if (rl == K + k) { val = c.ph[k]; dst = a.ph + kv; }
            if (rl == 2 * K + k) { val = c.f[k]; dst = a.f + kv; }
          }
          if (rl == 3 * K) { val = c.S0; dst = a.S0 + sv; }
          if (rl == 3 * K + 1) { val = c.d; dst = a.d + sv; }
          if (rl == 3 * K + 2 && a.dstd) { val = c.dstd; dst = a.dstd + sv; }
          if (rl == 3 * K + 3 && a.R) { val = c.Rr; dst = a.R + sv; }
          if (rl == 3 * K + 4 && a.f0s) { val = c.f0; dst = a.f0s + sv; }
          if (rl == 3 * K + 5 && a.taus) { val = c.tau; dst = a.taus + sv; }
          if (dst) *dst = val;
        }
      }
    }
    if ((j + 1) % a.upe == 0) c.update_proposals(RIC);
  }
  if (a.energy && rl == 0) {
    float e = RIC ? (float)(-(double)a.N * (double)flog(c.tau) - c.like) : (float)c.like;
    a.energy[vox] = e;
  }
}

}  // namespace fdt

// One translation unit per VPL value (-DFDT_VPL=...), instantiating K=1..3
// for Gaussian and Rician noise.
#ifndef FDT_VPL
#define FDT_VPL 18
#endif
#define FDT_CAT2_(a, b) a##b
#define FDT_CAT2(a, b) FDT_CAT2_(a, b)
extern "C" hipError_t FDT_CAT2(fdt_launch_mcmc_vpl, FDT_VPL)(const fdt::McmcArgs* a, int K, int rician, hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  dim3 grid((a->nvox + 15) / 16), block(256);
#define FDT_MC_CASE(KK, RR) \
  if (K == KK && (rician != 0) == RR) { hipLaunchKernelGGL((fdt::mcmc_kernel<KK, FDT_VPL, RR>), grid, block, 0, s, *a); return hipGetLastError(); }
  FDT_MC_CASE(1, false) FDT_MC_CASE(2, false) FDT_MC_CASE(3, false)
  FDT_MC_CASE(1, true) FDT_MC_CASE(2, true) FDT_MC_CASE(3, true)
#undef FDT_MC_CASE
  return hipErrorInvalidValue;
}
