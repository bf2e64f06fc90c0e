// fdt_device.h -- device helpers shared by the xfibres kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>

#define FDT_DEV __device__ __forceinline__

namespace fdt {

// diffmodels.h:30-31 / CUDA/diffmodels_utils.h:13-16 (two_pi is 2/pi there)
constexpr float TWO_PI_F = 0.636619772f;
constexpr double TWO_PI_D = 0.636619772;
constexpr float FSMALL_F = 0.001f;
constexpr float UPPERDIFF_F = 0.005f;  // fibre.h:24
constexpr float MAXPROP_F = 1e10f;     // fibre.h:19

// compile-time loop: f(integral_constant<int, i>) for i = 0..N-1
template <int N, class F>
FDT_DEV void sfor(F&& f) {
  [&]<int... is>(std::integer_sequence<int, is...>) { (f(std::integral_constant<int, is>{}), ...); }(
      std::make_integer_sequence<int, N>{});
}

// ----------------------------------------------------------------- lanes
// All-reduce inside a 16-lane DPP row: quad_perm xor1, quad_perm xor2,
// row_half_mirror, row_mirror.  Every lane of the row ends with the same sum
// (each step adds a partner value, and float addition is commutative).
// DPP lane moves with bound_ctrl set and every row and bank enabled: every
// control used here reads a lane inside the row, so no lane keeps an old
// value and the compiler needs no register initialisation for it
template <int CTRL>
FDT_DEV float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
FDT_DEV float row_sum16(float v) {
  v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);  // row_half_mirror
  v += dpp<0x140>(v);  // row_mirror
  return v;
}
FDT_DEV float row_max16(float v) {
  v = fmaxf(v, dpp<0xB1>(v));
  v = fmaxf(v, dpp<0x4E>(v));
  v = fmaxf(v, dpp<0x141>(v));
  v = fmaxf(v, dpp<0x140>(v));
  return v;
}
FDT_DEV float row_min16(float v) {
  v = fminf(v, dpp<0xB1>(v));
  v = fminf(v, dpp<0x4E>(v));
  v = fminf(v, dpp<0x141>(v));
  v = fminf(v, dpp<0x140>(v));
  return v;
}
// Broadcast lane L of each 16-lane row to the whole row (row_newbcast:L).
template <int L>
FDT_DEV float row_bcast(float v) {
  static_assert(L >= 0 && L < 16, "row lane");
  return dpp<0x150 + L>(v);
}
// Whole-wave reductions: the DPP row reduction leaves each 16-lane row's
// total in all its lanes; the four row totals are read from lanes 0, 16, 32
// and 48 with v_readlane and combined in a fixed order, so the result is
// wave-uniform.  (The gfx950 permlane16/32 swaps would keep this on the VALU,
// but ROCm 7.2 clang folds the two results of a swap whose operands are the
// same value into one, so they are not used.)
FDT_DEV float rl(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
FDT_DEV float wave_sum(float v) {
  v = row_sum16(v);
  return (rl(v, 0) + rl(v, 16)) + (rl(v, 32) + rl(v, 48));
}
FDT_DEV float wave_min(float v) {
  v = row_min16(v);
  return fminf(fminf(rl(v, 0), rl(v, 16)), fminf(rl(v, 32), rl(v, 48)));
}

template <int CTRL>
FDT_DEV double dppd(double v) {
  int2 b = __builtin_bit_cast(int2, v);
  b.x = __builtin_amdgcn_mov_dpp(b.x, CTRL, 0xF, 0xF, true);
  b.y = __builtin_amdgcn_mov_dpp(b.y, CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, b);
}
FDT_DEV double rld(double v, int lane) {
  int2 b = __builtin_bit_cast(int2, v);
  b.x = __builtin_amdgcn_readlane(b.x, lane);
  b.y = __builtin_amdgcn_readlane(b.y, lane);
  return __builtin_bit_cast(double, b);
}
FDT_DEV double wave_sum_d(double v) {
  v += dppd<0xB1>(v);
  v += dppd<0x4E>(v);
  v += dppd<0x141>(v);
  v += dppd<0x140>(v);
  return (rld(v, 0) + rld(v, 16)) + (rld(v, 32) + rld(v, 48));
}
// 16-lane row versions for doubles (every lane of the row ends with the
// same value: each step combines a lane with its partner symmetrically)
FDT_DEV double row_sum16_d(double v) {
  v += dppd<0xB1>(v);
  v += dppd<0x4E>(v);
  v += dppd<0x141>(v);
  v += dppd<0x140>(v);
  return v;
}
FDT_DEV double row_max16_d(double v) {
  v = fmax(v, dppd<0xB1>(v));
  v = fmax(v, dppd<0x4E>(v));
  v = fmax(v, dppd<0x141>(v));
  v = fmax(v, dppd<0x140>(v));
  return v;
}
template <int L>
FDT_DEV double row_bcast_d(double v) {
  static_assert(L >= 0 && L < 16, "row lane");
  return dppd<0x150 + L>(v);
}
template <int L>
FDT_DEV int row_bcast_i(int v) {
  static_assert(L >= 0 && L < 16, "row lane");
  return __builtin_amdgcn_mov_dpp(v, 0x150 + L, 0xF, 0xF, true);
}
// value of v in lane `src` of the wave (ds_bpermute; src may differ per lane)
FDT_DEV double bperm_d(int src, double v) {
  int2 b = __builtin_bit_cast(int2, v);
  b.x = __builtin_amdgcn_ds_bpermute(src * 4, b.x);
  b.y = __builtin_amdgcn_ds_bpermute(src * 4, b.y);
  return __builtin_bit_cast(double, b);
}
// make a per-lane-identical value provably wave-uniform
FDT_DEV float uni(float v) { return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v))); }
FDT_DEV double uni(double v) {
  int2 b = __builtin_bit_cast(int2, v);
  b.x = __builtin_amdgcn_readfirstlane(b.x);
  b.y = __builtin_amdgcn_readfirstlane(b.y);
  return __builtin_bit_cast(double, b);
}

// ----------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Draw layout:
// key = (lo32 seed, hi32 seed), counter = (lo32 voxel, hi32 voxel, jump, slot).
struct u4 { uint32_t x, y, z, w; };
FDT_DEV u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    u4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
// Standard normal by Box-Muller (cosine branch) and log of an acceptance
// uniform in (0,1], from one Philox block (24-bit uniforms).  With the
// hardware log2 and cos (v_cos_f32 takes its argument in revolutions, so
// cos(2 pi u2) needs no reduction and no rounding of 2 pi u2) the normal is
// within 7.6e-7 of the double-precision mapping and log u within 1.7e-7
// relative, over all 2^24 values of u2 (libm's float forms: 1.7e-6, from
// rounding 2 pi u2 before the cosine).
FDT_DEV void draw(uint64_t seed, uint64_t vox, uint32_t jump, uint32_t slot, float& n, float& logu) {
  u4 c{(uint32_t)vox, (uint32_t)(vox >> 32), jump, slot};
  u4 x = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const float s24 = 1.0f / 16777216.0f, LN2 = 0.69314718055994531f;
  float u1 = (float)((x.x >> 8) + 1u) * s24;
  float u2 = (float)(x.y >> 8) * s24;
  float u3 = (float)((x.z >> 8) + 1u) * s24;
  n = sqrtf(-2.0f * LN2 * __builtin_amdgcn_logf(u1)) * __builtin_amdgcn_cosf(u2);
  logu = LN2 * __builtin_amdgcn_logf(u3);
}

// ------------------------------------------------------------ FP64 exp/log
// exp and log in FP64 for the LM's model evaluations (the cost of the mixed
// mode, everything in the FP64 mode), without the special-case handling of
// libm's (98 and 42 VALU instructions there against ~33 and ~18 here).
// exp_d: x = n ln2 + r with |r| <= ln2/2 (n ln2 subtracted by two FMAs in the
// hi/lo split of ln2), Taylor series of e^r to r^13 (truncation below 5e-18),
// scaled by 2^n (underflows to 0 below -745).  log_d, for positive normal x:
// x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
// s = (m-1)/(m+1), |s| <= 0.172, series to s^21 (truncation below 3e-17).
// Measured over 4M arguments against long double: exp 1.4e-16 and log 3.3e-16
// largest relative error (libm: 1.4e-16 and 1.3e-16).
FDT_DEV double exp_d(double x) {
  const double L2E = 1.4426950408889634, LN2_HI = 0x1.62e42fefa39efp-1, LN2_LO = 0x1.abc9e3b39803fp-56;
  const double n = __builtin_rint(x * L2E);
  double r = fma(-n, LN2_HI, x);
  r = fma(-n, LN2_LO, r);
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return __builtin_amdgcn_ldexp(p, (int)n);
}
FDT_DEV double log_d(double x) {
  const double LN2_HI = 0x1.62e42fefa39efp-1, LN2_LO = 0x1.abc9e3b39803fp-56;
  double m = __builtin_amdgcn_frexp_mant(x);
  int e = __builtin_amdgcn_frexp_exp(x);
  if (m < 0.70710678118654752440) { m += m; e -= 1; }
  const double f = m - 1.0;  // exact
  const double s = f / (2.0 + f);
  const double z = s * s;
  double q = 1.0 / 21.0;
  q = fma(q, z, 1.0 / 19.0);
  q = fma(q, z, 1.0 / 17.0);
  q = fma(q, z, 1.0 / 15.0);
  q = fma(q, z, 1.0 / 13.0);
  q = fma(q, z, 1.0 / 11.0);
  q = fma(q, z, 1.0 / 9.0);
  q = fma(q, z, 1.0 / 7.0);
  q = fma(q, z, 1.0 / 5.0);
  q = fma(q, z, 1.0 / 3.0);
  const double s2 = s + s;
  const double ed = (double)e;
  return fma(ed, LN2_HI, fma(ed, LN2_LO, fma(s2 * z, q, s2)));
}

// ----------------------------------------------------------- transforms
template <class T> FDT_DEV T f2x(T f) { return tan(f / T(TWO_PI_D)); }
template <class T> FDT_DEV T x2f(T x) { return fabs(T(TWO_PI_D) * atan(x)); }
template <class T> FDT_DEV T f2beta(T f) { return asin(sqrt(f)); }
template <class T> FDT_DEV T beta2f(T b) { T s = sin(b); return s * s; }
template <class T> FDT_DEV T sgn(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }

}  // namespace fdt
