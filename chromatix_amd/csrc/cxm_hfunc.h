// On-the-fly evaluation of the propagation transfer function H(ky, kx) for
// one output plane, in FFT (ifftshifted) bin order.  No H array ever lives
// in HBM.
//
// ref32 mode follows the float32 operation order of
//   compute_transfer_propagator  (src/chromatix/functional/propagation.py:489-496)
//   compute_asm_propagator       (src/chromatix/functional/propagation.py:537-580)
//   freq_grid                    (src/chromatix/core/field.py:36-46)
// with every rounding explicit (__f*_rn, no contraction), so the phase
// argument carries the same float32 rounding as the reference; its
// sine/cosine is then evaluated accurately: the f32 phase is converted to
// turns by a double-float product with 1/(2 pi), reduced exactly mod 1 and
// passed to sincospif (Cody-Waite style reduction with an extended constant).
//
// accurate mode evaluates the phase in turns with double-float arithmetic,
// reduces it exactly mod 1 and calls sincospif, so the error is that of a
// correctly rounded f32 H.
#pragma once
#include "../../include/cxm.h"
#include "cxm_common.h"
#include "cxm_params.h"

namespace cxm {

constexpr float kTwoPiF = 6.28318530717958647692f;  // fl(2*pi)
// 1/(2 pi) as a double-float pair
constexpr float kInv2PiHi = 0.159154936671257019f;
constexpr float kInv2PiLo = (float)(0.15915494309189533577 - (double)0.159154936671257019f);

// sin and cos of an exact f32 angle (radians), accurate for |x| < 3e7.
CXM_DEV void sincos_exact(float x, float* s, float* c) {
  const float turns = df_frac(df_mul_f(df_make(kInv2PiHi, kInv2PiLo), x));
  sincospif(2.0f * turns, s, c);
}

struct XTerms {  // column-constant part (one kx per thread)
  float fx, fxk;  // ref32 f_grid x and (f_x - kx)
  df fx_df, fxk_df;
};

// fl(k / n) for an integer k (|k| < 2^24): for a power-of-two n the quotient
// is exact, so a multiplication by the exact reciprocal gives the same value
// as the correctly rounded division at a fraction of its cost.
CXM_DEV float bin_over_n(float k, int n) {
  if ((n & (n - 1)) == 0) return __fmul_rn(k, 1.0f / (float)n);
  return __fdiv_rn(k, (float)n);
}

CXM_DEV XTerms x_terms(const HParam& hp, int kx, int Nx, int mode) {
  XTerms t;
  const float kxs = (float)signed_bin(kx, Nx);
  if (mode == CXM_PHASE_REF32) {
    t.fx = __fmul_rn(hp.inv_dx, bin_over_n(kxs, Nx));
    t.fxk = __fsub_rn(t.fx, hp.kx0);
  } else {
    t.fx_df = df_mul_f(df_make(hp.rx_hi, hp.rx_lo), kxs);
    t.fxk_df = df_sub(t.fx_df, df_make(hp.kxc_hi, hp.kxc_lo));
    t.fx = t.fx_df.hi;
  }
  return t;
}

// H at (ky, kx) for kind TRANSFER / ASM; multiplied by `scale`.
CXM_DEV float2 eval_H(const HParam& hp, const XTerms& xt, int ky, int Ny, int kind,
                      int mode, int remove_evanescent, int bandlimit, float scale) {
  const float kys = (float)signed_bin(ky, Ny);
  float fy, re, amp = scale;
  float s, c;
  if (mode == CXM_PHASE_REF32) {
    fy = __fmul_rn(hp.inv_dy, bin_over_n(kys, Ny));
    const float fyk = __fsub_rn(fy, hp.ky0);
    const float l2 = __fadd_rn(__fmul_rn(fyk, fyk), __fmul_rn(xt.fxk, xt.fxk));
    if (kind == CXM_H_TRANSFER) {
      re = __fmul_rn(hp.c_l2, l2);
    } else {
      const float kern = __fsub_rn(1.0f, __fmul_rn(hp.c_l2, l2));
      const float osh = __fmul_rn(kTwoPiF, __fadd_rn(__fmul_rn(fy, hp.sy), __fmul_rn(xt.fx, hp.sx)));
      if (kern >= 0.0f || remove_evanescent) {
        const float d = sqrt_rn(fmaxf(kern, 0.0f));
        re = __fadd_rn(__fmul_rn(hp.c_ph, d), osh);
      } else {
        const float dim = sqrt_rn(-kern);
        amp = scale * expf(-__fmul_rn(hp.c_ph, dim));
        re = osh;
      }
    }
    sincos_exact(re, &s, &c);
  } else {
    const df fy_df = df_mul_f(df_make(hp.ry_hi, hp.ry_lo), kys);
    fy = fy_df.hi;
    const df fyk = df_sub(fy_df, df_make(hp.kyc_hi, hp.kyc_lo));
    const df l2 = df_add(df_mul(fyk, fyk), df_mul(xt.fxk_df, xt.fxk_df));
    float turns;
    if (kind == CXM_H_TRANSFER) {
      turns = df_frac(df_mul(df_make(hp.q_hi, hp.q_lo), l2));
    } else {
      const df kern = df_sub(df_make(1.0f, 0.0f), df_mul(df_make(hp.q_hi, hp.q_lo), l2));
      const df sh = df_add(df_mul_f(fy_df, hp.sy), df_mul_f(xt.fx_df, hp.sx));
      if (kern.hi >= 0.0f || remove_evanescent) {
        const df d = df_sqrt(kern);
        turns = df_frac(df_add(df_mul(df_make(hp.a_hi, hp.a_lo), d), sh));
      } else {
        const float dim = sqrtf(-(kern.hi + kern.lo));
        amp = scale * expf(-kTwoPiF * hp.a_hi * dim);
        turns = df_frac(sh);
      }
    }
    sincospif(2.0f * turns, &s, &c);
  }
  float2 h = make_float2(amp * c, amp * s);
  if (kind == CXM_H_ASM) {
    if (hp.conj) h.y = -h.y;
    if (bandlimit) {
      const bool in = (fabsf(__fsub_rn(fy, hp.by0)) <= hp.bym) &&
                      (fabsf(__fsub_rn(xt.fx, hp.bx0)) <= hp.bxm);
      if (!in) h = make_float2(0.0f, 0.0f);
    }
  }
  return h;
}

// Numerical-aperture mask |f|^2 <= na_r2 on the unshifted frequency grid,
// f = fl(fl(1/d) * fl(k/N)) as the reference's f_grid
// (src/chromatix/functional/samples.py:271-280).  na_fx2 is the
// column-constant part.
CXM_DEV float na_fx2(const HParam& hp, int kx, int Nx) {
  const float fx = __fmul_rn(hp.inv_dx, __fdiv_rn((float)signed_bin(kx, Nx), (float)Nx));
  return __fmul_rn(fx, fx);
}
CXM_DEV bool na_keep(const HParam& hp, int ky, int Ny, float fx2, float na_r2) {
  const float fy = __fmul_rn(hp.inv_dy, __fdiv_rn((float)signed_bin(ky, Ny), (float)Ny));
  return !(__fadd_rn(__fmul_rn(fy, fy), fx2) > na_r2);
}

}  // namespace cxm
