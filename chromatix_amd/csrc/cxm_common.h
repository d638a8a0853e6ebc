// Shared device helpers for the chromatix-on-MI355X kernels: complex64
// arithmetic, double-float (f32 pair) arithmetic and launch checks.
//
// Everything here is gfx950 device code compiled by hipcc; there is no
// host fallback.  Complex values are float2 (x = real, y = imaginary), the
// same in-memory layout as complex64 in torch, numpy and XLA.
//
// The library is compiled with -ffp-contract=off: rounding-exact sequences
// (the reference-order float32 phase, double-float arithmetic) must not be
// fused, so every FMA the arithmetic wants is written explicitly.
//
// CXM_CPU_CHECK builds these headers for the host with the shims of
// tests/cpu_check (a test-only build that checks the phase arithmetic on the
// CPU); the library itself is always the gfx950 build.
#pragma once
#ifdef CXM_CPU_CHECK
#include "cxm_cpu_shim.h"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#define CXM_DEV __device__ __forceinline__

CXM_DEV float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
CXM_DEV float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
CXM_DEV float2 cmul(float2 a, float2 b) {
  return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
CXM_DEV float2 cmulc(float2 a, float2 b) {
  return make_float2(__builtin_fmaf(a.x, b.x, a.y * b.y), __builtin_fmaf(a.y, b.x, -(a.x * b.y)));
}
CXM_DEV float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
CXM_DEV float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse): DIR = -1 forward, +1 inverse.
template <int DIR> CXM_DEV float2 mul_i(float2 a) {
  if constexpr (DIR < 0) return make_float2(a.y, -a.x);
  else return make_float2(-a.y, a.x);
}
CXM_DEV float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }

#ifndef CXM_CPU_CHECK
// ---------------------------------------------------------------------------
// Barriers for LDS exchanges.  s_barrier and the workgroup fences around it
// carry no memory operand, and LLVM's machine sinking moves a load below
// them into a successor block when nothing in between may store (seen: the
// last exchange reads of a transform issued after the barrier that lets the
// other waves write the next transform's stage into the same slots -- a
// cross-wave race that corrupted whole column strips now and then).  An asm
// statement that clobbers memory counts as a store there, so loads stay on
// their side.  It emits no instruction; the marker comment lets the built
// assembly be checked for LDS operations on the wrong side.
// ---------------------------------------------------------------------------
CXM_DEV void sink_fence() { asm volatile("; cxm sink fence" ::: "memory"); }
// __syncthreads with no load moved across it in either direction
CXM_DEV void block_sync() {
  sink_fence();
  __syncthreads();
  sink_fence();
}
// The same for the waves of one wavefront only (a transform that one wave
// computes alone): the LDS writes before it are visible to every lane after.
CXM_DEV void wave_sync() {
  sink_fence();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  sink_fence();
}
#endif
// Correctly rounded float32 square root.  (HIP's __fsqrt_rn maps to the
// approximate native sqrt unless OCML rounded operations are enabled; the
// builtin is correctly rounded under HIP's default
// -fhip-fp32-correctly-rounded-divide-sqrt.)
CXM_DEV float sqrt_rn(float x) { return __builtin_sqrtf(x); }

// ---------------------------------------------------------------------------
// Double-float ("df") arithmetic: a value is hi + lo with |lo| <= ulp(hi)/2.
// Error-free transformations after Dekker (1971) and Knuth (TAOCP vol. 2,
// 4.2.2): two_sum and an FMA-based two_prod.  Used for the propagator phase
// in accurate mode, where the phase reaches ~1.6e3 rad and a plain f32
// product loses ~1e-4 rad.
// ---------------------------------------------------------------------------
struct df { float hi, lo; };

CXM_DEV df df_make(float hi, float lo) { df r; r.hi = hi; r.lo = lo; return r; }
CXM_DEV df two_sum(float a, float b) {
  float s = __fadd_rn(a, b);
  float bb = __fsub_rn(s, a);
  float e = __fadd_rn(__fsub_rn(a, __fsub_rn(s, bb)), __fsub_rn(b, bb));
  return df_make(s, e);
}
CXM_DEV df fast_two_sum(float a, float b) {  // requires |a| >= |b|
  float s = __fadd_rn(a, b);
  float e = __fsub_rn(b, __fsub_rn(s, a));
  return df_make(s, e);
}
CXM_DEV df two_prod(float a, float b) {
  float p = __fmul_rn(a, b);
  float e = __builtin_fmaf(a, b, -p);
  return df_make(p, e);
}
CXM_DEV df df_add(df a, df b) {
  df s = two_sum(a.hi, b.hi);
  df t = two_sum(a.lo, b.lo);
  float lo = __fadd_rn(s.lo, t.hi);
  df u = fast_two_sum(s.hi, lo);
  lo = __fadd_rn(u.lo, t.lo);
  return fast_two_sum(u.hi, lo);
}
CXM_DEV df df_neg(df a) { return df_make(-a.hi, -a.lo); }
CXM_DEV df df_sub(df a, df b) { return df_add(a, df_neg(b)); }
CXM_DEV df df_mul(df a, df b) {
  df p = two_prod(a.hi, b.hi);
  float lo = __builtin_fmaf(a.hi, b.lo, __builtin_fmaf(a.lo, b.hi, p.lo));
  return fast_two_sum(p.hi, lo);
}
CXM_DEV df df_mul_f(df a, float b) {
  df p = two_prod(a.hi, b);
  float lo = __builtin_fmaf(a.lo, b, p.lo);
  return fast_two_sum(p.hi, lo);
}
// sqrt of a non-negative df: one Newton/Karp step on the f32 root.
CXM_DEV df df_sqrt(df a) {
  if (a.hi <= 0.0f) return df_make(0.0f, 0.0f);
  float s = sqrt_rn(a.hi);
  df s2 = two_prod(s, s);
  float r = __fadd_rn(__fsub_rn(__fsub_rn(a.hi, s2.hi), s2.lo), a.lo);
  float c = __fdiv_rn(r, __fmul_rn(2.0f, s));
  return fast_two_sum(s, c);
}
// Fractional part of a df value, as a single f32 in [-0.5, 0.5].  Exact
// reduction because |hi| < 2^23 for every phase (in turns) used here.
CXM_DEV float df_frac(df a) {
  float r = __fsub_rn(a.hi, rintf(a.hi));
  float f = __fadd_rn(r, a.lo);
  return __fsub_rn(f, rintf(f));
}

// ---------------------------------------------------------------------------
// Index helpers.
// ---------------------------------------------------------------------------
// numpy ifftshift along an axis of length n: y[i] = x[(i + n/2) % n].
CXM_DEV int ishift_src(int i, int n) { int j = i + (n >> 1); return j >= n ? j - n : j; }
// numpy fftshift: out[i] = X[(i - n/2) mod n] (n/2 = floor), so bin j lands
// at dst = (j + n/2) % n.
CXM_DEV int fshift_dst(int j, int n) { int d = j + (n >> 1); return d >= n ? d - n : d; }
// Twiddle tables: one compact table per implemented FFT length N, holding
// W_N[i] = exp(-2 pi i i / N) for i < N (FP64-evaluated, rounded once to
// f32), so that a transform's twiddle reads stay within N * 8 bytes (L1
// resident).  The tables are concatenated in the order of kTwLengths.
namespace cxm {
constexpr int kTwLengths[] = {2,  4,  8,  16,  32,  64,  128,  256,  512,  1024, 2048, 4096, 8192,
                              5,  10, 20, 40,  80,  160, 320,  640,  1280, 2560, 5120, 10240,
                              3,  6,  12, 24,  48,  96,  192,  384,  768,  1536, 3072};
constexpr int kTwCount = (int)(sizeof(kTwLengths) / sizeof(kTwLengths[0]));
constexpr int tw_offset(int N) {
  int off = 0;
  for (int i = 0; i < kTwCount; ++i) {
    if (kTwLengths[i] == N) return off;
    off += kTwLengths[i];
  }
  return -1;
}
constexpr int tw_total() {
  int off = 0;
  for (int i = 0; i < kTwCount; ++i) off += kTwLengths[i];
  return off;
}
}  // namespace cxm

// signed FFT bin for natural index k (numpy fftfreq numerator).
CXM_DEV int signed_bin(int k, int n) { return (k < (n + 1) / 2) ? k : k - n; }
