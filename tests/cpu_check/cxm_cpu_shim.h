// Host definitions of the device intrinsics used by cxm_common.h,
// cxm_hfunc.h and cxm_ops.h, for the CPU check build (-DCXM_CPU_CHECK).
// Test infrastructure only: it lets g++ compile the exact device arithmetic
// of the phase evaluation so that it can be compared with the oracle here.
#pragma once
#include <math.h>
#include <stdint.h>

#define __device__
#define __host__
#define __forceinline__ inline

struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
inline float2 make_float2(float x, float y) { float2 r; r.x = x; r.y = y; return r; }
inline float4 make_float4(float x, float y, float z, float w) { float4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }

// x86-64 SSE arithmetic is IEEE single precision with round-to-nearest;
// volatile keeps every rounding step observable.
inline float __fadd_rn(float a, float b) { volatile float r = a + b; return r; }
inline float __fsub_rn(float a, float b) { volatile float r = a - b; return r; }
inline float __fmul_rn(float a, float b) { volatile float r = a * b; return r; }
inline float __fdiv_rn(float a, float b) { volatile float r = a / b; return r; }
inline float __fsqrt_rn(float a) { return sqrtf(a); }
inline void sincospif(float x, float* s, float* c) {
  const double a = M_PI * (double)x;
  *s = (float)sin(a);
  *c = (float)cos(a);
}
