// Pointwise operators applied inside the FFT pipeline passes (on load, on
// the spectrum between two transforms, on store), so that chirps, masks,
// phase screens and scales never cost an extra HBM pass.
//
// Each operator multiplies the complex value at a grid position by a factor
// (or conjugates it).  Positions are FFT-grid indices (iy, ix) of an (ny, nx)
// grid; physical coordinates follow field.grid of the reference,
// g = fl(d * (i - N/2)) + origin  (src/chromatix/core/field.py:23-33).
// Spectrum-stage operators read (iy, ix) as FFT-order bins.
//
// Per-plane parameters live in one float4 table per launch (kernel
// arguments); operator k reads slot q = (plane / pdiv) % pmod at
// table[pbase + q * pstride + j].  Phase-evaluating operators follow the
// reference's float32 operation order (ref32) or evaluate the phase in
// double-float turns (accurate).
#pragma once
#include "cxm_common.h"
#include "cxm_hfunc.h"

namespace cxm {

constexpr int kMaxOps = 4;   // operators per stage
constexpr int kPTab = 128;   // float4 parameter slots per launch

enum OpKind : int {
  OP_NONE = 0,
  OP_SCALE = 1,   // v *= (p0.x + i p0.y)
  OP_CONJ = 2,    // v = conj(v)
  OP_CHIRP = 3,   // v *= exp(i ph); ph = fl(a*l2) [/ b if b != 0], l2 = |g - off|^2
                  //   p0 = (a, b, a_lo, -), p1 = (dy, dx, oy, ox), p2 = (off_y, off_x, -, -)
  OP_PHASE = 4,   // v *= exp(i fl(phi[y,x] * p0.x))                 (f32 array)
  OP_THIN = 5,    // v *= exp(i 2 pi (dn + i abs) t / p0.x)           (f32 arrays)
  OP_CARRAY = 6,  // v *= A[y,x]                                      (complex64 array)
  OP_CIRC = 7,    // v *= (|g - off|^2 <= p0.x); p1, p2 as CHIRP
  OP_VEC = 8,     // v *= V[q][i] along one axis, q the parameter slot (complex64 vectors)
  OP_SASDH = 9,   // SAS precompensation dH at bins (ky, kx)
                  //   ref32:    p0 = (kz, -, lam, n), p1 = (1/dy, 1/dx, ty0, tx0)
                  //   accurate: p0 = (turns hi, lo, s_step_y hi, lo),
                  //             p1 = (s_step_x hi, lo, ty0, tx0)
  OP_NAMASK = 10, // v *= (fy^2 + fx^2 <= p0.x) at bins; p0 = (r2, 1/dy, 1/dx, -)
  OP_IGRAD = 11,  // v = 2 g v, the cotangent of v from dL/dI with I = |v|^2:
                  //   mode 0: g = arr0[y', x'] (dL/dI given, f32)
                  //   mode 1: g = p0.x * (|v|^2 - arr0[y', x']) (mean-squared error
                  //           against a target, p0.x = 2 / element count)
                  //   mode 2: v = arr0[y', x'] (complex64 cotangent given)
                  //   (y', x') = (iy, ix), each fftshifted when bit 0 / 1 of `shift` is
                  //   set, minus (arr_row_off, arr_col_off); outside the array g = 0
  OP_THINADJ = 12,// adjoint of OP_THIN at the stored forward field u = arr2 (after the
                  //   sample): out0 (+)= k Im(v conj u), out1 (+)= -k Re(v conj u) (d/d dn,
                  //   d/d absorption, k = 2 pi t / lam), v = conj(T) v; p0 = (lam, k, -, -)
  OP_FLUOR = 13,  // v += f[y,x] exp(i r[y,x])  (f32 arrays: f = arr0, r = arr1; outside:
                  //   unchanged) -- the fluorescence added before each slice's
                  //   propagation (samples.py:388-392, complex64 order f cos r, f sin r)
};

struct OpDesc {
  int kind;
  int mode;                    // CXM_PHASE_* for CHIRP / SASDH
  int pdiv, pmod;              // slot q = (plane / pdiv) % pmod (pmod <= 1: slot 0)
  int pbase, pstride;          // parameter table addressing
  const void* arr0;            // PHASE: phi; THIN: dn; CARRAY / VEC: complex data
  const void* arr1;            // THIN: absorption (may be null)
  const void* arr2;            // THINADJ: stored forward field (complex64, grid layout)
  void* out0;                  // THINADJ: d/d dn   (f32, array layout)
  void* out1;                  // THINADJ: d/d absorption (f32, may be null)
  long long arr2_plane_stride; // THINADJ: elements between planes of arr2
  int arr2_row_stride;         // THINADJ: elements between rows of arr2
  int accum;                   // THINADJ: add to out0 / out1 instead of storing
  long long arr_plane_stride;  // elements between planes (0: broadcast)
  int arr_row_stride;          // elements between array rows
  int arr_rows, arr_cols;      // array extent (VEC: length in arr_cols)
  int arr_row_off, arr_col_off;// grid index of array element (0, 0)
  int outside_zero;            // positions outside the array: 0 (1) or unchanged (0)
  int vec_axis;                // VEC: 0 indexes by iy, 1 by ix
  float t;                     // THIN: thickness
  int shift;                   // IGRAD: fftshift of the row (bit 0) / column (bit 1) index
};

struct OpList {
  int n;
  int pad_;
  OpDesc op[kMaxOps];
};

struct OpTable {
  float4 p[kPTab];
};

CXM_DEV float grid_coord(int i, int n, float d, float o) {
  // fl(d * (i - n/2)) + o, as field.grid computes it in float32
  return __fadd_rn(__fmul_rn(d, (float)i - 0.5f * (float)n), o);
}

CXM_DEV float bin_freq(int k, int n, float inv_d) {
  // reference f_grid: fl(inv_d * fl(k_signed / n))
  return __fmul_rn(inv_d, __fdiv_rn((float)signed_bin(k, n), (float)n));
}

// The thin-sample factor exp(i 2 pi (dn + i ab) t / lam) = m (cos ang + i sin ang)
// of thin_sample (samples.py:126-133), in the reference's float32 order;
// ang and m only (the caller evaluates the sine and cosine).
CXM_DEV void thin_angle(float dn, float ab, float t, float lam, float* ang, float* m) {
  const float re = __fdiv_rn(__fmul_rn(-__fmul_rn(kTwoPiF, ab), t), lam);
  *ang = __fdiv_rn(__fmul_rn(__fmul_rn(kTwoPiF, dn), t), lam);
  *m = (re == 0.0f) ? 1.0f : expf(re);
}

CXM_DEV float2 apply_op(const OpDesc& o, const OpTable& T, float2 v, int plane, int iy, int ix,
                        int ny, int nx) {
  const int q = (o.pmod > 1) ? (plane / o.pdiv) % o.pmod : 0;
  const float4* P = &T.p[o.pbase + q * o.pstride];
  switch (o.kind) {
    case OP_SCALE:
      return cmul(v, make_float2(P[0].x, P[0].y));
    case OP_CONJ:
      return cconj(v);
    case OP_CHIRP: {
      float s, c;
      const float4 p0 = P[0], p1 = P[1], p2 = P[2];
      if (o.mode == CXM_PHASE_REF32) {
        const float gy = __fsub_rn(grid_coord(iy, ny, p1.x, p1.z), p2.x);
        const float gx = __fsub_rn(grid_coord(ix, nx, p1.y, p1.w), p2.y);
        const float l2 = __fadd_rn(__fmul_rn(gy, gy), __fmul_rn(gx, gx));
        float ph = __fmul_rn(p0.x, l2);
        if (p0.y != 0.0f) ph = __fdiv_rn(ph, p0.y);
        sincos_exact(ph, &s, &c);
      } else {
        // turns = (a_hi + a_lo) * |g - off|^2, g in double-float
        const df gy = df_sub(df_add(df_mul_f(df_make(p1.x, 0.f), (float)iy - 0.5f * (float)ny),
                                    df_make(p1.z, 0.f)), df_make(p2.x, 0.f));
        const df gx = df_sub(df_add(df_mul_f(df_make(p1.y, 0.f), (float)ix - 0.5f * (float)nx),
                                    df_make(p1.w, 0.f)), df_make(p2.y, 0.f));
        const df l2 = df_add(df_mul(gy, gy), df_mul(gx, gx));
        const float tr = df_frac(df_mul(df_make(p0.x, p0.z), l2));
        sincospif(2.0f * tr, &s, &c);
      }
      return cmul(v, make_float2(c, s));
    }
    case OP_CIRC: {
      const float4 p0 = P[0], p1 = P[1], p2 = P[2];
      const float gy = __fsub_rn(grid_coord(iy, ny, p1.x, p1.z), p2.x);
      const float gx = __fsub_rn(grid_coord(ix, nx, p1.y, p1.w), p2.y);
      const float l2 = __fadd_rn(__fmul_rn(gy, gy), __fmul_rn(gx, gx));
      return (l2 <= p0.x) ? v : make_float2(0.f, 0.f);
    }
    case OP_PHASE:
    case OP_THIN:
    case OP_CARRAY:
    case OP_FLUOR: {
      if (o.kind == OP_CARRAY && o.mode == 1) {
        // an array in FFT order at fftshifted coordinates: back to its index
        iy += ny - ny / 2;
        if (iy >= ny) iy -= ny;
        ix += nx - nx / 2;
        if (ix >= nx) ix -= nx;
      }
      const int ay = iy - o.arr_row_off, ax = ix - o.arr_col_off;
      if (ay < 0 || ay >= o.arr_rows || ax < 0 || ax >= o.arr_cols)
        return o.outside_zero ? make_float2(0.f, 0.f) : v;
      const long long idx =
          (long long)plane * o.arr_plane_stride + (long long)ay * o.arr_row_stride + ax;
      if (o.kind == OP_CARRAY) return cmul(v, ((const float2*)o.arr0)[idx]);
      // one sine/cosine site for the three phase-factor operators (another
      // inlined copy tips the pipeline kernels into a stack frame in scratch)
      float ang, m = 1.0f, f = 0.0f;
      if (o.kind == OP_PHASE) {
        ang = __fmul_rn(((const float*)o.arr0)[idx], P[0].x);
      } else if (o.kind == OP_THIN) {
        // thin sample (samples.py:126-133): exp(i 2 pi (dn + i abs) t / lam)
        const float dn = ((const float*)o.arr0)[idx];
        const float ab = o.arr1 ? ((const float*)o.arr1)[idx] : 0.0f;
        thin_angle(dn, ab, o.t, P[0].x, &ang, &m);
      } else {  // fluorescence f exp(i r), added
        f = ((const float*)o.arr0)[idx];
        ang = ((const float*)o.arr1)[idx];
      }
      float sn, cs;
      sincos_exact(ang, &sn, &cs);
      if (o.kind == OP_FLUOR)
        return make_float2(__fadd_rn(__fmul_rn(f, cs), v.x), __fadd_rn(__fmul_rn(f, sn), v.y));
      return cmul(v, make_float2(m * cs, m * sn));
    }
    case OP_VEC: {
      const int i = o.vec_axis ? ix - o.arr_col_off : iy - o.arr_row_off;
      if (i < 0 || i >= o.arr_cols) return o.outside_zero ? make_float2(0.f, 0.f) : v;
      return cmul(v, ((const float2*)o.arr0)[(long long)q * o.arr_plane_stride + i]);
    }
    case OP_NAMASK: {
      // ifftshift(|f_grid|^2 <= (NA/lam)^2) of samples.py:276-279
      const float4 p0 = P[0];
      const float fy = bin_freq(iy, ny, p0.y), fx = bin_freq(ix, nx, p0.z);
      const float l2 = __fadd_rn(__fmul_rn(fy, fy), __fmul_rn(fx, fx));
      return (l2 <= p0.x) ? v : make_float2(0.f, 0.f);
    }
    case OP_SASDH: {
      // compute_sas_precompensation (src/chromatix/functional/propagation.py:85-103),
      // ifftshifted: W * exp(i kz (H_AS - H_Fr)), s = fl(fl(lam f) / n)
      const float4 p0 = P[0], p1 = P[1];
      float s, c;
      if (o.mode != CXM_PHASE_REF32) {
        // FP64 path: s = k * lam / (n d N) per axis (step split hi + lo),
        // turns = (z n / lam) * (H_AS - H_Fr) with the difference written
        // -(q / (1 + sqrt(1 - q)))^2 / 2 to avoid cancellation; for q > 1
        // H_AS is clamped to 0 as in the reference
        const double sy = (double)signed_bin(iy, ny) * ((double)p0.z + (double)p0.w);
        const double sx = (double)signed_bin(ix, nx) * ((double)p1.x + (double)p1.y);
        const double sy2 = sy * sy, sx2 = sx * sx;
        const double ty = (double)p1.z + fabs(sy), tx = (double)p1.w + fabs(sx);
        if (!(sy2 * (2.0 + 1.0 / (ty * ty)) <= 1.0 && sx2 * (2.0 + 1.0 / (tx * tx)) <= 1.0))
          return make_float2(0.f, 0.f);
        const double q = sy2 + sx2;
        double d;
        if (q <= 1.0) {
          const double u = q / (1.0 + sqrt(1.0 - q));
          d = -0.5 * u * u;
        } else {
          d = -(1.0 - 0.5 * q);
        }
        const double tr = ((double)p0.x + (double)p0.y) * d;
        sincospif((float)(2.0 * (tr - rint(tr))), &s, &c);
        return cmul(v, make_float2(c, s));
      }
      const float fy = bin_freq(iy, ny, p1.x), fx = bin_freq(ix, nx, p1.y);
      const float sy = __fdiv_rn(__fmul_rn(p0.z, fy), p0.w);
      const float sx = __fdiv_rn(__fmul_rn(p0.z, fx), p0.w);
      const float sy2 = __fmul_rn(sy, sy), sx2 = __fmul_rn(sx, sx);
      const float ty = __fadd_rn(p1.z, fabsf(sy)), tx = __fadd_rn(p1.w, fabsf(sx));
      const bool wy = __fmul_rn(sy2, __fadd_rn(2.0f, __fdiv_rn(1.0f, __fmul_rn(ty, ty)))) <= 1.0f;
      const bool wx = __fmul_rn(sx2, __fadd_rn(2.0f, __fdiv_rn(1.0f, __fmul_rn(tx, tx)))) <= 1.0f;
      if (!(wy && wx)) return make_float2(0.f, 0.f);
      const float ssum = __fadd_rn(sy2, sx2);
      const float has = sqrt_rn(fmaxf(0.0f, __fsub_rn(1.0f, ssum)));
      const float hfr = __fsub_rn(1.0f, __fdiv_rn(ssum, 2.0f));
      sincos_exact(__fmul_rn(p0.x, __fsub_rn(has, hfr)), &s, &c);
      return cmul(v, make_float2(c, s));
    }
    case OP_IGRAD: {
      const int yy = ((o.shift & 1) ? fshift_dst(iy, ny) : iy) - o.arr_row_off;
      const int xx = ((o.shift & 2) ? fshift_dst(ix, nx) : ix) - o.arr_col_off;
      if (yy < 0 || yy >= o.arr_rows || xx < 0 || xx >= o.arr_cols) return make_float2(0.f, 0.f);
      const long long ai = (long long)plane * o.arr_plane_stride + (long long)yy * o.arr_row_stride + xx;
      if (o.mode == 2) return ((const float2*)o.arr0)[ai];
      const float a = ((const float*)o.arr0)[ai];
      const float g = (o.mode == 1) ? P[0].x * (cabs2(v) - a) : a;
      return cscale(v, 2.0f * g);
    }
    case OP_THINADJ: {
      const int ay = iy - o.arr_row_off, ax = ix - o.arr_col_off;
      if (ay < 0 || ay >= o.arr_rows || ax < 0 || ax >= o.arr_cols) return v;
      const long long idx = (long long)ay * o.arr_row_stride + ax;
      const float dn = ((const float*)o.arr0)[idx];
      const float ab = o.arr1 ? ((const float*)o.arr1)[idx] : 0.0f;
      const float k = P[0].y;
      float im, m, s, c;
      thin_angle(dn, ab, o.t, P[0].x, &im, &m);
      sincos_exact(im, &s, &c);
      const float2 u = ((const float2*)o.arr2)[(long long)plane * o.arr2_plane_stride +
                                                (long long)iy * o.arr2_row_stride + ix];
      const float gdn = k * (v.y * u.x - v.x * u.y);
      float* g0 = (float*)o.out0 + idx;
      *g0 = o.accum ? *g0 + gdn : gdn;
      if (o.out1) {
        const float gab = -k * (v.x * u.x + v.y * u.y);
        float* g1 = (float*)o.out1 + idx;
        *g1 = o.accum ? *g1 + gab : gab;
      }
      return cmul(v, make_float2(m * c, -m * s));
    }
    default:
      return v;
  }
}

// Operators whose factor is always finite map 0 to 0: once a value is 0
// (outside a pupil, in the zero padding) they are skipped, which saves
// their sine/cosine.
CXM_DEV bool maps_zero_to_zero(int kind) {
  return kind == OP_SCALE || kind == OP_CONJ || kind == OP_CHIRP || kind == OP_PHASE ||
         kind == OP_CIRC || kind == OP_SASDH || kind == OP_NAMASK;
}

CXM_DEV float2 apply_ops(const OpList& L, const OpTable& T, float2 v, int plane, int iy, int ix,
                         int ny, int nx) {
  for (int i = 0; i < L.n; ++i) {
    if (v.x == 0.0f && v.y == 0.0f && maps_zero_to_zero(L.op[i].kind)) continue;
    v = apply_op(L.op[i], T, v, plane, iy, ix, ny, nx);
  }
  return v;
}

}  // namespace cxm
