// Generic one-pass FFT pipelines over rows or column strips of a batch of
// planes.  A pass loads one line (zero padding and ifftshift folded into the
// addressing), applies `pre` operators, runs one transform or a transform
// pair with `mid` operators in between, applies `post` operators and stores
// the requested window (fftshift and crop folded into the addressing), as
// complex64 or |u|^2.  Every multi-step operation of the propagation path
// (optical FFT, transform/SAS propagation, chirp-z transforms, the fused
// inverse-multiply-forward step of the multislice loop) is a chain of these
// passes; see the entry points in cxm_pipe.hip.
#pragma once
#include <string>

#include "cxm_fft.h"
#include "cxm_internal.h"
#include "cxm_ops.h"
#include "cxm_sizes.h"

namespace cxm {

enum Seq : int { SEQ_F = 0, SEQ_I = 1, SEQ_FI = 2, SEQ_IF = 3 };

struct PipeArgs {
  const float2* in;           // null: the loaded value is 1 (a generated source)
  void* out;                  // complex64, or float32 when intensity
  const float2* tw;
  long long in_plane_stride, out_plane_stride;  // elements
  int in_line_stride;         // row kernel: elements between input rows; col kernel: between rows too
  int out_line_stride;
  int n_lines;                // rows (row pass) or columns (column pass) per plane
  int n_planes, plane0;       // planes in this launch; global index of the first
  int grid_ny, grid_nx;       // FFT grid for operator coordinates
  int line_off;               // grid index of line 0 (row index / column index)
  int in_off, in_len, in_shift;     // along the transform axis
  int out_off, out_len, out_shift;  // along the transform axis
  int out_line_shift;         // column pass: store grid column c at fshift_dst(c)
  int out_line_off, out_line_len;   // lines kept: grid lines [off, off + len)
  int seq, mid_shift;
  int intensity;              // 0: complex store; 1: |x|^2; 2: out += iscale |x|^2
  float iscale;
  int accum;                  // phase-gradient row pass: add to `out` instead of storing
                              // (gridDim.y > 1: plane group y writes its partial sum to
                              // out + y * out_plane_stride)
  int partial_complex;        // phase-gradient row pass: store the complex plane sum
                              // sum_p g conj(pre(1)) (complex64 out) and leave the
                              // `mid` operators and Im() to the caller's reduction
  // optional store of the value after the `mid` operators (row pass only),
  // grid position (y, x) -> mid_out[plane * mid_plane_stride + (y - mid_y0) *
  // mid_row_stride + (x - mid_x0)] for y - mid_y0 in [0, mid_h), x - mid_x0 in [0, mid_w)
  float2* mid_out;
  long long mid_plane_stride;
  int mid_row_stride, mid_y0, mid_x0, mid_h, mid_w;
  // optional store of |x|^2 of the value entering the `mid` operators
  // (column pass only), at the fftshifted grid position:
  // mid_int[plane * mid_int_plane_stride + fshift(y) * grid_nx + fshift(x)]
  float* mid_int;
  long long mid_int_plane_stride;
  float scale, mid_scale;
  // optional: the pass does nothing once *stop != 0 (a solver's device-side
  // stop flag, so that iterations enqueued past convergence cost no work)
  const double* stop;
  OpList pre, mid, post;
  OpTable tab;
};

// kernels take PipeArgs by value: stay within the 4 KiB kernel-argument limit
static_assert(sizeof(PipeArgs) <= 4096, "PipeArgs exceeds the kernel-argument limit");

// A value the caller knows to be the same in every lane of the wavefront
// (UNIFORM), made visibly so to the compiler: scalar registers, scalar loads
// and branches for what depends on it.
template <bool UNIFORM>
CXM_DEV int wave_uniform(int x) {
  if constexpr (UNIFORM) return __builtin_amdgcn_readfirstlane(x);
  return x;
}

// Store of one output value: complex, |x|^2, or |x|^2 accumulated.
CXM_DEV void store_value(const PipeArgs& a, long long idx, float2 x) {
  if (a.intensity == 2) {
    float* q = (float*)a.out + idx;
    *q = __fadd_rn(*q, __fmul_rn(a.iscale, cabs2(x)));
  } else if (a.intensity) {
    ((float*)a.out)[idx] = cabs2(x);
  } else {
    ((float2*)a.out)[idx] = x;
  }
}

// Rotate the register layout by N/2 (even N, even E): applies ifftshift to
// the sequence held in the v[k] <-> X[t + kT] layout.
template <int E>
CXM_DEV void rotate_half(float2 (&v)[E]) {
  static_assert(E % 2 == 0, "rotate_half needs an even E");
  float2 w[E];
#pragma unroll
  for (int k = 0; k < E; ++k) w[k] = v[(k + E / 2) % E];
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = w[k];
}

// Operator stages.  The operator code (a switch over every operator kind)
// is too large to be unrolled E times, and a rolled loop over the register
// array v[] would put v in scratch memory.  So when a stage has operators,
// each thread parks its E values in its own LDS slots (the slots the FFT's
// first exchange reads), applies the operators in a rolled loop over those
// slots, and takes the values back: one copy of the operator code, v stays
// in registers.  `st` is the transform's LDS base, slot(e) = lds_at<N, S>(e).
template <int N, int E, int S>
CXM_DEV void stage_out(const float2 (&v)[E], float2* st, int t) {
  constexpr int T = N / E;
#pragma unroll
  for (int k = 0; k < E; ++k) st[lds_at_split<N, S>(t, k * T)] = v[k];
}
template <int N, int E, int S>
CXM_DEV void stage_in(float2 (&v)[E], const float2* st, int t) {
  constexpr int T = N / E;
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = st[lds_at_split<N, S>(t, k * T)];
}

template <int N, int E, int S, bool SYNC>
CXM_DEV void run_seq(const PipeArgs& a, float2 (&v)[E], float2* lds, int t, int plane, int line_grid,
                     bool is_row) {
  constexpr int T = N / E;
  if (a.seq == SEQ_F || a.seq == SEQ_FI) fft<N, E, -1, S, SYNC>(v, lds, t, a.tw);
  else fft<N, E, 1, S, SYNC>(v, lds, t, a.tw);
  if (a.seq == SEQ_FI || a.seq == SEQ_IF) {
    if (a.mid.n || (a.mid_out && is_row)) {
      block_sync();  // the FFT's last exchange has been read by every thread
      stage_out<N, E, S>(v, lds, t);
#pragma unroll 1
      for (int k = 0; k < E; ++k) {
        const int j = t + k * T;
        float2* slot = lds + lds_at<N, S>(j);
        float2 x = cscale(*slot, a.mid_scale);
        if (a.mid_int && !is_row)
          a.mid_int[(long long)(plane - a.plane0) * a.mid_int_plane_stride +
                    (long long)fshift_dst(j, a.grid_ny) * a.grid_nx + fshift_dst(line_grid, a.grid_nx)] =
              cabs2(x);
        if (a.mid.n)
          x = is_row ? apply_ops(a.mid, a.tab, x, plane, line_grid, j, a.grid_ny, a.grid_nx)
                     : apply_ops(a.mid, a.tab, x, plane, j, line_grid, a.grid_ny, a.grid_nx);
        if (a.mid_out && is_row) {
          const int yl = line_grid - a.mid_y0, xl = j - a.mid_x0;
          if (yl >= 0 && yl < a.mid_h && xl >= 0 && xl < a.mid_w)
            a.mid_out[(long long)(plane - a.plane0) * a.mid_plane_stride +
                      (long long)yl * a.mid_row_stride + xl] = x;
        }
        *slot = x;
      }
      stage_in<N, E, S>(v, lds, t);
      block_sync();  // before the next FFT scatters into other threads' slots
    } else {
#pragma unroll
      for (int k = 0; k < E; ++k) v[k] = cscale(v[k], a.mid_scale);
    }
    if constexpr (E % 2 == 0) {
      if (a.mid_shift) rotate_half<E>(v);
    }
    if (a.seq == SEQ_FI) fft<N, E, 1, S, SYNC>(v, lds, t, a.tw);
    else fft<N, E, -1, S, SYNC>(v, lds, t, a.tw);
  }
}

// ---- row pass -----------------------------------------------------------------
// Rows of at most one wavefront run 64 threads (one wave) per workgroup:
// their LDS per workgroup is then one row, so more workgroups fit per CU
// than with 256-thread groups of 4 rows (LDS-limited to 3 per CU at 1280
// points), and a workgroup barrier waits for one wave.  11 planes of 1280^2:
// generated-source row pass 187 -> 171 us, phase-gradient pass 308 -> 273 us
// (128 threads: 178 / 280 us).
template <int N, int E>
struct PipeRowCfg {
  static constexpr int T = N / E;
  static constexpr int G = (T >= 64) ? 1 : 64 / T;
  static constexpr int THREADS = G * T;
};

template <int N, int E>
__global__ void __launch_bounds__((PipeRowCfg<N, E>::THREADS))
pipe_row_kernel(PipeArgs a) {
  if (a.stop && *a.stop != 0.0) return;
  using Cfg = PipeRowCfg<N, E>;
  constexpr int T = Cfg::T;
  constexpr int SPAN = lds_span<N, 1>();
  __shared__ float2 lds[Cfg::G * SPAN];
  const int g = threadIdx.x / T;
  const int t = threadIdx.x - g * T;
  const long long line = (long long)blockIdx.x * Cfg::G + g;
  const bool ok = line < (long long)a.n_planes * a.n_lines;
  const int p = ok ? (int)(line / a.n_lines) : 0;
  const int r = ok ? (int)(line - (long long)p * a.n_lines) : 0;
  // whole wavefronts per row (T a multiple of 64): plane and row are uniform
  // per wave, so the operators' parameter lookups stay scalar
  const int plane = wave_uniform<T % 64 == 0>(a.plane0 + p);
  const int yg = wave_uniform<T % 64 == 0>(a.line_off + r);
  const float2* src = a.in ? a.in + p * a.in_plane_stride + (long long)r * a.in_line_stride : nullptr;
  float2* st = lds + g * SPAN;
  float2 v[E];
  if (a.pre.n) {
#pragma unroll 1
    for (int k = 0; k < E; ++k) {
      const int i = t + k * T;
      const int xf = a.in_shift ? ishift_src(i, N) : i;
      const int xs = xf - a.in_off;
      const bool in = ok && xs >= 0 && xs < a.in_len;
      float2 x = in ? (src ? src[xs] : make_float2(1.f, 0.f)) : make_float2(0.f, 0.f);
      if (in) x = apply_ops(a.pre, a.tab, x, plane, yg, xf, a.grid_ny, a.grid_nx);
      st[lds_at<N, 1>(i)] = x;
    }
    stage_in<N, E, 1>(v, st, t);
    block_sync();
  } else {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = t + k * T;
      const int xs = (a.in_shift ? ishift_src(i, N) : i) - a.in_off;
      const bool in = ok && xs >= 0 && xs < a.in_len;
      v[k] = in ? (src ? src[xs] : make_float2(1.f, 0.f)) : make_float2(0.f, 0.f);
    }
  }
  run_seq<N, E, 1, true>(a, v, st, t, plane, yg, true);
  const long long obase = p * a.out_plane_stride + (long long)r * a.out_line_stride;
  if (a.post.n) {
    block_sync();
    stage_out<N, E, 1>(v, st, t);
    if (!ok) return;
#pragma unroll 1
    for (int k = 0; k < E; ++k) {
      const int j = t + k * T;
      const int xo = a.out_shift ? fshift_dst(j, N) : j;
      const int xl = xo - a.out_off;
      if (xl >= 0 && xl < a.out_len) {
        float2 x = cscale(st[lds_at<N, 1>(j)], a.scale);
        x = apply_ops(a.post, a.tab, x, plane, yg, xo, a.grid_ny, a.grid_nx);
        store_value(a, obase + xl, x);
      }
    }
    return;
  }
  if (!ok) return;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int j = t + k * T;
    const int xo = a.out_shift ? fshift_dst(j, N) : j;
    const int xl = xo - a.out_off;
    if (xl >= 0 && xl < a.out_len) {
      const float2 x = cscale(v[k], a.scale);
      store_value(a, obase + xl, x);
    }
  }
}

// ---- phase-gradient row pass ----------------------------------------------------
// Adjoint of a row pass whose input was generated (in == null) and multiplied
// by a phase screen exp(i phi): for every row, every plane p of the launch,
//   g = IFFT_row(in[p, row, :]) * scale          (unnormalised inverse, the
//                                                  adjoint of the forward FFT)
//   out[row, x] += sum_p Im(g[j] * conj(src_p(row, x))),  x = ishift(j) - in_off
// where src_p = the `pre` operators applied to 1, then the `mid` operators,
// i.e. the regenerated input of the forward pass, so Im(g conj(src)) =
// dL/dphi (d src / d phi = i src).  `mid` holds the operators that are the
// same for every plane (all operators multiply): with w = mid(1),
//   sum_p Im(g_p conj(w pre_p(1))) = Im(conj(w) sum_p g_p conj(pre_p(1))),
// so they are evaluated once per element, after the plane sum.  The sum is
// kept in LDS; `accum` adds to an existing out.
template <int N, int E>
__global__ void __launch_bounds__((PipeRowCfg<N, E>::THREADS))
pipe_row_phasegrad_kernel(PipeArgs a) {
  using Cfg = PipeRowCfg<N, E>;
  constexpr int T = Cfg::T;
  constexpr int SPAN = lds_span<N, 1>();
  __shared__ float2 lds[Cfg::G * SPAN];
  const int g = threadIdx.x / T;
  const int t = threadIdx.x - g * T;
  const int r = blockIdx.x * Cfg::G + g;
  const bool ok = r < a.n_lines;
  const int yg = wave_uniform<T % 64 == 0>(a.line_off + (ok ? r : 0));
  float2* st = lds + g * SPAN;
  // the plane sum of g conj(pre(1)) stays in registers, element t + k T in
  // acc[k]; each plane's products pass through this thread's own LDS slots
  // (the operator loop is rolled), so LDS holds one row per thread group
  float2 acc[E];
#pragma unroll
  for (int k = 0; k < E; ++k) acc[k] = make_float2(0.0f, 0.0f);
  const int per = (a.n_planes + (int)gridDim.y - 1) / (int)gridDim.y;
  const int p_lo = (int)blockIdx.y * per;
  const int p_hi = min(a.n_planes, p_lo + per);
  for (int p = p_lo; p < p_hi; ++p) {
    const float2* src = a.in + p * a.in_plane_stride + (long long)(ok ? r : 0) * a.in_line_stride;
    float2 v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) v[k] = ok ? src[t + k * T] : make_float2(0.f, 0.f);
    fft<N, E, 1, 1, true>(v, st, t, a.tw);
    block_sync();
    stage_out<N, E, 1>(v, st, t);
#pragma unroll 1
    for (int k = 0; k < E; ++k) {
      const int j = t + k * T;
      const int xf = a.in_shift ? ishift_src(j, N) : j;
      const int xs = xf - a.in_off;
      float2 prod = make_float2(0.f, 0.f);
      if (ok && xs >= 0 && xs < a.in_len) {
        const float2 gv = cscale(st[lds_at<N, 1>(j)], a.scale);
        const float2 sv = apply_ops(a.pre, a.tab, make_float2(1.f, 0.f), a.plane0 + p, yg, xf,
                                    a.grid_ny, a.grid_nx);
        prod = make_float2(gv.x * sv.x + gv.y * sv.y, gv.y * sv.x - gv.x * sv.y);
      }
      st[lds_at<N, 1>(j)] = prod;
    }
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const float2 q = st[lds_at<N, 1>(t + k * T)];
      acc[k] = make_float2(acc[k].x + q.x, acc[k].y + q.y);
    }
    block_sync();  // before the next plane's FFT reuses the slots
  }
  if (!ok) return;
  if (a.partial_complex) {
    float2* cdst = (float2*)a.out + (long long)blockIdx.y * a.out_plane_stride +
                   (long long)r * a.out_line_stride;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int j = t + k * T;
      const int xs = (a.in_shift ? ishift_src(j, N) : j) - a.in_off;
      if (xs >= 0 && xs < a.in_len) cdst[xs] = acc[k];
    }
    return;
  }
  // own slots only from here on
#pragma unroll
  for (int k = 0; k < E; ++k) st[lds_at<N, 1>(t + k * T)] = acc[k];
  float* dst = (float*)a.out + (long long)blockIdx.y * a.out_plane_stride +
               (long long)r * a.out_line_stride;
  const bool add = a.accum && gridDim.y == 1;
#pragma unroll 1
  for (int k = 0; k < E; ++k) {
    const int j = t + k * T;
    const int xf = a.in_shift ? ishift_src(j, N) : j;
    const int xs = xf - a.in_off;
    if (xs >= 0 && xs < a.in_len) {
      const float2 c = st[lds_at<N, 1>(j)];
      const float2 w = a.mid.n ? apply_ops(a.mid, a.tab, make_float2(1.f, 0.f), a.plane0, yg, xf,
                                           a.grid_ny, a.grid_nx)
                               : make_float2(1.f, 0.f);
      const float v = w.x * c.y - w.y * c.x;  // Im(conj(w) c)
      dst[xs] = add ? dst[xs] + v : v;
    }
  }
}

// ---- column pass --------------------------------------------------------------
template <int N, int E, int C>
struct PipeColCfg {
  static constexpr int T = N / E;
  static constexpr int THREADS = T * C;
};

template <int N, int E, int C>
__global__ void __launch_bounds__((PipeColCfg<N, E, C>::THREADS))
pipe_col_kernel(PipeArgs a) {
  if (a.stop && *a.stop != 0.0) return;
  constexpr int T = N / E;
  __shared__ float2 lds[lds_span<N, C>()];
  const int c = threadIdx.x % C;
  const int ty = threadIdx.x / C;
  const int nstrips = (a.n_lines + C - 1) / C;
  const int nb = gridDim.x;
  const int L = ((nb & 7) == 0) ? (int)((blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3)) : (int)blockIdx.x;
  const int p = L / nstrips;
  const int col = (L - p * nstrips) * C + c;
  const bool colok = col < a.n_lines;
  const int plane = a.plane0 + p;
  const int xg = a.line_off + (colok ? col : 0);
  const float2* src = a.in ? a.in + p * a.in_plane_stride + col : nullptr;
  float2* st = lds + c;
  float2 v[E];
  if (a.pre.n) {
#pragma unroll 1
    for (int k = 0; k < E; ++k) {
      const int i = ty + k * T;
      const int yf = a.in_shift ? ishift_src(i, N) : i;
      const int ys = yf - a.in_off;
      const bool in = colok && ys >= 0 && ys < a.in_len;
      float2 x = in ? (src ? src[(long long)ys * a.in_line_stride] : make_float2(1.f, 0.f))
                    : make_float2(0.f, 0.f);
      if (in) x = apply_ops(a.pre, a.tab, x, plane, yf, xg, a.grid_ny, a.grid_nx);
      st[lds_at<N, C>(i)] = x;
    }
    stage_in<N, E, C>(v, st, ty);
    block_sync();
  } else {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = ty + k * T;
      const int ys = (a.in_shift ? ishift_src(i, N) : i) - a.in_off;
      const bool in = colok && ys >= 0 && ys < a.in_len;
      v[k] = in ? (src ? src[(long long)ys * a.in_line_stride] : make_float2(1.f, 0.f))
                : make_float2(0.f, 0.f);
    }
  }
  run_seq<N, E, C, true>(a, v, st, ty, plane, xg, false);
  const int xo = a.out_line_shift ? fshift_dst(xg, a.grid_nx) : xg;
  const int xl = xo - a.out_line_off;
  const bool store = colok && xl >= 0 && xl < a.out_line_len;
  const long long obase = p * a.out_plane_stride + xl;
  if (a.post.n) {
    block_sync();
    stage_out<N, E, C>(v, st, ty);
    if (!store) return;
#pragma unroll 1
    for (int k = 0; k < E; ++k) {
      const int j = ty + k * T;
      const int yo = a.out_shift ? fshift_dst(j, N) : j;
      const int yl = yo - a.out_off;
      if (yl >= 0 && yl < a.out_len) {
        float2 x = cscale(st[lds_at<N, C>(j)], a.scale);
        x = apply_ops(a.post, a.tab, x, plane, yo, xo, a.grid_ny, a.grid_nx);
        store_value(a, obase + (long long)yl * a.out_line_stride, x);
      }
    }
    return;
  }
  if (!store) return;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int j = ty + k * T;
    const int yo = a.out_shift ? fshift_dst(j, N) : j;
    const int yl = yo - a.out_off;
    if (yl >= 0 && yl < a.out_len) {
      const float2 x = cscale(v[k], a.scale);
      store_value(a, obase + (long long)yl * a.out_line_stride, x);
    }
  }
}

template <int N, int E>
int launch_pipe_row_t(const PipeArgs& a, hipStream_t s) {
  using Cfg = PipeRowCfg<N, E>;
  const long long lines = (long long)a.n_planes * a.n_lines;
  const long long blocks = (lines + Cfg::G - 1) / Cfg::G;
  if (blocks == 0) return CXM_OK;
  if (blocks > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "row pass grid too large");
  void* t0 = timing_start(s);
  hipLaunchKernelGGL((pipe_row_kernel<N, E>), dim3((unsigned)blocks), dim3(Cfg::THREADS), 0, s, a);
  CXM_LAUNCH_CHECK();
  if (t0) timing_stop(t0, s, "pipe_row_" + std::to_string(N));
  return CXM_OK;
}

template <int N, int E>
int launch_pipe_phasegrad_t(const PipeArgs& a, hipStream_t s, int groups) {
  using Cfg = PipeRowCfg<N, E>;
  const long long blocks = (a.n_lines + Cfg::G - 1) / Cfg::G;
  if (blocks == 0 || a.n_planes == 0) return CXM_OK;
  void* t0 = timing_start(s);
  hipLaunchKernelGGL((pipe_row_phasegrad_kernel<N, E>), dim3((unsigned)blocks, (unsigned)groups),
                     dim3(Cfg::THREADS), 0, s, a);
  CXM_LAUNCH_CHECK();
  if (t0) timing_stop(t0, s, "pipe_phasegrad_" + std::to_string(N));
  return CXM_OK;
}

template <int N, int E, int C>
int launch_pipe_col_t(const PipeArgs& a, hipStream_t s) {
  using Cfg = PipeColCfg<N, E, C>;
  const long long nstrips = (a.n_lines + C - 1) / C;
  const long long blocks = nstrips * a.n_planes;
  if (blocks == 0) return CXM_OK;
  if (blocks > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "column pass grid too large");
  void* t0 = timing_start(s);
  hipLaunchKernelGGL((pipe_col_kernel<N, E, C>), dim3((unsigned)blocks), dim3(Cfg::THREADS), 0, s, a);
  CXM_LAUNCH_CHECK();
  if (t0) timing_stop(t0, s, "pipe_col_" + std::to_string(N));
  return CXM_OK;
}

}  // namespace cxm

#define CXM_DEFINE_PIPE_GROUP(K, SIZES)                                       \
  namespace cxm {                                                             \
  int launch_pipe_row_g##K(int N, const PipeArgs& a, hipStream_t s) {         \
    switch (N) {                                                              \
      SIZES(CXM_PROW_CASE)                                                    \
      default: return -1;                                                     \
    }                                                                         \
  }                                                                           \
  int launch_pipe_col_g##K(int N, const PipeArgs& a, hipStream_t s) {         \
    switch (N) {                                                              \
      SIZES(CXM_PCOL_CASE)                                                    \
      default: return -1;                                                     \
    }                                                                         \
  }                                                                           \
  int launch_pipe_phasegrad_g##K(int N, const PipeArgs& a, hipStream_t s, int groups) { \
    switch (N) {                                                              \
      SIZES(CXM_PGRAD_CASE)                                                   \
      default: return -1;                                                     \
    }                                                                         \
  }                                                                           \
  }
#define CXM_PROW_CASE(NN, ER, EC, CC) case NN: return launch_pipe_row_t<NN, ER>(a, s);
#define CXM_PCOL_CASE(NN, ER, EC, CC) case NN: return launch_pipe_col_t<NN, EC, CC>(a, s);
#define CXM_PGRAD_CASE(NN, ER, EC, CC) case NN: return launch_pipe_phasegrad_t<NN, ER>(a, s, groups);
