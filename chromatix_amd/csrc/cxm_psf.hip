// 4f widefield PSF stack (objective point source -> phase mask in the
// pupil plane -> tube-lens far-field FFT -> intensity), its gradient with
// respect to the phase mask, and the two elementwise building blocks
// (objective_point_source, phase_change) on their own.
//
// Forward, per defocus plane (two passes, nothing materialised but one
// row-transformed chunk):
//   row pass:    source generated in registers (defocus chirp, pupil,
//                amplitude / power normalisation, phase screen, lens norm),
//                ifftshift folded into the positions, row FFT -> workspace
//   column pass: column FFT, fftshift folded into the store, |u|^2 or u
// Backward (three passes):
//   row pass again -> workspace
//   column pass: FFT, cotangent 2 dL/dI u at the output position (dL/dI
//                read, or formed from a target for the mean-squared error),
//                inverse FFT (the adjoint), in place
//   row pass:    inverse FFT (adjoint), times conj(regenerated source),
//                imaginary part summed over the defocus planes in registers
//                -> dL/dphi
// Reference: Optical4FSystemPSF.__call__ (src/chromatix/systems/microscopes.py:334-355)
// composing objective_point_source (src/chromatix/functional/sources.py:122-209),
// phase_change (src/chromatix/functional/phase_masks.py:29-68) and ff_lens
// (src/chromatix/functional/lenses.py:64-95).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "cxm_pipe_host.h"

namespace cxm {
namespace {

constexpr int kMaxZ = 32;

int check_psf(const cxm_psf4f_desc* d) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->height <= 0 || d->width <= 0) return fail(CXM_ERR_INVALID, "invalid grid shape");
  if (d->num_z < 1 || d->num_z > kMaxZ)
    return fail(CXM_ERR_INVALID, "num_z must be in [1, " + std::to_string(kMaxZ) + "]");
  if (!d->z) return fail(CXM_ERR_INVALID, "null z list");
  if (d->phase_mode != CXM_PHASE_REF32 && d->phase_mode != CXM_PHASE_ACCURATE)
    return fail(CXM_ERR_INVALID, "invalid phase mode");
  if (!(d->dy > 0 && d->dx > 0 && d->wavelength > 0 && d->n > 0 && d->f > 0 && d->NA > 0))
    return fail(CXM_ERR_INVALID, "spacing, wavelength, n, f and NA must be positive");
  if (d->normalize && !(d->power >= 0)) return fail(CXM_ERR_INVALID, "power must be >= 0");
  return CXM_OK;
}

// Samples inside the circular pupil |g|^2 <= r2, evaluated with the float32
// arithmetic of OP_CIRC so the count matches the kernel's mask exactly.
long long pupil_count(int ny, int nx, float dy, float dx, float r2) {
  static std::mutex mu;
  static int cny = -1, cnx = -1;
  static float cdy = 0, cdx = 0, cr2 = 0;
  static long long cm = 0;
  std::lock_guard<std::mutex> lk(mu);
  if (ny == cny && nx == cnx && dy == cdy && dx == cdx && r2 == cr2) return cm;
  long long m = 0;
  for (int iy = 0; iy < ny; ++iy) {
    const float gy = dy * ((float)iy - 0.5f * (float)ny) + 0.0f;
    const float gy2 = gy * gy;
    for (int ix = 0; ix < nx; ++ix) {
      const float gx = dx * ((float)ix - 0.5f * (float)nx) + 0.0f;
      if (gy2 + gx * gx <= r2) ++m;
    }
  }
  cny = ny; cnx = nx; cdy = dy; cdx = dx; cr2 = r2; cm = m;
  return m;
}

// The point source of objective_point_source (sources.py:196-208) as
// operators: CHIRP (defocus phase -pi (z/f) |g - offset|^2 / L^2, one slot
// per plane), CIRC (pupil of diameter 2 f NA / n, pupils.py:17-23) and a
// constant SCALE, -i A / L^2 times the power normalisation
// sqrt(P / (dy dx sum |u|^2)); |u| = |A| / L^2 inside the pupil, so the sum
// is M A^2 / L^4 for M pupil samples and the scale is independent of z.
// `extra` multiplies the scale (the tube-lens norm of the fused PSF).
int source_ops(Ops& B, OpList& L, const cxm_psf4f_desc* d, float2 extra) {
  const int nz = d->num_z;
  const int mode = d->phase_mode;
  // the pupil first: outside it the value is 0 and the chirp, scale and
  // phase operators are skipped (the product is the same, the mask is exact)
  OpDesc* ci = B.add(L, OP_CIRC, 1, 3);
  OpDesc* ch = B.add(L, OP_CHIRP, nz, 3, mode);
  OpDesc* sc = B.add(L, OP_SCALE, 1, 1);
  if (!ch || !ci || !sc) return fail(CXM_ERR_INVALID, "operator table overflow");
  const float dyf = (float)d->dy, dxf = (float)d->dx;
  // float32 chain of the reference: L = sqrt(lam f / n), L^2 = L * L
  const float Lf = std::sqrt(((float)d->wavelength * (float)d->f) / (float)d->n);
  const float L2f = Lf * Lf;
  const double L2 = d->wavelength * d->f / d->n;
  for (int j = 0; j < nz; ++j) {
    float4& p0 = B.p(ch, j, 0);
    if (mode == CXM_PHASE_REF32) {
      p0 = make_float4((float)(-M_PI) * ((float)d->z[j] / (float)d->f), L2f, 0.f, 0.f);
    } else {
      float hi, lo;
      split(-(d->z[j] / d->f) / (2.0 * L2), hi, lo);
      p0 = make_float4(hi, 0.f, lo, 0.f);
    }
    B.p(ch, j, 1) = make_float4(dyf, dxf, 0.f, 0.f);
    B.p(ch, j, 2) = make_float4((float)d->offset_y, (float)d->offset_x, 0.f, 0.f);
  }
  const float r = (float)(2.0 * d->f * d->NA / d->n) / 2.0f;
  const float r2 = r * r;
  B.p(ci, 0, 0) = make_float4(r2, 0.f, 0.f, 0.f);
  B.p(ci, 0, 1) = make_float4(dyf, dxf, 0.f, 0.f);
  B.p(ci, 0, 2) = make_float4(0.f, 0.f, 0.f, 0.f);
  double amp;  // imaginary part of the (purely imaginary) source amplitude
  if (d->normalize) {
    const long long m = pupil_count(d->height, d->width, dyf, dxf, r2);
    if (m == 0) return fail(CXM_ERR_INVALID, "the pupil contains no sample: power is undefined");
    const double sgn = d->amplitude < 0 ? -1.0 : 1.0;
    amp = -sgn * std::sqrt(d->power / (d->dy * d->dx * (double)m));
  } else {
    amp = -d->amplitude / L2;
  }
  // (i amp) * extra
  B.p(sc, 0, 0) = make_float4((float)(-amp * extra.y), (float)(amp * extra.x), 0.f, 0.f);
  return CXM_OK;
}

// Tube-lens norm of ff_lens / optical_fft for z = f_tube > 0:
// -i dy dx / (lam f_tube / n) (convenience.py:43-48).
float2 tube_norm(const cxm_psf4f_desc* d) {
  if (d->phase_mode == CXM_PHASE_REF32) {
    const float L2t = ((float)d->wavelength * (float)d->f_tube) / (float)d->n;
    return make_float2(0.f, -(((float)d->dy * (float)d->dx) / L2t));
  }
  return make_float2(0.f, (float)(-(d->dy * d->dx) / (d->wavelength * d->f_tube / d->n)));
}

float phase_scale(const cxm_psf4f_desc* d) { return d->phase_scale == 0.0 ? 1.0f : (float)d->phase_scale; }

int psf_ops(Ops& B, PipeArgs& ra, const cxm_psf4f_desc* d, const float* phase) {
  int rc = source_ops(B, ra.pre, d, tube_norm(d));
  if (rc) return rc;
  OpDesc* ph = B.add(ra.pre, OP_PHASE, 1, 1);
  if (!ph) return fail(CXM_ERR_INVALID, "operator table overflow");
  ph->arr0 = phase;
  ph->arr_plane_stride = 0;
  ph->arr_row_stride = d->width;
  ph->arr_rows = d->height;
  ph->arr_cols = d->width;
  // the phase in float32 as fl(phase * m), m the spectral modulation of the
  // wavelength (phase_masks.py:105-124; 0 in the descriptor means 1)
  B.p(ph, 0, 0) = make_float4(phase_scale(d), 0.f, 0.f, 0.f);
  return CXM_OK;
}

// Planes per chunk: at most 160 MiB of complex64 intermediates (they stay in
// the 256 MiB Infinity Cache), split evenly so that no chunk is a lone
// latency-bound plane (11 planes of 1280^2: one chunk, not 10 + 1).
long long psf_chunk(const cxm_psf4f_desc* d) {
  const long long per = (long long)d->height * d->width;
  const long long budget = (160ll << 20) / 8;
  const long long cap = std::max(1ll, std::min<long long>(d->num_z, budget / std::max(1ll, per)));
  const long long n = (d->num_z + cap - 1) / cap;
  return (d->num_z + n - 1) / n;
}

void row_args(PipeArgs& ra, const cxm_psf4f_desc* d) {
  ra.tw = twiddles_for(d->width);
  ra.in = nullptr;  // generated source
  ra.out_plane_stride = (long long)d->height * d->width;
  ra.out_line_stride = d->width;
  ra.n_lines = d->height;
  ra.grid_ny = d->height;
  ra.grid_nx = d->width;
  ra.in_len = d->width;
  ra.in_shift = 1;
  ra.out_len = d->width;
  ra.seq = SEQ_F;
  ra.scale = 1.0f;
  ra.mid_scale = 1.0f;
}

void col_args(PipeArgs& ca, const cxm_psf4f_desc* d) {
  ca.tw = twiddles_for(d->height);
  ca.in_plane_stride = (long long)d->height * d->width;
  ca.in_line_stride = d->width;
  ca.out_plane_stride = (long long)d->height * d->width;
  ca.out_line_stride = d->width;
  ca.n_lines = d->width;
  ca.grid_ny = d->height;
  ca.grid_nx = d->width;
  ca.in_len = d->height;
  ca.in_shift = 1;
  ca.out_shift = 1;
  ca.out_len = d->height;
  ca.out_line_len = d->width;
  ca.seq = SEQ_F;
  ca.scale = 1.0f;
  ca.mid_scale = 1.0f;
}

int psf_geom(const cxm_psf4f_desc* d) {
  int rc = check_psf(d);
  if (rc) return rc;
  if (!(d->f_tube > 0)) return fail(CXM_ERR_INVALID, "f_tube must be positive");
  if (!fft_len_supported(d->height) || !fft_len_supported(d->width))
    return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT grid " + std::to_string(d->height) + "x" +
                                         std::to_string(d->width));
  return CXM_OK;
}

// ---- dedicated passes over tables -------------------------------------------
// The generated source of plane j at grid position (y, x) is
//   src_j(y, x) = P(y, x) c_j(y, x),
// P the plane-invariant factor (pupil, amplitude / power scale, tube-lens
// norm, phase mask exp(i phase)) and c_j the defocus chirp
// exp(i a_j |g - o|^2 / L^2).  P is evaluated once per call into a table by
// the pipeline's own operators; the chirp is a product of a row factor
// cy_j(y) = exp(i a_j (gy - oy)^2 / L^2) and a column factor cx_j(x) (equal in
// real arithmetic; in float32 each factor's phase is rounded on its own, as
// the per-element phase is).  A pass then costs two complex products per
// value instead of a pupil test, two sine/cosine pairs and the operator
// stage.
struct PsfTables {
  float2* P;   // height x width
  float2* cy;  // num_z x height
  float2* cx;  // num_z x width
};

// One axis of the CHIRP operator (cxm_ops.h OP_CHIRP, parameters of plane
// slot q): exp(i a (g - o)^2 / b) (ref32) or exp(2 pi i frac((a_hi + a_lo)
// (g - o)^2)) (double-float turns), g the axis' grid coordinate.
CXM_DEV float2 chirp_axis(const OpDesc& o, const OpTable& T, int q, int i, int n, int axis) {
  const float4* P = &T.p[o.pbase + q * o.pstride];
  const float4 p0 = P[0], p1 = P[1], p2 = P[2];
  const float d = axis == 0 ? p1.x : p1.y, org = axis == 0 ? p1.z : p1.w, off = axis == 0 ? p2.x : p2.y;
  float s, c;
  if (o.mode == CXM_PHASE_REF32) {
    const float g = __fsub_rn(grid_coord(i, n, d, org), off);
    float ph = __fmul_rn(p0.x, __fmul_rn(g, g));
    if (p0.y != 0.0f) ph = __fdiv_rn(ph, p0.y);
    sincos_exact(ph, &s, &c);
  } else {
    const df g = df_sub(df_add(df_mul_f(df_make(d, 0.f), (float)i - 0.5f * (float)n), df_make(org, 0.f)),
                        df_make(off, 0.f));
    const float tr = df_frac(df_mul(df_make(p0.x, p0.z), df_mul(g, g)));
    sincospif(2.0f * tr, &s, &c);
  }
  return make_float2(c, s);
}

// cy[j][y] (i < ny) and cx[j][x] (i >= ny) for plane slot j = blockIdx.y
__global__ void psf_chirp_kernel(float2* __restrict__ cy, float2* __restrict__ cx, int ny, int nx,
                                 OpDesc ch, OpTable tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < ny) cy[(long long)j * ny + i] = chirp_axis(ch, tab, j, i, ny, 0);
  else if (i < ny + nx) cx[(long long)j * nx + i - ny] = chirp_axis(ch, tab, j, i - ny, nx, 1);
}

// P(y, x) = the plane-invariant operators applied to 1 (grid order)
__global__ void psf_p_kernel(float2* __restrict__ P, int ny, int nx, OpList L, OpTable T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)ny * nx) return;
  const int iy = (int)(i / nx), ix = (int)(i - (long long)iy * nx);
  P[i] = apply_ops(L, T, make_float2(1.f, 0.f), 0, iy, ix, ny, nx);
}

struct PsfRowArgs {
  PsfTables t;
  const float2* tw;
  float2* out;  // chunk planes (row-transformed)
  long long plane_stride;
  int ny, plane0;
};

// Row pass: src_j at (r, ishift(i)) for the row's positions i, row FFT.  One
// row per workgroup of N / E threads.
template <int N, int E>
__global__ void __launch_bounds__(N / E) psf_row_kernel(PsfRowArgs a) {
  constexpr int T = N / E;
  constexpr bool SYNC = T > 64;
  __shared__ float2 lds[lds_span<N, 1>()];
  const int p = blockIdx.x / a.ny;
  const int r = blockIdx.x - p * a.ny;
  const int pg = a.plane0 + p;
  const int t = threadIdx.x;
  const float2 fy = a.t.cy[(long long)pg * a.ny + r];
  const float2* Prow = a.t.P + (long long)r * N;
  const float2* cxp = a.t.cx + (long long)pg * N;
  float2 v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int xf = ishift_src(t + k * T, N);
    v[k] = cmul(cmul(Prow[xf], cxp[xf]), fy);
  }
  fft<N, E, -1, 1, SYNC>(v, lds, t, a.tw);
  float2* dst = a.out + p * a.plane_stride + (long long)r * N;
#pragma unroll
  for (int k = 0; k < E; ++k) dst[t + k * T] = v[k];
}

struct PsfGradArgs {
  PsfTables t;
  const float2* tw;
  const float2* in;  // chunk planes after the column adjoint
  float2* partial;   // groups x height x width
  float* grad;       // one group: the gradient itself, m Im(conj(P) sum)
  int accum;         // one group: add to grad
  float m;           // the phase's spectral modulation (d/dphase = m d/d(m phase))
  long long plane_stride;
  int ny, plane0, n_planes;
};

// Phase-gradient row pass: for row r and the planes of group blockIdx.y,
// g_p = IFFT_row(in[p, r, :]) (the adjoint of the forward row FFT) and
//   sum[x] = sum_p g_p(x) conj(c_p(r, x)),  x = ishift(i),
// the plane-invariant factor applied once per element to the sum.  W row
// transforms per workgroup: transform w takes planes pa + w, pa + w + W, ...
// and keeps its sum in registers; the W sums meet in LDS and are added in
// the order w = 0 .. W-1.  Every transform group runs the same number of
// transforms (the last with zeros where it has no plane), so the
// block-wide barriers of transforms longer than a wave match.
template <int N, int E, int W>
__global__ void __launch_bounds__((N / E) * W) psf_grad_kernel(PsfGradArgs a) {
  constexpr int T = N / E;
  // a transform of one wave exchanges within its own LDS region; only the
  // final reduction of the W sums needs the whole workgroup
  constexpr bool SYNC = T > 64;
  constexpr int SPAN = lds_span<N, 1>();
  static_assert(W * SPAN >= W * N, "the partial sums reuse the transforms' LDS");
  __shared__ float2 lds[W * SPAN];
  const int r = blockIdx.x;
  const int grp = blockIdx.y, groups = gridDim.y;
  const int pa = (int)((long long)grp * a.n_planes / groups);
  const int pb = (int)((long long)(grp + 1) * a.n_planes / groups);
  const int w = threadIdx.x / T;
  const int t = threadIdx.x - w * T;
  float2* st = lds + w * SPAN;
  float2 acc[E];
#pragma unroll
  for (int k = 0; k < E; ++k) acc[k] = make_float2(0.f, 0.f);
  const int rounds = (pb - pa + W - 1) / W;
  for (int i = 0; i < rounds; ++i) {
    // the thread index laundered per plane: the transform's twiddle loads
    // are then not hoisted out of the loop (hoisted, every stage's twiddles
    // stay live next to the plane sum: 256 VGPRs, one wave per SIMD)
    int tl = t;
    asm volatile("" : "+v"(tl));
    const int p = pa + i * W + w;
    const bool have = p < pb;
    const int pg = a.plane0 + (have ? p : pa);
    const float2* src = a.in + (have ? p : pa) * a.plane_stride + (long long)r * N;
    float2 v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) v[k] = have ? src[tl + k * T] : make_float2(0.f, 0.f);
    fft<N, E, 1, 1, SYNC>(v, st, tl, a.tw);
    if (have) {
      const float2 fy = a.t.cy[(long long)pg * a.ny + r];
      const float2* cxp = a.t.cx + (long long)pg * N;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const float2 c = cmul(cxp[ishift_src(tl + k * T, N)], fy);
        acc[k] = cadd(acc[k], cmulc(v[k], c));
      }
    }
  }
  float2* part = lds;  // W partial sums in grid order, over the transforms' LDS
  if constexpr (W > 1) {
    block_sync();  // every transform's last exchange has been read
#pragma unroll
    for (int k = 0; k < E; ++k) part[w * N + ishift_src(t + k * T, N)] = acc[k];
    block_sync();
  }
  const float2* Prow = a.t.P + (long long)r * N;
  float* g = a.grad + (long long)r * N;
  float2* dst = a.partial + (long long)grp * a.ny * N + (long long)r * N;
  // element x of the row: thread threadIdx.x takes x = threadIdx.x + j * (T W)
#pragma unroll
  for (int j = 0; j < E; ++j) {
    float2 c;
    int xf;
    if constexpr (W > 1) {
      xf = threadIdx.x + j * T * W;
      if (xf >= N) break;
      c = part[xf];
#pragma unroll
      for (int u = 1; u < W; ++u) c = cadd(c, part[u * N + xf]);
    } else {
      xf = ishift_src(t + j * T, N);
      c = acc[j];
    }
    if (groups == 1) {
      // the plane-invariant factor applied here: grad (+)= m Im(conj(P) sum)
      const float2 q = Prow[xf];
      const float v = a.m * (q.x * c.y - q.y * c.x);
      g[xf] = a.accum ? g[xf] + v : v;
    } else {
      dst[xf] = c;
    }
  }
}

// grad[i] (+)= Im(conj(P[i]) sum_g partial[g][i]), g in order
__global__ void psf_sum_kernel(const float2* __restrict__ partial, const float2* __restrict__ P,
                               float* __restrict__ grad, long long n, int groups, int accum, float m) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float2 c = make_float2(0.f, 0.f);
  for (int g = 0; g < groups; ++g) c = cadd(c, partial[g * n + i]);
  const float2 w = P[i];
  const float v = m * (w.x * c.y - w.y * c.x);
  grad[i] = accum ? grad[i] + v : v;
}

struct PsfColArgs {
  float2* ws;  // chunk planes, row-transformed
  const float2* tw;
  long long plane_stride;
  int nx, plane0;
  void* out;        // forward: chunk output (fftshifted), float32 |u|^2 or complex64
  int intensity;
  const float* cot; // backward: dL/dI or the target, all planes (fftshifted)
  float* inten;     // backward: optional forward intensity of the chunk
  int mse;
  float p0;         // MSE: 2 / element count
};

// Column pass on strips of C columns (transform length N = height): the
// column FFT (ifftshift folded into the load); forward: |u|^2 or u stored at
// the fftshifted position; backward: at that position the cotangent of u,
// 2 g u with g = dL/dI or 2 (|u|^2 - target) / count (the intensity stored
// on the way when asked), then the inverse column FFT back into the strip.
// The arithmetic of the pipeline's column pass with OP_IGRAD, in registers.
template <int N, int E, int C, bool BWD>
__global__ void __launch_bounds__((N / E) * C) psf_col_kernel(PsfColArgs a) {
  constexpr int T = N / E;
  __shared__ float2 lds[lds_span<N, C>()];
  const int c = threadIdx.x % C;
  const int ty = threadIdx.x / C;
  const int nstrips = a.nx / C;
  const int nb = gridDim.x;
  const int L = ((nb & 7) == 0) ? (int)((blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3)) : (int)blockIdx.x;
  const int p = L / nstrips;
  const int col = (L - p * nstrips) * C + c;
  float2* colp = a.ws + p * a.plane_stride + col;
  const int xo = fshift_dst(col, a.nx);
  float2 v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = colp[(long long)ishift_src(ty + k * T, N) * a.nx];
  // backward: the cotangents are loaded with the strip, so that their
  // latency overlaps the forward transform instead of following it
  float cv[BWD ? E : 1];
  if constexpr (BWD) {
    const long long cbase = (long long)(a.plane0 + p) * a.plane_stride + xo;
#pragma unroll
    for (int k = 0; k < E; ++k) cv[k] = a.cot[cbase + (long long)fshift_dst(ty + k * T, N) * a.nx];
  }
  fft<N, E, -1, C, true>(v, lds + c, ty, a.tw);
  if constexpr (!BWD) {
    const long long obase = p * a.plane_stride + xo;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const long long idx = obase + (long long)fshift_dst(ty + k * T, N) * a.nx;
      if (a.intensity) ((float*)a.out)[idx] = cabs2(v[k]);
      else ((float2*)a.out)[idx] = v[k];
    }
  } else {
    const long long ibase = p * a.plane_stride + xo;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const long long yo = (long long)fshift_dst(ty + k * T, N) * a.nx;
      const float I = cabs2(v[k]);
      if (a.inten) a.inten[ibase + yo] = I;
      const float g = a.mse ? a.p0 * (I - cv[k]) : cv[k];
      v[k] = cscale(v[k], 2.0f * g);
    }
    // the inverse transform's twiddle loads stay after the cotangents
    int tyl = ty;
    asm volatile("" : "+v"(tyl));
    fft<N, E, 1, C, true>(v, lds + c, tyl, a.tw);
#pragma unroll
    for (int k = 0; k < E; ++k) colp[(long long)fshift_dst(ty + k * T, N) * a.nx] = v[k];
  }
}

constexpr int kPsfColC = 4;  // columns per column-pass workgroup

// (row length, values per thread) of the dedicated passes: whole wavefronts
// per row
#define CXM_PSF_SIZES(X) X(512, 8) X(1024, 16) X(1280, 20) X(1536, 24) X(2048, 16)

int launch_psf_col(int n, bool bwd, const PsfColArgs& a, int planes, hipStream_t s) {
  const long long blocks = (long long)planes * (a.nx / kPsfColC);
  if (blocks > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "column pass grid too large");
  switch (n) {
#define CXM_PS_COL(N, E)                                                                          \
  case N:                                                                                         \
    if (bwd)                                                                                      \
      hipLaunchKernelGGL((psf_col_kernel<N, E, kPsfColC, true>), dim3((unsigned)blocks),          \
                         dim3((N / E) * kPsfColC), 0, s, a);                                      \
    else                                                                                          \
      hipLaunchKernelGGL((psf_col_kernel<N, E, kPsfColC, false>), dim3((unsigned)blocks),         \
                         dim3((N / E) * kPsfColC), 0, s, a);                                      \
    CXM_LAUNCH_CHECK();                                                                           \
    return CXM_OK;
    CXM_PSF_SIZES(CXM_PS_COL)
#undef CXM_PS_COL
    default:
      return fail(CXM_ERR_UNSUPPORTED, "no dedicated PSF column pass for " + std::to_string(n));
  }
}

bool psf_dedicated_size(int n) {
  switch (n) {
#define CXM_PS_CASE(N, E) case N:
    CXM_PSF_SIZES(CXM_PS_CASE)
#undef CXM_PS_CASE
    return true;
    default:
      return false;
  }
}

int launch_psf_row(int n, const PsfRowArgs& a, int planes, hipStream_t s) {
  const unsigned blocks = (unsigned)((long long)planes * a.ny);
  switch (n) {
#define CXM_PS_ROW(N, E)                                                                      \
  case N:                                                                                     \
    hipLaunchKernelGGL((psf_row_kernel<N, E>), dim3(blocks), dim3(N / E), 0, s, a);           \
    CXM_LAUNCH_CHECK();                                                                       \
    return CXM_OK;
    CXM_PSF_SIZES(CXM_PS_ROW)
#undef CXM_PS_ROW
    default:
      return fail(CXM_ERR_UNSUPPORTED, "no dedicated PSF row pass for " + std::to_string(n));
  }
}

// Row transforms per workgroup of the phase-gradient pass
// (CXM_OPT_PSF_GRAD_WAVES; 0 = the default).  1280^2 x 11 planes, per
// value-and-gradient step: W = 1 0.363 ms, 2 0.394, 4 0.371 -- more waves per
// row do not shorten the pass, so one wave per row sums every plane.
constexpr int kPsfGradWavesDefault = 1;

template <int N, int E>
int launch_psf_grad_n(const PsfGradArgs& a, int groups, int waves, hipStream_t s) {
  const dim3 grid((unsigned)a.ny, (unsigned)groups);
  switch (waves) {
    case 1:
      hipLaunchKernelGGL((psf_grad_kernel<N, E, 1>), grid, dim3(N / E), 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((psf_grad_kernel<N, E, 2>), grid, dim3(2 * (N / E)), 0, s, a);
      break;
    case 4:
      hipLaunchKernelGGL((psf_grad_kernel<N, E, 4>), grid, dim3(4 * (N / E)), 0, s, a);
      break;
    default:
      return fail(CXM_ERR_INVALID, "phase-gradient transforms per workgroup must be 1, 2 or 4");
  }
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

int launch_psf_grad(int n, const PsfGradArgs& a, int groups, hipStream_t s) {
  const int ow = option(CXM_OPT_PSF_GRAD_WAVES);
  const int waves = ow > 0 ? ow : kPsfGradWavesDefault;
  switch (n) {
#define CXM_PS_GRAD(N, E) \
  case N:                 \
    return launch_psf_grad_n<N, E>(a, groups, waves, s);
    CXM_PSF_SIZES(CXM_PS_GRAD)
#undef CXM_PS_GRAD
    default:
      return fail(CXM_ERR_UNSUPPORTED, "no dedicated PSF gradient pass for " + std::to_string(n));
  }
}

// The dedicated passes where the row length has them (CXM_OPT_PSF_PASSES 0),
// and the dedicated column pass where the column length has one too
bool psf_dedicated(const cxm_psf4f_desc* d) {
  return option(CXM_OPT_PSF_PASSES) != 1 && psf_dedicated_size(d->width);
}
bool psf_dedicated_col(const cxm_psf4f_desc* d) {
  return psf_dedicated(d) && psf_dedicated_size(d->height) && d->width % kPsfColC == 0;
}

// Fill P and the chirp factors for the call: the operators of psf_ops split
// into the chirp (its parameter slots) and the rest (applied to 1).
int psf_tables(const PipeArgs& ra, const cxm_psf4f_desc* d, const PsfTables& t, hipStream_t s) {
  OpList inv;
  memset(&inv, 0, sizeof(inv));
  const OpDesc* ch = nullptr;
  for (int i = 0; i < ra.pre.n; ++i) {
    if (ra.pre.op[i].kind == OP_CHIRP) ch = &ra.pre.op[i];
    else inv.op[inv.n++] = ra.pre.op[i];
  }
  if (!ch) return fail(CXM_ERR_INVALID, "no chirp operator");
  const long long n = (long long)d->height * d->width;
  hipLaunchKernelGGL(psf_p_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t.P, d->height,
                     d->width, inv, ra.tab);
  CXM_LAUNCH_CHECK();
  hipLaunchKernelGGL(psf_chirp_kernel, dim3((unsigned)((d->height + d->width + 255) / 256), (unsigned)d->num_z),
                     dim3(256), 0, s, t.cy, t.cx, d->height, d->width, *ch, ra.tab);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

}  // namespace
}  // namespace cxm

using namespace cxm;

// Plane groups of the phase-gradient pass (run in parallel, complex partial
// sums reduced in a fixed order afterwards, where the plane-invariant
// operators are applied once per element): at most kGradGroups, by default
// kGradGroupsDefault.  11 planes of 1280^2, per gradient: 1 group 0.760 ms,
// 2: 0.710, 3: 0.686, 4: 0.689, 11: 0.751 (fewer groups: too few waves for
// a latency-bound pass; more: partial planes written and summed).
constexpr int kGradGroups = 12;
constexpr int kGradGroupsDefault = 4;

namespace cxm {
namespace {
// workspace: one chunk of planes | kGradGroups partial planes | the tables
// of the dedicated passes (P: one plane; cy, cx: num_z rows and columns)
size_t psf_ws_bytes(const cxm_psf4f_desc* d) {
  const size_t plane = (size_t)d->height * d->width;
  return ((size_t)psf_chunk(d) * plane + kGradGroups * plane + plane +
          (size_t)d->num_z * (d->height + d->width)) * sizeof(float2);
}

PsfTables psf_table_ptrs(const cxm_psf4f_desc* d, void* workspace) {
  const long long plane = (long long)d->height * d->width;
  PsfTables t;
  t.P = (float2*)workspace + (psf_chunk(d) + kGradGroups) * plane;
  t.cy = t.P + plane;
  t.cx = t.cy + (long long)d->num_z * d->height;
  return t;
}

// grad[i] (+)= Im(conj(w) sum_g partial[g][i]), g in order, w = the
// plane-invariant operators at grid position i
__global__ void sum_groups_kernel(const float2* __restrict__ partial, float* __restrict__ grad,
                                  int ny, int nx, int groups, int accum, int plane0, OpList mid,
                                  OpTable tab, float m) {
  const long long n = (long long)ny * nx;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float2 c = make_float2(0.f, 0.f);
  for (int g = 0; g < groups; ++g) {
    const float2 q = partial[g * n + i];
    c = make_float2(c.x + q.x, c.y + q.y);
  }
  const int iy = (int)(i / nx), ix = (int)(i - (long long)iy * nx);
  const float2 w = mid.n ? apply_ops(mid, tab, make_float2(1.f, 0.f), plane0, iy, ix, ny, nx)
                         : make_float2(1.f, 0.f);
  const float v = m * (w.x * c.y - w.y * c.x);
  grad[i] = accum ? grad[i] + v : v;
}
}  // namespace
}  // namespace cxm

extern "C" size_t cxm_psf4f_workspace_size(const cxm_psf4f_desc* d) {
  if (psf_geom(d)) return 0;
  return psf_ws_bytes(d);
}

extern "C" int cxm_psf4f_fwd(void* stream, const float* phase, void* out, void* workspace,
                             size_t workspace_bytes, const cxm_psf4f_desc* d) {
  int rc = psf_geom(d);
  if (rc) return rc;
  if (d->out_kind != CXM_OUT_COMPLEX && d->out_kind != CXM_OUT_INTENSITY)
    return fail(CXM_ERR_INVALID, "invalid output kind");
  if (!phase || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const long long chunk = psf_chunk(d);
  const size_t need = psf_ws_bytes(d);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  PipeArgs ra, ca;
  memset(&ra, 0, sizeof(ra));
  memset(&ca, 0, sizeof(ca));
  Ops rb(ra);
  if ((rc = psf_ops(rb, ra, d, phase))) return rc;
  row_args(ra, d);
  const bool ded = psf_dedicated(d);
  PsfRowArgs pr;
  memset(&pr, 0, sizeof(pr));
  if (ded) {
    pr.t = psf_table_ptrs(d, workspace);
    if ((rc = psf_tables(ra, d, pr.t, s))) return rc;
    pr.tw = twiddles_for(d->width);
    pr.out = (float2*)workspace;
    pr.plane_stride = (long long)d->height * d->width;
    pr.ny = d->height;
  }
  col_args(ca, d);
  ca.out_line_shift = 1;
  ca.intensity = d->out_kind == CXM_OUT_INTENSITY;
  const size_t esz = ca.intensity ? sizeof(float) : sizeof(float2);
  for (long long p0 = 0; p0 < d->num_z; p0 += chunk) {
    const int kc = (int)std::min<long long>(chunk, d->num_z - p0);
    ra.out = workspace;
    ra.n_planes = kc;
    ra.plane0 = (int)p0;
    if (ded) {
      pr.plane0 = (int)p0;
      if ((rc = launch_psf_row(d->width, pr, kc, s))) return rc;
    } else if ((rc = launch_pipe_row(d->width, ra, s))) {
      return rc;
    }
    ca.in = (const float2*)workspace;
    ca.out = (char*)out + (size_t)p0 * ca.out_plane_stride * esz;
    ca.n_planes = kc;
    ca.plane0 = (int)p0;
    if (psf_dedicated_col(d)) {
      PsfColArgs pc;
      memset(&pc, 0, sizeof(pc));
      pc.ws = (float2*)workspace;
      pc.tw = twiddles_for(d->height);
      pc.plane_stride = (long long)d->height * d->width;
      pc.nx = d->width;
      pc.plane0 = (int)p0;
      pc.out = ca.out;
      pc.intensity = ca.intensity;
      if ((rc = launch_psf_col(d->height, false, pc, kc, s))) return rc;
      continue;
    }
    if ((rc = launch_pipe_col(d->height, ca, s))) return rc;
  }
  return CXM_OK;
}

namespace cxm {
namespace {
// The adjoint passes; with `intensity`, the column pass also stores the
// forward intensity stack it forms on the way (|u|^2 entering the cotangent).
int psf_bwd(void* stream, const float* phase, const float* cot, float* grad_phase, float* intensity,
            void* workspace, size_t workspace_bytes, const cxm_psf4f_desc* d) {
  int rc = psf_geom(d);
  if (rc) return rc;
  if (d->grad_kind != CXM_GRAD_UPSTREAM && d->grad_kind != CXM_GRAD_MSE)
    return fail(CXM_ERR_INVALID, "invalid gradient kind");
  if (!phase || !cot || !grad_phase) return fail(CXM_ERR_INVALID, "null buffer");
  const long long chunk = psf_chunk(d);
  const size_t need = psf_ws_bytes(d);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  const long long plane_px = (long long)d->height * d->width;
  float2* partial = (float2*)((char*)workspace + (size_t)chunk * plane_px * sizeof(float2));
  PipeArgs ra, ca, ga;
  memset(&ra, 0, sizeof(ra));
  memset(&ca, 0, sizeof(ca));
  Ops rb(ra), cb(ca);
  if ((rc = psf_ops(rb, ra, d, phase))) return rc;
  row_args(ra, d);
  // column pass: forward FFT, cotangent at the output position, inverse FFT
  col_args(ca, d);
  ca.seq = SEQ_FI;
  OpDesc* ig = cb.add(ca.mid, OP_IGRAD, 1, 1, d->grad_kind == CXM_GRAD_MSE ? 1 : 0);
  if (!ig) return fail(CXM_ERR_INVALID, "operator table overflow");
  ig->arr0 = cot;
  ig->arr_plane_stride = (long long)d->height * d->width;
  ig->arr_row_stride = d->width;
  ig->arr_rows = d->height;
  ig->arr_cols = d->width;
  ig->shift = 3;
  const double count = (double)d->num_z * d->height * d->width;
  cb.p(ig, 0, 0) = make_float4((float)(2.0 / count), 0.f, 0.f, 0.f);
  // phase-gradient pass: same generated source as the row pass, split into
  // the operators that change with the plane (pupil, defocus chirp; `pre`)
  // and those that do not (scale, phase mask; `mid`, once per element)
  ga = ra;
  ga.pre.n = 0;
  ga.mid.n = 0;
  for (int i = 0; i < ra.pre.n; ++i) {
    const OpDesc& o = ra.pre.op[i];
    OpList& dst = (o.kind == OP_CIRC || o.kind == OP_CHIRP) ? ga.pre : ga.mid;
    dst.op[dst.n++] = o;
  }
  ga.in_plane_stride = (long long)d->height * d->width;
  ga.in_line_stride = d->width;
  ga.partial_complex = 1;
  ga.out = partial;
  ga.out_plane_stride = plane_px;
  ga.out_line_stride = d->width;
  const bool ded = psf_dedicated(d);
  PsfRowArgs pr;
  PsfGradArgs pg;
  memset(&pr, 0, sizeof(pr));
  memset(&pg, 0, sizeof(pg));
  if (ded) {
    pr.t = psf_table_ptrs(d, workspace);
    if ((rc = psf_tables(ra, d, pr.t, s))) return rc;
    pr.tw = twiddles_for(d->width);
    pr.out = (float2*)workspace;
    pr.plane_stride = plane_px;
    pr.ny = d->height;
    pg.t = pr.t;
    pg.tw = pr.tw;
    pg.in = (const float2*)workspace;
    pg.partial = partial;
    pg.plane_stride = plane_px;
    pg.ny = d->height;
  }
  for (long long p0 = 0; p0 < d->num_z; p0 += chunk) {
    const int kc = (int)std::min<long long>(chunk, d->num_z - p0);
    ra.out = workspace;
    ra.n_planes = kc;
    ra.plane0 = (int)p0;
    if (ded) {
      pr.plane0 = (int)p0;
      if ((rc = launch_psf_row(d->width, pr, kc, s))) return rc;
    } else if ((rc = launch_pipe_row(d->width, ra, s))) {
      return rc;
    }
    ca.in = (const float2*)workspace;
    ca.out = workspace;  // in place: a strip is fully loaded before any store, rows only permute
    ca.n_planes = kc;
    ca.plane0 = (int)p0;
    if (intensity) {
      ca.mid_int = intensity + p0 * plane_px;
      ca.mid_int_plane_stride = plane_px;
    }
    if (psf_dedicated_col(d)) {
      PsfColArgs pc;
      memset(&pc, 0, sizeof(pc));
      pc.ws = (float2*)workspace;
      pc.tw = twiddles_for(d->height);
      pc.plane_stride = plane_px;
      pc.nx = d->width;
      pc.plane0 = (int)p0;
      pc.cot = cot;
      pc.inten = intensity ? intensity + p0 * plane_px : nullptr;
      pc.mse = d->grad_kind == CXM_GRAD_MSE;
      pc.p0 = (float)(2.0 / count);
      if ((rc = launch_psf_col(d->height, true, pc, kc, s))) return rc;
    } else if ((rc = launch_pipe_col(d->height, ca, s))) {
      return rc;
    }
    ga.in = (const float2*)workspace;
    ga.n_planes = kc;
    ga.plane0 = (int)p0;
    const int og = option(CXM_OPT_GRAD_GROUPS);
    const int groups = std::min<int>(og > 0 ? std::min(og, kGradGroups) : kGradGroupsDefault, kc);
    if (ded) {
      // one plane group by default (measured 355 us per 1280^2 x 11 step,
      // 358 with 3 or 4 groups and their partial-sum pass)
      const int dg = std::min<int>(og > 0 ? std::min(og, kGradGroups) : 1, kc);
      pg.plane0 = (int)p0;
      pg.n_planes = kc;
      pg.grad = grad_phase;
      pg.accum = p0 > 0 ? 1 : 0;
      pg.m = phase_scale(d);
      if ((rc = launch_psf_grad(d->width, pg, dg, s))) return rc;
      if (dg == 1) continue;
      hipLaunchKernelGGL(psf_sum_kernel, dim3((unsigned)((plane_px + 255) / 256)), dim3(256), 0, s,
                         (const float2*)partial, (const float2*)pr.t.P, grad_phase, plane_px, dg,
                         p0 > 0 ? 1 : 0, phase_scale(d));
      CXM_LAUNCH_CHECK();
      continue;
    }
    if ((rc = launch_pipe_phasegrad(d->width, ga, s, groups))) return rc;
    hipLaunchKernelGGL(sum_groups_kernel, dim3((unsigned)((plane_px + 255) / 256)), dim3(256), 0, s,
                       (const float2*)partial, grad_phase, d->height, d->width, groups, p0 > 0 ? 1 : 0,
                       (int)p0, ga.mid, ga.tab, phase_scale(d));
    CXM_LAUNCH_CHECK();
  }
  return CXM_OK;
}
}  // namespace
}  // namespace cxm

extern "C" int cxm_psf4f_bwd(void* stream, const float* phase, const float* cot, float* grad_phase,
                             void* workspace, size_t workspace_bytes, const cxm_psf4f_desc* d) {
  return psf_bwd(stream, phase, cot, grad_phase, nullptr, workspace, workspace_bytes, d);
}

extern "C" int cxm_psf4f_value_and_grad(void* stream, const float* phase, const float* cot,
                                        float* grad_phase, float* intensity, void* workspace,
                                        size_t workspace_bytes, const cxm_psf4f_desc* d) {
  if (!intensity) return fail(CXM_ERR_INVALID, "null buffer");
  return psf_bwd(stream, phase, cot, grad_phase, intensity, workspace, workspace_bytes, d);
}

// ---------------------------------------------------------------------------
// objective_point_source and phase_change on their own
// ---------------------------------------------------------------------------
namespace cxm {
__global__ void source_grid_kernel(float2* out, int ny, int nx, OpList L, OpTable T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int plane = blockIdx.y;
  if (i >= (long long)ny * nx) return;
  const int iy = (int)(i / nx), ix = (int)(i - (long long)iy * nx);
  out[(long long)plane * ny * nx + i] = apply_ops(L, T, make_float2(1.f, 0.f), plane, iy, ix, ny, nx);
}

constexpr int kMaxWvPhase = 64;
struct Modulation { int nwv; float m[kMaxWvPhase]; };

// out[p, y, x, w] = in[p, y, x, w] * exp(i fl(phi[y, x] * m[w]))
__global__ void phase_change_kernel(const float2* __restrict__ in, const float* __restrict__ phi,
                                    float2* __restrict__ out, long long n, long long plane_px,
                                    Modulation mod) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long px = i / mod.nwv;
  const int w = (int)(i - px * mod.nwv);
  const float ph = __fmul_rn(phi[px % plane_px], mod.m[w]);
  float s, c;
  sincos_exact(ph, &s, &c);
  out[i] = cmul(in[i], make_float2(c, s));
}
struct Wavelengths { int nwv; float lam[kMaxWvPhase]; };

// out = in * exp(i 2 pi (dn + i abs) t / lam) in the reference's complex64
// order (samples.py:126-133), wavelength innermost
__global__ void thin_sample_kernel(const float2* __restrict__ in, const float* __restrict__ dn,
                                   const float* __restrict__ ab, float2* __restrict__ out,
                                   long long n, long long plane_px, float t, Wavelengths wl) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long px = i / wl.nwv;
  const int w = (int)(i - px * wl.nwv);
  const long long q = px % plane_px;
  const float d = dn[q];
  const float a = ab ? ab[q] : 0.0f;
  const float lam = wl.lam[w];
  const float re = __fdiv_rn(__fmul_rn(-__fmul_rn(kTwoPiF, a), t), lam);
  const float im = __fdiv_rn(__fmul_rn(__fmul_rn(kTwoPiF, d), t), lam);
  float s, c;
  sincos_exact(im, &s, &c);
  const float m = (re == 0.0f) ? 1.0f : expf(re);
  out[i] = cmul(in[i], make_float2(m * c, m * s));
}

// The modulation factor of phase_change (kind 0: exp(i fl(phi m_w))) or
// thin_sample (kind 1: exp(i 2 pi (dn + i abs) t / lam_w)) at element i, as
// the forward kernels evaluate it.
CXM_DEV float2 modulation_at(int kind, const float* a, const float* b, long long q, int w,
                             const Modulation& m, const Wavelengths& wl, float t) {
  float s, c;
  if (kind == 0) {
    sincos_exact(__fmul_rn(a[q], m.m[w]), &s, &c);
    return make_float2(c, s);
  }
  const float d = a[q], ab = b ? b[q] : 0.0f, lam = wl.lam[w];
  const float re = __fdiv_rn(__fmul_rn(-__fmul_rn(kTwoPiF, ab), t), lam);
  const float im = __fdiv_rn(__fmul_rn(__fmul_rn(kTwoPiF, d), t), lam);
  sincos_exact(im, &s, &c);
  const float mg = (re == 0.0f) ? 1.0f : expf(re);
  return make_float2(mg * c, mg * s);
}

// grad_u[i] = g[i] conj(T[i]) (the input cotangent of out = in T)
__global__ void modulation_vjp_u_kernel(int kind, const float2* __restrict__ g,
                                        const float* __restrict__ a, const float* __restrict__ b,
                                        float2* __restrict__ gu, long long n, long long plane_px,
                                        Modulation m, Wavelengths wl, float t) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long px = i / m.nwv;
  const int w = (int)(i - px * m.nwv);
  const float2 T = modulation_at(kind, a, b, px % plane_px, w, m, wl, t);
  const float2 v = g[i];
  gu[i] = make_float2(v.x * T.x + v.y * T.y, v.y * T.x - v.x * T.y);
}

// Per pixel q of the (H, W) parameter arrays, over every plane and
// wavelength in order: ga[q] = sum k_w Im(g conj(y)), gb[q] = -sum k_w
// Re(g conj(y)), with k_w = m_w (phase) or 2 pi t / lam_w (thin sample):
// d out / d phi = i m_w out, d out / d dn = i k_w out, d out / d abs = -k_w out.
__global__ void modulation_vjp_param_kernel(const float2* __restrict__ g,
                                            const float2* __restrict__ y, float* __restrict__ ga,
                                            float* __restrict__ gb, long long n_planes,
                                            long long plane_px, Modulation k) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= plane_px) return;
  double si = 0.0, sr = 0.0;
  for (long long p = 0; p < n_planes; ++p) {
    const long long base = (p * plane_px + q) * k.nwv;
    for (int w = 0; w < k.nwv; ++w) {
      const float2 gv = g[base + w], yv = y[base + w];
      // g conj(y)
      const float zr = gv.x * yv.x + gv.y * yv.y, zi = gv.y * yv.x - gv.x * yv.y;
      si += (double)k.m[w] * zi;
      sr += (double)k.m[w] * zr;
    }
  }
  if (ga) ga[q] = (float)si;
  if (gb) gb[q] = (float)(-sr);
}

// out[p, w] = 2 density[w] g[p] u[p, w]: the vector-Jacobian product of the
// density-weighted intensity
__global__ void intensity_vjp_kernel(const float2* __restrict__ u, const float* __restrict__ g,
                                     float2* __restrict__ out, long long n, Modulation dens) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long px = i / dens.nwv;
  const int w = (int)(i - px * dens.nwv);
  out[i] = cscale(u[i], 2.0f * dens.m[w] * g[px]);
}
// out[p, y, x] = a[p, y, x] * b[p % b_planes, y', x'], (y', x') = (y, x) or,
// with shift_b, the fftshifted position ((y - H/2) mod H, (x - W/2) mod W)
__global__ void cmul_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                            float2* __restrict__ out, long long n, int h, int w, int b_planes,
                            int shift_b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long plane_px = (long long)h * w;
  const long long p = i / plane_px;
  const int r = (int)(i - p * plane_px);
  int y = r / w, x = r - (r / w) * w;
  if (shift_b) {
    y = (y + h - h / 2) % h;
    x = (x + w - w / 2) % w;
  }
  out[i] = cmul(a[i], b[(p % b_planes) * plane_px + (long long)y * w + x]);
}
}  // namespace cxm

extern "C" int cxm_cmul_c64(void* stream, const void* a, const void* b, void* out,
                            int64_t n_planes, int32_t height, int32_t width, int32_t b_planes,
                            int32_t shift_b) {
  if (n_planes < 0 || height <= 0 || width <= 0 || b_planes < 1)
    return fail(CXM_ERR_INVALID, "invalid shape");
  if (n_planes == 0) return CXM_OK;
  if (!a || !b || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const long long n = n_planes * (long long)height * width;
  if ((n + 255) / 256 > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "array too large");
  hipLaunchKernelGGL(cmul_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)a, (const float2*)b, (float2*)out, n,
                     height, width, b_planes, shift_b);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

extern "C" int cxm_thin_sample_c64(void* stream, const void* in, const float* dn,
                                   const float* absorption, void* out, int64_t n_planes,
                                   int32_t height, int32_t width, int32_t num_wv,
                                   const double* wavelengths, double thickness) {
  if (n_planes < 0 || height <= 0 || width <= 0) return fail(CXM_ERR_INVALID, "invalid shape");
  if (num_wv < 1 || num_wv > kMaxWvPhase) return fail(CXM_ERR_INVALID, "num_wv must be in [1, 64]");
  if (n_planes == 0) return CXM_OK;
  if (!in || !dn || !out || !wavelengths) return fail(CXM_ERR_INVALID, "null buffer");
  Wavelengths wl;
  wl.nwv = num_wv;
  for (int w = 0; w < kMaxWvPhase; ++w) wl.lam[w] = 1.0f;
  for (int w = 0; w < num_wv; ++w) {
    if (!(wavelengths[w] > 0)) return fail(CXM_ERR_INVALID, "wavelength must be positive");
    wl.lam[w] = (float)wavelengths[w];
  }
  const long long plane_px = (long long)height * width;
  const long long n = n_planes * plane_px * num_wv;
  if ((n + 255) / 256 > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "array too large");
  hipLaunchKernelGGL(thin_sample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)in, dn, absorption, (float2*)out, n,
                     plane_px, (float)thickness, wl);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

extern "C" int cxm_intensity_vjp_c64(void* stream, const void* u, const float* g, void* out,
                                     int64_t n_pixels, int32_t num_wv, const double* density) {
  if (n_pixels < 0) return fail(CXM_ERR_INVALID, "negative size");
  if (num_wv < 1 || num_wv > kMaxWvPhase) return fail(CXM_ERR_INVALID, "num_wv must be in [1, 64]");
  if (n_pixels == 0) return CXM_OK;
  if (!u || !g || !out) return fail(CXM_ERR_INVALID, "null buffer");
  Modulation d;
  d.nwv = num_wv;
  for (int w = 0; w < kMaxWvPhase; ++w) d.m[w] = 1.0f;
  for (int w = 0; w < num_wv; ++w) d.m[w] = density ? (float)density[w] : 1.0f;
  const long long n = n_pixels * num_wv;
  if ((n + 255) / 256 > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "array too large");
  hipLaunchKernelGGL(intensity_vjp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)u, g, (float2*)out, n, d);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

extern "C" int cxm_point_source_c64(void* stream, void* out, const cxm_psf4f_desc* d) {
  int rc = check_psf(d);
  if (rc) return rc;
  if (!out) return fail(CXM_ERR_INVALID, "null buffer");
  if ((rc = ensure_init())) return rc;
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  Ops b(a);
  if ((rc = source_ops(b, a.pre, d, make_float2(1.f, 0.f)))) return rc;
  const long long n = (long long)d->height * d->width;
  hipLaunchKernelGGL(source_grid_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)d->num_z),
                     dim3(256), 0, (hipStream_t)stream, (float2*)out, d->height, d->width, a.pre,
                     a.tab);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

extern "C" int cxm_phase_change_c64(void* stream, const void* in, const float* phase, void* out,
                                    int64_t n_planes, int32_t height, int32_t width,
                                    int32_t num_wv, const double* modulation) {
  if (n_planes < 0 || height <= 0 || width <= 0) return fail(CXM_ERR_INVALID, "invalid shape");
  if (num_wv < 1 || num_wv > kMaxWvPhase) return fail(CXM_ERR_INVALID, "num_wv must be in [1, 64]");
  if (n_planes == 0) return CXM_OK;
  if (!in || !phase || !out) return fail(CXM_ERR_INVALID, "null buffer");
  Modulation m;
  m.nwv = num_wv;
  for (int w = 0; w < kMaxWvPhase; ++w) m.m[w] = 1.0f;
  for (int w = 0; w < num_wv; ++w) m.m[w] = modulation ? (float)modulation[w] : 1.0f;
  const long long plane_px = (long long)height * width;
  const long long n = n_planes * plane_px * num_wv;
  if ((n + 255) / 256 > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "array too large");
  hipLaunchKernelGGL(phase_change_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)in, phase, (float2*)out, n, plane_px, m);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

// Vector-Jacobian product of cxm_phase_change_c64 (kind 0) and
// cxm_thin_sample_c64 (kind 1): with g the cotangent of out (dL/dRe + i
// dL/dIm) and y = out the forward result,
//   grad_in = g conj(T),  grad_a = sum k_w Im(g conj y),  grad_b = -sum k_w Re(g conj y),
// where a is the phase (k_w = modulation) or dn (k_w = 2 pi t / lam_w) and b
// the absorption.  Any output pointer may be null.  params: modulation per
// wavelength (kind 0) or wavelengths (kind 1).
extern "C" int cxm_modulation_vjp_c64(void* stream, int32_t kind, const void* g, const void* y,
                                      const float* a, const float* b, void* grad_in,
                                      float* grad_a, float* grad_b, int64_t n_planes,
                                      int32_t height, int32_t width, int32_t num_wv,
                                      const double* params, double thickness) {
  if (kind != 0 && kind != 1) return fail(CXM_ERR_INVALID, "kind must be 0 (phase) or 1 (thin)");
  if (n_planes < 0 || height <= 0 || width <= 0) return fail(CXM_ERR_INVALID, "invalid shape");
  if (num_wv < 1 || num_wv > kMaxWvPhase) return fail(CXM_ERR_INVALID, "num_wv must be in [1, 64]");
  if (!g || !a || (!y && (grad_a || grad_b)) || (kind == 1 && !params))
    return fail(CXM_ERR_INVALID, "null buffer");
  Modulation m, k;
  Wavelengths wl;
  m.nwv = k.nwv = wl.nwv = num_wv;
  for (int w = 0; w < kMaxWvPhase; ++w) m.m[w] = k.m[w] = wl.lam[w] = 1.0f;
  for (int w = 0; w < num_wv; ++w) {
    if (kind == 0) {
      m.m[w] = k.m[w] = params ? (float)params[w] : 1.0f;
    } else {
      if (!(params[w] > 0)) return fail(CXM_ERR_INVALID, "wavelength must be positive");
      wl.lam[w] = (float)params[w];
      k.m[w] = (float)(2.0 * M_PI * thickness / params[w]);
    }
  }
  const long long plane_px = (long long)height * width;
  const long long n = n_planes * plane_px * num_wv;
  if ((n + 255) / 256 > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "array too large");
  hipStream_t s = (hipStream_t)stream;
  if (grad_in && n > 0) {
    hipLaunchKernelGGL(modulation_vjp_u_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       kind, (const float2*)g, a, b, (float2*)grad_in, n, plane_px, m, wl,
                       (float)thickness);
    CXM_LAUNCH_CHECK();
  }
  if (grad_a || grad_b) {
    hipLaunchKernelGGL(modulation_vjp_param_kernel, dim3((unsigned)((plane_px + 255) / 256)),
                       dim3(256), 0, s, (const float2*)g, (const float2*)y, grad_a, grad_b,
                       (long long)n_planes, plane_px, k);
    CXM_LAUNCH_CHECK();
  }
  return CXM_OK;
}
