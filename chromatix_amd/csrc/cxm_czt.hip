// Chirp-z transform along one axis of a batch of planes (Bluestein).
//
//   y_k = post_k * scale * sum_{j<n} x_j a^{-j} w^{jk},   k = 0 .. m-1
//
// as one pass over the planes (rows, or column strips): load the n samples
// times a^{-j} w^{j^2/2}, FFT of length L >= n + m - 1 (the next supported
// length; the reference uses exactly n + m - 1), times the spectrum of the
// chirp kernel 1 / w^{(t-(n-1))^2/2}, inverse FFT, keep the window
// [n-1, n-1+m) times w^{k^2/2} * post_k * scale.  The chirp vectors are
// evaluated in FP64 on the device from log a, log w (the reference builds
// them in complex64) and the kernel spectrum by one row FFT.
// Reference: czt (src/chromatix/utils/czt.py:7-52); the CZT branch of
// kernel_propagate (src/chromatix/functional/propagation.py:395-443) calls it
// per axis with its modulation as `post`.
//
// The same pass also evaluates one axis of kernel_propagate's scaled
// convolution (Yu et al. 2012, propagation.py:444-458): pre-chirp, linear
// convolution with a finite chirp kernel (FFT of length >= 2n - 1, window
// [s, s + n) of the full result), post-chirp.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>

#include "cxm_pipe_host.h"

namespace cxm {
namespace {

using cd = std::complex<double>;

int czt_len(int n, int m) {
  const long long need = (long long)n + m - 1;
  static const int lens[] = {
#define CXM_LEN(N, ER, EC, C) N,
      CXM_FOR_EACH_SIZE(CXM_LEN)
#undef CXM_LEN
  };
  int best = -1;
  for (int L : lens)
    if (L >= need && (best < 0 || L < best)) best = L;
  return best;
}

struct CztGeom { int n, m, L, P; };

int czt_geom(const cxm_czt_desc* d, CztGeom& g) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->batch < 0 || d->height <= 0 || d->width <= 0 || d->m <= 0)
    return fail(CXM_ERR_INVALID, "invalid shape parameters");
  if (d->axis != 0 && d->axis != 1) return fail(CXM_ERR_INVALID, "axis must be 0 or 1");
  if (d->num_param < 1 || d->num_param > 64)
    return fail(CXM_ERR_INVALID, "num_param must be in [1, 64]");
  if (!d->a || !d->w) return fail(CXM_ERR_INVALID, "null a / w");
  g.n = d->axis == 0 ? d->height : d->width;
  g.m = d->m;
  g.P = d->num_param;
  g.L = czt_len(g.n, g.m);
  if (g.L < 0)
    return fail(CXM_ERR_UNSUPPORTED, "CZT needs an FFT of length >= " + std::to_string(g.n + g.m - 1));
  return CXM_OK;
}

struct ChirpParams {
  double la[128], lw[128];  // log a, log w per parameter set (re, im)
  double scale;
  int n, m, L, P, mod_n, pad_;
};

CXM_DEV float2 cexp_d(double re, double im) {
  const double r = exp(re);
  return make_float2((float)(r * cos(im)), (float)(r * sin(im)));
}

// Awk2[j] = a^-j w^(j^2/2); h[t] = w^-((t-(n-1))^2/2) on [0, n+m-1), 0 beyond;
// post[k] = scale w^(k^2/2) (w^(-N k) a^N when mod_n = N >= 0)
__global__ void czt_chirps_kernel(float2* __restrict__ out, ChirpParams c) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nA = (long long)c.P * c.n, nH = (long long)c.P * c.L, nP = (long long)c.P * c.m;
  if (i >= nA + nH + nP) return;
  if (i < nA) {
    const int p = (int)(i / c.n), j = (int)(i - (long long)p * c.n);
    const double q = 0.5 * (double)j * (double)j;
    out[i] = cexp_d(-c.la[2 * p] * j + c.lw[2 * p] * q, -c.la[2 * p + 1] * j + c.lw[2 * p + 1] * q);
  } else if (i < nA + nH) {
    const long long r = i - nA;
    const int p = (int)(r / c.L), t = (int)(r - (long long)p * c.L);
    if (t >= c.n + c.m - 1) {
      out[i] = make_float2(0.f, 0.f);
    } else {
      const double k = (double)(t - (c.n - 1));
      const double q = 0.5 * k * k;
      out[i] = cexp_d(-c.lw[2 * p] * q, -c.lw[2 * p + 1] * q);
    }
  } else {
    const long long r = i - nA - nH;
    const int p = (int)(r / c.m), k = (int)(r - (long long)p * c.m);
    const double q = 0.5 * (double)k * (double)k;
    double er = c.lw[2 * p] * q, ei = c.lw[2 * p + 1] * q;
    if (c.mod_n >= 0) {
      const double N = (double)c.mod_n;
      er += -c.lw[2 * p] * N * k + c.la[2 * p] * N;
      ei += -c.lw[2 * p + 1] * N * k + c.la[2 * p + 1] * N;
    }
    const double mag = exp(er) * c.scale;
    out[i] = make_float2((float)(mag * cos(ei)), (float)(mag * sin(ei)));
  }
}

size_t czt_ws(const CztGeom& g) {
  return (size_t)g.P * ((size_t)g.n + g.L + g.m) * sizeof(float2);
}

// ---- dedicated chirp-convolution passes ---------------------------------------
// The pass both entry points run, without the operator interpreter: per line
// (a row, or one column of a strip) of plane p, with chirp set q = p mod P,
//   v_i = x_i wA[q][i] (i < n; zero to length L),  v = FFT(v),
//   v_k = (v_k / L) wH[q][k],  v = IFFT(v),
//   out_j = v_{j + out_off} wP[q][j]  (j < m).
// The multiplies sit in registers around the two transforms; the arithmetic
// is the operator pipeline's (the same products in the same order), so both
// give the same bits.
struct ChirpConvArgs {
  const float2* in;
  float2* out;
  const float2* tw;
  const float2* wA;  // P x n
  const float2* wH;  // P x L (the spectrum of the chirp kernel)
  const float2* wP;  // P x m
  long long in_plane_stride, out_plane_stride;
  int lines;            // rows (row pass) or columns (column pass) of a plane
  int width;            // elements between consecutive rows of in (and of out, column pass)
  int out_line_stride;  // row pass: elements between output rows
  int n, m, out_off, P;
  float mid_scale;      // 1 / L
};

template <int L, int E>
__global__ void __launch_bounds__(L / E) chirp_row_kernel(ChirpConvArgs a) {
  constexpr int T = L / E;
  constexpr bool SYNC = T > 64;
  __shared__ float2 lds[lds_span<L, 1>()];
  const int t = threadIdx.x;
  const int p = blockIdx.x / a.lines;
  const int r = blockIdx.x - p * a.lines;
  const int q = a.P > 1 ? p % a.P : 0;
  const float2* src = a.in + p * a.in_plane_stride + (long long)r * a.width;
  const float2* A = a.wA + (long long)q * a.n;
  float2 v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = t + k * T;
    v[k] = i < a.n ? cmul(src[i], A[i]) : make_float2(0.f, 0.f);
  }
  fft<L, E, -1, 1, SYNC>(v, lds, t, a.tw);
  // the thread index laundered after the first transform: the spectrum and
  // the second transform's twiddle loads are then not hoisted above it (held
  // across it they take the register count to 256)
  int tl = t;
  asm volatile("" : "+v"(tl));
  const float2* H = a.wH + (long long)q * L;
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = cmul(cscale(v[k], a.mid_scale), H[tl + k * T]);
  fft<L, E, 1, 1, SYNC>(v, lds, tl, a.tw);
  const float2* Pq = a.wP + (long long)q * a.m;
  float2* dst = a.out + p * a.out_plane_stride + (long long)r * a.out_line_stride;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int j = tl + k * T - a.out_off;
    if (j >= 0 && j < a.m) dst[j] = cmul(v[k], Pq[j]);
  }
}

template <int L, int E, int C>
__global__ void __launch_bounds__((L / E) * C) chirp_col_kernel(ChirpConvArgs a) {
  constexpr int T = L / E;
  __shared__ float2 lds[lds_span<L, C>()];
  const int c = threadIdx.x % C;
  const int ty = threadIdx.x / C;
  const int nstrips = (a.lines + C - 1) / C;
  const int nb = gridDim.x;
  // consecutive strips of a plane on one XCD
  const int b = ((nb & 7) == 0) ? (int)((blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3)) : (int)blockIdx.x;
  const int p = b / nstrips;
  const int col = (b - p * nstrips) * C + c;
  const bool colok = col < a.lines;
  const int q = a.P > 1 ? p % a.P : 0;
  const float2* src = a.in + p * a.in_plane_stride + col;
  const float2* A = a.wA + (long long)q * a.n;
  float2 v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = ty + k * T;
    v[k] = (colok && i < a.n) ? cmul(src[(long long)i * a.width], A[i]) : make_float2(0.f, 0.f);
  }
  fft<L, E, -1, C, true>(v, lds + c, ty, a.tw);
  int tyl = ty;  // laundered as in the row pass
  asm volatile("" : "+v"(tyl));
  const float2* H = a.wH + (long long)q * L;
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = cmul(cscale(v[k], a.mid_scale), H[tyl + k * T]);
  fft<L, E, 1, C, true>(v, lds + c, tyl, a.tw);
  if (!colok) return;
  const float2* Pq = a.wP + (long long)q * a.m;
  float2* dst = a.out + p * a.out_plane_stride + col;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int j = tyl + k * T - a.out_off;
    if (j >= 0 && j < a.m) dst[(long long)j * a.width] = cmul(v[k], Pq[j]);
  }
}

// (length, values per thread) of the dedicated row passes and (length,
// values per thread, columns per workgroup) of the column passes: the
// shapes of the operator pipeline's passes at these lengths
#define CXM_CHIRP_ROW_SIZES(X) \
  X(1024, 16) X(1280, 20) X(1536, 24) X(2048, 16) X(2560, 20) X(3072, 24) X(4096, 16) X(5120, 20)
#define CXM_CHIRP_COL_SIZES(X) \
  X(1024, 16, 4) X(1280, 20, 4) X(1536, 24, 4) X(2048, 16, 4) X(2560, 20, 4) X(3072, 24, 4)

// The dedicated pass along `axis` (0: columns, 1: rows) for `planes` planes:
// 1 if there is none for length L (the caller runs the operator pipeline),
// else the launch status.
int chirp_conv_pass(int axis, int L, int planes, const ChirpConvArgs& a, hipStream_t s) {
  if (option(CXM_OPT_CZT_PASSES) == 1) return 1;
  const long long lines = (long long)planes * a.lines;
  if (axis == 1) {
    if (lines > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "chirp-z row pass grid too large");
    if (L == 1024 && lines <= 4096) {
      // few rows: 256 threads per row, four times the waves of the default
      // shape (the operator pipeline's choice for these launches too)
      hipLaunchKernelGGL((chirp_row_kernel<1024, 4>), dim3((unsigned)lines), dim3(256), 0, s, a);
      CXM_LAUNCH_CHECK();
      return CXM_OK;
    }
    switch (L) {
#define CXM_CR(NN, EE)                                                                          \
  case NN:                                                                                      \
    hipLaunchKernelGGL((chirp_row_kernel<NN, EE>), dim3((unsigned)lines), dim3(NN / EE), 0, s, a); \
    CXM_LAUNCH_CHECK();                                                                         \
    return CXM_OK;
      CXM_CHIRP_ROW_SIZES(CXM_CR)
#undef CXM_CR
      default:
        return 1;
    }
  }
  switch (L) {
#define CXM_CC(NN, EE, CC)                                                                        \
  case NN: {                                                                                      \
    const long long blocks = (long long)planes * ((a.lines + CC - 1) / CC);                       \
    if (blocks > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "chirp-z column pass grid too large"); \
    hipLaunchKernelGGL((chirp_col_kernel<NN, EE, CC>), dim3((unsigned)blocks), dim3((NN / EE) * CC), \
                       0, s, a);                                                                  \
    CXM_LAUNCH_CHECK();                                                                           \
    return CXM_OK;                                                                                \
  }
    CXM_CHIRP_COL_SIZES(CXM_CC)
#undef CXM_CC
    default:
      return 1;
  }
}

}  // namespace
}  // namespace cxm

using namespace cxm;

extern "C" size_t cxm_czt_axis_workspace_size(const cxm_czt_desc* d) {
  CztGeom g;
  if (czt_geom(d, g)) return 0;
  return czt_ws(g);
}

extern "C" int cxm_czt_axis_c64(void* stream, const void* in, void* out, void* workspace,
                                size_t workspace_bytes, const cxm_czt_desc* d) {
  CztGeom g;
  int rc = czt_geom(d, g);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const size_t need = czt_ws(g);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int n = g.n, m = g.m, L = g.L, P = g.P;
  // chirps in FP64 on the device: [Awk2 (P x n) | h (P x L) | post (P x m)]
  ChirpParams cp;
  memset(&cp, 0, sizeof(cp));
  for (int p = 0; p < P; ++p) {
    const cd a(d->a[2 * p], d->a[2 * p + 1]), w(d->w[2 * p], d->w[2 * p + 1]);
    if (std::abs(a) == 0.0 || std::abs(w) == 0.0) return fail(CXM_ERR_INVALID, "a and w must be non-zero");
    const cd la = std::log(a), lw = std::log(w);
    cp.la[2 * p] = la.real(); cp.la[2 * p + 1] = la.imag();
    cp.lw[2 * p] = lw.real(); cp.lw[2 * p + 1] = lw.imag();
  }
  cp.n = n; cp.m = m; cp.L = L; cp.P = P;
  cp.mod_n = d->mod_n;
  cp.scale = d->scale;
  float2* wA = (float2*)workspace;
  float2* wH = wA + (size_t)P * n;
  float2* wP = wH + (size_t)P * L;
  const long long tot = (long long)P * (n + L + m);
  hipLaunchKernelGGL(czt_chirps_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                     (float2*)workspace, cp);
  CXM_LAUNCH_CHECK();
  // kernel spectra: P rows of length L, forward FFT in place
  {
    PipeArgs f;
    memset(&f, 0, sizeof(f));
    f.in = wH;
    f.out = wH;
    f.tw = twiddles_for(L);
    f.in_plane_stride = f.out_plane_stride = (long long)P * L;
    f.in_line_stride = f.out_line_stride = L;
    f.n_lines = P;
    f.n_planes = 1;
    f.grid_ny = P;
    f.grid_nx = L;
    f.in_len = f.out_len = L;
    f.seq = SEQ_F;
    f.scale = f.mid_scale = 1.0f;
    if ((rc = launch_pipe_row(L, f, s))) return rc;
  }
  const int ax = d->axis == 1 ? 1 : 0;  // VEC index axis: x for rows, y for columns
  {
    ChirpConvArgs c;
    c.in = (const float2*)in;
    c.out = (float2*)out;
    c.tw = twiddles_for(L);
    c.wA = wA;
    c.wH = wH;
    c.wP = wP;
    c.in_plane_stride = (long long)d->height * d->width;
    c.out_plane_stride = ax ? (long long)d->height * m : (long long)m * d->width;
    c.lines = ax ? d->height : d->width;
    c.width = d->width;
    c.out_line_stride = ax ? m : d->width;
    c.n = n;
    c.m = m;
    c.out_off = n - 1;
    c.P = P;
    c.mid_scale = (float)(1.0 / L);
    if ((rc = chirp_conv_pass(ax, L, d->batch, c, s)) != 1) return rc;
  }
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  Ops b(a);
  OpDesc* oA = b.add(a.pre, OP_VEC, P, 0);
  OpDesc* oH = b.add(a.mid, OP_VEC, P, 0);
  OpDesc* oP = b.add(a.post, OP_VEC, P, 0);
  if (!oA || !oH || !oP) return fail(CXM_ERR_INVALID, "operator table overflow");
  oA->arr0 = wA; oA->arr_plane_stride = n; oA->arr_cols = n; oA->vec_axis = ax;
  oH->arr0 = wH; oH->arr_plane_stride = L; oH->arr_cols = L; oH->vec_axis = ax;
  oP->arr0 = wP; oP->arr_plane_stride = m; oP->arr_cols = m; oP->vec_axis = ax;
  if (ax) oP->arr_col_off = n - 1; else oP->arr_row_off = n - 1;
  a.tw = twiddles_for(L);
  a.in = (const float2*)in;
  a.out = out;
  a.in_plane_stride = (long long)d->height * d->width;
  a.n_planes = d->batch;
  a.grid_ny = ax ? d->height : L;
  a.grid_nx = ax ? L : d->width;
  a.in_len = n;
  a.out_off = n - 1;
  a.out_len = m;
  a.seq = SEQ_FI;
  a.scale = 1.0f;
  a.mid_scale = (float)(1.0 / L);
  if (ax) {
    a.in_line_stride = d->width;
    a.n_lines = d->height;
    a.out_plane_stride = (long long)d->height * m;
    a.out_line_stride = m;
    return launch_pipe_row(L, a, s);
  }
  a.in_line_stride = d->width;
  a.n_lines = d->width;
  a.out_plane_stride = (long long)m * d->width;
  a.out_line_stride = d->width;
  a.out_line_len = d->width;
  return launch_pipe_col(L, a, s);
}

// ---------------------------------------------------------------------------
// Scaled convolution (Yu et al.): [pre (n) | kernel (L) | post (n)] in FP64.
// ---------------------------------------------------------------------------
namespace cxm {
namespace {
struct YuParams {
  double alpha, df, out_dx, x0;
  int n, L;
};

__global__ void yu_chirps_kernel(float2* __restrict__ out, YuParams c) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = c.n, L = c.L, h = c.n / 2;
  if (i >= 2LL * n + L) return;
  constexpr double kPi = 3.14159265358979323846;
  if (i < n) {  // (1/alpha) exp(i pi alpha f_k^2)
    const double f = (double)((int)i - h) * c.df;
    const double ph = kPi * c.alpha * f * f;
    out[i] = make_float2((float)(cos(ph) / c.alpha), (float)(sin(ph) / c.alpha));
  } else if (i < n + L) {  // exp(-i pi alpha f_t^2), t < n; 0 beyond
    const int t = (int)(i - n);
    if (t >= n) {
      out[i] = make_float2(0.f, 0.f);
    } else {
      const double f = (double)(t - h) * c.df;
      const double ph = -kPi * c.alpha * f * f;
      out[i] = make_float2((float)cos(ph), (float)sin(ph));
    }
  } else {  // out_dx exp(i pi x_j^2 / alpha)
    const int j = (int)(i - n - L);
    const double x = c.x0 + (double)j * c.out_dx;
    const double ph = kPi * x * x / c.alpha;
    out[i] = make_float2((float)(c.out_dx * cos(ph)), (float)(c.out_dx * sin(ph)));
  }
}

// The n x n matrix of one Yu axis pass, element (j, k) at out[j * ld_r +
// k * ld_c] (conjugated on request), in FP64:
//   out_dx / alpha exp(i pi (x_j^2 / alpha + alpha f_k^2 - alpha f_t^2)),
//   t = j + s - k in [0, n) (0 otherwise), s = (n - 1) / 2
__global__ void yu_matrix_kernel(float2* __restrict__ out, YuParams c, long long ld_r,
                                 long long ld_c, double sg) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = c.n, h = c.n / 2, sw = (c.n - 1) / 2;
  if (i >= (long long)n * n) return;
  const int j = (int)(i / n), k = (int)(i - (long long)j * n);
  const int t = j + sw - k;
  float2 v = make_float2(0.f, 0.f);
  if (t >= 0 && t < n) {
    constexpr double kPi = 3.14159265358979323846;
    const double x = c.x0 + (double)j * c.out_dx;
    const double fk = (double)(k - h) * c.df, ft = (double)(t - h) * c.df;
    const double ph = kPi * (x * x / c.alpha + c.alpha * (fk * fk - ft * ft));
    const double m = c.out_dx / c.alpha;
    v = make_float2((float)(m * cos(ph)), (float)(sg * m * sin(ph)));
  }
  out[(long long)j * ld_r + (long long)k * ld_c] = v;
}

int yu_geom(const cxm_yu_desc* d, int& n, int& L) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->batch < 0 || d->height <= 0 || d->width <= 0)
    return fail(CXM_ERR_INVALID, "invalid shape parameters");
  if (d->axis != 0 && d->axis != 1) return fail(CXM_ERR_INVALID, "axis must be 0 or 1");
  if (!(d->alpha > 0.0) || !(d->df > 0.0) || !(d->out_dx > 0.0))
    return fail(CXM_ERR_INVALID, "alpha, df and out_dx must be positive");
  n = d->axis == 0 ? d->height : d->width;
  L = czt_len(n, n);
  if (L < 0) return fail(CXM_ERR_UNSUPPORTED, "scaled convolution needs an FFT of length >= " + std::to_string(2 * n - 1));
  return CXM_OK;
}
}  // namespace
}  // namespace cxm

extern "C" size_t cxm_yu_axis_workspace_size(const cxm_yu_desc* d) {
  int n, L;
  if (yu_geom(d, n, L)) return 0;
  return ((size_t)2 * n + L) * sizeof(float2);
}

extern "C" int cxm_yu_axis_c64(void* stream, const void* in, void* out, void* workspace,
                               size_t workspace_bytes, const cxm_yu_desc* d) {
  int n, L;
  int rc = yu_geom(d, n, L);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  if (in == out) return fail(CXM_ERR_INVALID, "in and out must differ");
  const size_t need = ((size_t)2 * n + L) * sizeof(float2);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  YuParams yp;
  yp.alpha = d->alpha; yp.df = d->df; yp.out_dx = d->out_dx; yp.x0 = d->x0;
  yp.n = n; yp.L = L;
  float2* wA = (float2*)workspace;
  float2* wH = wA + n;
  float2* wP = wH + L;
  const long long tot = 2LL * n + L;
  hipLaunchKernelGGL(yu_chirps_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, wA, yp);
  CXM_LAUNCH_CHECK();
  {  // kernel spectrum: one row of length L, forward FFT in place
    PipeArgs f;
    memset(&f, 0, sizeof(f));
    f.in = wH;
    f.out = wH;
    f.tw = twiddles_for(L);
    f.in_plane_stride = f.out_plane_stride = L;
    f.in_line_stride = f.out_line_stride = L;
    f.n_lines = 1;
    f.n_planes = 1;
    f.grid_ny = 1;
    f.grid_nx = L;
    f.in_len = f.out_len = L;
    f.seq = SEQ_F;
    f.scale = f.mid_scale = 1.0f;
    if ((rc = launch_pipe_row(L, f, s))) return rc;
  }
  const int ax = d->axis == 1 ? 1 : 0;
  const int sw = (n - 1) / 2;  // "same" window of the full convolution
  {
    ChirpConvArgs c;
    c.in = (const float2*)in;
    c.out = (float2*)out;
    c.tw = twiddles_for(L);
    c.wA = wA;
    c.wH = wH;
    c.wP = wP;
    c.in_plane_stride = c.out_plane_stride = (long long)d->height * d->width;
    c.lines = ax ? d->height : d->width;
    c.width = d->width;
    c.out_line_stride = d->width;
    c.n = n;
    c.m = n;
    c.out_off = sw;
    c.P = 1;
    c.mid_scale = (float)(1.0 / L);
    if ((rc = chirp_conv_pass(ax, L, d->batch, c, s)) != 1) return rc;
  }
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  Ops b(a);
  OpDesc* oA = b.add(a.pre, OP_VEC, 1, 0);
  OpDesc* oH = b.add(a.mid, OP_VEC, 1, 0);
  OpDesc* oP = b.add(a.post, OP_VEC, 1, 0);
  if (!oA || !oH || !oP) return fail(CXM_ERR_INVALID, "operator table overflow");
  oA->arr0 = wA; oA->arr_plane_stride = n; oA->arr_cols = n; oA->vec_axis = ax;
  oH->arr0 = wH; oH->arr_plane_stride = L; oH->arr_cols = L; oH->vec_axis = ax;
  oP->arr0 = wP; oP->arr_plane_stride = n; oP->arr_cols = n; oP->vec_axis = ax;
  if (ax) oP->arr_col_off = sw; else oP->arr_row_off = sw;
  a.tw = twiddles_for(L);
  a.in = (const float2*)in;
  a.out = out;
  a.in_plane_stride = (long long)d->height * d->width;
  a.out_plane_stride = (long long)d->height * d->width;
  a.n_planes = d->batch;
  a.grid_ny = ax ? d->height : L;
  a.grid_nx = ax ? L : d->width;
  a.in_len = n;
  a.out_off = sw;
  a.out_len = n;
  a.seq = SEQ_FI;
  a.scale = 1.0f;
  a.mid_scale = (float)(1.0 / L);
  a.in_line_stride = d->width;
  if (ax) {
    a.n_lines = d->height;
    a.out_line_stride = d->width;
    return launch_pipe_row(L, a, s);
  }
  a.n_lines = d->width;
  a.out_line_stride = d->width;
  a.out_line_len = d->width;
  return launch_pipe_col(L, a, s);
}

extern "C" int cxm_yu_matrix_c64(void* stream, const cxm_yu_desc* d, void* out, int64_t ld_r,
                                 int64_t ld_c, int32_t conj) {
  int n, L;
  int rc = yu_geom(d, n, L);
  if (rc) return rc;
  if (!out) return fail(CXM_ERR_INVALID, "null buffer");
  if ((long long)n * n > (1ll << 31)) return fail(CXM_ERR_UNSUPPORTED, "matrix too large");
  if ((rc = ensure_init())) return rc;
  YuParams yp;
  yp.alpha = d->alpha; yp.df = d->df; yp.out_dx = d->out_dx; yp.x0 = d->x0;
  yp.n = n; yp.L = L;
  const long long tot = (long long)n * n;
  hipLaunchKernelGGL(yu_matrix_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (float2*)out, yp, (long long)ld_r, (long long)ld_c,
                     conj ? -1.0 : 1.0);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}
