// Modified Born series (convergent Born series) solver for the vector
// Helmholtz problem, after the preconditioned fixed-point iteration of
// src/chromatix/experimental/modified_born_series/solver.py:42-146.
//
// Fields are three complex64 component volumes (component-planar
// 3 x nx x ny x nz in the kernels; the boundary converts from the
// reference's nx x ny x nz x 3).  With s = scale, b = bias, kb = b s + 1:
//   V(x)  = (s eps - kb) x    (eps a scalar per voxel, or a 3x3 matrix per
//                              voxel: the tensorial extension)
//   G(y)  = ifftn( P_T y^ / (kb - s |K|^2) + P_L y^ / kb )   (K = k / k0)
//   prec(y)          = -V(G(s y))
//   prec_forward(x)  = -V(G(V(x)) + x)
//   x <- x + prec_y - prec_forward(x) = x + prec_y + V(G(V(x)) + x)
// One iteration: V pass, three forward FFT passes (z rows, y and x column
// strips of the pipe kernels), the Green multiply (3x3 projection per k),
// three inverse FFT passes, and the update pass, which also writes the FP64
// partial sums of |r|^2 for the residual norm.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "cxm_pipe_host.h"

namespace cxm {
namespace {

struct MbsGeom { int nx, ny, nz; long long vox; };

int mbs_geom(const cxm_mbs_desc* d, MbsGeom& g) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0)
    return fail(CXM_ERR_INVALID, "invalid shape");
  if (!(d->spacing > 0 && d->k0 > 0)) return fail(CXM_ERR_INVALID, "spacing and k0 must be positive");
  if (d->eps_tensor != 0 && d->eps_tensor != 1) return fail(CXM_ERR_INVALID, "eps_tensor must be 0 or 1");
  for (int n : {d->nx, d->ny, d->nz})
    if (!fft_len_supported(n))
      return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(n));
  g.nx = d->nx; g.ny = d->ny; g.nz = d->nz;
  g.vox = (long long)d->nx * d->ny * d->nz;
  return CXM_OK;
}

struct C2 { float re, im; };

// A solver's device-side stop flag (null: never stopped); every kernel of an
// iteration returns at once when it is set.
CXM_DEV bool stopped(const double* stop) { return stop && *stop != 0.0; }

CXM_DEV float2 cmulf(float2 a, C2 b) { return make_float2(a.x * b.re - a.y * b.im, a.x * b.im + a.y * b.re); }

// V applied to the field vector v at voxel i: o = (s eps - kb) v, eps a
// scalar (eps[i]) or, with `tensor`, a 3x3 matrix (eps[(3 a + b) vox + i]).
CXM_DEV void apply_v(const float2* __restrict__ eps, int tensor, long long i, long long vox, C2 s,
                     C2 kb, const float2 (&v)[3], float2 (&o)[3]) {
  if (!tensor) {
    const float2 e = eps[i];
    const C2 f{s.re * e.x - s.im * e.y - kb.re, s.re * e.y + s.im * e.x - kb.im};
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = cmulf(v[c], f);
    return;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float2 acc = cmulf(v[a], C2{-kb.re, -kb.im});
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const float2 e = eps[(long long)(3 * a + b) * vox + i];
      const C2 se{s.re * e.x - s.im * e.y, s.re * e.y + s.im * e.x};
      const float2 t = cmulf(v[b], se);
      acc = make_float2(acc.x + t.x, acc.y + t.y);
    }
    o[a] = acc;
  }
}

// out[c][i] = pre * V(src)[c][i]; src interleaved (i*3 + c) when
// interleaved, else planar; eps may be null (out = pre * src then, used for
// the plain scaling of the source)
__global__ void mbs_v_kernel(const float2* __restrict__ src, const float2* __restrict__ eps,
                             float2* __restrict__ out, long long vox, int interleaved, int tensor, C2 s,
                             C2 kb, C2 pre, const double* __restrict__ stop) {
  if (stopped(stop)) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vox) return;
  float2 v[3], o[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = interleaved ? src[i * 3 + c] : src[c * vox + i];
  if (eps) {
    apply_v(eps, tensor, i, vox, s, kb, v, o);
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = v[c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c * vox + i] = cmulf(o[c], pre);
}

// K along an axis of n bins: bin b (signed) times kscale / n
CXM_DEV float k_step(float kscale, int n) { return kscale / (float)n; }
// 1/x by the hardware reciprocal (within 1 ulp; the Green's operator's
// factors are the only place it is used, where IEEE division cost a third
// of the fused x pass's arithmetic)
CXM_DEV float rcp_approx(float x) { return __builtin_amdgcn_rcpf(x); }

// In-place Green multiply of the spectrum Y (3 planar components):
// Y <- P_T Y / (kb - s |K|^2) + P_L Y / kb, P_L = K K^T / |K|^2 (0 at K = 0)
__global__ void mbs_green_kernel(float2* __restrict__ y, long long vox, int nx, int ny, int nz,
                                 float kscale, C2 s, C2 kb, const double* __restrict__ stop) {
  if (stopped(stop)) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vox) return;
  const int iz = (int)(i % nz);
  const int iy = (int)((i / nz) % ny);
  const int ix = (int)(i / ((long long)nz * ny));
  // K = 2 pi fftfreq / spacing / k0
  const float kx = (float)signed_bin(ix, nx) * k_step(kscale, nx);
  const float ky = (float)signed_bin(iy, ny) * k_step(kscale, ny);
  const float kz = (float)signed_bin(iz, nz) * k_step(kscale, nz);
  const float k2 = kx * kx + ky * ky + kz * kz;
  const float2 y0 = y[i], y1 = y[vox + i], y2 = y[2 * vox + i];
  float2 p = make_float2(kx * y0.x + ky * y1.x + kz * y2.x, kx * y0.y + ky * y1.y + kz * y2.y);
  const float inv = k2 != 0.0f ? rcp_approx(k2) : 0.0f;
  p = make_float2(p.x * inv, p.y * inv);
  // 1 / (kb - s k2) and 1 / kb
  const float dre = kb.re - s.re * k2, dim = kb.im - s.im * k2;
  const float rdn = rcp_approx(dre * dre + dim * dim);
  const C2 gt{dre * rdn, -dim * rdn};
  const float bn = kb.re * kb.re + kb.im * kb.im;
  const C2 gl{kb.re / bn, -kb.im / bn};
  const float kk[3] = {kx, ky, kz};
  const float2 yy[3] = {y0, y1, y2};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float2 yl = make_float2(kk[c] * p.x, kk[c] * p.y);
    const float2 yt = make_float2(yy[c].x - yl.x, yy[c].y - yl.y);
    const float2 a = cmulf(yt, gt), b = cmulf(yl, gl);
    y[c * vox + i] = make_float2(a.x + b.x, a.y + b.y);
  }
}

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

// x pass of the Green's operator, fused: for a strip of C columns of the
// (nx x ny nz) matrix, the three components of each column are transformed
// along x by three groups of threads (3 C column transforms interleaved in
// LDS), exchanged through LDS, multiplied by the Green's projection at
// every (kx, ky, kz) (the arithmetic of mbs_green_kernel, each thread
// writing its own component), transformed back along x and stored with
// `scale`; the y and z passes stay pipeline passes.  One HBM round trip
// instead of three (forward x pass, Green multiply, inverse x pass).
template <int N, int E, int C>
__global__ void __launch_bounds__((N / E) * 3 * C)
mbs_xgreen_kernel(float2* __restrict__ v, const float2* __restrict__ tw, long long vox, int ncols,
                  int ny, int nz, float kscale, C2 s, C2 kb, float scale, const double* __restrict__ stop) {
  if (stopped(stop)) return;
  constexpr int T = N / E;
  constexpr int CC = 3 * C;
  __shared__ float2 lds[lds_span<N, CC>()];
  const int cc = threadIdx.x % CC;
  const int ty = threadIdx.x / CC;
  const int q = cc / C;  // component of this thread's column
  const int c = cc - q * C;
  const int nb = gridDim.x;
  const int L = ((nb & 7) == 0) ? (int)((blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3)) : (int)blockIdx.x;
  const int col = L * C + c;
  const bool colok = col < ncols;
  float2* vq = v + q * vox + col;
  float2 u[E];
#pragma unroll
  for (int k = 0; k < E; ++k)
    u[k] = colok ? vq[(long long)(ty + k * T) * ncols] : make_float2(0.f, 0.f);
  fft<N, E, -1, CC, true>(u, lds + cc, ty, tw);
  block_sync();  // the transform's last exchange has been read by every thread
  stage_out<N, E, CC>(u, lds + cc, ty);
  block_sync();
  const int iy = colok ? col / nz : 0;
  const int iz = colok ? col - iy * nz : 0;
  const float ky = (float)signed_bin(iy, ny) * k_step(kscale, ny);
  const float kz = (float)signed_bin(iz, nz) * k_step(kscale, nz);
  const float kxs = k_step(kscale, N);
  const float bn = kb.re * kb.re + kb.im * kb.im;
  const C2 gl{kb.re / bn, -kb.im / bn};
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = ty + k * T;
    const float kx = (float)signed_bin(i, N) * kxs;
    const float k2 = kx * kx + ky * ky + kz * kz;
    const float2 y0 = lds[c + lds_at<N, CC>(i)];
    const float2 y1 = lds[C + c + lds_at<N, CC>(i)];
    const float2 y2 = lds[2 * C + c + lds_at<N, CC>(i)];
    float2 p = make_float2(kx * y0.x + ky * y1.x + kz * y2.x, kx * y0.y + ky * y1.y + kz * y2.y);
    const float inv = k2 != 0.0f ? rcp_approx(k2) : 0.0f;
    p = make_float2(p.x * inv, p.y * inv);
    const float dre = kb.re - s.re * k2, dim = kb.im - s.im * k2;
    const float rdn = rcp_approx(dre * dre + dim * dim);
    const C2 gt{dre * rdn, -dim * rdn};
    const float kq = q == 0 ? kx : (q == 1 ? ky : kz);
    const float2 yq = q == 0 ? y0 : (q == 1 ? y1 : y2);
    const float2 yl = make_float2(kq * p.x, kq * p.y);
    const float2 yt = make_float2(yq.x - yl.x, yq.y - yl.y);
    const float2 ga = cmulf(yt, gt), gb = cmulf(yl, gl);
    u[k] = make_float2(ga.x + gb.x, ga.y + gb.y);
  }
  // the inverse transform's first stage writes this thread's column before
  // any barrier: every thread must have read the other components first
  block_sync();
  fft<N, E, 1, CC, true>(u, lds + cc, ty, tw);
  if (!colok) return;
#pragma unroll
  for (int k = 0; k < E; ++k)
    vq[(long long)(ty + k * T) * ncols] = make_float2(u[k].x * scale, u[k].y * scale);
}

// (length, values per thread, columns per workgroup; 3x as many column
// transforms): at most 32 KiB of LDS per workgroup (at 160: 8 columns
// measured 0.377 ms per c4 iteration, 16 columns 0.431, the three
// separate launches 0.423)
#define CXM_XGREEN_SIZES(X)                                                   \
  X(8, 8, 32) X(16, 16, 32) X(32, 16, 32) X(64, 16, 16) X(128, 16, 8)          \
  X(256, 16, 4) X(20, 20, 32) X(40, 20, 32) X(80, 20, 16) X(160, 20, 8)        \
  X(12, 12, 32) X(24, 12, 32) X(48, 12, 16) X(96, 12, 8) X(192, 12, 4)

bool xgreen_supported(int n) {
  switch (n) {
#define CXM_XG_CASE(N, E, C) case N:
    CXM_XGREEN_SIZES(CXM_XG_CASE)
#undef CXM_XG_CASE
    return true;
    default:
      return false;
  }
}

int launch_xgreen(int n, float2* v, long long vox, int ncols, int ny, int nz, float kscale, C2 s, C2 kb,
                  float scale, const double* stop, hipStream_t st) {
  switch (n) {
#define CXM_XG_LAUNCH(N, E, C)                                                                        \
  case N: {                                                                                           \
    const long long blocks = ((long long)ncols + (C)-1) / (C);                                        \
    if (blocks > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "fused Green pass grid too large");     \
    hipLaunchKernelGGL((mbs_xgreen_kernel<N, E, C>), dim3((unsigned)blocks), dim3((N / E) * 3 * (C)), 0, \
                       st, v, twiddles_for(N), vox, ncols, ny, nz, kscale, s, kb, scale, stop);       \
    CXM_LAUNCH_CHECK();                                                                               \
    return CXM_OK;                                                                                    \
  }
    CXM_XGREEN_SIZES(CXM_XG_LAUNCH)
#undef CXM_XG_LAUNCH
    default:
      return fail(CXM_ERR_UNSUPPORTED, "no fused Green pass for length " + std::to_string(n));
  }
}

// ---- z passes fused with V and with the update -------------------------------
// z is the innermost axis, so the z rows (ix, iy) of a component are
// consecutive in memory and R rows form one contiguous run of R*N values.  A
// workgroup takes R rows of all three components: it reads the runs linearly
// (coalesced, every voxel's three components and eps in one thread), places
// the values in LDS as 3R interleaved transforms (transform c*R + r, stride
// S = 3R rounded up to an odd count so that consecutive z land on different
// banks), transforms each (row, component) along z with T = N/E threads, and
// goes back through LDS for a linear pass over the voxels.
constexpr int pow2_ceil(int n) {
  int p = 1;
  while (p < n) p *= 2;
  return p;
}

template <int N, int E>
struct ZCfg {
  static constexpr int T = N / E;
  static constexpr int R = (T >= 64) ? 1 : 64 / T;
  static constexpr int S = (3 * R) | 1;
  static constexpr int THREADS = 3 * R * T;
};

// LDS slot of value z of row r, component c
template <int N, int E>
CXM_DEV int zslot(int z, int r, int c) {
  using Cfg = ZCfg<N, E>;
  return lds_at<N, Cfg::S>(z) + c * Cfg::R + r;
}

// transform of (row, component) `tr` along z, values from and back to LDS
template <int N, int E, int DIR>
CXM_DEV void ztransform(float2* lds, const float2* __restrict__ tw) {
  using Cfg = ZCfg<N, E>;
  constexpr int T = Cfg::T;
  const int tr = threadIdx.x / T;
  const int t = threadIdx.x - tr * T;
  float2 u[E];
#pragma unroll
  for (int k = 0; k < E; ++k) u[k] = lds[lds_at<N, Cfg::S>(t + k * T) + tr];
  block_sync();  // every initial value has been read before the first scatter
  fft<N, E, DIR, Cfg::S, true>(u, lds + tr, t, tw);
#pragma unroll
  for (int k = 0; k < E; ++k) lds[lds_at<N, Cfg::S>(t + k * T) + tr] = u[k];
  block_sync();
}

// gb[c] = FFT_z( V(x)[c] ) for the rows of this workgroup (the first pass of
// G(V x); unnormalised, like pass_z)
template <int N, int E>
__global__ void __launch_bounds__((ZCfg<N, E>::THREADS))
mbs_zv_kernel(const float2* __restrict__ x, const float2* __restrict__ eps, float2* __restrict__ gb,
              const float2* __restrict__ tw, long long vox, long long nrows, int tensor, C2 s, C2 kb,
              const double* __restrict__ stop) {
  if (stopped(stop)) return;
  using Cfg = ZCfg<N, E>;
  constexpr int R = Cfg::R;
  __shared__ float2 lds[lds_span<N, Cfg::S>()];
  const long long row0 = (long long)blockIdx.x * R;
  const int nv = (int)min((long long)R, nrows - row0) * N;
  const long long base = row0 * N;
  for (int j = threadIdx.x; j < R * N; j += Cfg::THREADS) {
    const int r = j / N, z = j - r * N;
    float2 o[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    if (j < nv) {
      float2 v[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = x[c * vox + base + j];
      apply_v(eps, tensor, base + j, vox, s, kb, v, o);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) lds[zslot<N, E>(z, r, c)] = o[c];
  }
  block_sync();
  ztransform<N, E, -1>(lds, tw);
  for (int j = threadIdx.x; j < nv; j += Cfg::THREADS) {
    const int r = j / N, z = j - r * N;
#pragma unroll
    for (int c = 0; c < 3; ++c) gb[c * vox + base + j] = lds[zslot<N, E>(z, r, c)];
  }
}

// The last pass of G(V x) fused with the fixed-point update: g = IFFT_z(gb)
// (unnormalised; the 1/(nx ny nz) is applied by the x pass), then per voxel
// r = prec_y + V(g + x), x <- x + r, and the FP64 sum of |r|^2 of the
// workgroup's rows in partial[block].
template <int N, int E>
__global__ void __launch_bounds__((ZCfg<N, E>::THREADS))
mbs_zupdate_kernel(const float2* __restrict__ gb, const float2* __restrict__ eps,
                   const float2* __restrict__ py, float2* __restrict__ x, const float2* __restrict__ tw,
                   long long vox, long long nrows, int tensor, C2 s, C2 kb, double* __restrict__ partial,
                   const double* __restrict__ stop) {
  if (stopped(stop)) return;
  using Cfg = ZCfg<N, E>;
  constexpr int R = Cfg::R;
  __shared__ float2 lds[lds_span<N, Cfg::S>()];
  const long long row0 = (long long)blockIdx.x * R;
  const int nv = (int)min((long long)R, nrows - row0) * N;
  const long long base = row0 * N;
  for (int j = threadIdx.x; j < R * N; j += Cfg::THREADS) {
    const int r = j / N, z = j - r * N;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      lds[zslot<N, E>(z, r, c)] = j < nv ? gb[c * vox + base + j] : make_float2(0.f, 0.f);
  }
  block_sync();
  ztransform<N, E, 1>(lds, tw);
  double acc = 0.0;
  for (int j = threadIdx.x; j < nv; j += Cfg::THREADS) {
    const int r = j / N, z = j - r * N;
    const long long i = base + j;
    float2 xv[3], t[3], v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xv[c] = x[c * vox + i];
      const float2 gv = lds[zslot<N, E>(z, r, c)];
      t[c] = make_float2(gv.x + xv[c].x, gv.y + xv[c].y);
    }
    apply_v(eps, tensor, i, vox, s, kb, t, v);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long jj = c * vox + i;
      const float2 rr = make_float2(py[jj].x + v[c].x, py[jj].y + v[c].y);
      x[jj] = make_float2(xv[c].x + rr.x, xv[c].y + rr.y);
      acc += (double)rr.x * rr.x + (double)rr.y * rr.y;
    }
  }
  // fixed-order tree over the threads (padded to a power of two), in the
  // transform's LDS once every thread has read its values
  constexpr int P = pow2_ceil(Cfg::THREADS);
  static_assert(P * sizeof(double) <= lds_span<N, Cfg::S>() * sizeof(float2), "reduction exceeds LDS");
  block_sync();
  double* sh = reinterpret_cast<double*>(lds);
  sh[threadIdx.x] = acc;
  if (threadIdx.x + Cfg::THREADS < P) sh[threadIdx.x + Cfg::THREADS] = 0.0;
  block_sync();
  for (int w = P / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    block_sync();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// (length, values per thread) of the fused z passes: the lengths of the
// fused Green pass, with its thread split
#define CXM_ZFUSE_SIZES(X)                                                     \
  X(8, 8) X(16, 16) X(32, 16) X(64, 16) X(128, 16) X(256, 16) X(20, 20)         \
  X(40, 20) X(80, 20) X(160, 20) X(12, 12) X(24, 12) X(48, 12) X(96, 12) X(192, 12)

bool zfuse_supported(int n) {
  switch (n) {
#define CXM_ZF_CASE(N, E) case N:
    CXM_ZFUSE_SIZES(CXM_ZF_CASE)
#undef CXM_ZF_CASE
    return true;
    default:
      return false;
  }
}

// z rows per workgroup of the fused passes (0: no fused pass for n)
int zfuse_rows(int n) {
  switch (n) {
#define CXM_ZF_ROWS(N, E) \
  case N:                 \
    return ZCfg<N, E>::R;
    CXM_ZFUSE_SIZES(CXM_ZF_ROWS)
#undef CXM_ZF_ROWS
    default:
      return 0;
  }
}

// x -> gb = FFT_z(V x)
int launch_zv(int n, const float2* x, const float2* eps, float2* gb, long long vox, long long nrows,
              int tensor, C2 s, C2 kb, const double* stop, hipStream_t st) {
  switch (n) {
#define CXM_ZV_LAUNCH(N, E)                                                                         \
  case N: {                                                                                         \
    using Cfg = ZCfg<N, E>;                                                                         \
    const long long blocks = (nrows + Cfg::R - 1) / Cfg::R;                                         \
    if (blocks > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "fused z pass grid too large");        \
    hipLaunchKernelGGL((mbs_zv_kernel<N, E>), dim3((unsigned)blocks), dim3(Cfg::THREADS), 0, st, x, \
                       eps, gb, twiddles_for(N), vox, nrows, tensor, s, kb, stop);                  \
    CXM_LAUNCH_CHECK();                                                                             \
    return CXM_OK;                                                                                  \
  }
    CXM_ZFUSE_SIZES(CXM_ZV_LAUNCH)
#undef CXM_ZV_LAUNCH
    default:
      return fail(CXM_ERR_UNSUPPORTED, "no fused z pass for length " + std::to_string(n));
  }
}

// gb -> g = IFFT_z(gb); x <- x + prec_y + V(g + x); partial sums of |r|^2 in
// partial[0, *nblocks)
int launch_zupdate(int n, const float2* gb, const float2* eps, const float2* py, float2* x, long long vox,
                   long long nrows, int tensor, C2 s, C2 kb, double* partial, int* nblocks, const double* stop,
                   hipStream_t st) {
  switch (n) {
#define CXM_ZU_LAUNCH(N, E)                                                                          \
  case N: {                                                                                          \
    using Cfg = ZCfg<N, E>;                                                                          \
    const long long groups = (nrows + Cfg::R - 1) / Cfg::R;                                          \
    if (groups > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "fused z pass grid too large");         \
    const int blocks = (int)groups;                                                                  \
    hipLaunchKernelGGL((mbs_zupdate_kernel<N, E>), dim3(blocks), dim3(Cfg::THREADS), 0, st, gb, eps, \
                       py, x, twiddles_for(N), vox, nrows, tensor, s, kb, partial, stop);            \
    CXM_LAUNCH_CHECK();                                                                              \
    *nblocks = blocks;                                                                               \
    return CXM_OK;                                                                                   \
  }
    CXM_ZFUSE_SIZES(CXM_ZU_LAUNCH)
#undef CXM_ZU_LAUNCH
    default:
      return fail(CXM_ERR_UNSUPPORTED, "no fused z pass for length " + std::to_string(n));
  }
}

// r = prec_y + V(g + x);  x <- x + r;  partial[block] = sum |r|^2 (FP64)
__global__ void __launch_bounds__(kRedThreads)
mbs_update_kernel(const float2* __restrict__ g, const float2* __restrict__ eps,
                  const float2* __restrict__ py, float2* __restrict__ x, long long vox, int tensor,
                  C2 s, C2 kb, double* __restrict__ partial, const double* __restrict__ stop) {
  if (stopped(stop)) return;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < vox;
       i += (long long)gridDim.x * blockDim.x) {
    float2 xv[3], t[3], v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xv[c] = x[c * vox + i];
      const float2 gv = g[c * vox + i];
      t[c] = make_float2(gv.x + xv[c].x, gv.y + xv[c].y);
    }
    apply_v(eps, tensor, i, vox, s, kb, t, v);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * vox + i;
      const float2 r = make_float2(py[j].x + v[c].x, py[j].y + v[c].y);
      x[j] = make_float2(xv[c].x + r.x, xv[c].y + r.y);
      acc += (double)r.x * r.x + (double)r.y * r.y;
    }
  }
  __shared__ double sh[kRedThreads];
  sh[threadIdx.x] = acc;
  block_sync();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    block_sync();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out[0] = sum partial (fixed order), out[1] (optional max) untouched
// (one block of kRedThreads: strided partial sums, then a fixed tree)
__global__ void __launch_bounds__(kRedThreads)
sum_partials_kernel(const double* __restrict__ partial, int n, double* __restrict__ out) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) acc += partial[i];
  sh[threadIdx.x] = acc;
  block_sync();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    block_sync();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// The stopping test of cxm_mbs_iterate_stop after every iteration, on the
// device: status = [|r|^2 of the last iteration, iterations performed,
// stopped (0/1), |prec_y|^2]; |r|^2 is the fixed-order sum of the update
// pass's partials; stop once |r| <= target (rtol >= 0: rtol |prec_y|, else
// tol) or after maxiter iterations, the test of jaxopt.FixedPointIteration
// (solver.py:136-142).  Nothing changes once stopped.
__global__ void __launch_bounds__(kRedThreads)
mbs_stop_kernel(const double* __restrict__ partial, int n, double* __restrict__ status, int maxiter,
                double tol, double rtol) {
  if (status[2] != 0.0) return;
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) acc += partial[i];
  sh[threadIdx.x] = acc;
  block_sync();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    block_sync();
  }
  if (threadIdx.x == 0) {
    const double rs = sh[0];
    const double k = status[1] + 1.0;
    const double target = rtol >= 0.0 ? rtol * sqrt(status[3]) : tol;
    status[0] = rs;
    status[1] = k;
    if (sqrt(rs) <= target || k >= (double)maxiter) status[2] = 1.0;
  }
}

// partial[block] = max |v| over 3 planar components (as FP64)
__global__ void __launch_bounds__(kRedThreads)
maxabs_kernel(const float2* __restrict__ v, long long n, double* __restrict__ partial) {
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double a = sqrt((double)v[i].x * v[i].x + (double)v[i].y * v[i].y);
    m = a > m ? a : m;
  }
  __shared__ double sh[kRedThreads];
  sh[threadIdx.x] = m;
  block_sync();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    block_sync();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kRedThreads)
max_partials_kernel(const double* __restrict__ partial, int n, double* __restrict__ out) {
  __shared__ double sh[kRedThreads];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) m = fmax(m, partial[i]);
  sh[threadIdx.x] = m;
  block_sync();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    block_sync();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// v[c][i] *= 1 / max (normalisation by max |prec_y|, read on the device)
__global__ void scale_by_inv_kernel(float2* __restrict__ v, long long n, const double* __restrict__ m) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float f = (float)(1.0 / m[0]);
  v[i] = make_float2(v[i].x * f, v[i].y * f);
}

// E[i*3 + c] = x[c][i] * num_scale (interleaved output)
__global__ void mbs_output_kernel(const float2* __restrict__ x, float2* __restrict__ e, long long vox,
                                  const double* __restrict__ num_scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vox) return;
  const float f = (float)num_scale[0];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float2 v = x[c * vox + i];
    e[i * 3 + c] = make_float2(v.x * f, v.y * f);
  }
}

unsigned blocks_for(long long n) { return (unsigned)std::max(1LL, std::min<long long>((n + 255) / 256, 0x7fffffffLL)); }

// One axis of the 3D FFT of the 3 planar component volumes, in place
// (dir -1 forward, +1 inverse, unnormalised; `scale` on the output).
int pass_z(const MbsGeom& g, float2* v, int dir, float scale, hipStream_t s,
           const double* stop = nullptr) {
  // z: contiguous rows
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  a.in = v; a.out = v;
  a.tw = twiddles_for(g.nz);
  a.in_plane_stride = a.out_plane_stride = g.vox;
  a.in_line_stride = a.out_line_stride = g.nz;
  a.n_lines = g.nx * g.ny;
  a.n_planes = 3;
  a.grid_ny = g.nx * g.ny;
  a.grid_nx = g.nz;
  a.in_len = a.out_len = g.nz;
  a.seq = dir < 0 ? SEQ_F : SEQ_I;
  a.scale = scale;
  a.mid_scale = 1.0f;
  a.stop = stop;
  return launch_pipe_row(g.nz, a, s);
}

int pass_y(const MbsGeom& g, float2* v, int dir, float scale, hipStream_t s,
           const double* stop = nullptr) {
  // y: column strips of each (ny x nz) slab, 3 nx planes
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  a.in = v; a.out = v;
  a.tw = twiddles_for(g.ny);
  a.in_plane_stride = a.out_plane_stride = (long long)g.ny * g.nz;
  a.in_line_stride = a.out_line_stride = g.nz;
  a.n_lines = g.nz;
  a.n_planes = 3 * g.nx;
  a.grid_ny = g.ny;
  a.grid_nx = g.nz;
  a.in_len = a.out_len = g.ny;
  a.out_line_len = g.nz;
  a.seq = dir < 0 ? SEQ_F : SEQ_I;
  a.scale = scale;
  a.mid_scale = 1.0f;
  a.stop = stop;
  return launch_pipe_col(g.ny, a, s);
}

int pass_x(const MbsGeom& g, float2* v, int dir, float scale, hipStream_t s,
           const double* stop = nullptr) {
  // x: column strips of the (nx x ny nz) matrix, 3 planes
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  a.in = v; a.out = v;
  a.tw = twiddles_for(g.nx);
  a.in_plane_stride = a.out_plane_stride = g.vox;
  a.in_line_stride = a.out_line_stride = g.ny * g.nz;
  a.n_lines = g.ny * g.nz;
  a.n_planes = 3;
  a.grid_ny = g.nx;
  a.grid_nx = g.ny * g.nz;
  a.in_len = a.out_len = g.nx;
  a.out_line_len = g.ny * g.nz;
  a.seq = dir < 0 ? SEQ_F : SEQ_I;
  a.scale = scale;
  a.mid_scale = 1.0f;
  a.stop = stop;
  return launch_pipe_col(g.nx, a, s);
}

// g <- G(g) = ifftn(P_T g^ / (kb - s |K|^2) + P_L g^ / kb): z and y passes
// forward, the fused x pass (x FFT, Green multiply, inverse x FFT, the
// 1/(nx ny nz) scale) where its length has one, else the x pass, the Green
// kernel and the inverse x pass; then the y and z passes inverse.
// The passes between the first and the last z pass of G: y forward, x
// (fused with the Green multiply where possible), y inverse.
int green_inner(const MbsGeom& g, const cxm_mbs_desc* d, float2* gb, C2 sc, C2 kb, hipStream_t s,
                const double* stop = nullptr) {
  int rc;
  const float inv = (float)(1.0 / ((double)g.nx * g.ny * g.nz));
  const float kscale = (float)(2.0 * M_PI / d->spacing / d->k0);
  if ((rc = pass_y(g, gb, -1, 1.0f, s, stop))) return rc;
  if (xgreen_supported(g.nx) && option(CXM_OPT_MBS_XGREEN) != 2) {
    if ((rc = launch_xgreen(g.nx, gb, g.vox, g.ny * g.nz, g.ny, g.nz, kscale, sc, kb, inv, stop, s))) return rc;
  } else {
    if ((rc = pass_x(g, gb, -1, 1.0f, s, stop))) return rc;
    hipLaunchKernelGGL(mbs_green_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, gb, g.vox, g.nx, g.ny,
                       g.nz, kscale, sc, kb, stop);
    CXM_LAUNCH_CHECK();
    if ((rc = pass_x(g, gb, 1, inv, s, stop))) return rc;
  }
  return pass_y(g, gb, 1, 1.0f, s, stop);
}

int green_apply(const MbsGeom& g, const cxm_mbs_desc* d, float2* gb, C2 sc, C2 kb, hipStream_t s,
                const double* stop = nullptr) {
  int rc;
  if ((rc = pass_z(g, gb, -1, 1.0f, s, stop))) return rc;
  if ((rc = green_inner(g, d, gb, sc, kb, s, stop))) return rc;
  return pass_z(g, gb, 1, 1.0f, s, stop);
}

// The fixed-point iteration runs its first and last z passes fused with V
// and with the update where the z length has them (CXM_OPT_MBS_ZFUSE 0).
bool zfused(const MbsGeom& g) { return zfuse_supported(g.nz) && option(CXM_OPT_MBS_ZFUSE) != 1; }

C2 c2(double re, double im) { return C2{(float)re, (float)im}; }

// ---- BiCGStab on the preconditioned system ---------------------------------
// Solves prec_forward(x) = prec_y with the stabilised bi-conjugate gradient
// iteration of jax.scipy.sparse.linalg.bicgstab (no preconditioner M; the
// solver calls it as spa.bicgstab(prec_forward, prec_y, x, maxiter, tol),
// solver.py:141-143).  With A = prec_forward, b = prec_y, per iteration:
//   rho' = <rh, r>, beta = rho'/rho alpha/omega, p = r + beta (p - omega q),
//   q = A p, alpha' = rho' / <rh, q>, s = r - alpha' q,
//   early = |s|^2 < atol2, t = A s, omega' = <t, s> / <t, t>,
//   x += alpha' p (+ omega' s unless early), r = s (- omega' t unless early),
//   k = -11 if omega' == 0 or alpha' == 0, -10 if rho' == 0, else k + 1,
// while |r|^2 > atol2 = max(tol^2 |b|^2, atol^2) and k >= 0 (<a, b> conjugates
// a).  The scalars live on the device (FP64), so chunks of iterations are
// enqueued without host round trips; once stopped, every kernel of a later
// iteration leaves the vectors untouched.
struct BiScal {
  double rho[2], alpha[2], omega[2];  // values of the last completed iteration
  double beta[2], rho_n[2], alpha_n[2], omega_n[2];  // this iteration
  double rr[2];                       // <rh, r> of the current r
  double rs, atol2, bs, k, early, done;
};

CXM_DEV void cmul_d(const double* a, const double* b, double* o) {
  const double re = a[0] * b[0] - a[1] * b[1], im = a[0] * b[1] + a[1] * b[0];
  o[0] = re;
  o[1] = im;
}
CXM_DEV void cdiv_d(const double* a, const double* b, double* o) {
  const double d = b[0] * b[0] + b[1] * b[1];
  const double re = (a[0] * b[0] + a[1] * b[1]) / d, im = (a[1] * b[0] - a[0] * b[1]) / d;
  o[0] = re;
  o[1] = im;
}
CXM_DEV float2 cf(const double* a) { return make_float2((float)a[0], (float)a[1]); }
CXM_DEV float2 cmul2(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
CXM_DEV float2 cadd2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
CXM_DEV float2 csub2(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// conj(a) b accumulated in FP64
CXM_DEV void cdot_acc(float2 a, float2 b, double& re, double& im) {
  re += (double)a.x * b.x + (double)a.y * b.y;
  im += (double)a.x * b.y - (double)a.y * b.x;
}

// Block sums of NV doubles to partial[block * 4 + j] (fixed order).
template <int NV>
CXM_DEV void block_sums(double (&v)[NV], double* partial) {
  __shared__ double sh[NV][kRedThreads];
#pragma unroll
  for (int j = 0; j < NV; ++j) sh[j][threadIdx.x] = v[j];
  block_sync();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int j = 0; j < NV; ++j) sh[j][threadIdx.x] += sh[j][threadIdx.x + w];
    }
    block_sync();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) partial[blockIdx.x * 4 + j] = sh[j][0];
  }
}

#define BI_LOOP(i) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < vox; \
       i += (long long)gridDim.x * blockDim.x)

// r = b - A x = b + V(g + x); rh = p = q = r; partials |r|^2, |b|^2
__global__ void __launch_bounds__(kRedThreads)
bi_r0_kernel(const float2* __restrict__ g, const float2* __restrict__ eps, const float2* __restrict__ b,
             const float2* __restrict__ x, float2* __restrict__ r, float2* __restrict__ rh,
             float2* __restrict__ p, float2* __restrict__ q, long long vox, int tensor, C2 s, C2 kb,
             double* __restrict__ partial) {
  double acc[2] = {0.0, 0.0};
  BI_LOOP(i) {
    float2 t[3], v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) t[c] = cadd2(g[c * vox + i], x[c * vox + i]);
    apply_v(eps, tensor, i, vox, s, kb, t, v);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * vox + i;
      const float2 bv = b[j];
      const float2 rv = cadd2(bv, v[c]);
      r[j] = rv;
      rh[j] = rv;
      p[j] = rv;
      q[j] = rv;
      acc[0] += (double)rv.x * rv.x + (double)rv.y * rv.y;
      acc[1] += (double)bv.x * bv.x + (double)bv.y * bv.y;
    }
  }
  block_sums<2>(acc, partial);
}

// partial sums of NV values -> out[0..NV) (one block, fixed order)
template <int NV>
CXM_DEV void sum_partials(const double* partial, double (&out)[NV]) {
  __shared__ double sh[NV][kRedThreads];
  double a[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) a[j] = 0.0;
  for (int b = threadIdx.x; b < kRedBlocks; b += kRedThreads) {
#pragma unroll
    for (int j = 0; j < NV; ++j) a[j] += partial[b * 4 + j];
  }
#pragma unroll
  for (int j = 0; j < NV; ++j) sh[j][threadIdx.x] = a[j];
  block_sync();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int j = 0; j < NV; ++j) sh[j][threadIdx.x] += sh[j][threadIdx.x + w];
    }
    block_sync();
  }
#pragma unroll
  for (int j = 0; j < NV; ++j) out[j] = sh[j][0];
}

__global__ void __launch_bounds__(kRedThreads)
bi_init_scalars_kernel(const double* __restrict__ partial, BiScal* st, double tol, double atol,
                       double* __restrict__ status) {
  double v[2];
  sum_partials<2>(partial, v);
  if (threadIdx.x != 0) return;
  st->rho[0] = st->alpha[0] = st->omega[0] = 1.0;
  st->rho[1] = st->alpha[1] = st->omega[1] = 0.0;
  st->rs = v[0];
  st->bs = v[1];
  st->rr[0] = v[0];  // rh = r
  st->rr[1] = 0.0;
  st->atol2 = fmax(tol * tol * v[1], atol * atol);
  st->k = 0.0;
  st->early = 0.0;
  st->done = (st->rs > st->atol2) ? 0.0 : 1.0;
  status[0] = st->rs;
  status[1] = st->k;
  status[2] = st->done;
  status[3] = st->bs;
}

// rho' = <rh, r>, beta = rho'/rho alpha/omega
__global__ void bi_beta_kernel(BiScal* st) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || st->done != 0.0) return;
  st->rho_n[0] = st->rr[0];
  st->rho_n[1] = st->rr[1];
  double t[2], u[2];
  cdiv_d(st->rho_n, st->rho, t);
  cdiv_d(st->alpha, st->omega, u);
  cmul_d(t, u, st->beta);
}

// p = r + beta (p - omega q);  g = V(p)
__global__ void __launch_bounds__(kRedThreads)
bi_p_kernel(const float2* __restrict__ r, float2* __restrict__ p, const float2* __restrict__ q,
            const float2* __restrict__ eps, float2* __restrict__ g, long long vox, int tensor, C2 s,
            C2 kb, const BiScal* st) {
  if (st->done != 0.0) return;
  const float2 beta = cf(st->beta), omega = cf(st->omega);
  BI_LOOP(i) {
    float2 pv[3], v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * vox + i;
      pv[c] = cadd2(r[j], cmul2(beta, csub2(p[j], cmul2(omega, q[j]))));
      p[j] = pv[c];
    }
    apply_v(eps, tensor, i, vox, s, kb, pv, v);
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c * vox + i] = v[c];
  }
}

// q = A p = -V(g + p); partial <rh, q>
__global__ void __launch_bounds__(kRedThreads)
bi_q_kernel(const float2* __restrict__ g, const float2* __restrict__ p, const float2* __restrict__ rh,
            const float2* __restrict__ eps, float2* __restrict__ q, long long vox, int tensor, C2 s,
            C2 kb, const BiScal* st, double* __restrict__ partial) {
  if (st->done != 0.0) return;
  double acc[2] = {0.0, 0.0};
  BI_LOOP(i) {
    float2 t[3], w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) t[c] = cadd2(g[c * vox + i], p[c * vox + i]);
    apply_v(eps, tensor, i, vox, s, kb, t, w);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * vox + i;
      const float2 qv = make_float2(-w[c].x, -w[c].y);
      q[j] = qv;
      cdot_acc(rh[j], qv, acc[0], acc[1]);
    }
  }
  block_sums<2>(acc, partial);
}

// alpha' = rho' / <rh, q>
__global__ void __launch_bounds__(kRedThreads)
bi_alpha_kernel(const double* __restrict__ partial, BiScal* st) {
  if (st->done != 0.0) return;
  double v[2];
  sum_partials<2>(partial, v);
  if (threadIdx.x == 0) cdiv_d(st->rho_n, v, st->alpha_n);
}

// s = r - alpha' q (into r);  g = V(s);  partial |s|^2
__global__ void __launch_bounds__(kRedThreads)
bi_s_kernel(float2* __restrict__ r, const float2* __restrict__ q, const float2* __restrict__ eps,
            float2* __restrict__ g, long long vox, int tensor, C2 s, C2 kb, const BiScal* st,
            double* __restrict__ partial) {
  if (st->done != 0.0) return;
  const float2 alpha = cf(st->alpha_n);
  double acc[1] = {0.0};
  BI_LOOP(i) {
    float2 sv[3], v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * vox + i;
      sv[c] = csub2(r[j], cmul2(alpha, q[j]));
      r[j] = sv[c];
      acc[0] += (double)sv[c].x * sv[c].x + (double)sv[c].y * sv[c].y;
    }
    apply_v(eps, tensor, i, vox, s, kb, sv, v);
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c * vox + i] = v[c];
  }
  block_sums<1>(acc, partial);
}

__global__ void __launch_bounds__(kRedThreads)
bi_early_kernel(const double* __restrict__ partial, BiScal* st) {
  if (st->done != 0.0) return;
  double v[1];
  sum_partials<1>(partial, v);
  if (threadIdx.x == 0) st->early = (v[0] < st->atol2) ? 1.0 : 0.0;
}

// t = A s = -V(g + s) (into g); partials <t, s>, <t, t>
__global__ void __launch_bounds__(kRedThreads)
bi_t_kernel(float2* __restrict__ g, const float2* __restrict__ sv_, const float2* __restrict__ eps,
            long long vox, int tensor, C2 s, C2 kb, const BiScal* st, double* __restrict__ partial) {
  if (st->done != 0.0) return;
  double acc[3] = {0.0, 0.0, 0.0};
  BI_LOOP(i) {
    float2 u[3], w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c] = cadd2(g[c * vox + i], sv_[c * vox + i]);
    apply_v(eps, tensor, i, vox, s, kb, u, w);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * vox + i;
      const float2 sv = sv_[j];
      const float2 tv = make_float2(-w[c].x, -w[c].y);
      g[j] = tv;
      cdot_acc(tv, sv, acc[0], acc[1]);
      acc[2] += (double)tv.x * tv.x + (double)tv.y * tv.y;
    }
  }
  block_sums<3>(acc, partial);
}

__global__ void __launch_bounds__(kRedThreads)
bi_omega_kernel(const double* __restrict__ partial, BiScal* st) {
  if (st->done != 0.0) return;
  double v[3];
  sum_partials<3>(partial, v);
  if (threadIdx.x == 0) {
    const double tt[2] = {v[2], 0.0};
    cdiv_d(v, tt, st->omega_n);
  }
}

// x += alpha' p (+ omega' s);  r = s (- omega' t);  partials |r|^2, <rh, r>
__global__ void __launch_bounds__(kRedThreads)
bi_x_kernel(float2* __restrict__ x, const float2* __restrict__ p, float2* __restrict__ r,
            const float2* __restrict__ t, const float2* __restrict__ rh, long long vox,
            const BiScal* st, double* __restrict__ partial) {
  if (st->done != 0.0) return;
  const float2 alpha = cf(st->alpha_n), omega = cf(st->omega_n);
  const bool early = st->early != 0.0;
  double acc[3] = {0.0, 0.0, 0.0};
  BI_LOOP(i) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * vox + i;
      const float2 sv = r[j];
      float2 xv = cadd2(x[j], cmul2(alpha, p[j]));
      float2 rv = sv;
      if (!early) {
        xv = cadd2(xv, cmul2(omega, sv));
        rv = csub2(sv, cmul2(omega, t[j]));
      }
      x[j] = xv;
      r[j] = rv;
      acc[0] += (double)rv.x * rv.x + (double)rv.y * rv.y;
      cdot_acc(rh[j], rv, acc[1], acc[2]);
    }
  }
  block_sums<3>(acc, partial);
}

__global__ void __launch_bounds__(kRedThreads)
bi_end_kernel(const double* __restrict__ partial, BiScal* st, double* __restrict__ status) {
  if (st->done == 0.0) {
    double v[3];
    sum_partials<3>(partial, v);
    if (threadIdx.x == 0) {
      st->rs = v[0];
      st->rr[0] = v[1];
      st->rr[1] = v[2];
      double k = st->k + 1.0;
      if ((st->omega_n[0] == 0.0 && st->omega_n[1] == 0.0) ||
          (st->alpha_n[0] == 0.0 && st->alpha_n[1] == 0.0))
        k = -11.0;
      if (st->rho_n[0] == 0.0 && st->rho_n[1] == 0.0) k = -10.0;
      st->k = k;
      for (int j = 0; j < 2; ++j) {
        st->rho[j] = st->rho_n[j];
        st->alpha[j] = st->alpha_n[j];
        st->omega[j] = st->omega_n[j];
      }
      st->done = (st->rs > st->atol2 && k >= 0.0) ? 0.0 : 1.0;
    }
  }
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    status[0] = st->rs;
    status[1] = st->k;
    status[2] = st->done;
    status[3] = st->bs;
  }
}
#undef BI_LOOP

}  // namespace
}  // namespace cxm

using namespace cxm;

// workspace: gbuf (3 vox c64) | partial (f64: kRedBlocks, or one per row
// group of the fused z pass if more) | scalars (4 f64)
extern "C" size_t cxm_mbs_workspace_size(const cxm_mbs_desc* d) {
  MbsGeom g;
  if (mbs_geom(d, g)) return 0;
  long long parts = kRedBlocks;
  const int zr = zfuse_rows(g.nz);
  if (zr > 0) parts = std::max(parts, ((long long)g.nx * g.ny + zr - 1) / zr);
  return (size_t)3 * g.vox * sizeof(float2) + (size_t)(parts + 4) * sizeof(double);
}

extern "C" int cxm_mbs_precondition(void* stream, const void* eps, const void* source,
                                    void* prec_y, double* num_scale, void* workspace,
                                    size_t workspace_bytes, const cxm_mbs_desc* d) {
  MbsGeom g;
  int rc = mbs_geom(d, g);
  if (rc) return rc;
  if (!eps || !source || !prec_y || !num_scale) return fail(CXM_ERR_INVALID, "null buffer");
  if (!workspace || workspace_bytes < cxm_mbs_workspace_size(d))
    return fail(CXM_ERR_WORKSPACE, "workspace too small");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  const C2 sc = c2(d->scale_re, d->scale_im);
  const C2 kb = c2(d->bias_re * d->scale_re - d->bias_im * d->scale_im + 1.0,
                   d->bias_re * d->scale_im + d->bias_im * d->scale_re);
  float2* py = (float2*)prec_y;
  double* partial = (double*)((char*)workspace + (size_t)3 * g.vox * sizeof(float2));
  // t = s * source (planar), G(t), prec_y = -V(G(t))
  hipLaunchKernelGGL(mbs_v_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, (const float2*)source,
                     nullptr, py, g.vox, 1, 0, sc, kb, sc, nullptr);
  CXM_LAUNCH_CHECK();
  if ((rc = green_apply(g, d, py, sc, kb, s))) return rc;
  hipLaunchKernelGGL(mbs_v_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, (const float2*)py,
                     (const float2*)eps, py, g.vox, 0, d->eps_tensor, sc, kb, C2{-1.0f, 0.0f}, nullptr);
  CXM_LAUNCH_CHECK();
  // num_scale = max |prec_y|; prec_y /= num_scale
  hipLaunchKernelGGL(maxabs_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, (const float2*)py,
                     3 * g.vox, partial);
  CXM_LAUNCH_CHECK();
  hipLaunchKernelGGL(max_partials_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, kRedBlocks, num_scale);
  CXM_LAUNCH_CHECK();
  hipLaunchKernelGGL(scale_by_inv_kernel, dim3(blocks_for(3 * g.vox)), dim3(256), 0, s, py, 3 * g.vox,
                     (const double*)num_scale);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

namespace cxm {
namespace {
// `iterations` fixed-point steps; with `status` (cxm_mbs_iterate_stop) each
// step is followed by the device-side stopping test and every kernel of a
// step returns at once after the stop, else |r|^2 of the last step goes to
// residual_sq.
int mbs_steps(const MbsGeom& g, const cxm_mbs_desc* d, const void* eps, const void* prec_y, void* x,
              int iterations, void* workspace, double* residual_sq, double* status, int maxiter, double tol,
              double rtol, hipStream_t s) {
  int rc;
  const C2 sc = c2(d->scale_re, d->scale_im);
  const C2 kb = c2(d->bias_re * d->scale_re - d->bias_im * d->scale_im + 1.0,
                   d->bias_re * d->scale_im + d->bias_im * d->scale_re);
  float2* gb = (float2*)workspace;
  double* partial = (double*)((char*)workspace + (size_t)3 * g.vox * sizeof(float2));
  const double* stop = status ? status + 2 : nullptr;
  const bool fz = zfused(g);
  const long long nrows = (long long)g.nx * g.ny;
  int nparts = kRedBlocks;
  for (int it = 0; it < iterations; ++it) {
    if (fz) {
      if ((rc = launch_zv(g.nz, (const float2*)x, (const float2*)eps, gb, g.vox, nrows, d->eps_tensor, sc,
                          kb, stop, s)))
        return rc;
      if ((rc = green_inner(g, d, gb, sc, kb, s, stop))) return rc;
      if ((rc = launch_zupdate(g.nz, gb, (const float2*)eps, (const float2*)prec_y, (float2*)x, g.vox, nrows,
                               d->eps_tensor, sc, kb, partial, &nparts, stop, s)))
        return rc;
    } else {
      hipLaunchKernelGGL(mbs_v_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, (const float2*)x,
                         (const float2*)eps, gb, g.vox, 0, d->eps_tensor, sc, kb, C2{1.0f, 0.0f}, stop);
      CXM_LAUNCH_CHECK();
      if ((rc = green_apply(g, d, gb, sc, kb, s, stop))) return rc;
      hipLaunchKernelGGL(mbs_update_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s,
                         (const float2*)gb, (const float2*)eps, (const float2*)prec_y, (float2*)x,
                         g.vox, d->eps_tensor, sc, kb, partial, stop);
      CXM_LAUNCH_CHECK();
    }
    if (status) {
      hipLaunchKernelGGL(mbs_stop_kernel, dim3(1), dim3(kRedThreads), 0, s, (const double*)partial, nparts,
                         status, maxiter, tol, rtol);
      CXM_LAUNCH_CHECK();
    }
  }
  if (!status && iterations > 0) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, nparts,
                       residual_sq);
    CXM_LAUNCH_CHECK();
  }
  return CXM_OK;
}
}  // namespace
}  // namespace cxm

extern "C" int cxm_mbs_iterate(void* stream, const void* eps, const void* prec_y, void* x,
                               int32_t iterations, double* residual_sq, void* workspace,
                               size_t workspace_bytes, const cxm_mbs_desc* d) {
  MbsGeom g;
  int rc = mbs_geom(d, g);
  if (rc) return rc;
  if (iterations < 0) return fail(CXM_ERR_INVALID, "iterations must be >= 0");
  if (!eps || !prec_y || !x || !residual_sq) return fail(CXM_ERR_INVALID, "null buffer");
  if (!workspace || workspace_bytes < cxm_mbs_workspace_size(d))
    return fail(CXM_ERR_WORKSPACE, "workspace too small");
  if ((rc = ensure_init())) return rc;
  return mbs_steps(g, d, eps, prec_y, x, iterations, workspace, residual_sq, nullptr, 0, 0.0, -1.0,
                   (hipStream_t)stream);
}

extern "C" int cxm_mbs_iterate_stop(void* stream, const void* eps, const void* prec_y, void* x,
                                    int32_t iterations, int32_t maxiter, double tol, double rtol,
                                    double* status, void* workspace, size_t workspace_bytes,
                                    const cxm_mbs_desc* d) {
  MbsGeom g;
  int rc = mbs_geom(d, g);
  if (rc) return rc;
  if (iterations < 0 || maxiter < 0) return fail(CXM_ERR_INVALID, "iterations and maxiter must be >= 0");
  if (!(tol >= 0.0)) return fail(CXM_ERR_INVALID, "tol must be >= 0");
  if (!eps || !prec_y || !x || !status) return fail(CXM_ERR_INVALID, "null buffer");
  if (!workspace || workspace_bytes < cxm_mbs_workspace_size(d))
    return fail(CXM_ERR_WORKSPACE, "workspace too small");
  if ((rc = ensure_init())) return rc;
  return mbs_steps(g, d, eps, prec_y, x, iterations, workspace, nullptr, status, maxiter, tol,
                   rtol >= 0.0 ? rtol : -1.0, (hipStream_t)stream);
}

extern "C" int cxm_mbs_output(void* stream, const void* x, const double* num_scale, void* e,
                              const cxm_mbs_desc* d) {
  MbsGeom g;
  int rc = mbs_geom(d, g);
  if (rc) return rc;
  if (!x || !num_scale || !e) return fail(CXM_ERR_INVALID, "null buffer");
  hipLaunchKernelGGL(mbs_output_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)x, (float2*)e, g.vox, num_scale);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

// ---- BiCGStab entry points -----------------------------------------------------
namespace cxm {
namespace {
// state: r | rh | p | q (3 vox c64 each) | BiScal (256 B) | partials (kRedBlocks x 4 f64)
struct BiLayout {
  float2 *r, *rh, *p, *q;
  BiScal* st;
  double* partial;
};
size_t bi_state_bytes(const MbsGeom& g) {
  return (size_t)4 * 3 * g.vox * sizeof(float2) + 256 + (size_t)kRedBlocks * 4 * sizeof(double);
}
BiLayout bi_layout(const MbsGeom& g, void* state) {
  BiLayout b;
  const long long n = 3 * g.vox;
  b.r = (float2*)state;
  b.rh = b.r + n;
  b.p = b.rh + n;
  b.q = b.p + n;
  b.st = (BiScal*)(b.q + n);
  b.partial = (double*)((char*)b.st + 256);
  return b;
}
static_assert(sizeof(BiScal) <= 256, "BiCGStab scalars exceed their slot");


int bi_check(const cxm_mbs_desc* d, MbsGeom& g, const void* eps, const void* x, void* state,
             size_t state_bytes, double* status, void* workspace, size_t workspace_bytes) {
  int rc = mbs_geom(d, g);
  if (rc) return rc;
  if (!eps || !x || !state || !status) return fail(CXM_ERR_INVALID, "null buffer");
  if (state_bytes < bi_state_bytes(g)) return fail(CXM_ERR_WORKSPACE, "BiCGStab state too small");
  if (!workspace || workspace_bytes < cxm_mbs_workspace_size(d))
    return fail(CXM_ERR_WORKSPACE, "workspace too small");
  return ensure_init();
}
}  // namespace
}  // namespace cxm

extern "C" size_t cxm_mbs_bicgstab_state_size(const cxm_mbs_desc* d) {
  MbsGeom g;
  if (mbs_geom(d, g)) return 0;
  return bi_state_bytes(g);
}

extern "C" int cxm_mbs_bicgstab_init(void* stream, const void* eps, const void* prec_y, const void* x,
                                     void* state, size_t state_bytes, double tol, double atol,
                                     double* status, void* workspace, size_t workspace_bytes,
                                     const cxm_mbs_desc* d) {
  MbsGeom g;
  int rc = bi_check(d, g, eps, x, state, state_bytes, status, workspace, workspace_bytes);
  if (rc) return rc;
  if (!prec_y) return fail(CXM_ERR_INVALID, "null buffer");
  if (!(tol >= 0.0) || !(atol >= 0.0)) return fail(CXM_ERR_INVALID, "tol and atol must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  const C2 sc = c2(d->scale_re, d->scale_im);
  const C2 kb = c2(d->bias_re * d->scale_re - d->bias_im * d->scale_im + 1.0,
                   d->bias_re * d->scale_im + d->bias_im * d->scale_re);
  float2* gb = (float2*)workspace;
  const BiLayout b = bi_layout(g, state);
  // r0 = b - A x0 = b + V(G(V(x0)) + x0)
  hipLaunchKernelGGL(mbs_v_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, (const float2*)x,
                     (const float2*)eps, gb, g.vox, 0, d->eps_tensor, sc, kb, C2{1.0f, 0.0f}, nullptr);
  CXM_LAUNCH_CHECK();
  if ((rc = green_apply(g, d, gb, sc, kb, s))) return rc;
  hipLaunchKernelGGL(bi_r0_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, (const float2*)gb,
                     (const float2*)eps, (const float2*)prec_y, (const float2*)x, b.r, b.rh, b.p, b.q,
                     g.vox, d->eps_tensor, sc, kb, b.partial);
  CXM_LAUNCH_CHECK();
  hipLaunchKernelGGL(bi_init_scalars_kernel, dim3(1), dim3(kRedThreads), 0, s, (const double*)b.partial,
                     b.st, tol, atol, status);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

extern "C" int cxm_mbs_bicgstab_iterate(void* stream, const void* eps, void* x, void* state,
                                        size_t state_bytes, int32_t iterations, double* status,
                                        void* workspace, size_t workspace_bytes, const cxm_mbs_desc* d) {
  MbsGeom g;
  int rc = bi_check(d, g, eps, x, state, state_bytes, status, workspace, workspace_bytes);
  if (rc) return rc;
  if (iterations < 0) return fail(CXM_ERR_INVALID, "iterations must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  const C2 sc = c2(d->scale_re, d->scale_im);
  const C2 kb = c2(d->bias_re * d->scale_re - d->bias_im * d->scale_im + 1.0,
                   d->bias_re * d->scale_im + d->bias_im * d->scale_re);
  float2* gb = (float2*)workspace;
  const BiLayout b = bi_layout(g, state);
  const float2* ep = (const float2*)eps;
  const dim3 rg(kRedBlocks), rt(kRedThreads);
  for (int it = 0; it < iterations; ++it) {
    hipLaunchKernelGGL(bi_beta_kernel, dim3(1), dim3(1), 0, s, b.st);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(bi_p_kernel, rg, rt, 0, s, (const float2*)b.r, b.p, (const float2*)b.q, ep, gb,
                       g.vox, d->eps_tensor, sc, kb, (const BiScal*)b.st);
    CXM_LAUNCH_CHECK();
    if ((rc = green_apply(g, d, gb, sc, kb, s, &b.st->done))) return rc;
    hipLaunchKernelGGL(bi_q_kernel, rg, rt, 0, s, (const float2*)gb, (const float2*)b.p,
                       (const float2*)b.rh, ep, b.q, g.vox, d->eps_tensor, sc, kb, (const BiScal*)b.st,
                       b.partial);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(bi_alpha_kernel, dim3(1), rt, 0, s, (const double*)b.partial, b.st);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(bi_s_kernel, rg, rt, 0, s, b.r, (const float2*)b.q, ep, gb, g.vox, d->eps_tensor, sc,
                       kb, (const BiScal*)b.st, b.partial);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(bi_early_kernel, dim3(1), rt, 0, s, (const double*)b.partial, b.st);
    CXM_LAUNCH_CHECK();
    if ((rc = green_apply(g, d, gb, sc, kb, s, &b.st->done))) return rc;
    hipLaunchKernelGGL(bi_t_kernel, rg, rt, 0, s, gb, (const float2*)b.r, ep, g.vox, d->eps_tensor, sc,
                       kb, (const BiScal*)b.st, b.partial);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(bi_omega_kernel, dim3(1), rt, 0, s, (const double*)b.partial, b.st);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(bi_x_kernel, rg, rt, 0, s, (float2*)x, (const float2*)b.p, b.r, (const float2*)gb,
                       (const float2*)b.rh, g.vox, (const BiScal*)b.st, b.partial);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(bi_end_kernel, dim3(1), rt, 0, s, (const double*)b.partial, b.st, status);
    CXM_LAUNCH_CHECK();
  }
  return CXM_OK;
}

// ---- smallest enclosing circle of the permittivity values --------------------
// get_shift_and_scale (solver.py:10-39) minimises max |eps - shift| over the
// complex shift: the centre of the smallest circle enclosing the values.  The
// device finds the extreme values in 8 directions (their octagon lies inside
// the convex hull), keeps only the values outside that octagon as candidates,
// and the host runs Welzl's algorithm on the candidates and the 8 extremes;
// a last device pass takes the radius as max |eps - centre| over every value.
namespace cxm {
namespace {
constexpr int kEcBlocks = 1024;
constexpr int kEcDirs = 8;

CXM_DEV double ec_proj(double x, double y, int d) {
  switch (d) {
    case 0: return x;
    case 1: return x + y;
    case 2: return y;
    case 3: return y - x;
    case 4: return -x;
    case 5: return -x - y;
    case 6: return -y;
    default: return x - y;
  }
}

// partial[(block * 8 + d) * 2] = best projection, [.. + 1] = its index (as double)
__global__ void __launch_bounds__(256)
ec_extreme_kernel(const float2* __restrict__ v, long long n, double* __restrict__ partial) {
  double best[kEcDirs];
  long long arg[kEcDirs];
#pragma unroll
  for (int d = 0; d < kEcDirs; ++d) {
    best[d] = -INFINITY;
    arg[d] = -1;
  }
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = v[i].x, y = v[i].y;
#pragma unroll
    for (int d = 0; d < kEcDirs; ++d) {
      const double p = ec_proj(x, y, d);
      if (p > best[d]) {
        best[d] = p;
        arg[d] = i;
      }
    }
  }
  __shared__ double sb[256];
  __shared__ long long sa[256];
  for (int d = 0; d < kEcDirs; ++d) {
    sb[threadIdx.x] = best[d];
    sa[threadIdx.x] = arg[d];
    block_sync();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w && sb[threadIdx.x + w] > sb[threadIdx.x]) {
        sb[threadIdx.x] = sb[threadIdx.x + w];
        sa[threadIdx.x] = sa[threadIdx.x + w];
      }
      block_sync();
    }
    if (threadIdx.x == 0) {
      partial[(blockIdx.x * kEcDirs + d) * 2] = sb[0];
      partial[(blockIdx.x * kEcDirs + d) * 2 + 1] = (double)sa[0];
    }
    block_sync();
  }
}

// ext[2 d], ext[2 d + 1]: the extreme value in direction d; count = 0
// (256 threads: 32 per direction scan the block partials, then the best of
// each 32 with ties to the earlier block, as a serial scan would pick)
__global__ void __launch_bounds__(256)
ec_extreme_final_kernel(const float2* __restrict__ v, const double* __restrict__ partial,
                        double* __restrict__ ext, int* __restrict__ count) {
  const int d = threadIdx.x / 32, l = threadIdx.x % 32;
  double best = -INFINITY;
  int bb = -1;
  for (int b = l; b < kEcBlocks; b += 32) {
    const double p = partial[(b * kEcDirs + d) * 2];
    if (p > best) {
      best = p;
      bb = b;
    }
  }
  __shared__ double sv[256];
  __shared__ int sb[256];
  sv[threadIdx.x] = best;
  sb[threadIdx.x] = bb;
  block_sync();
  for (int w = 16; w > 0; w >>= 1) {
    if (l < w) {
      const double o = sv[threadIdx.x + w];
      const int ob = sb[threadIdx.x + w];
      const int mb = sb[threadIdx.x];
      if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && ob >= 0 && (mb < 0 || ob < mb))) {
        sv[threadIdx.x] = o;
        sb[threadIdx.x] = ob;
      }
    }
    block_sync();
  }
  if (l == 0) {
    const int b = sb[threadIdx.x];
    const long long arg = b >= 0 ? (long long)partial[(b * kEcDirs + d) * 2 + 1] : 0;
    ext[2 * d] = v[arg < 0 ? 0 : arg].x;
    ext[2 * d + 1] = v[arg < 0 ? 0 : arg].y;
  }
  if (threadIdx.x == 0) *count = 0;
}

// values strictly outside the octagon of the extremes (counter-clockwise
// in direction order) are appended to cand
__global__ void __launch_bounds__(256)
ec_filter_kernel(const float2* __restrict__ v, long long n, const double* __restrict__ ext,
                 float2* __restrict__ cand, int* __restrict__ count) {
  double px[kEcDirs], py[kEcDirs];
#pragma unroll
  for (int d = 0; d < kEcDirs; ++d) {
    px[d] = ext[2 * d];
    py[d] = ext[2 * d + 1];
  }
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = v[i].x, y = v[i].y;
    bool outside = false;
#pragma unroll
    for (int d = 0; d < kEcDirs; ++d) {
      const int e = (d + 1) % kEcDirs;
      const double ex = px[e] - px[d], ey = py[e] - py[d];
      if (ex == 0.0 && ey == 0.0) continue;
      if (ex * (y - py[d]) - ey * (x - px[d]) < 0.0) outside = true;
    }
    if (outside) {
      const int k = atomicAdd(count, 1);
      cand[k] = v[i];
    }
  }
}

// partial[block] = max |v - c|^2
__global__ void __launch_bounds__(256)
ec_radius_kernel(const float2* __restrict__ v, long long n, double cx, double cy,
                 double* __restrict__ partial) {
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double dx = (double)v[i].x - cx, dy = (double)v[i].y - cy;
    m = fmax(m, dx * dx + dy * dy);
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = m;
  block_sync();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    block_sync();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

struct Pt { double x, y; };
struct Circ { Pt c; double r2; };

bool in_circ(const Circ& c, const Pt& p) {
  const double dx = p.x - c.c.x, dy = p.y - c.c.y;
  return dx * dx + dy * dy <= c.r2 * (1.0 + 1e-12) + 1e-300;
}
Circ circ2(const Pt& a, const Pt& b) {
  const Pt c{(a.x + b.x) / 2, (a.y + b.y) / 2};
  const double dx = a.x - c.x, dy = a.y - c.y;
  return Circ{c, dx * dx + dy * dy};
}
Circ circ3(const Pt& a, const Pt& b, const Pt& c) {
  const double d = 2 * (a.x * (b.y - c.y) + b.x * (c.y - a.y) + c.x * (a.y - b.y));
  if (std::fabs(d) < 1e-300) {  // collinear: the widest pair
    Circ best = circ2(a, b), t = circ2(a, c);
    if (t.r2 > best.r2) best = t;
    t = circ2(b, c);
    if (t.r2 > best.r2) best = t;
    return best;
  }
  const double a2 = a.x * a.x + a.y * a.y, b2 = b.x * b.x + b.y * b.y, c2_ = c.x * c.x + c.y * c.y;
  const Pt o{(a2 * (b.y - c.y) + b2 * (c.y - a.y) + c2_ * (a.y - b.y)) / d,
             (a2 * (c.x - b.x) + b2 * (a.x - c.x) + c2_ * (b.x - a.x)) / d};
  const double dx = a.x - o.x, dy = a.y - o.y;
  return Circ{o, dx * dx + dy * dy};
}

// Welzl's algorithm, iterative form, on points in a fixed shuffled order
Circ welzl(std::vector<Pt>& p) {
  uint64_t s = 0x9e3779b97f4a7c15ull;
  for (size_t i = p.size(); i > 1; --i) {  // deterministic shuffle
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    std::swap(p[i - 1], p[s % i]);
  }
  Circ c{p[0], 0.0};
  for (size_t i = 1; i < p.size(); ++i) {
    if (in_circ(c, p[i])) continue;
    c = Circ{p[i], 0.0};
    for (size_t j = 0; j < i; ++j) {
      if (in_circ(c, p[j])) continue;
      c = circ2(p[i], p[j]);
      for (size_t k = 0; k < j; ++k)
        if (!in_circ(c, p[k])) c = circ3(p[i], p[j], p[k]);
    }
  }
  return c;
}
}  // namespace
}  // namespace cxm

extern "C" size_t cxm_enclosing_circle_workspace_size(int64_t n) {
  if (n <= 0) return 0;
  return (size_t)n * sizeof(float2) + (size_t)kEcBlocks * kEcDirs * 2 * sizeof(double) +
         (size_t)kEcDirs * 2 * sizeof(double) + 256;
}

extern "C" int cxm_enclosing_circle_c64(void* stream, const void* values, int64_t n, double* out,
                                        void* workspace, size_t workspace_bytes) {
  if (n <= 0) return fail(CXM_ERR_INVALID, "no values");
  if (n > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "more than 2^31 - 1 values");
  if (!values || !out) return fail(CXM_ERR_INVALID, "null buffer");
  if (!workspace || workspace_bytes < cxm_enclosing_circle_workspace_size(n))
    return fail(CXM_ERR_WORKSPACE, "workspace too small");
  int rc;
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  const float2* v = (const float2*)values;
  float2* cand = (float2*)workspace;
  double* partial = (double*)(cand + n);
  double* ext = partial + (size_t)kEcBlocks * kEcDirs * 2;
  int* count = (int*)(ext + kEcDirs * 2);
  hipLaunchKernelGGL(ec_extreme_kernel, dim3(kEcBlocks), dim3(256), 0, s, v, (long long)n, partial);
  CXM_LAUNCH_CHECK();
  hipLaunchKernelGGL(ec_extreme_final_kernel, dim3(1), dim3(256), 0, s, v, (const double*)partial, ext,
                     count);
  CXM_LAUNCH_CHECK();
  hipLaunchKernelGGL(ec_filter_kernel, dim3(kEcBlocks), dim3(256), 0, s, v, (long long)n,
                     (const double*)ext, cand, count);
  CXM_LAUNCH_CHECK();
  int nc = 0;
  double hext[kEcDirs * 2];
  CXM_HIP_CHECK(hipMemcpyAsync(&nc, count, sizeof(int), hipMemcpyDeviceToHost, s));
  CXM_HIP_CHECK(hipMemcpyAsync(hext, ext, sizeof(hext), hipMemcpyDeviceToHost, s));
  CXM_HIP_CHECK(hipStreamSynchronize(s));
  std::vector<float2> hc((size_t)nc);
  if (nc > 0) {
    CXM_HIP_CHECK(hipMemcpyAsync(hc.data(), cand, (size_t)nc * sizeof(float2), hipMemcpyDeviceToHost, s));
    CXM_HIP_CHECK(hipStreamSynchronize(s));
  }
  std::vector<Pt> pts;
  pts.reserve((size_t)nc + kEcDirs);
  for (int d = 0; d < kEcDirs; ++d) pts.push_back(Pt{hext[2 * d], hext[2 * d + 1]});
  for (const float2& q : hc) pts.push_back(Pt{(double)q.x, (double)q.y});
  const Circ c = welzl(pts);
  hipLaunchKernelGGL(ec_radius_kernel, dim3(kEcBlocks), dim3(256), 0, s, v, (long long)n, c.c.x, c.c.y,
                     partial);
  CXM_LAUNCH_CHECK();
  std::vector<double> hp(kEcBlocks);
  CXM_HIP_CHECK(hipMemcpyAsync(hp.data(), partial, kEcBlocks * sizeof(double), hipMemcpyDeviceToHost, s));
  CXM_HIP_CHECK(hipStreamSynchronize(s));
  double r2 = 0.0;
  for (double x : hp) r2 = std::max(r2, x);
  out[0] = c.c.x;
  out[1] = c.c.y;
  out[2] = std::sqrt(std::max(r2, c.r2));
  return CXM_OK;
}

// ---------------------------------------------------------------------------
// Reverse-mode rule of the solve (the implicit differentiation of
// solver.py:136-142, jaxopt FixedPointIteration with implicit_diff=True).
// With A = prec_forward = -V (G V + I) and b = -V G(s S), E = A^-1 b (the
// num_scale normalisation cancels).  Given lam solving A^H lam = Ebar (the
// same iteration on the conjugated problem: A^H = -V^H (G^H V^H + I), V^H
// and G^H the operators of conj(eps), conj(s), conj(kb)):
//   mu = G^H(V^H lam),  w = G(V E) + E - G(s S),
//   eps_bar = conj(s sum_c (conj(lam_c) w_c + conj(mu_c) E_c)),
//   S_bar = -conj(s) mu.
// ---------------------------------------------------------------------------
namespace cxm {
namespace {
// out[c][i] = conj(s eps[i] - kb) lam[c][i] (planar)
__global__ void mbs_vconj_kernel(const float2* __restrict__ lam, const float2* __restrict__ eps,
                                 float2* __restrict__ out, long long vox, C2 s, C2 kb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vox) return;
  const float2 e = eps[i];
  const C2 f{s.re * e.x - s.im * e.y - kb.re, -(s.re * e.y + s.im * e.x - kb.im)};
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c * vox + i] = cmulf(lam[c * vox + i], f);
}

// grad_eps[i] = conj(s sum_c (conj(lam_c) (gve_c + E_c - gss_c) + conj(mu_c) E_c)),
// grad_src[i][c] = -conj(s) mu_c; E and grad_src interleaved, the rest planar
__global__ void mbs_vjp_kernel(const float2* __restrict__ gve, const float2* __restrict__ gss,
                               const float2* __restrict__ E, const float2* __restrict__ lam,
                               const float2* __restrict__ mu, float2* __restrict__ grad_eps,
                               float2* __restrict__ grad_src, long long vox, C2 s) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vox) return;
  float ar = 0.f, ai = 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float2 e = E[i * 3 + c], l = lam[c * vox + i], m = mu[c * vox + i];
    const float2 a = gve[c * vox + i], b = gss[c * vox + i];
    const float wr = a.x + e.x - b.x, wi = a.y + e.y - b.y;
    // conj(l) w + conj(m) e
    ar += l.x * wr + l.y * wi + m.x * e.x + m.y * e.y;
    ai += l.x * wi - l.y * wr + m.x * e.y - m.y * e.x;
    if (grad_src) {
      // -conj(s) m
      grad_src[i * 3 + c] = make_float2(-(s.re * m.x + s.im * m.y), -(s.re * m.y - s.im * m.x));
    }
  }
  if (grad_eps) {
    // conj(s acc)
    grad_eps[i] = make_float2(s.re * ar - s.im * ai, -(s.re * ai + s.im * ar));
  }
}
}  // namespace
}  // namespace cxm

extern "C" size_t cxm_mbs_vjp_workspace_size(const cxm_mbs_desc* d) {
  MbsGeom g;
  if (mbs_geom(d, g)) return 0;
  return (size_t)9 * g.vox * sizeof(float2);
}

extern "C" int cxm_mbs_vjp(void* stream, const void* eps, const void* source, const void* E,
                           const void* lam, void* grad_eps, void* grad_source, void* workspace,
                           size_t workspace_bytes, const cxm_mbs_desc* d) {
  MbsGeom g;
  int rc = mbs_geom(d, g);
  if (rc) return rc;
  if (d->eps_tensor) return fail(CXM_ERR_UNSUPPORTED, "the solve's gradient: scalar permittivity only");
  if (!eps || !source || !E || !lam) return fail(CXM_ERR_INVALID, "null buffer");
  if (!workspace || workspace_bytes < cxm_mbs_vjp_workspace_size(d))
    return fail(CXM_ERR_WORKSPACE, "workspace too small");
  if (!grad_eps && !grad_source) return CXM_OK;
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  const C2 sc = c2(d->scale_re, d->scale_im);
  const C2 kb = c2(d->bias_re * d->scale_re - d->bias_im * d->scale_im + 1.0,
                   d->bias_re * d->scale_im + d->bias_im * d->scale_re);
  const C2 scc{sc.re, -sc.im}, kbc{kb.re, -kb.im};
  float2* gve = (float2*)workspace;
  float2* gss = gve + 3 * g.vox;
  float2* mu = gss + 3 * g.vox;
  // G(V E)
  hipLaunchKernelGGL(mbs_v_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, (const float2*)E,
                     (const float2*)eps, gve, g.vox, 1, 0, sc, kb, C2{1.0f, 0.0f}, nullptr);
  CXM_LAUNCH_CHECK();
  if ((rc = green_apply(g, d, gve, sc, kb, s))) return rc;
  // G(s S)
  hipLaunchKernelGGL(mbs_v_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, (const float2*)source,
                     nullptr, gss, g.vox, 1, 0, sc, kb, sc, nullptr);
  CXM_LAUNCH_CHECK();
  if ((rc = green_apply(g, d, gss, sc, kb, s))) return rc;
  // mu = G^H(V^H lam)
  hipLaunchKernelGGL(mbs_vconj_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, (const float2*)lam,
                     (const float2*)eps, mu, g.vox, sc, kb);
  CXM_LAUNCH_CHECK();
  if ((rc = green_apply(g, d, mu, scc, kbc, s))) return rc;
  hipLaunchKernelGGL(mbs_vjp_kernel, dim3(blocks_for(g.vox)), dim3(256), 0, s, (const float2*)gve,
                     (const float2*)gss, (const float2*)E, (const float2*)lam, (const float2*)mu,
                     (float2*)grad_eps, (float2*)grad_source, g.vox, sc);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}
