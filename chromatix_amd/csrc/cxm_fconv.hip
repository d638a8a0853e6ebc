// Fourier convolution over the last two or three axes (fourier_convolution,
// src/chromatix/ops/ops.py:9-86) on the propagation passes.  Lx, Ly, Ld are
// the implemented FFT lengths >= n + k - 1 per axis (the linear
// convolution, which the reference's zero-padded transforms also compute).
//
// Two axes (depth == k_depth == 1), per chunk of image planes:
//   row FFT (zero-padded) -> column FFT * K * column IFFT (the propagation's
//   column pass with a per-plane spectrum table) -> row IFFT, cropped.
// Three axes, per chunk of image volumes:
//   row FFT -> column FFT -> depth FFT * K * depth IFFT (the same column pass
//   over the volume seen as Ld rows of Ly * Lx columns) -> column IFFT ->
//   row IFFT, cropped.
// The kernel spectra K = fftn(kernel) on the (Ld x) Ly x Lx grid are
// computed once per call by the same row / column / depth passes.
//
// Real input over two axes with an even row length (the default) takes the
// paired path instead: image and kernel planes as one complex transform on
// pair-tiled planes (see the paired passes below), optionally summed over
// the batch before the inverse row transform (sum_batch).
#include <algorithm>
#include <cstring>
#include <string>

#include "cxm_pipe_host.h"
#include "cxm_prop2d_kernels.h"

namespace cxm {
namespace {

int next_len(long long need) {
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

struct Axis {
  int L;      // FFT length (1: no transform along this axis)
  int keep;   // convolution indices written
  int off;    // first convolution index written
};

struct FcGeom {
  bool vol;                // three-axis convolution
  Axis d, y, x;
  int out_d, out_h, out_w; // output volume
  int Rws;                 // rows per workspace plane
  long long vol_ws;        // workspace elements per image volume
  long long kvol;          // elements per kernel spectrum
  int chunk;               // image volumes per chunk
};

int axis_geom(int n, int k, int mode, int out_n, bool transform, Axis& a) {
  if (!transform) {
    a.L = 1;
    a.keep = 1;
    a.off = 0;
    return CXM_OK;
  }
  a.L = next_len((long long)n + k - 1);
  if (a.L < 0) return fail(CXM_ERR_UNSUPPORTED, "convolution needs FFT lengths beyond the implemented sizes");
  if (mode == 0) {
    a.keep = n;
    a.off = (k - 1) / 2;
  } else {
    a.keep = std::min(out_n, a.L);
    a.off = 0;
  }
  return CXM_OK;
}

int fc_geom(const cxm_fconv_desc* d, FcGeom& g) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->batch < 0 || d->depth <= 0 || d->height <= 0 || d->width <= 0 || d->k_planes < 1 ||
      d->k_depth <= 0 || d->k_height <= 0 || d->k_width <= 0)
    return fail(CXM_ERR_INVALID, "invalid shape parameters");
  if (d->mode != 0 && d->mode != 1) return fail(CXM_ERR_INVALID, "mode must be 0 (same) or 1 (full)");
  if (d->mode == 1 && (d->out_depth <= 0 || d->out_height <= 0 || d->out_width <= 0))
    return fail(CXM_ERR_INVALID, "invalid full-mode output shape");
  g.vol = !(d->depth == 1 && d->k_depth == 1);
  if (d->sum_batch != 0 && d->sum_batch != 1) return fail(CXM_ERR_INVALID, "sum_batch must be 0 or 1");
  int rc;
  if ((rc = axis_geom(d->depth, d->k_depth, d->mode, d->out_depth, g.vol, g.d))) return rc;
  if ((rc = axis_geom(d->height, d->k_height, d->mode, d->out_height, true, g.y))) return rc;
  if ((rc = axis_geom(d->width, d->k_width, d->mode, d->out_width, true, g.x))) return rc;
  g.out_d = d->mode == 0 ? d->depth : (g.vol ? d->out_depth : 1);
  g.out_h = d->mode == 0 ? d->height : d->out_height;
  g.out_w = d->mode == 0 ? d->width : d->out_width;
  // two axes: the column pass runs in place over `height` input rows and
  // leaves `keep` rows; three axes: whole Ly-row planes
  g.Rws = g.vol ? g.y.L : std::max(d->height, g.y.keep);
  g.vol_ws = (long long)g.d.L * g.Rws * g.x.L;
  g.kvol = (long long)g.d.L * g.y.L * g.x.L;
  // 32-bit element offsets inside one volume (the column pass addresses a
  // strip with unsigned byte offsets)
  if (g.kvol * (long long)sizeof(float2) >= (1LL << 32))
    return fail(CXM_ERR_UNSUPPORTED, "convolution volume exceeds 4 GiB");
  g.chunk = (int)std::max(1LL, std::min<long long>(std::max(1, d->batch), ((128LL << 20) / 8) / g.vol_ws));
  return CXM_OK;
}

size_t fc_ws(const cxm_fconv_desc* d, const FcGeom& g) {
  return ((size_t)d->k_planes * g.kvol + (size_t)g.chunk * g.vol_ws) * sizeof(float2);
}

// ---- real image and kernel in one complex plane ------------------------------
// For real image a and real kernel k of a plane, z = a + i s k (s a power of
// two that brings s k to the scale of a, so that neither spectrum loses
// relative precision to the other) has Z = A + i s K with A, K Hermitian, so
// at every frequency f, with Z2 = conj(Z(-f)):
//   A = (Z + Z2) / 2,  s K = (Z - Z2) / (2 i),  A K = (Z^2 - Z2^2) / (4 i s).
// One forward transform of z replaces those of a and k (and the kernel
// spectra table); the product is Hermitian, so the real part of its inverse
// is the convolution.  The column pass pairs columns kx and -kx in one
// workgroup to reach Z(-f).

constexpr int kSsBlocks = 64;  // partial sums per plane

// partial[(plane * kSsBlocks + b)] = sum of squares over a slice of the plane
__global__ void __launch_bounds__(256)
plane_ss_kernel(const float* __restrict__ x, long long plane_elems, double* __restrict__ partial) {
  const int plane = blockIdx.y;
  const float* px = x + (long long)plane * plane_elems;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < plane_elems; i += (long long)kSsBlocks * 256)
    acc += (double)px[i] * px[i];
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  block_sync();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    block_sync();
  }
  if (threadIdx.x == 0) partial[(long long)plane * kSsBlocks + blockIdx.x] = sh[0];
}

// scale[b] = 2^round(log2(|a_b| / |k_{b % k_planes}|)) and, after the batch's
// scales, scale[batch + b] = 1 / (4 scale[b]) (exact: a power of two).  Both
// are 0 when either plane is all zeros: the product A K of that pair is
// exactly 0, and the column pass then stores exact zeros (the shared
// transform would leave a rounding residual of the nonzero plane there).
__global__ void pair_scale_kernel(const double* __restrict__ pa, const double* __restrict__ pk,
                                  int batch, int k_planes, float* __restrict__ scale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  double sa = 0.0, sk = 0.0;
  for (int j = 0; j < kSsBlocks; ++j) sa += pa[(long long)b * kSsBlocks + j];
  const int kb = b % k_planes;
  for (int j = 0; j < kSsBlocks; ++j) sk += pk[(long long)kb * kSsBlocks + j];
  float s = 0.0f, q = 0.0f;
  if (sa > 0.0 && sk > 0.0) {
    const int e = max(-100, min(100, (int)lrint(0.25 * log2(sa / sk) * 2.0)));  // round(log2(sqrt(sa / sk)))
    s = ldexpf(1.0f, e);
    q = ldexpf(1.0f, -e - 2);
  }
  scale[b] = s;
  scale[batch + b] = q;
}

// Pair-tiled planes.  Slot 2 j + side of a row holds column j (side 0) and
// column L - j (side 1) for pair j > 0; pair 0 is columns 0 and L / 2.  The
// slots go in strips of kTileW: element (y, slot) at
// (slot / 4) * strip_stride + y * 4 + slot % 4, so a column pass over the
// slots of a strip reads and writes one contiguous block, and the row
// passes move 4 rows per workgroup (whole 128-byte lines of 4 rows x 4
// slots) through LDS.
CXM_DEV int slot_col(int slot, int L) {
  const int j = slot >> 1;
  if (j == 0) return (slot & 1) ? L / 2 : 0;
  return (slot & 1) ? L - j : j;
}

// 4 rows of N-point row transforms per workgroup fit the LDS
template <int N, int E>
constexpr bool pair_rows_fit() {
  return 4 * (N / E) <= 1024 && 4 * RowCfg<N, E>::PITCH * (int)sizeof(float2) <= 160 * 1024;
}

// LDS rows (natural column order) <-> the pair tiles of 4 rows [r0, r0+4):
// piece c = 16 bytes = row (c % 8) / 2, slots 4 (c / 8) + 2 (c % 2) + {0, 1}
// (columns j and its mirror of pair j = 2 (c / 8) + c % 2).
template <int N, int PITCH, bool TO_TILES>
CXM_DEV void pair_tiles_exchange(float2* lds, float2* base, long long strip_stride, int nthreads) {
  constexpr int PIECES = (N + 3) / 4 * 8;
  for (int c = threadIdx.x; c < PIECES; c += nthreads) {
    const int q = c & 7, j = 2 * (c >> 3) + (q & 1);
    if (j >= N / 2) continue;  // the second half of the last strip when N % 4 == 2
    const int x0 = j == 0 ? 0 : j, x1 = j == 0 ? N / 2 : N - j;
    float2* d = lds + (q >> 1) * PITCH;
    float4* g = (float4*)(base + (c >> 3) * strip_stride + 2 * q);
    if constexpr (TO_TILES) {
      const float2 a = d[lds_at<N, 1>(x0)], b = d[lds_at<N, 1>(x1)];
      *g = make_float4(a.x, a.y, b.x, b.y);
    } else {
      const float4 w = *g;
      d[lds_at<N, 1>(x0)] = make_float2(w.x, w.y);
      d[lds_at<N, 1>(x1)] = make_float2(w.z, w.w);
    }
  }
}

struct PairRowArgs {
  const float* img;          // batch x height x width
  const float* ker;          // k_planes x k_height x k_width
  const float* scale;        // per image plane
  float2* out;               // pair-tiled planes of `rows` rows
  const float2* tw;
  long long img_plane, ker_plane, out_plane, strip_stride;
  int height, width, k_height, k_width, k_planes;
  int rows;                  // rows per plane (a multiple of 4, >= height, k_height)
  int plane0, n_rows;        // first image plane, rows in this launch (a multiple of 4)
};

// Forward row transforms of z = a + i s k, 4 rows per workgroup, stored
// pair-tiled.
template <int N, int E>
__global__ void __launch_bounds__(4 * (N / E))
pair_row_kernel(PairRowArgs a) {
  constexpr int T = N / E;
  constexpr int PITCH = RowCfg<N, E>::PITCH;
  constexpr bool SYNC = T > 64;
  __shared__ float2 lds[4 * PITCH];
  const int g = threadIdx.x / T;
  const int t = threadIdx.x - g * T;
  const int row0 = blockIdx.x * 4;
  const int p = row0 / a.rows;
  const int r = row0 - p * a.rows + g;
  const int b = a.plane0 + p;
  const float s = a.scale[b];
  const float* ir = a.img + (long long)b * a.img_plane + (long long)r * a.width;
  const float* kr = a.ker + (long long)(b % a.k_planes) * a.ker_plane + (long long)r * a.k_width;
  const bool iok = r < a.height, kok = r < a.k_height;
  float2* st = lds + g * PITCH;
  float2 v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int x = t + k * T;
    const float re = (iok && x < a.width) ? __builtin_nontemporal_load(ir + x) : 0.f;
    const float im = (kok && x < a.k_width) ? kr[x] * s : 0.f;
    v[k] = make_float2(re, im);
  }
  fft<N, E, -1, 1, SYNC>(v, st, t, a.tw);
  block_sync();  // every wave of the row has read the transform's last exchange
#pragma unroll
  for (int k = 0; k < E; ++k) st[lds_at_split<N, 1>(t, k * T)] = v[k];
  block_sync();
  pair_tiles_exchange<N, PITCH, true>(lds, a.out + p * a.out_plane + (long long)(row0 - p * a.rows) * kTileW,
                                      a.strip_stride, 4 * T);
}

struct PairColArgs {
  float2* buf;               // pair-tiled planes, in place
  const float2* tw;
  const float* inv4s;        // per image plane: 1 / (4 scale), 0 for a zero plane
  long long plane_stride, strip_stride;
  int lx;                    // columns (even) = slots
  int in_len;                // rows holding data
  int out_off, out_len;      // rows kept
  int n_planes, plane0;
  float norm;                // 1 / (L_y L_x)
  float2* acc;               // sum mode: pair-tiled plane of the kept rows
  int first;                 // sum mode: acc is written, not added to
};

// Column pass of one plane p up to the spectrum: forward transform of the
// slot's column and the product A K of the pair, in u (scaled by the
// plane's 1 / (4 s) and the transform norm); ends with a barrier, so the
// next transform may write the LDS.
template <int N, int E, int C>
CXM_DEV void pair_col_product(const PairColArgs& a, int p, const float2* colp, bool valid, int c,
                              int ty, int partner, float2* lds, float2 (&u)[E]) {
  constexpr int T = N / E;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = ty + k * T;
    u[k] = (valid && i < a.in_len) ? colp[(long long)i * kTileW] : make_float2(0.f, 0.f);
  }
  fft<N, E, -1, C, true>(u, lds + c, ty, a.tw);
  block_sync();  // the transform's last exchange has been read by every thread
  stage_out<N, E, C>(u, lds + c, ty);
  block_sync();
  const float f = a.norm * a.inv4s[a.plane0 + p];  // 0 for a zero plane: exact zeros
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = ty + k * T;
    const int mi = i == 0 ? 0 : N - i;
    const float2 z = u[k];
    const float2 w = lds[partner + lds_at<N, C>(mi)];
    const float2 z2 = make_float2(w.x, -w.y);  // conj(Z(-f))
    // D = z^2 - z2^2;  A K = D / (4 i s) = (D.y, -D.x) / (4 s)
    const float dx = (z.x * z.x - z.y * z.y) - (z2.x * z2.x - z2.y * z2.y);
    const float dy = 2.0f * (z.x * z.y - z2.x * z2.y);
    u[k] = make_float2(dy * f, -dx * f);
  }
  block_sync();  // the partner values are read before the next transform writes LDS
}

// The whole column pass of one plane p: the product, inverse transform, in u.
template <int N, int E, int C>
CXM_DEV void pair_col_plane(const PairColArgs& a, int p, const float2* colp, bool valid, int c,
                            int ty, int partner, float2* lds, float2 (&u)[E]) {
  pair_col_product<N, E, C>(a, p, colp, valid, c, ty, partner, lds, u);
  fft<N, E, 1, C, true>(u, lds + c, ty, a.tw);
}

// C consecutive slots per workgroup (C even: both columns of each pair);
// slot 2 j + side pairs with its neighbour, pair 0's columns with themselves.
// One workgroup per (plane, C slots), the kept rows stored in place; SUM:
// one workgroup per C slots of every plane of the launch: the planes'
// spectral products are added in registers in plane order (the inverse
// transform is linear), one inverse transform, and the kept rows written
// to acc (first) or added to it, so that launches in chunk order sum the
// planes in order.
template <int N, int E, int C, bool SUM>
__global__ void __launch_bounds__((N / E) * C)
pair_col_kernel(PairColArgs a) {
  static_assert(C % 2 == 0, "column pairs");
  constexpr int T = N / E;
  __shared__ float2 lds[lds_span<N, C>()];
  const int c = threadIdx.x % C;
  const int ty = threadIdx.x / C;
  const int nblk = (a.lx + C - 1) / C;
  const int slot = (SUM ? blockIdx.x : blockIdx.x % nblk) * C + c;
  const bool valid = slot < a.lx;
  const int partner = (slot >> 1) == 0 ? c : (c ^ 1);
  const long long soff = (long long)(slot >> 2) * a.strip_stride + (slot & 3);
  float2 u[E];
  if constexpr (SUM) {
    float2 s[E];
#pragma unroll
    for (int k = 0; k < E; ++k) s[k] = make_float2(0.f, 0.f);
    for (int p = 0; p < a.n_planes; ++p) {
      // the thread index laundered per plane: the transforms' twiddle loads
      // are then not hoisted out of the plane loop
      int tl = ty;
      asm volatile("" : "+v"(tl));
      pair_col_product<N, E, C>(a, p, a.buf + (long long)p * a.plane_stride + soff, valid, c, tl, partner,
                                lds, u);
#pragma unroll
      for (int k = 0; k < E; ++k) s[k] = make_float2(s[k].x + u[k].x, s[k].y + u[k].y);
    }
    fft<N, E, 1, C, true>(s, lds + c, ty, a.tw);
#pragma unroll
    for (int k = 0; k < E; ++k) u[k] = s[k];
  } else {
    const int p = blockIdx.x / nblk;
    pair_col_plane<N, E, C>(a, p, a.buf + (long long)p * a.plane_stride + soff, valid, c, ty, partner, lds,
                            u);
  }
  if (!valid) return;
  if constexpr (SUM) {
    float2* accp = a.acc + soff;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int yl = ty + k * T - a.out_off;
      if (yl >= 0 && yl < a.out_len) {
        float2* q = accp + (long long)yl * kTileW;
        *q = a.first ? u[k] : make_float2(q->x + u[k].x, q->y + u[k].y);
      }
    }
  } else {
    float2* colp = a.buf + (long long)(blockIdx.x / nblk) * a.plane_stride + soff;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int yl = ty + k * T - a.out_off;
      if (yl >= 0 && yl < a.out_len) colp[(long long)yl * kTileW] = u[k];
    }
  }
}

struct PairInvArgs {
  const float2* in;          // pair-tiled planes of `rows` rows
  float* out;                // out_len-wide rows, `keep` rows per plane
  const float2* tw;
  long long in_plane, strip_stride, out_plane;
  int rows;                  // rows per workspace plane group (a multiple of 4)
  int keep;                  // rows stored per plane
  int out_w, out_off, out_len;
  int n_rows;                // planes in this launch x rows (a multiple of 4)
};

// Inverse row transforms of pair-tiled rows, 4 rows per workgroup: the real
// part of columns [out_off, out_off + out_len), streaming stores.
template <int N, int E>
__global__ void __launch_bounds__(4 * (N / E))
pair_inv_kernel(PairInvArgs a) {
  constexpr int T = N / E;
  constexpr int PITCH = RowCfg<N, E>::PITCH;
  constexpr bool SYNC = T > 64;
  __shared__ float2 lds[4 * PITCH];
  const int g = threadIdx.x / T;
  const int t = threadIdx.x - g * T;
  const int row0 = blockIdx.x * 4;
  const int p = row0 / a.rows;
  const int r = row0 - p * a.rows + g;
  pair_tiles_exchange<N, PITCH, false>(lds, (float2*)a.in + p * a.in_plane + (long long)(row0 - p * a.rows) * kTileW,
                                       a.strip_stride, 4 * T);
  block_sync();
  float2* st = lds + g * PITCH;
  float2 v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = st[lds_at_split<N, 1>(t, k * T)];
  fft<N, E, 1, 1, SYNC>(v, st, t, a.tw);
  if (r >= a.keep) return;
  float* o = a.out + p * a.out_plane + (long long)r * a.out_w - a.out_off;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int x = t + k * T;
    if (x >= a.out_off && x < a.out_off + a.out_len) __builtin_nontemporal_store(v[k].x, o + x);
  }
}

enum PairPass { PAIR_ROW, PAIR_COL, PAIR_INV };

// Values per thread of the paired column pass: half the propagation
// column pass's for lengths 3 * 2^a >= 512 (twice the threads at half the
// registers each: 13.7 against 14.7 ms for 128 planes of 3072-point
// transforms), where that stays within 1024 threads; else 0 (the table's).
template <int N, int EC, int C>
constexpr int pair_col_half_e() {
  if constexpr (EC % 2 != 0 || EC / 2 < 2 || N % (EC / 2) != 0 || (N / (EC / 2)) * C > 1024 || N < 512 ||
                N % 3 != 0)
    return 0;
  else
    return schedule_ok(N, EC / 2) ? EC / 2 : 0;
}

// Values per thread of the paired row passes: half the propagation row
// pass's for lengths 3 * 2^a >= 512 (13.5 against 13.75 ms for 128 planes
// of 3072-point transforms), where 4 rows still fit; else 0 (the table's).
template <int N, int ER>
constexpr int pair_row_half_e() {
  if constexpr (ER % 2 != 0 || ER / 2 < 2 || N % (ER / 2) != 0 || N < 512 || N % 3 != 0)
    return 0;
  else if constexpr (!pair_rows_fit<N, ER / 2>())
    return 0;
  else
    return schedule_ok(N, ER / 2) ? ER / 2 : 0;
}

template <int N, int E>
int launch_pair_rows(PairPass pass, const PairRowArgs* ra, const PairInvArgs* ia, hipStream_t s) {
  if (pass == PAIR_ROW)
    hipLaunchKernelGGL((pair_row_kernel<N, E>), dim3((unsigned)(ra->n_rows / 4)), dim3(4 * (N / E)), 0, s, *ra);
  else
    hipLaunchKernelGGL((pair_inv_kernel<N, E>), dim3((unsigned)(ia->n_rows / 4)), dim3(4 * (N / E)), 0, s, *ia);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

template <int N, int E, int C>
int launch_pair_col(const PairColArgs* ca, hipStream_t s) {
  const long long nblk = (ca->lx + C - 1) / C;
  if (ca->acc)
    hipLaunchKernelGGL((pair_col_kernel<N, E, C, true>), dim3((unsigned)nblk), dim3((N / E) * C), 0, s, *ca);
  else
    hipLaunchKernelGGL((pair_col_kernel<N, E, C, false>), dim3((unsigned)(ca->n_planes * nblk)),
                       dim3((N / E) * C), 0, s, *ca);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

#define CXM_PAIR_LAUNCH(N, ER, EC, C)                                                           \
  case N: {                                                                                     \
    if (pass == PAIR_COL) {                                                                     \
      if constexpr ((C) % 2 == 0) {                                                             \
        constexpr int EH = pair_col_half_e<N, EC, C>();                                         \
        return launch_pair_col<N, (EH > 0 ? EH : EC), C>(ca, s);                                \
      }                                                                                         \
    } else if constexpr (pair_rows_fit<N, ER>() && N % 2 == 0) {                                \
      constexpr int RH = pair_row_half_e<N, ER>();                                              \
      return launch_pair_rows<N, (RH > 0 ? RH : ER)>(pass, ra, ia, s);                          \
    }                                                                                           \
    return fail(CXM_ERR_UNSUPPORTED, "no paired pass for length " + std::to_string(N));         \
  }

int launch_pair(int n, PairPass pass, const PairRowArgs* ra, const PairColArgs* ca,
                const PairInvArgs* ia, hipStream_t s) {
  switch (n) {
    CXM_FOR_EACH_SIZE(CXM_PAIR_LAUNCH)
    default:
      return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(n));
  }
}
#undef CXM_PAIR_LAUNCH

// paired passes need an even row length whose 4-row transforms fit the LDS
// and an even number of columns per column-pass workgroup at the column length
bool pair_supported(int ly, int lx) {
  bool ok_y = false, ok_x = false;
#define CXM_PAIR_OK(N, ER, EC, C)                     \
  if (N == ly && (C) % 2 == 0) ok_y = true;           \
  if (N == lx && N % 2 == 0 && pair_rows_fit<N, ER>()) ok_x = true;
  CXM_FOR_EACH_SIZE(CXM_PAIR_OK)
#undef CXM_PAIR_OK
  return ok_y && ok_x;
}

// Row FFT of `planes` input planes of `rows` x `cols` (float32 or complex64,
// zero-padded to Lx) into rows of Lx elements, planes `out_plane` apart.
int row_forward(const void* in, int real, long long in_plane, int rows, int cols, float2* out,
                long long out_plane, int Lx, int planes, hipStream_t s, int nt = 0) {
  RowArgs r;
  memset(&r, 0, sizeof(r));
  r.nt_in = nt;
  r.in = (const float2*)in;
  r.in_real = real;
  r.out_c = out;
  r.tw = twiddles_for(Lx);
  r.in_plane_stride = in_plane;
  r.out_plane_stride = out_plane;
  r.in_row_stride = cols;
  r.out_row_stride = Lx;
  r.n_rows = planes * rows;
  r.rows_per_plane = rows;
  r.in_len = cols;
  r.out_len = Lx;
  r.scale = 1.0f;
  return launch_row(Lx, -1, r, s);
}

// In-place column transform (SEQ_F or SEQ_I) of `planes` planes: `lines`
// columns of `line_stride` pitch, FFT length N over the first `in_len`
// rows; rows [out_off, out_off + out_len) stored at rows [0, out_len).
int col_pass(float2* buf, long long plane_stride, int lines, int line_stride, int N, int in_len,
             int out_off, int out_len, int planes, int seq, hipStream_t s) {
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  a.in = buf;
  a.out = buf;
  a.tw = twiddles_for(N);
  a.in_plane_stride = a.out_plane_stride = plane_stride;
  a.in_line_stride = a.out_line_stride = line_stride;
  a.n_lines = lines;
  a.n_planes = planes;
  a.grid_ny = N;
  a.grid_nx = lines;
  a.in_len = in_len;
  a.out_off = out_off;
  a.out_len = out_len;
  a.out_line_len = lines;
  a.seq = seq;
  a.scale = a.mid_scale = 1.0f;
  return launch_pipe_col(N, a, s);
}
}  // namespace
}  // namespace cxm

using namespace cxm;

namespace cxm {
namespace {
// The paired path: real two-axis convolution with an even row length.
bool use_pair(const cxm_fconv_desc* d, const FcGeom& g) {
  return d->real_input && !g.vol && option(CXM_OPT_FCONV_PAIR) != 1 && d->batch <= 65535 &&
         d->k_planes <= 65535 && pair_supported(g.y.L, g.x.L);
}
struct PairGeom {
  int rz;         // rows holding data: max(height, k_height)
  int rows_f;     // rows per plane of the forward row pass (rz rounded up to 4)
  int rows_i;     // rows per plane of the inverse row pass (kept rows rounded up to 4)
  int rws;        // rows per workspace plane
  long long strip, plane;  // elements per strip and per workspace plane
  int chunk;
  size_t ws_off;  // bytes before the planes: scales and partial sums
};
PairGeom pair_geom(const cxm_fconv_desc* d, const FcGeom& g) {
  PairGeom q;
  q.rz = std::max(d->height, d->k_height);
  q.rows_f = (q.rz + 3) / 4 * 4;
  q.rows_i = (g.y.keep + 3) / 4 * 4;
  q.rws = std::max(q.rows_f, q.rows_i);
  q.strip = (long long)q.rws * kTileW;
  q.plane = q.strip * ((g.x.L + kTileW - 1) / kTileW);
  // planes per chunk: at most 160 MiB of intermediates (they stay in the
  // 256 MiB Infinity Cache between the row and column passes), split evenly
  // (128 planes of 1536 x 3072 slots: chunks of 4, 13.0 ms, against 13.5 ms
  // with chunks of 3 under 128 MiB)
  const long long nb = std::max(1, d->batch);
  const long long cap = std::max(1LL, std::min<long long>(nb, ((160LL << 20) / 8) / q.plane));
  const long long nchunks = (nb + cap - 1) / cap;
  q.chunk = (int)((nb + nchunks - 1) / nchunks);
  const size_t scales = ((size_t)std::max(1, d->batch) * 2 * sizeof(float) + 255) / 256 * 256;
  const size_t partial = (size_t)(std::max(1, d->batch) + d->k_planes) * kSsBlocks * sizeof(double);
  q.ws_off = scales + (partial + 255) / 256 * 256;
  return q;
}
size_t pair_ws(const cxm_fconv_desc* d, const FcGeom& g) {
  const PairGeom q = pair_geom(d, g);
  return q.ws_off + (size_t)(q.chunk + (d->sum_batch ? 1 : 0)) * q.plane * sizeof(float2);
}
// the batch sum runs on the paired path only
int check_sum(const cxm_fconv_desc* d, const FcGeom& g) {
  if (d->sum_batch && !use_pair(d, g))
    return fail(CXM_ERR_UNSUPPORTED,
                "sum_batch needs real two-axis input with paired transform lengths (and the paired "
                "path enabled)");
  return CXM_OK;
}
}  // namespace
}  // namespace cxm

extern "C" size_t cxm_fconv_workspace_size(const cxm_fconv_desc* d) {
  FcGeom g;
  if (fc_geom(d, g) || check_sum(d, g)) return 0;
  return use_pair(d, g) ? pair_ws(d, g) : fc_ws(d, g);
}

extern "C" int cxm_fconv_c64(void* stream, const void* image, const void* kernel, void* out,
                             void* workspace, size_t workspace_bytes, const cxm_fconv_desc* d) {
  FcGeom g;
  int rc = fc_geom(d, g);
  if (rc || (rc = check_sum(d, g))) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!image || !kernel || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const bool pair = use_pair(d, g);
  const size_t need = pair ? pair_ws(d, g) : fc_ws(d, g);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (pair) {
    const PairGeom q = pair_geom(d, g);
    float* scale = (float*)workspace;
    double* pa = (double*)((char*)workspace +
                           ((size_t)d->batch * 2 * sizeof(float) + 255) / 256 * 256);
    double* pk = pa + (size_t)d->batch * kSsBlocks;
    float2* ws = (float2*)((char*)workspace + q.ws_off);
    float2* acc = ws + (size_t)q.chunk * q.plane;  // sum_batch: the summed kept rows
    const long long in_plane = (long long)d->height * d->width;
    const long long k_plane = (long long)d->k_height * d->k_width;
    hipLaunchKernelGGL(plane_ss_kernel, dim3(kSsBlocks, d->batch), dim3(256), 0, s, (const float*)image,
                       in_plane, pa);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(plane_ss_kernel, dim3(kSsBlocks, d->k_planes), dim3(256), 0, s,
                       (const float*)kernel, k_plane, pk);
    CXM_LAUNCH_CHECK();
    hipLaunchKernelGGL(pair_scale_kernel, dim3((d->batch + 255) / 256), dim3(256), 0, s,
                       (const double*)pa, (const double*)pk, d->batch, d->k_planes, scale);
    CXM_LAUNCH_CHECK();
    const int Ly = g.y.L, Lx = g.x.L;
    const long long out_plane = (long long)g.out_h * g.out_w;
    for (int p0 = 0; p0 < d->batch; p0 += q.chunk) {
      const int kc = std::min(q.chunk, d->batch - p0);
      PairRowArgs ra;
      memset(&ra, 0, sizeof(ra));
      ra.img = (const float*)image;
      ra.ker = (const float*)kernel;
      ra.scale = scale;
      ra.out = ws;
      ra.tw = twiddles_for(Lx);
      ra.img_plane = in_plane;
      ra.ker_plane = k_plane;
      ra.out_plane = q.plane;
      ra.strip_stride = q.strip;
      ra.height = d->height;
      ra.width = d->width;
      ra.k_height = d->k_height;
      ra.k_width = d->k_width;
      ra.k_planes = d->k_planes;
      ra.rows = q.rows_f;
      ra.plane0 = p0;
      ra.n_rows = kc * q.rows_f;
      if ((rc = launch_pair(Lx, PAIR_ROW, &ra, nullptr, nullptr, s))) return rc;
      PairColArgs ca;
      memset(&ca, 0, sizeof(ca));
      ca.buf = ws;
      ca.tw = twiddles_for(Ly);
      ca.inv4s = scale + d->batch;
      ca.plane_stride = q.plane;
      ca.strip_stride = q.strip;
      ca.lx = Lx;
      ca.in_len = q.rz;
      ca.out_off = g.y.off;
      ca.out_len = g.y.keep;
      ca.n_planes = kc;
      ca.plane0 = p0;
      ca.norm = (float)(1.0 / ((double)Ly * (double)Lx));
      if (d->sum_batch) {
        // one launch per chunk, the planes summed in plane order; chunks in
        // order (the sum's order)
        ca.acc = acc;
        ca.first = p0 == 0;
        if ((rc = launch_pair(Ly, PAIR_COL, nullptr, &ca, nullptr, s))) return rc;
        continue;
      }
      if ((rc = launch_pair(Ly, PAIR_COL, nullptr, &ca, nullptr, s))) return rc;
      PairInvArgs ia;
      memset(&ia, 0, sizeof(ia));
      ia.in = ws;
      ia.out = (float*)out + (size_t)p0 * out_plane;
      ia.tw = twiddles_for(Lx);
      ia.in_plane = q.plane;
      ia.strip_stride = q.strip;
      ia.out_plane = out_plane;
      ia.rows = q.rows_i;
      ia.keep = g.y.keep;
      ia.out_w = g.out_w;
      ia.out_off = g.x.off;
      ia.out_len = g.x.keep;
      ia.n_rows = kc * q.rows_i;
      if ((rc = launch_pair(Lx, PAIR_INV, nullptr, nullptr, &ia, s))) return rc;
    }
    if (d->sum_batch) {
      PairInvArgs ia;
      memset(&ia, 0, sizeof(ia));
      ia.in = acc;
      ia.out = (float*)out;
      ia.tw = twiddles_for(Lx);
      ia.in_plane = q.plane;
      ia.strip_stride = q.strip;
      ia.out_plane = out_plane;
      ia.rows = q.rows_i;
      ia.keep = g.y.keep;
      ia.out_w = g.out_w;
      ia.out_off = g.x.off;
      ia.out_len = g.x.keep;
      ia.n_rows = q.rows_i;
      if ((rc = launch_pair(Lx, PAIR_INV, nullptr, nullptr, &ia, s))) return rc;
    }
    return CXM_OK;
  }
  const int Ld = g.d.L, Ly = g.y.L, Lx = g.x.L;
  const long long lplane = (long long)Ly * Lx;
  const size_t esz = d->real_input ? sizeof(float) : sizeof(float2);
  float2* K = (float2*)workspace;
  float2* ws = K + (long long)d->k_planes * g.kvol;

  // Kernel spectra: volume k at K + k * kvol, plane z at + z * Ly * Lx.
  const long long k_in_plane = (long long)d->k_height * d->k_width;
  if (!g.vol) {
    if ((rc = row_forward(kernel, d->real_input, k_in_plane, d->k_height, d->k_width, K, lplane, Lx,
                          d->k_planes, s)))
      return rc;
    if ((rc = col_pass(K, lplane, Lx, Lx, Ly, d->k_height, 0, Ly, d->k_planes, SEQ_F, s))) return rc;
  } else {
    for (int k = 0; k < d->k_planes; ++k) {
      const char* kin = (const char*)kernel + (size_t)k * d->k_depth * k_in_plane * esz;
      if ((rc = row_forward(kin, d->real_input, k_in_plane, d->k_height, d->k_width, K + k * g.kvol,
                            lplane, Lx, d->k_depth, s)))
        return rc;
      if ((rc = col_pass(K + k * g.kvol, lplane, Lx, Lx, Ly, d->k_height, 0, Ly, d->k_depth, SEQ_F, s)))
        return rc;
    }
    // depth axis: the volume as Ld rows of Ly * Lx columns
    if ((rc = col_pass(K, g.kvol, (int)lplane, (int)lplane, Ld, d->k_depth, 0, Ld, d->k_planes, SEQ_F, s)))
      return rc;
  }

  const long long in_plane = (long long)d->height * d->width;
  const long long in_vol = in_plane * d->depth;
  const long long out_plane = (long long)g.out_h * g.out_w;
  const long long out_vol = out_plane * g.out_d;
  const float inv = (float)(1.0 / ((double)Ld * (double)Ly * (double)Lx));
  for (int v0 = 0; v0 < d->batch; v0 += g.chunk) {
    const int kc = std::min(g.chunk, d->batch - v0);
    const char* img = (const char*)image + (size_t)v0 * in_vol * esz;
    // forward transforms of the rows (and, for volumes, the columns)
    if (!g.vol) {
      // image planes are read once: streaming loads
      if ((rc = row_forward(img, d->real_input, in_plane, d->height, d->width, ws, g.vol_ws, Lx, kc, s, 1)))
        return rc;
    } else {
      for (int v = 0; v < kc; ++v) {
        float2* wv = ws + v * g.vol_ws;
        if ((rc = row_forward(img + (size_t)v * in_vol * esz, d->real_input, in_plane, d->height,
                              d->width, wv, lplane, Lx, d->depth, s)))
          return rc;
        if ((rc = col_pass(wv, lplane, Lx, Lx, Ly, d->height, 0, Ly, d->depth, SEQ_F, s))) return rc;
      }
    }
    // last axis forward * K * inverse: columns of the planes (two axes) or
    // the depth axis of the volumes (three axes)
    ColArgs ca;
    memset(&ca, 0, sizeof(ca));
    ca.in = ws;
    ca.out = ws;
    ca.H = K;
    ca.h_planes = d->k_planes;
    ca.h_plane_stride = g.kvol;
    ca.in_plane_stride = ca.out_plane_stride = g.vol_ws;
    ca.n_planes = kc;
    ca.plane0 = v0;
    ca.nz = 1;
    ca.nwv = 1;
    ca.kind = CXM_H_ARRAY;
    ca.scale = inv;
    int N;
    if (!g.vol) {
      N = Ly;
      ca.row_stride = Lx;
      ca.ncols = Lx;
      ca.in_len = d->height;
      ca.out_off = g.y.off;
      ca.out_len = g.y.keep;
    } else {
      N = Ld;
      ca.row_stride = (int)lplane;
      ca.ncols = (int)lplane;
      ca.h_row_stride = (int)lplane;
      ca.in_len = d->depth;
      ca.out_off = g.d.off;
      ca.out_len = g.d.keep;
    }
    ca.tw = twiddles_for(N);
    if ((rc = launch_col_prop(N, ca, s))) return rc;
    // inverse transforms of the columns (volumes) and rows, cropped
    RowArgs r;
    memset(&r, 0, sizeof(r));
    r.tw = twiddles_for(Lx);
    r.in_row_stride = Lx;
    r.out_row_stride = g.out_w;
    r.in_len = Lx;
    r.out_off = g.x.off;
    r.out_len = g.x.keep;
    r.scale = 1.0f;
    r.intensity = d->real_input ? 2 : 0;
    r.nt_out = 1;  // written once
    r.out_plane_stride = out_plane;
    r.rows_per_plane = g.y.keep;
    if (!g.vol) {
      r.in = ws;
      r.in_plane_stride = g.vol_ws;
      r.n_rows = kc * g.y.keep;
      r.out_c = (float2*)out + (size_t)v0 * out_plane;
      r.out_i = (float*)out + (size_t)v0 * out_plane;
      if ((rc = launch_row(Lx, 1, r, s))) return rc;
    } else {
      for (int v = 0; v < kc; ++v) {
        float2* wv = ws + v * g.vol_ws;
        if ((rc = col_pass(wv, lplane, Lx, Lx, Ly, Ly, g.y.off, g.y.keep, g.d.keep, SEQ_I, s))) return rc;
        r.in = wv;
        r.in_plane_stride = lplane;
        r.n_rows = g.d.keep * g.y.keep;
        r.out_c = (float2*)out + (size_t)(v0 + v) * out_vol;
        r.out_i = (float*)out + (size_t)(v0 + v) * out_vol;
        if ((rc = launch_row(Lx, 1, r, s))) return rc;
      }
    }
  }
  return CXM_OK;
}
