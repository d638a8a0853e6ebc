// Matrix-DFT zoom on MFMA: the chirp-z transform of every plane along both
// axes as two complex GEMMs,
//     Y = A_y . X . A_x^T,   A_d[k][j] = a_d^-j w_d^(j k)  (k < m_d, j < n_d),
// times the CZT branch's output modulation w_d^(-N_d k) a_d^(N_d) and scale
// when requested -- the same values as two cxm_czt_axis_c64 passes
// (src/chromatix/utils/czt.py:7-52 per axis, :55-79 over both, :82-124 for
// zoomed_fft; the modulation of src/chromatix/functional/propagation.py:
// 427-441).  For a small output window the transform is a genuine
// contraction (8 (m n^2 + m^2 n) flops for an m x m window of an n x n
// plane), which FP32 matrix cores run faster than the three FFTs per line
// of Bluestein's method.
//
// The matrices are built on the device in FP64 (phase j (k log w - log a)
// evaluated in double, rounded once to complex64).  The GEMM keeps complex64
// operands split into real and imaginary parts in LDS and accumulates
// Re = Ar Br - Ai Bi and Im = Ar Bi + Ai Br with v_mfma_f32_16x16x4f32 in
// FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "cxm_internal.h"

namespace cxm {

// ---- matrices -----------------------------------------------------------------
// out[k * ld_k + j * ld_j] = exp((j - mod_n) (k Lw - La)) * scale for k < m,
// j < n, where La = log a, Lw = log w (complex logs); mod_n < 0 means no
// modulation (0).
__global__ void zoom_matrix_kernel(float2* __restrict__ out, int m, int n, long long ld_k,
                                   long long ld_j, double la_re, double la_im, double lw_re,
                                   double lw_im, int mod_n, double scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)m * n) return;
  const int k = (int)(i / n), j = (int)(i - (long long)k * n);
  const double e_re = (double)k * lw_re - la_re, e_im = (double)k * lw_im - la_im;
  // A[k][j] = a^-j w^(jk); with the modulation w^(-N k) a^N this is
  // exp((j - N)(k Lw - La))
  const double q = (double)(j - (mod_n > 0 ? mod_n : 0));
  const double mag = exp(q * e_re) * scale;
  double s, c;
  sincos(q * e_im, &s, &c);
  out[k * ld_k + j * ld_j] = make_float2((float)(mag * c), (float)(mag * s));
}

// ---- complex GEMM on MFMA -----------------------------------------------------
// C[z] = A[z] B[z]: A (M x K), B (K x N), C (M x N), complex64 row-major with
// leading dimensions lda, ldb, ldc and batch strides sa, sb, sc (0: shared).
// Split K (ksplit > 1): block z + batch * q multiplies the K range
// [q * kchunk, (q + 1) * kchunk) and writes its partial product to
// C + q * sp; cgemm_sum_kernel adds the partials in order of q.
struct CgemmArgs {
  const float2* A;
  const float2* B;
  float2* C;
  int M, N, K;
  long long lda, ldb, ldc, sa, sb, sc;
  int batch, kchunk;
  long long sp;
};

constexpr int kBM = 64, kBN = 64, kBK = 16;
constexpr int kAP = kBK + 1;  // LDS pitch of the A tile (rows of K)
constexpr int kBP = kBN + 1;  // LDS pitch of the B tile (rows of N)

using f32x4 = __attribute__((__vector_size__(4 * sizeof(float)))) float;

// 256 threads, four waves in 2 x 2; each wave computes a 32 x 32 block of C
// as 2 x 2 tiles of 16 x 16.  v_mfma_f32_16x16x4f32 operand layout: lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; it holds
// D[4 (l / 16) + r][l % 16] in result register r.
__global__ void __launch_bounds__(256) cgemm_mfma_kernel(CgemmArgs g) {
  __shared__ float ar[kBM * kAP], ai[kBM * kAP];
  __shared__ float br[kBK * kBP], bi[kBK * kBP];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wy = wave >> 1, wx = wave & 1;
  const int row0 = blockIdx.y * kBM, col0 = blockIdx.x * kBN;
  const int ks = blockIdx.z / g.batch;
  const long long z = blockIdx.z - (long long)ks * g.batch;
  const int kbeg = ks * g.kchunk, kend = min(g.K, kbeg + g.kchunk);
  const float2* A = g.A + z * g.sa;
  const float2* B = g.B + z * g.sb;
  float2* C = g.C + ks * g.sp + z * g.sc;
  f32x4 acc_r[2][2], acc_i[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc_r[ti][tj][r] = acc_i[ti][tj][r] = 0.0f;
  // tile loads: A rows of 16 k (4 threads per row, 4 values each), B rows of
  // 64 columns (16 threads per row, 4 values each); the next K tile is
  // loaded into registers while the current one is multiplied
  const int a_row = tid >> 2, a_k = (tid & 3) * 4;
  const int b_k = tid >> 4, b_col = (tid & 15) * 4;
  float2 pa[4], pb[4];
  auto fetch = [&](int k0) {
    const int r = row0 + a_row;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + a_k + q;
      pa[q] = (r < g.M && k < kend) ? A[(long long)r * g.lda + k] : make_float2(0.f, 0.f);
    }
    const int k = k0 + b_k;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = col0 + b_col + q;
      pb[q] = (k < kend && c < g.N) ? B[(long long)k * g.ldb + c] : make_float2(0.f, 0.f);
    }
  };
  fetch(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += kBK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ar[a_row * kAP + a_k + q] = pa[q].x;
      ai[a_row * kAP + a_k + q] = pa[q].y;
      br[b_k * kBP + b_col + q] = pb[q].x;
      bi[b_k * kBP + b_col + q] = pb[q].y;
    }
    block_sync();
    if (k0 + kBK < kend) fetch(k0 + kBK);
#pragma unroll
    for (int s = 0; s < kBK / 4; ++s) {
      const int kk = 4 * s + (lane >> 4);
      float xa_r[2], xa_i[2], xb_r[2], xb_i[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int ra = wy * 32 + t * 16 + (lane & 15);
        xa_r[t] = ar[ra * kAP + kk];
        xa_i[t] = ai[ra * kAP + kk];
        const int cb = wx * 32 + t * 16 + (lane & 15);
        xb_r[t] = br[kk * kBP + cb];
        xb_i[t] = bi[kk * kBP + cb];
      }
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
          acc_r[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa_r[ti], xb_r[tj], acc_r[ti][tj], 0, 0, 0);
          acc_r[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(-xa_i[ti], xb_i[tj], acc_r[ti][tj], 0, 0, 0);
          acc_i[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa_r[ti], xb_i[tj], acc_i[ti][tj], 0, 0, 0);
          acc_i[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa_i[ti], xb_r[tj], acc_i[ti][tj], 0, 0, 0);
        }
    }
    block_sync();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const int c = col0 + wx * 32 + tj * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wy * 32 + ti * 16 + 4 * (lane >> 4) + r;
        if (row < g.M && c < g.N)
          C[(long long)row * g.ldc + c] = make_float2(acc_r[ti][tj][r], acc_i[ti][tj][r]);
      }
    }
}

// C[z] (M x N, leading dimension ldc) = sum over q < nsplit of P + q * sp + z * sc
__global__ void cgemm_sum_kernel(const float2* __restrict__ P, float2* __restrict__ C, int M,
                                 int N, long long ldc, long long sc, long long sp, int nsplit,
                                 int batch) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)M * N;
  if (i >= per * batch) return;
  const long long z = i / per, e = i - z * per;
  const long long off = z * sc + (e / N) * ldc + e % N;
  float2 acc = P[off];
  for (int q = 1; q < nsplit; ++q) {
    const float2 v = P[q * sp + off];
    acc = make_float2(acc.x + v.x, acc.y + v.y);
  }
  C[off] = acc;
}

// Output tiles fewer than the CUs can hold: split K over up to `max_split`
// slices (partials in `partial`, max_split * batch * sc elements, then
// summed in order).  Returns the split used.
int cgemm_split(const CgemmArgs& g, int batch, int max_split) {
  const long long tiles = (long long)((g.N + kBN - 1) / kBN) * ((g.M + kBM - 1) / kBM) * batch;
  int q = 1;
  while (q * 2 <= max_split && tiles * q * 2 <= 512 && g.K / (q * 2) >= 128) q *= 2;
  return q;
}

int launch_cgemm(const CgemmArgs& g0, int batch, hipStream_t s, float2* partial = nullptr,
                 int max_split = 1) {
  if (g0.M <= 0 || g0.N <= 0 || batch <= 0) return CXM_OK;
  CgemmArgs g = g0;
  g.batch = batch;
  const int nsplit = partial ? cgemm_split(g0, batch, max_split) : 1;
  g.kchunk = (g.K + nsplit - 1) / nsplit;
  g.kchunk = (g.kchunk + kBK - 1) / kBK * kBK;
  if (nsplit > 1) {
    g.C = partial;
    g.sp = (long long)batch * g0.sc;
  }
  const dim3 grid((unsigned)((g.N + kBN - 1) / kBN), (unsigned)((g.M + kBM - 1) / kBM),
                  (unsigned)(batch * nsplit));
  if (grid.y > 65535u || grid.z > 65535u) return fail(CXM_ERR_INVALID, "GEMM grid too large");
  void* t0 = timing_start(s);
  hipLaunchKernelGGL(cgemm_mfma_kernel, grid, dim3(256), 0, s, g);
  CXM_LAUNCH_CHECK();
  if (nsplit > 1) {
    const long long n = (long long)g0.M * g0.N * batch;
    hipLaunchKernelGGL(cgemm_sum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       (const float2*)partial, g0.C, g0.M, g0.N, g0.ldc, g0.sc, g.sp, nsplit, batch);
    CXM_LAUNCH_CHECK();
  }
  if (t0) timing_stop(t0, s, "cgemm_mfma");
  return CXM_OK;
}

}  // namespace cxm

using namespace cxm;

namespace {
struct ZoomGeom {
  long long mat_y, mat_x;  // elements of A_y (m_y x n_y) and A_x^T (n_x x m_x)
  long long t_plane;       // elements of one intermediate plane (n_y x m_x)
  int chunk;               // planes per GEMM pair
  long long part;          // elements of the split-K partial products
};

int zoom_geom(const cxm_zoom_desc* d, ZoomGeom& g) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->batch < 0 || d->height <= 0 || d->width <= 0 || d->m_y <= 0 || d->m_x <= 0)
    return fail(CXM_ERR_INVALID, "invalid zoom-DFT shape");
  if (d->mod_n_y >= d->m_y || d->mod_n_x >= d->m_x)
    return fail(CXM_ERR_INVALID, "mod_n must be smaller than the output length");
  if ((long long)d->height * d->width > (1ll << 31) || (long long)d->m_y * d->height > (1ll << 31) ||
      (long long)d->m_x * d->width > (1ll << 31))
    return fail(CXM_ERR_UNSUPPORTED, "zoom-DFT matrices too large");
  g.mat_y = (long long)d->m_y * d->height;
  g.mat_x = (long long)d->width * d->m_x;
  g.t_plane = (long long)d->height * d->m_x;
  const long long budget = (64ll << 20) / 8;
  g.chunk = (int)std::max(1ll, std::min<long long>(std::max(1, d->batch),
                                                   std::min(4096ll, budget / g.t_plane)));
  // up to 16 K slices of either product, within 32 MiB
  const long long per_split = (long long)g.chunk * std::max(g.t_plane, (long long)d->m_y * d->m_x);
  g.part = std::min(16 * per_split, std::max(per_split, (32ll << 20) / 8));
  return CXM_OK;
}

size_t zoom_ws_elems(const ZoomGeom& g) {
  auto up = [](long long e) { return (e + 31) / 32 * 32; };  // 256-byte aligned pieces
  return (size_t)(up(g.mat_y) + up(g.mat_x) + up((long long)g.chunk * g.t_plane) + up(g.part));
}

void complex_log(const double* v, double& re, double& im) {
  re = std::log(std::hypot(v[0], v[1]));
  im = std::atan2(v[1], v[0]);
}
}  // namespace

extern "C" size_t cxm_zoom_dft_workspace_size(const cxm_zoom_desc* d) {
  ZoomGeom g;
  if (zoom_geom(d, g)) return 0;
  return zoom_ws_elems(g) * sizeof(float2);
}

extern "C" int cxm_zoom_dft_mfma_c64(void* stream, const void* in, void* out, void* workspace,
                                     size_t workspace_bytes, const cxm_zoom_desc* d) {
  ZoomGeom g;
  int rc = zoom_geom(d, g);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const size_t need = zoom_ws_elems(g) * sizeof(float2);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  for (const double* v : {d->a_y, d->w_y, d->a_x, d->w_x})
    if (!(std::hypot(v[0], v[1]) > 0.0)) return fail(CXM_ERR_INVALID, "a and w must be non-zero");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  auto up = [](long long e) { return (e + 31) / 32 * 32; };
  float2* Ay = (float2*)workspace;
  float2* AxT = Ay + up(g.mat_y);
  float2* T = AxT + up(g.mat_x);
  float2* part = T + up((long long)g.chunk * g.t_plane);
  double la[2], lw[2];
  // A_y: m_y x n_y, row k, column j
  complex_log(d->a_y, la[0], la[1]);
  complex_log(d->w_y, lw[0], lw[1]);
  hipLaunchKernelGGL(zoom_matrix_kernel, dim3((unsigned)((g.mat_y + 255) / 256)), dim3(256), 0, s,
                     Ay, d->m_y, d->height, (long long)d->height, 1ll, la[0], la[1], lw[0], lw[1],
                     d->mod_n_y, d->scale_y);
  CXM_LAUNCH_CHECK();
  // A_x^T: n_x x m_x, element (j, k) = A_x[k][j]
  complex_log(d->a_x, la[0], la[1]);
  complex_log(d->w_x, lw[0], lw[1]);
  hipLaunchKernelGGL(zoom_matrix_kernel, dim3((unsigned)((g.mat_x + 255) / 256)), dim3(256), 0, s,
                     AxT, d->m_x, d->width, 1ll, (long long)d->m_x, la[0], la[1], lw[0], lw[1],
                     d->mod_n_x, d->scale_x);
  CXM_LAUNCH_CHECK();
  const long long in_plane = (long long)d->height * d->width;
  const long long out_plane = (long long)d->m_y * d->m_x;
  for (int p0 = 0; p0 < d->batch; p0 += g.chunk) {
    const int kc = std::min(g.chunk, d->batch - p0);
    // T = X A_x^T  (n_y x n_x) (n_x x m_x)
    CgemmArgs g1;
    memset(&g1, 0, sizeof(g1));
    g1.A = (const float2*)in + (long long)p0 * in_plane;
    g1.B = AxT;
    g1.C = T;
    g1.M = d->height;
    g1.N = d->m_x;
    g1.K = d->width;
    g1.lda = d->width;
    g1.ldb = d->m_x;
    g1.ldc = d->m_x;
    g1.sa = in_plane;
    g1.sb = 0;
    g1.sc = g.t_plane;
    if ((rc = launch_cgemm(g1, kc, s, part, (int)(g.part / ((long long)kc * g1.sc))))) return rc;
    // Y = A_y T  (m_y x n_y) (n_y x m_x)
    CgemmArgs g2;
    memset(&g2, 0, sizeof(g2));
    g2.A = Ay;
    g2.B = T;
    g2.C = (float2*)out + (long long)p0 * out_plane;
    g2.M = d->m_y;
    g2.N = d->m_x;
    g2.K = d->height;
    g2.lda = d->height;
    g2.ldb = d->m_x;
    g2.ldc = d->m_x;
    g2.sa = 0;
    g2.sb = g.t_plane;
    g2.sc = out_plane;
    if ((rc = launch_cgemm(g2, kc, s, part, (int)(g.part / ((long long)kc * g2.sc))))) return rc;
  }
  return CXM_OK;
}

// Complex GEMM entry (batched), C[z] = A[z] B[z]; exposed for tests of the
// MFMA layout and for callers with their own DFT matrices.
extern "C" int cxm_cgemm_mfma_c64(void* stream, const void* A, const void* B, void* C, int32_t M,
                                  int32_t N, int32_t K, int64_t lda, int64_t ldb, int64_t ldc,
                                  int64_t stride_a, int64_t stride_b, int64_t stride_c,
                                  int32_t batch) {
  if (M < 0 || N < 0 || K < 0 || batch < 0) return fail(CXM_ERR_INVALID, "negative GEMM size");
  if (M == 0 || N == 0 || batch == 0) return CXM_OK;
  if (!A || !B || !C) return fail(CXM_ERR_INVALID, "null buffer");
  if (lda < K || ldb < N || ldc < N) return fail(CXM_ERR_INVALID, "leading dimension too small");
  int rc;
  if ((rc = ensure_init())) return rc;
  CgemmArgs g;
  memset(&g, 0, sizeof(g));
  g.A = (const float2*)A;
  g.B = (const float2*)B;
  g.C = (float2*)C;
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.sa = stride_a;
  g.sb = stride_b;
  g.sc = stride_c;
  return launch_cgemm(g, batch, (hipStream_t)stream);
}

// Matrix of a cxm_czt_axis_c64 pass (parameter set `param`), element (k, j)
// at out[k * ld_r + j * ld_c], conjugated on request: exp((j - N)(k Lw - La))
// scale with the logs conjugated is its conjugate (j, k, N and scale real).
extern "C" int cxm_czt_matrix_c64(void* stream, const cxm_czt_desc* d, int32_t param, void* out,
                                  int64_t ld_r, int64_t ld_c, int32_t conj) {
  if (!d || !out || !d->a || !d->w) return fail(CXM_ERR_INVALID, "null descriptor or buffer");
  if (d->height <= 0 || d->width <= 0 || d->m <= 0 || (d->axis != 0 && d->axis != 1))
    return fail(CXM_ERR_INVALID, "invalid shape parameters");
  if (param < 0 || param >= d->num_param) return fail(CXM_ERR_INVALID, "param out of range");
  if (d->mod_n >= d->m) return fail(CXM_ERR_INVALID, "mod_n must be smaller than m");
  const int n = d->axis == 0 ? d->height : d->width;
  const long long total = (long long)d->m * n;
  if (total > (1ll << 31)) return fail(CXM_ERR_UNSUPPORTED, "CZT matrix too large");
  const double* a = d->a + 2 * param;
  const double* w = d->w + 2 * param;
  if (!(std::hypot(a[0], a[1]) > 0.0) || !(std::hypot(w[0], w[1]) > 0.0))
    return fail(CXM_ERR_INVALID, "a and w must be non-zero");
  int rc;
  if ((rc = ensure_init())) return rc;
  double la[2], lw[2];
  complex_log(a, la[0], la[1]);
  complex_log(w, lw[0], lw[1]);
  const double sg = conj ? -1.0 : 1.0;
  hipLaunchKernelGGL(zoom_matrix_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (float2*)out, d->m, n, (long long)ld_r, (long long)ld_c,
                     la[0], sg * la[1], lw[0], sg * lw[1], d->mod_n, d->scale);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}
