// Elementwise and per-plane reduction kernels around the propagation path:
//   intensity  |u|^2, spectrally weighted over an innermost wavelength axis
//              (src/chromatix/core/field.py:663-665, 735-740)
//   plane sums of |u|^2 for power (field.py:656-661, 728-733), reduced in a
//              fixed order (two stages, FP64 partials) so results are
//              bitwise reproducible and independent of what else runs.
//   copy       plain complex64 copy with the 8-byte-per-lane accesses of the
//              FFT passes: the bandwidth reference and the known-size kernel
//              that scales the HBM traffic counters.
//   high-NA    the defocus factor of high_na_ff_lens
//              (src/chromatix/functional/lenses.py:98-156), in FP64.
//   sensor     sum over an axis (reduce_axis, functional/sensors.py:79-80)
//              and approximate shot noise (ops/noise.py:10-16).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "cxm_internal.h"

namespace cxm {

constexpr int kMaxWv = 64;
struct Density { int nwv; float w[kMaxWv]; };

__global__ void intensity_kernel(const float2* __restrict__ u, float* __restrict__ out,
                                 long long n_pix, Density d) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pix) return;
  if (d.nwv == 1) {
    const float2 v = u[p];
    out[p] = v.x * v.x + v.y * v.y;
  } else {
    float acc = 0.f;
    const float2* q = u + p * d.nwv;
    for (int w = 0; w < d.nwv; ++w) acc += d.w[w] * (q[w].x * q[w].x + q[w].y * q[w].y);
    out[p] = acc;
  }
}

__global__ void __launch_bounds__(256)
copy_kernel(const float2* __restrict__ in, float2* __restrict__ out, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = in[i];
}

constexpr int kSumThreads = 256;
constexpr int kSumChunks = 64;  // blocks per plane in stage 1

__global__ void __launch_bounds__(kSumThreads)
plane_sumsq_stage1(const float2* __restrict__ u, double* __restrict__ partial,
                   long long plane_elems, Density d) {
  const int plane = blockIdx.y;
  const int chunk = blockIdx.x;
  const long long per = (plane_elems + kSumChunks - 1) / kSumChunks;
  const long long b = chunk * per;
  const long long e = min(plane_elems, b + per);
  const float2* src = u + (long long)plane * plane_elems;
  double acc = 0.0;
  for (long long i = b + threadIdx.x; i < e; i += kSumThreads) {
    const float2 v = src[i];
    const double m = (double)v.x * v.x + (double)v.y * v.y;
    acc += (d.nwv > 1) ? (double)d.w[i % d.nwv] * m : m;
  }
  __shared__ double s[kSumThreads];
  s[threadIdx.x] = acc;
  block_sync();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    block_sync();
  }
  if (threadIdx.x == 0) partial[(long long)plane * kSumChunks + chunk] = s[0];
}

__global__ void plane_sumsq_stage2(const double* __restrict__ partial, double* __restrict__ out,
                                   int n_planes) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_planes) return;
  double acc = 0.0;
  for (int c = 0; c < kSumChunks; ++c) acc += partial[(long long)p * kSumChunks + c];
  out[p] = acc;
}

static int make_density(int num_wv, const double* density, Density& d) {
  if (num_wv < 1 || num_wv > kMaxWv) return fail(CXM_ERR_INVALID, "num_wv must be in [1, 64]");
  d.nwv = num_wv;
  for (int w = 0; w < kMaxWv; ++w) d.w[w] = 0.f;
  for (int w = 0; w < num_wv; ++w) d.w[w] = density ? (float)density[w] : 1.0f;
  return CXM_OK;
}

}  // namespace cxm

using namespace cxm;

extern "C" int cxm_intensity_c64(void* stream, const void* u, void* out, int64_t n_pixels,
                                 int32_t num_wv, const double* density) {
  Density d;
  int rc = make_density(num_wv, density, d);
  if (rc) return rc;
  if (n_pixels < 0) return fail(CXM_ERR_INVALID, "negative size");
  if (n_pixels == 0) return CXM_OK;
  if (!u || !out) return fail(CXM_ERR_INVALID, "null buffer");
  if (num_wv == 1 && density == nullptr) d.w[0] = 1.0f;
  const long long blocks = (n_pixels + 255) / 256;
  hipLaunchKernelGGL(intensity_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)u, (float*)out, (long long)n_pixels, d);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

extern "C" int cxm_copy_c64(void* stream, const void* in, void* out, int64_t n) {
  if (n < 0) return fail(CXM_ERR_INVALID, "negative size");
  if (n == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  // 8 workgroups of 256 lanes per CU, grid-stride over the rest
  const long long blocks = std::min<long long>((n + 255) / 256, 256LL * 8);
  void* t0 = timing_start(s);
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const float2*)in,
                     (float2*)out, (long long)n);
  CXM_LAUNCH_CHECK();
  if (t0) timing_stop(t0, s, "copy_c64");
  return CXM_OK;
}

extern "C" size_t cxm_plane_sumsq_workspace_size(int64_t n_planes) {
  return (size_t)(n_planes > 0 ? n_planes : 0) * kSumChunks * sizeof(double);
}

extern "C" int cxm_plane_sumsq_c64(void* stream, const void* u, void* out, int64_t n_planes,
                                   int64_t plane_elems, int32_t num_wv, const double* density,
                                   void* workspace, size_t workspace_bytes) {
  Density d;
  int rc = make_density(num_wv, density, d);
  if (rc) return rc;
  if (n_planes < 0 || plane_elems < 0) return fail(CXM_ERR_INVALID, "negative size");
  if (n_planes == 0) return CXM_OK;
  if (n_planes > 65535) return fail(CXM_ERR_INVALID, "at most 65535 planes per call");
  if (!u || !out) return fail(CXM_ERR_INVALID, "null buffer");
  if (num_wv > 1 && plane_elems % num_wv != 0)
    return fail(CXM_ERR_INVALID, "plane_elems must be a multiple of num_wv");
  const size_t need = cxm_plane_sumsq_workspace_size(n_planes);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(plane_sumsq_stage1, dim3(kSumChunks, (unsigned)n_planes), dim3(kSumThreads), 0, s,
                     (const float2*)u, (double*)workspace, (long long)plane_elems, d);
  CXM_LAUNCH_CHECK();
  hipLaunchKernelGGL(plane_sumsq_stage2, dim3((unsigned)((n_planes + 63) / 64)), dim3(64), 0, s,
                     (const double*)workspace, (double*)out, (int)n_planes);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

// ---------------------------------------------------------------------------
// high_na_ff_lens defocus: u * exp(i k sz f) / sz (0 where sz == 0) on the
// centred frequency grid, sz = sqrt(max(1 - NA^2 |f_grid lambda / n|^2, 0)).
// ---------------------------------------------------------------------------
namespace cxm {
struct HighNaParams {
  double fy_step, fx_step;  // 1 / (N d) per axis
  double s_scale;           // lambda / n
  double NA2, k, f;
  int ny, nx;
};

// One thread per grid position: the factor depends on the position only, so
// it is evaluated once (FP64) and applied to that position of every plane.
__global__ void high_na_defocus_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                       long long n_planes, HighNaParams c) {
  const long long plane = (long long)c.ny * c.nx;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= plane) return;
  const int y = (int)(r / c.nx), x = (int)(r - (long long)y * c.nx);
  // fftshift(fftfreq(N)) / d = (index - N/2 rounded down) / (N d)
  const double sy = (double)(y - c.ny / 2) * c.fy_step * c.s_scale;
  const double sx = (double)(x - c.nx / 2) * c.fx_step * c.s_scale;
  const double sz2 = 1.0 - c.NA2 * (sy * sy + sx * sx);
  const double sz = sqrt(sz2 > 0.0 ? sz2 : 0.0);
  if (sz == 0.0) {
    for (long long p = 0; p < n_planes; ++p) out[p * plane + r] = make_float2(0.f, 0.f);
    return;
  }
  const double ph = c.k * sz * c.f;
  const double re = cos(ph) / sz, im = sin(ph) / sz;
  for (long long p = 0; p < n_planes; ++p) {
    const float2 v = in[p * plane + r];
    out[p * plane + r] = make_float2((float)(v.x * re - v.y * im), (float)(v.x * im + v.y * re));
  }
}
}  // namespace cxm

extern "C" int cxm_high_na_defocus_c64(void* stream, const void* in, void* out, int64_t n_planes,
                                       int32_t height, int32_t width, double dy, double dx,
                                       double wavelength, double n, double NA, double f) {
  if (n_planes < 0 || height <= 0 || width <= 0) return fail(CXM_ERR_INVALID, "invalid shape");
  if (!(dy > 0.0) || !(dx > 0.0) || !(wavelength > 0.0) || !(n > 0.0))
    return fail(CXM_ERR_INVALID, "spacing, wavelength and n must be positive");
  if (n_planes == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  HighNaParams c;
  c.fy_step = 1.0 / ((double)height * dy);
  c.fx_step = 1.0 / ((double)width * dx);
  c.s_scale = wavelength / n;
  c.NA2 = NA * NA;
  c.k = 2.0 * 3.14159265358979323846 * n / wavelength;
  c.f = f;
  c.ny = height;
  c.nx = width;
  const long long plane = (long long)height * width;
  if ((plane + 255) / 256 > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "plane too large for one launch");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(high_na_defocus_kernel, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s,
                     (const float2*)in, (float2*)out, (long long)n_planes, c);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

// ---------------------------------------------------------------------------
// Sensor: sum over the middle axis of (outer, n_sum, inner) in index order,
// and approximate shot noise max(x + sqrt(x) n, 0).
// ---------------------------------------------------------------------------
namespace cxm {
__global__ void sum_axis_kernel(const float* __restrict__ in, float* __restrict__ out, long long outer,
                                long long n_sum, long long inner) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= outer * inner) return;
  const long long o = i / inner, j = i - o * inner;
  const float* p = in + o * n_sum * inner + j;
  float acc = 0.0f;
  for (long long k = 0; k < n_sum; ++k) acc = __fadd_rn(acc, p[k * inner]);
  out[i] = acc;
}

__global__ void shot_noise_approx_kernel(const float* __restrict__ x, const float* __restrict__ nrm,
                                         float* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = x[i];
  const float y = __fadd_rn(v, __fmul_rn(sqrtf(v), nrm[i]));
  out[i] = y > 0.0f ? y : 0.0f;
}
}  // namespace cxm

extern "C" int cxm_sum_axis_f32(void* stream, const float* in, float* out, int64_t outer,
                                int64_t n_sum, int64_t inner) {
  if (outer < 0 || n_sum < 0 || inner < 0) return fail(CXM_ERR_INVALID, "negative size");
  const long long tot = outer * inner;
  if (tot == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  if ((tot + 255) / 256 > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "too many elements for one launch");
  hipLaunchKernelGGL(sum_axis_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, in, out, (long long)outer, (long long)n_sum, (long long)inner);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

extern "C" int cxm_shot_noise_approx_f32(void* stream, const float* image, const float* normal,
                                         float* out, int64_t n) {
  if (n < 0) return fail(CXM_ERR_INVALID, "negative size");
  if (n == 0) return CXM_OK;
  if (!image || !normal || !out) return fail(CXM_ERR_INVALID, "null buffer");
  if ((n + 255) / 256 > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "too many elements for one launch");
  hipLaunchKernelGGL(shot_noise_approx_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, image, normal, out, (long long)n);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}
