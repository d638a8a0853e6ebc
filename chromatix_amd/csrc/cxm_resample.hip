// Separable resampling of batches of float32 planes: the sensor's resampler
// (init_plane_resample, src/chromatix/ops/resample.py:10-128 -- sum pooling,
// or jax.image.scale_and_translate's antialiased interpolation divided by
// the scale product) and its adjoint, as two passes over per-output tap
// tables that the caller builds:
//   pass y: t[b, i, x] = sum_u wy[i][u] * m(y0[i] + u, x) * in[b, y0[i] + u, x]
//   pass x: out[b, i, j] = sum_v wx[j][v] * t[b, i, x0[j] + v]   (* m(i, j))
// m the optional mask, on the input (the Microscope's PSF taper before
// resampling) or on the output (the adjoint of an input mask).  Each pass
// reads its rows once, coalesced along the contiguous axis; FP32 sums over
// the taps in order.
#include <algorithm>
#include <string>

#include "cxm_internal.h"

namespace cxm {
namespace {

int check_desc(const cxm_resample_desc* d) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->batch < 0 || d->in_h <= 0 || d->in_w <= 0 || d->out_h <= 0 || d->out_w <= 0)
    return fail(CXM_ERR_INVALID, "invalid shape");
  if (d->ky < 1 || d->kx < 1) return fail(CXM_ERR_INVALID, "tap widths must be >= 1");
  if (d->ky > d->in_h || d->kx > d->in_w) return fail(CXM_ERR_INVALID, "tap width exceeds the input");
  if (!d->y_start || !d->wy || !d->x_start || !d->wx) return fail(CXM_ERR_INVALID, "null tap table");
  if (d->mask_on_output != 0 && d->mask_on_output != 1) return fail(CXM_ERR_INVALID, "invalid mask mode");
  return CXM_OK;
}

// t[b, i, x]: thread per (b, i, x), x fastest
__global__ void resample_y_kernel(const float* __restrict__ in, float* __restrict__ t,
                                  const int* __restrict__ y0, const float* __restrict__ wy,
                                  const float* __restrict__ mask, long long n, int in_h, int in_w,
                                  int out_h, int ky, int mask_in) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int x = (int)(idx % in_w);
  const long long bi = idx / in_w;
  const int i = (int)(bi % out_h);
  const long long b = bi / out_h;
  const int ys = y0[i];
  const float* w = wy + (long long)i * ky;
  float acc = 0.0f;
  for (int u = 0; u < ky; ++u) {
    const int y = ys + u;
    if (y < 0 || y >= in_h) continue;
    float v = in[(b * in_h + y) * (long long)in_w + x];
    if (mask_in) v = v * mask[(long long)y * in_w + x];
    acc = fmaf(w[u], v, acc);
  }
  t[idx] = acc;
}

// out[b, i, j]: thread per (b, i, j), j fastest
__global__ void resample_x_kernel(const float* __restrict__ t, float* __restrict__ out,
                                  const int* __restrict__ x0, const float* __restrict__ wx,
                                  const float* __restrict__ mask, long long n, int in_w, int out_h,
                                  int out_w, int kx, int mask_out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int j = (int)(idx % out_w);
  const long long bi = idx / out_w;
  const int xs = x0[j];
  const float* w = wx + (long long)j * kx;
  const float* row = t + bi * in_w;
  float acc = 0.0f;
  for (int v = 0; v < kx; ++v) {
    const int x = xs + v;
    if (x < 0 || x >= in_w) continue;
    acc = fmaf(w[v], row[x], acc);
  }
  if (mask_out) acc = acc * mask[(bi % out_h) * (long long)out_w + j];
  out[idx] = acc;
}

unsigned grid_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace cxm

using namespace cxm;

extern "C" size_t cxm_resample_workspace_size(const cxm_resample_desc* d) {
  if (check_desc(d)) return 0;
  return (size_t)std::max<long long>(1, d->batch * d->out_h * (long long)d->in_w) * sizeof(float);
}

extern "C" int cxm_resample_f32(void* stream, const float* in, float* out, void* workspace,
                                size_t workspace_bytes, const cxm_resample_desc* d) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  if (!workspace || workspace_bytes < cxm_resample_workspace_size(d))
    return fail(CXM_ERR_WORKSPACE, "workspace too small");
  const long long nt = d->batch * d->out_h * (long long)d->in_w;
  const long long no = d->batch * d->out_h * (long long)d->out_w;
  if ((nt + 255) / 256 > 0x7fffffffLL || (no + 255) / 256 > 0x7fffffffLL)
    return fail(CXM_ERR_INVALID, "batch too large");
  hipStream_t s = (hipStream_t)stream;
  float* t = (float*)workspace;
  const int mask_in = d->mask && !d->mask_on_output;
  const int mask_out = d->mask && d->mask_on_output;
  hipLaunchKernelGGL(resample_y_kernel, dim3(grid_for(nt)), dim3(256), 0, s, in, t, d->y_start, d->wy,
                     d->mask, nt, d->in_h, d->in_w, d->out_h, d->ky, mask_in);
  CXM_LAUNCH_CHECK();
  hipLaunchKernelGGL(resample_x_kernel, dim3(grid_for(no)), dim3(256), 0, s, (const float*)t, out,
                     d->x_start, d->wx, d->mask, no, d->in_w, d->out_h, d->out_w, d->kx, mask_out);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}
