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
#include <algorithm>
#include <cstring>
#include <string>

#include "cxm_pipe_host.h"

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

extern "C" size_t cxm_fconv_workspace_size(const cxm_fconv_desc* d) {
  FcGeom g;
  if (fc_geom(d, g)) return 0;
  return fc_ws(d, g);
}

extern "C" int cxm_fconv_c64(void* stream, const void* image, const void* kernel, void* out,
                             void* workspace, size_t workspace_bytes, const cxm_fconv_desc* d) {
  FcGeom g;
  int rc = fc_geom(d, g);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!image || !kernel || !out) return fail(CXM_ERR_INVALID, "null buffer");
  if (!workspace || workspace_bytes < fc_ws(d, g))
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(fc_ws(d, g)) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
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
