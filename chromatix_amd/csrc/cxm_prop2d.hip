// Batched 2D Fourier-convolution propagation on gfx950.
//
// One propagation of a plane is three passes over HBM:
//   row pass    : zero-padded load -> row FFT                  -> workspace
//   column pass : column FFT -> * H(ky,kx) (on the fly) -> column IFFT, in place
//   row pass    : row IFFT -> crop -> complex or |u|^2         -> output
// The column pass reads all rows of its strip before writing any, so it runs
// in place; the workspace holds one chunk of planes, sized to stay inside the
// 256 MB Infinity Cache so that the intermediate round trips are served by
// the cache rather than HBM.
//
// Replaces kernel_propagate's default branch together with the propagators
// of transfer_propagate / asm_propagate
// (src/chromatix/functional/propagation.py:168-352, 391-394, 466-580).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "cxm_hfunc.h"
#include "cxm_internal.h"
#include "cxm_sizes.h"

namespace cxm {

bool fft_len_supported(long long n) {
  switch (n) {
#define CXM_CASE(N, ER, EC, C) case N: return true;
    CXM_FOR_EACH_SIZE(CXM_CASE)
#undef CXM_CASE
    default: return false;
  }
}

int launch_row_variant_1024(int v, int dir, const RowArgs& a, hipStream_t s);

int launch_row(int N, int dir, const RowArgs& a, hipStream_t s) {
  int rc;
  const int v = option(CXM_OPT_ROW_VARIANT);
  if (v > 0 && N == 1024 && (rc = launch_row_variant_1024(v, dir, a, s)) >= 0) return rc;
  if ((rc = launch_row_g0(N, dir, a, s)) >= 0) return rc;
  if ((rc = launch_row_g1(N, dir, a, s)) >= 0) return rc;
  if ((rc = launch_row_g2(N, dir, a, s)) >= 0) return rc;
  if ((rc = launch_row_g3(N, dir, a, s)) >= 0) return rc;
  if ((rc = launch_row_g4(N, dir, a, s)) >= 0) return rc;
  return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(N));
}

int launch_col_variant_1024(int v, const ColArgs& a, hipStream_t s);
int launch_col_variant_2048(int v, const ColArgs& a, hipStream_t s);

int launch_col_prop(int N, const ColArgs& a, hipStream_t s) {
  int rc;
  int v = option(CXM_OPT_COL_VARIANT);
  // single planes (the multislice loop): 8 values per thread and 4 columns,
  // twice the waves of the batched default (20 -> 18 us per plane)
  if (v == 0 && N == 1024 && a.n_planes <= 4) v = 2;
  if (v > 0 && N == 1024 && (rc = launch_col_variant_1024(v, a, s)) >= 0) return rc;
  if (v > 0 && N == 2048 && (rc = launch_col_variant_2048(v, a, s)) >= 0) return rc;
  if ((rc = launch_col_g0(N, a, s)) >= 0) return rc;
  if ((rc = launch_col_g1(N, a, s)) >= 0) return rc;
  if ((rc = launch_col_g2(N, a, s)) >= 0) return rc;
  if ((rc = launch_col_g3(N, a, s)) >= 0) return rc;
  if ((rc = launch_col_g4(N, a, s)) >= 0) return rc;
  return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(N));
}

}  // namespace cxm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace cxm;

namespace {
struct Geom {
  int Ny, Nx, Hin, Win, Hout, Wout;
  long long plane_in, plane_ws1, plane_ws2;  // elements
  int chunk;                                 // input planes per chunk
};

int validate(const cxm_prop2d_desc* d, Geom& g) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->batch < 0 || d->height <= 0 || d->width <= 0 || d->pad_y < 0 || d->pad_x < 0 ||
      d->crop_y < 0 || d->crop_x < 0 || d->num_z < 1)
    return fail(CXM_ERR_INVALID, "invalid prop2d shape parameters");
  if (d->kind < CXM_H_NONE || d->kind > CXM_H_ARRAY)
    return fail(CXM_ERR_INVALID, "invalid propagator kind");
  if (d->phase_mode != CXM_PHASE_REF32 && d->phase_mode != CXM_PHASE_ACCURATE)
    return fail(CXM_ERR_INVALID, "invalid phase mode");
  if (d->out_kind != CXM_OUT_COMPLEX && d->out_kind != CXM_OUT_INTENSITY)
    return fail(CXM_ERR_INVALID, "invalid output kind");
  if ((d->kind == CXM_H_TRANSFER || d->kind == CXM_H_ASM) && !d->z)
    return fail(CXM_ERR_INVALID, "z list required");
  if (d->kind == CXM_H_ARRAY && !d->propagator)
    return fail(CXM_ERR_INVALID, "propagator array required");
  if (d->kind != CXM_H_TRANSFER && d->kind != CXM_H_ASM && d->num_z != 1)
    return fail(CXM_ERR_INVALID, "num_z must be 1 without an analytic propagator");
  if (d->z_zip != 0 && d->z_zip != 1) return fail(CXM_ERR_INVALID, "z_zip must be 0 or 1");
  const int nwv = d->num_wv > 0 ? d->num_wv : 1;
  if ((d->kind == CXM_H_TRANSFER || d->kind == CXM_H_ASM) && nwv * d->num_z > kMaxHP)
    return fail(CXM_ERR_INVALID, "num_wv * num_z exceeds " + std::to_string(kMaxHP));
  if (d->n <= 0 || (!d->wavelengths && d->wavelength <= 0) || (!d->spacing && (d->dx <= 0 || d->dy <= 0)))
    return fail(CXM_ERR_INVALID, "wavelength, spacing and n must be positive");
  g.Ny = d->height + 2 * d->pad_y;
  g.Nx = d->width + 2 * d->pad_x;
  g.Hin = d->height;
  g.Win = d->width;
  g.Hout = g.Ny - 2 * d->crop_y;
  g.Wout = g.Nx - 2 * d->crop_x;
  if (g.Hout <= 0 || g.Wout <= 0) return fail(CXM_ERR_INVALID, "crop removes the whole plane");
  if (!fft_len_supported(g.Ny) || !fft_len_supported(g.Nx))
    return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT grid " + std::to_string(g.Ny) + "x" +
                                         std::to_string(g.Nx) + " (lengths 2^a <= 8192, 5*2^a <= 5120, 3*2^a <= 3072)");
  g.plane_in = (long long)g.Hin * g.Win;
  g.plane_ws1 = (long long)g.Hin * g.Nx;
  g.plane_ws2 = (long long)g.Hout * g.Nx;
  // Chunk so that the workspace of one chunk stays <= 128 MiB, half the 256 MiB
  // Infinity Cache: the intermediate planes are re-read from it, not from HBM.
  const int nout = d->z_zip ? 1 : d->num_z;
  const long long per_plane = (nout == 1) ? std::max(g.plane_ws1, g.plane_ws2)
                                          : g.plane_ws1 + (long long)nout * g.plane_ws2;
  const long long budget = (128ll << 20) / 8;
  long long ch = std::max(1ll, budget / std::max(1ll, per_plane));
  if (option(CXM_OPT_CHUNK_PLANES) > 0) ch = option(CXM_OPT_CHUNK_PLANES);
  g.chunk = (int)std::min<long long>(ch, std::max(1, d->batch));
  return CXM_OK;
}

int out_per_plane(const cxm_prop2d_desc* d) { return d->z_zip ? 1 : d->num_z; }

size_t ws_elems(const cxm_prop2d_desc* d, const Geom& g) {
  const int nout = out_per_plane(d);
  if (nout == 1) return (size_t)g.chunk * std::max(g.plane_ws1, g.plane_ws2);
  return (size_t)g.chunk * (g.plane_ws1 + (size_t)nout * g.plane_ws2);
}
}  // namespace

extern "C" size_t cxm_prop2d_workspace_size(const cxm_prop2d_desc* d) {
  Geom g;
  if (validate(d, g) != CXM_OK) return 0;
  return ws_elems(d, g) * sizeof(float2);
}

namespace {
int prop2d_impl(void* stream, const void* in, void* out, void* workspace, size_t workspace_bytes,
                const cxm_prop2d_desc* d, const HParam* hp_dev) {
  Geom g;
  int rc = validate(d, g);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const size_t need = ws_elems(d, g) * sizeof(float2);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;

  ColArgs ca;
  memset(&ca, 0, sizeof(ca));
  if (hp_dev) {
    if (d->kind != CXM_H_TRANSFER && d->kind != CXM_H_ASM)
      return fail(CXM_ERR_INVALID, "a parameter table needs kind TRANSFER or ASM");
    if (d->z_zip) return fail(CXM_ERR_INVALID, "a parameter table replaces z_zip");
    ca.hp_dev = hp_dev;
  } else if (d->kind == CXM_H_TRANSFER || d->kind == CXM_H_ASM) {
    if ((rc = build_hparams(d, g.Ny, g.Nx, ca.hp, kMaxHP))) return rc;
    ca.nhp = (d->num_wv > 0 ? d->num_wv : 1) * d->num_z;
  }
  float2* ws1 = (float2*)workspace;
  const int nz = out_per_plane(d);
  float2* ws2 = (nz == 1) ? ws1 : ws1 + (size_t)g.chunk * g.plane_ws1;
  const float2* src = (const float2*)in;
  const long long out_plane = (long long)g.Hout * g.Wout;

  for (int p0 = 0; p0 < d->batch; p0 += g.chunk) {
    const int kc = std::min(g.chunk, d->batch - p0);
    // 1) rows: zero-padded forward FFT
    RowArgs r1;
    memset(&r1, 0, sizeof(r1));
    r1.in = src + (size_t)p0 * g.plane_in;
    r1.out_c = ws1;
    r1.tw = twiddles_for(g.Nx);
    r1.in_plane_stride = g.plane_in;
    r1.out_plane_stride = g.plane_ws1;
    r1.in_row_stride = g.Win;
    r1.out_row_stride = g.Nx;
    r1.n_rows = kc * g.Hin;
    r1.rows_per_plane = g.Hin;
    r1.in_off = d->pad_x;
    r1.in_len = g.Win;
    r1.out_off = 0;
    r1.out_len = g.Nx;
    r1.scale = 1.0f;
    if ((rc = launch_row(g.Nx, -1, r1, s))) return rc;
    // 2) columns: FFT * H * IFFT (in place when num_z == 1)
    ca.in = ws1;
    ca.out = ws2;
    ca.tw = twiddles_for(g.Ny);
    ca.H = (const float2*)d->propagator;
    ca.in_plane_stride = g.plane_ws1;
    ca.out_plane_stride = g.plane_ws2;
    ca.row_stride = g.Nx;
    ca.ncols = g.Nx;
    ca.n_planes = kc;
    ca.plane0 = p0;
    ca.nz = d->num_z;
    ca.z_zip = d->z_zip;
    ca.nwv = d->num_wv > 0 ? d->num_wv : 1;
    ca.in_off = d->pad_y;
    ca.in_len = g.Hin;
    ca.out_off = d->crop_y;
    ca.out_len = g.Hout;
    ca.kind = d->kind;
    ca.phase_mode = d->phase_mode;
    ca.remove_evanescent = d->remove_evanescent;
    ca.bandlimit = d->bandlimit;
    ca.scale = (float)(1.0 / ((double)g.Ny * (double)g.Nx));
    if ((rc = launch_col_prop(g.Ny, ca, s))) return rc;
    // 3) rows: inverse FFT, crop, complex or intensity
    RowArgs r2;
    memset(&r2, 0, sizeof(r2));
    r2.in = ws2;
    r2.tw = twiddles_for(g.Nx);
    r2.in_plane_stride = g.plane_ws2;
    r2.in_row_stride = g.Nx;
    r2.out_plane_stride = out_plane;
    r2.out_row_stride = g.Wout;
    r2.n_rows = kc * nz * g.Hout;
    r2.rows_per_plane = g.Hout;
    r2.in_off = 0;
    r2.in_len = g.Nx;
    r2.out_off = d->crop_x;
    r2.out_len = g.Wout;
    r2.scale = 1.0f;
    const size_t out_off_planes = (size_t)p0 * nz;
    if (d->out_kind == CXM_OUT_INTENSITY) {
      r2.intensity = 1;
      r2.out_i = (float*)out + out_off_planes * out_plane;
    } else {
      r2.out_c = (float2*)out + out_off_planes * out_plane;
    }
    if ((rc = launch_row(g.Nx, 1, r2, s))) return rc;
  }
  return CXM_OK;
}
}  // namespace

extern "C" int cxm_prop2d_c64(void* stream, const void* in, void* out, void* workspace,
                              size_t workspace_bytes, const cxm_prop2d_desc* d) {
  return prop2d_impl(stream, in, out, workspace, workspace_bytes, d, nullptr);
}

extern "C" size_t cxm_prop2d_param_bytes(void) { return sizeof(HParam); }

extern "C" int cxm_prop2d_build_params(const cxm_prop2d_desc* d, const double* z_planes,
                                       int64_t n_planes, void* host_out) {
  Geom g;
  cxm_prop2d_desc t = *d;
  double z0 = 1.0;
  t.num_z = 1;
  t.z_zip = 0;
  t.z = &z0;
  int rc = validate(&t, g);
  if (rc) return rc;
  if (n_planes < 0 || (n_planes > 0 && (!z_planes || !host_out)))
    return fail(CXM_ERR_INVALID, "null buffer");
  const int nwv = d->num_wv > 0 ? d->num_wv : 1;
  const int nout = d->num_z > 0 ? d->num_z : 1;
  HParam* o = (HParam*)host_out;
  for (int64_t i = 0; i < n_planes; ++i) {
    // output plane i = (input plane p, distance iz); wavelength p % nwv
    const int64_t p = i / nout;
    const int w = (int)(p % nwv);
    cxm_prop2d_desc one = t;
    one.num_wv = 1;
    one.z = &z_planes[i];
    one.wavelength = d->wavelengths ? d->wavelengths[w] : d->wavelength;
    one.wavelengths = nullptr;
    if (d->spacing) {
      one.dy = d->spacing[2 * w];
      one.dx = d->spacing[2 * w + 1];
    }
    one.spacing = nullptr;
    if ((rc = build_hparams(&one, g.Ny, g.Nx, &o[i], 1))) return rc;
  }
  return CXM_OK;
}

extern "C" int cxm_prop2d_c64_params(void* stream, const void* in, void* out, void* workspace,
                                     size_t workspace_bytes, const cxm_prop2d_desc* d,
                                     const void* params) {
  if (!params) return fail(CXM_ERR_INVALID, "null parameter table");
  cxm_prop2d_desc t = *d;
  double z0 = 1.0;
  t.z = &z0;  // distances come from the table
  return prop2d_impl(stream, in, out, workspace, workspace_bytes, &t, (const HParam*)params);
}

// ---------------------------------------------------------------------------
// Materialised propagator (compute_transfer_propagator / compute_asm_propagator
// of src/chromatix/functional/propagation.py:466-580), FFT order.
// ---------------------------------------------------------------------------
namespace cxm {
template <int KIND, int MODE>
__global__ void materialise_H_kernel(float2* __restrict__ out, int Ny, int Nx, int nz,
                                     int remove_evanescent, int bandlimit, ColArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;  // w * nz + iz
  if (i >= (long long)Ny * Nx) return;
  const int ky = (int)(i / Nx), kx = (int)(i - (long long)ky * Nx);
  const HParam& hp = a.hp[q];
  const XTerms xt = x_terms(hp, kx, Nx, MODE);
  out[(long long)q * Ny * Nx + i] = eval_H(hp, xt, ky, Ny, KIND, MODE, remove_evanescent, bandlimit, 1.0f);
}
}  // namespace cxm

extern "C" int cxm_propagator_c64(void* stream, void* out, const cxm_prop2d_desc* d) {
  Geom g;
  int rc = validate(d, g);
  if (rc) return rc;
  if (d->kind != CXM_H_TRANSFER && d->kind != CXM_H_ASM)
    return fail(CXM_ERR_INVALID, "propagator kind must be TRANSFER or ASM");
  if (!out) return fail(CXM_ERR_INVALID, "null buffer");
  if ((rc = ensure_init())) return rc;
  ColArgs ca;
  memset(&ca, 0, sizeof(ca));
  if ((rc = build_hparams(d, g.Ny, g.Nx, ca.hp, kMaxHP))) return rc;
  const int nq = (d->num_wv > 0 ? d->num_wv : 1) * d->num_z;
  const long long n = (long long)g.Ny * g.Nx;
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)nq), block(256);
  hipStream_t s = (hipStream_t)stream;
  float2* o = (float2*)out;
  if (d->kind == CXM_H_TRANSFER) {
    if (d->phase_mode == CXM_PHASE_REF32)
      hipLaunchKernelGGL((materialise_H_kernel<CXM_H_TRANSFER, CXM_PHASE_REF32>), grid, block, 0, s, o, g.Ny, g.Nx, d->num_z, d->remove_evanescent, d->bandlimit, ca);
    else
      hipLaunchKernelGGL((materialise_H_kernel<CXM_H_TRANSFER, CXM_PHASE_ACCURATE>), grid, block, 0, s, o, g.Ny, g.Nx, d->num_z, d->remove_evanescent, d->bandlimit, ca);
  } else {
    if (d->phase_mode == CXM_PHASE_REF32)
      hipLaunchKernelGGL((materialise_H_kernel<CXM_H_ASM, CXM_PHASE_REF32>), grid, block, 0, s, o, g.Ny, g.Nx, d->num_z, d->remove_evanescent, d->bandlimit, ca);
    else
      hipLaunchKernelGGL((materialise_H_kernel<CXM_H_ASM, CXM_PHASE_ACCURATE>), grid, block, 0, s, o, g.Ny, g.Nx, d->num_z, d->remove_evanescent, d->bandlimit, ca);
  }
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}
