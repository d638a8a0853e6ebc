// Batched 2D Fourier-convolution propagation on gfx950.
//
// One propagation of a plane is three passes over HBM:
//   row pass    : zero-padded load -> row FFT                  -> workspace
//   column pass : column FFT -> * H(ky,kx) -> column IFFT, in place
//   row pass    : row IFFT -> crop -> complex or |u|^2         -> output
// H is evaluated on the fly in the column pass, or, when the whole batch
// shares one parameter set, once per call into a table after the chunk
// workspace (the reference likewise builds one propagator and broadcasts it
// over the batch); the column pass then reads the table from L2.
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
int col_variant_width_1024(int v);
int col_variant_width_2048(int v);
int launch_col_variant_640(int v, const ColArgs& a, hipStream_t s);
int col_variant_width_640(int v);

// Launch-shape variant of the column pass: CXM_OPT_COL_VARIANT, or for a
// single plane or a few (the multislice loops) twice the waves of the
// batched default: 1024 points 8 values per thread and 4 columns (20 -> 18
// us per plane), 640 points 10 and 4 (16.7 -> 14.2 us).
int col_variant(int N, int n_planes) {
  int v = option(CXM_OPT_COL_VARIANT);
  if (v == 0 && N == 1024 && n_planes <= 4) v = 2;
  if (v == 0 && N == 640 && n_planes <= 4) v = 1;
  if (N == 1024 && col_variant_width_1024(v) > 0) return v;
  if (N == 2048 && col_variant_width_2048(v) > 0) return v;
  if (N == 640 && col_variant_width_640(v) > 0) return v;
  return 0;
}

// Columns per column-pass workgroup (the strip width C) for this launch.
int col_width(int N, int n_planes) {
  const int v = col_variant(N, n_planes);
  if (v > 0)
    return N == 1024 ? col_variant_width_1024(v)
                     : N == 2048 ? col_variant_width_2048(v) : col_variant_width_640(v);
  switch (N) {
#define CXM_CASE(NN, ER, EC, CC) case NN: return CC;
    CXM_FOR_EACH_SIZE(CXM_CASE)
#undef CXM_CASE
    default: return 0;
  }
}

// Rows per row-pass workgroup.
int row_group(int N) {
  if (option(CXM_OPT_ROW_VARIANT) != 0) return 0;
  switch (N) {
#define CXM_CASE(NN, ER, EC, CC) case NN: return (NN / ER >= 256) ? 1 : 256 / (NN / ER);
    CXM_FOR_EACH_SIZE(CXM_CASE)
#undef CXM_CASE
    default: return 0;
  }
}

int launch_col_prop(int N, const ColArgs& a, hipStream_t s) {
  int rc;
  const int v = col_variant(N, a.n_planes);
  if (v > 0 && N == 1024 && (rc = launch_col_variant_1024(v, a, s)) >= 0) return rc;
  if (v > 0 && N == 2048 && (rc = launch_col_variant_2048(v, a, s)) >= 0) return rc;
  if (v > 0 && N == 640 && (rc = launch_col_variant_640(v, a, s)) >= 0) return rc;
  if ((rc = launch_col_g0(N, a, s)) >= 0) return rc;
  if ((rc = launch_col_g1(N, a, s)) >= 0) return rc;
  if ((rc = launch_col_g2(N, a, s)) >= 0) return rc;
  if ((rc = launch_col_g3(N, a, s)) >= 0) return rc;
  if ((rc = launch_col_g4(N, a, s)) >= 0) return rc;
  return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(N));
}

// H(ky, kx) of one parameter set over the whole Ny x Nx frequency grid, the
// exact values col_prop_kernel would evaluate on the fly (same functions,
// same arguments), NA mask included.  Stored in strips of W columns (the
// column pass's strip width), each strip's rows contiguous: element (ky, kx)
// at (kx / W) * Ny * W + ky * W + kx % W, so that a strip is read as whole
// lines.
template <int KIND, int MODE>
__global__ void __launch_bounds__(256)
h_table_kernel(HParam hp, int Ny, int Nx, int W, int remove_evanescent, int bandlimit, float scale,
               float na_r2, float2* table) {
  // thread i writes table element i (contiguous stores)
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long strip = i / ((long long)Ny * W);
  const int rem = (int)(i - strip * Ny * W);
  const int ky = rem / W;
  const int kx = (int)strip * W + (rem - ky * W);
  if (kx >= Nx) return;
  const XTerms xt = x_terms(hp, kx, Nx, MODE);
  float2 h = eval_H(hp, xt, ky, Ny, KIND, MODE, remove_evanescent, bandlimit, scale);
  if (na_r2 > 0.0f && !na_keep(hp, ky, Ny, na_fx2(hp, kx, Nx), na_r2)) h = make_float2(0.f, 0.f);
  table[i] = h;
}

int launch_h_table(const ColArgs& a, int Ny, int W, float2* table, hipStream_t s) {
  const long long n = (long long)((a.ncols + W - 1) / W) * W * Ny;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  void* t0 = timing_start(s);
  const HParam& hp = a.hp[0];
#define CXM_HT(K, M)                                                                     \
  hipLaunchKernelGGL((h_table_kernel<K, M>), grid, block, 0, s, hp, Ny, a.ncols, W,      \
                     a.remove_evanescent, a.bandlimit, a.scale, a.na_r2, table)
  if (a.kind == CXM_H_TRANSFER) {
    if (a.phase_mode == CXM_PHASE_REF32) CXM_HT(CXM_H_TRANSFER, CXM_PHASE_REF32);
    else CXM_HT(CXM_H_TRANSFER, CXM_PHASE_ACCURATE);
  } else {
    if (a.phase_mode == CXM_PHASE_REF32) CXM_HT(CXM_H_ASM, CXM_PHASE_REF32);
    else CXM_HT(CXM_H_ASM, CXM_PHASE_ACCURATE);
  }
#undef CXM_HT
  CXM_LAUNCH_CHECK();
  if (t0) timing_stop(t0, s, "h_table_" + std::to_string(Ny));
  return CXM_OK;
}

}  // namespace cxm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace cxm;

namespace {
// Elements between the strips of a tiled plane with R rows: the strip plus
// one 128-byte line.  Without the line, strips of 1024 rows lie 32 KiB
// apart, and the 4-row groups a row-pass workgroup moves (one line in each
// strip) all fall on the same HBM channels: measured 1.25-1.4x slower
// passes.
long long tile_strip_stride(int R) { return (long long)R * kTileW + 16; }

struct Geom {
  int Ny, Nx, Hin, Win, Hout, Wout;
  int Rws;                                   // rows per workspace plane: max(Hin, Hout)
  long long plane_in, plane_ws;              // elements
  int chunk;                                 // input planes per chunk
  int W;                                     // column-pass strip width
  bool tiled;                                // workspace planes in strips of kTileW columns
  bool generic;                              // a grid length without a fused FFT: chirp-z passes
  long long gen_plane;                       // generic path: workspace elements per input plane
  size_t gen_czt_ws;                         // generic path: bytes of the chirp-z workspace
};

bool use_h_table(const cxm_prop2d_desc* d);
bool use_persist(const cxm_prop2d_desc* d);
size_t table_elems(const Geom& g);
int generic_geom(const cxm_prop2d_desc* d, Geom& g, size_t limit_bytes);

// limit_bytes > 0: the chunk is the largest whose workspace (planes and the
// H table) fits in limit_bytes; CXM_ERR_UNSUPPORTED when one plane does not.
int validate(const cxm_prop2d_desc* d, Geom& g, size_t limit_bytes = 0) {
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
  g.generic = !fft_len_supported(g.Ny) || !fft_len_supported(g.Nx);
  g.plane_in = (long long)g.Hin * g.Win;
  if (g.generic) return generic_geom(d, g, limit_bytes);
  // Every workspace plane (column-pass input and output alike) has
  // max(Hin, Hout) rows, so that with one distance the column pass works in
  // place: a strip's output rows lie where its input rows were.
  g.Rws = std::max(g.Hin, g.Hout);
  g.plane_ws = (long long)g.Rws * g.Nx;
  // Chunk so that the workspace of one chunk stays <= 128 MiB, half the 256 MiB
  // Infinity Cache: the intermediate planes are re-read from it, not from HBM.
  // With the persistent 1024-point column pass, 256 MiB (32 planes): its
  // workgroups then hold each strip for twice the planes, so H per strip is
  // evaluated half as often per plane, and the passes have half the tails
  // (64 planes: 0.584 ms per step against 0.610 at 16 planes per chunk and
  // 0.618 / 0.623 at 48 / 64).
  const int nout = d->z_zip ? 1 : d->num_z;
  const long long per_plane = (nout == 1 ? 1 : 1 + nout) * g.plane_ws;
  const bool big_chunk = use_persist(d) && d->height + 2 * d->pad_y == 1024;
  const long long budget = ((big_chunk ? 256ll : 128ll) << 20) / 8;
  long long ch = std::max(1ll, budget / std::max(1ll, per_plane));
  if (option(CXM_OPT_CHUNK_PLANES) > 0) ch = option(CXM_OPT_CHUNK_PLANES);
  const long long plane_rm = g.plane_ws;  // row-major workspace plane
  for (ch = std::min<long long>(ch, std::max(1, d->batch)); ch >= 1; --ch) {
    g.chunk = (int)ch;
    g.W = col_width(g.Ny, g.chunk);
    // Tiled intermediates: the row pass moves 4 rows per workgroup, the column
    // pass reads strips of 4 columns, and planes are whole 4-row groups.
    g.tiled = option(CXM_OPT_TILES) != 1 && row_group(g.Nx) == kTileW && g.W == kTileW &&
              g.Hin % kTileW == 0 && g.Hout % kTileW == 0 && g.Nx % kTileW == 0;
    g.plane_ws = g.tiled ? (long long)(g.Nx / kTileW) * tile_strip_stride(g.Rws) : plane_rm;
    if (limit_bytes == 0) return CXM_OK;
    const size_t need = ((size_t)g.chunk * (nout == 1 ? 1 : 1 + nout) * (size_t)g.plane_ws +
                         (use_h_table(d) ? table_elems(g) : 0)) * sizeof(float2);
    if (need <= limit_bytes) return CXM_OK;
  }
  return fail(CXM_ERR_UNSUPPORTED, "one plane of a " + std::to_string(g.Ny) + "x" +
                                       std::to_string(g.Nx) + " grid needs more workspace than the " +
                                       std::to_string(limit_bytes >> 20) + " MiB static scratch");
}

int out_per_plane(const cxm_prop2d_desc* d) { return d->z_zip ? 1 : d->num_z; }

// Analytic kind, a single parameter set, and enough planes to share H.
bool one_param_set(const cxm_prop2d_desc* d) {
  if (d->kind != CXM_H_TRANSFER && d->kind != CXM_H_ASM) return false;
  return (d->num_wv > 0 ? d->num_wv : 1) * d->num_z == 1;
}

// The persistent column pass (H per strip in registers) for such a call,
// where the column length has one and no launch-shape option is set.
bool fused_grid(const cxm_prop2d_desc* d) {
  return fft_len_supported(d->height + 2 * d->pad_y) && fft_len_supported(d->width + 2 * d->pad_x);
}

bool use_persist(const cxm_prop2d_desc* d) {
  if (!fused_grid(d) || !one_param_set(d) || d->batch < CXM_H_TABLE_MIN_PLANES) return false;
  if (option(CXM_OPT_COL_PERSIST) == 1 || option(CXM_OPT_H_TABLE) != 0 ||
      option(CXM_OPT_COL_VARIANT) != 0 || option(CXM_OPT_COL_ORDER) != 0)
    return false;
  return col_persist_available(d->height + 2 * d->pad_y);
}

// Otherwise one H table for the batch, evaluated once per call and read by
// the per-plane column pass (CXM_OPT_H_TABLE overrides).
bool use_h_table(const cxm_prop2d_desc* d) {
  if (!fused_grid(d) || !one_param_set(d) || use_persist(d)) return false;
  const int o = option(CXM_OPT_H_TABLE);
  if (o == 1) return false;
  return o == 2 || d->batch >= CXM_H_TABLE_MIN_PLANES;
}

size_t chunk_elems(const cxm_prop2d_desc* d, const Geom& g) {
  const int nout = out_per_plane(d);
  return (size_t)g.chunk * (nout == 1 ? 1 : 1 + nout) * (size_t)g.plane_ws;
}

size_t table_elems(const Geom& g) {  // whole strips of W columns
  return (size_t)((g.Nx + g.W - 1) / g.W) * g.W * g.Ny;
}

size_t ws_elems(const cxm_prop2d_desc* d, const Geom& g, bool table) {
  if (g.generic) return (size_t)g.chunk * g.gen_plane + (g.gen_czt_ws + 7) / 8;
  return chunk_elems(d, g) + (table ? table_elems(g) : 0);
}
}  // namespace

extern "C" size_t cxm_prop2d_workspace_size(const cxm_prop2d_desc* d) {
  Geom g;
  if (validate(d, g) != CXM_OK) return 0;
  return ws_elems(d, g, use_h_table(d)) * sizeof(float2);
}

// ---------------------------------------------------------------------------
// Grids with a length that has no fused FFT (any length the reference's
// jnp.fft accepts, e.g. the 4604- or 1856-point grids of its padding
// calculators, propagation.py:583-656).  The 2D DFT pair runs as chirp-z
// passes along each axis (cxm_czt_axis_c64, Bluestein with FP64 chirps):
//   forward, rows then columns:  X_k = sum_j x_j w^(j k), w = e^(-2 pi i / N),
//     over the unpadded input samples only (the zero padding adds nothing);
//   H(ky, kx) evaluated per bin (the functions of the fused column pass);
//   inverse, columns then rows:  y_j = sum_k Z_k a^(-k) w^(-k j), j < N - 2 c,
//     a = e^(-2 pi i (c - p) / N): the pad offset p of the forward sum and the
//     crop offset c of the output, both a factor w^(+-k .) per bin, meet in a.
// The 1/(Ny Nx) of the inverse is part of H's scale, as in the fused passes.
// Lengths are limited by the chirp-z FFT: n + N - 1 per axis must fit the
// largest supported transform (10240).
// ---------------------------------------------------------------------------
namespace cxm {
template <int KIND, int MODE>
__global__ void __launch_bounds__(256)
apply_H_kernel(const float2* __restrict__ in, float2* __restrict__ out, long long plane, int Ny,
               int Nx, int iz, ColArgs a) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= plane) return;
  const int ky = (int)(i / Nx), kx = (int)(i - (long long)ky * Nx);
  const float2 v = in[p * plane + i];
  float2 h;
  if constexpr (KIND == CXM_H_TRANSFER || KIND == CXM_H_ASM) {
    const int pg = a.plane0 + p;
    const int w = (a.nwv > 1) ? (pg % a.nwv) : 0;
    const int zi = a.z_zip ? (pg / a.nwv) % a.nz : iz;
    const HParam& hp = a.hp[w * a.nz + zi];
    h = eval_H(hp, x_terms(hp, kx, Nx, MODE), ky, Ny, KIND, MODE, a.remove_evanescent,
               a.bandlimit, a.scale);
  } else if constexpr (KIND == CXM_H_ARRAY) {
    h = cscale(a.H[i], a.scale);
  } else {
    h = make_float2(a.scale, 0.f);
  }
  out[p * plane + i] = cmul(v, h);
}
}  // namespace cxm

namespace {
// Chirp-z passes of one chunk: 0 rows forward, 1 columns forward,
// 2 columns inverse, 3 rows inverse.
struct GenCzt {
  double a[4][2], w[4][2];
  cxm_czt_desc c[4];
};

void gen_czt(const cxm_prop2d_desc* d, const Geom& g, int kc, GenCzt& z) {
  const double tau = 2.0 * M_PI;
  auto set = [&](int i, int h, int w, int axis, int m, double a_ang, double w_ang) {
    z.a[i][0] = std::cos(a_ang);
    z.a[i][1] = std::sin(a_ang);
    z.w[i][0] = std::cos(w_ang);
    z.w[i][1] = std::sin(w_ang);
    memset(&z.c[i], 0, sizeof(z.c[i]));
    z.c[i].batch = kc;
    z.c[i].height = h;
    z.c[i].width = w;
    z.c[i].axis = axis;
    z.c[i].m = m;
    z.c[i].num_param = 1;
    z.c[i].mod_n = -1;
    z.c[i].a = z.a[i];
    z.c[i].w = z.w[i];
    z.c[i].scale = 1.0;
  };
  set(0, g.Hin, g.Win, 1, g.Nx, 0.0, -tau / g.Nx);
  set(1, g.Hin, g.Nx, 0, g.Ny, 0.0, -tau / g.Ny);
  set(2, g.Ny, g.Nx, 0, g.Hout, -tau * (d->crop_y - d->pad_y) / g.Ny, tau / g.Ny);
  set(3, g.Hout, g.Nx, 1, g.Wout, -tau * (d->crop_x - d->pad_x) / g.Nx, tau / g.Nx);
}

int generic_geom(const cxm_prop2d_desc* d, Geom& g, size_t limit_bytes) {
  if (limit_bytes)
    return fail(CXM_ERR_UNSUPPORTED, "FFT grid " + std::to_string(g.Ny) + "x" +
                                         std::to_string(g.Nx) +
                                         " has no fused transform; the allocation-free entry "
                                         "covers fused grids only");
  const int nout = d->z_zip ? 1 : d->num_z;
  const bool inten = d->out_kind == CXM_OUT_INTENSITY;
  g.gen_plane = (long long)g.Hin * g.Nx + (long long)g.Ny * g.Nx * (nout > 1 ? 2 : 1) +
                (long long)g.Hout * g.Nx + (inten ? (long long)g.Hout * g.Wout : 0);
  const long long budget = (256ll << 20) / 8;
  g.chunk = nout > 1 ? 1
                     : (int)std::max(1ll, std::min<long long>(std::max(1, d->batch),
                                                              budget / g.gen_plane));
  g.W = 0;
  g.tiled = false;
  g.plane_ws = 0;
  GenCzt z;
  gen_czt(d, g, g.chunk, z);
  g.gen_czt_ws = 0;
  for (int i = 0; i < 4; ++i) {
    const size_t b = cxm_czt_axis_workspace_size(&z.c[i]);
    if (b == 0)
      return fail(CXM_ERR_UNSUPPORTED,
                  "FFT grid " + std::to_string(g.Ny) + "x" + std::to_string(g.Nx) +
                      ": the chirp-z passes of a grid without a fused transform need "
                      "n + N - 1 <= 10240 per axis");
    g.gen_czt_ws = std::max(g.gen_czt_ws, b);
  }
  return CXM_OK;
}

int launch_apply_H(const float2* in, float2* out, int kc, const Geom& g, int iz, const ColArgs& a,
                   hipStream_t s) {
  const long long plane = (long long)g.Ny * g.Nx;
  const dim3 grid((unsigned)((plane + 255) / 256), (unsigned)kc), block(256);
#define CXM_AH(K, M) hipLaunchKernelGGL((apply_H_kernel<K, M>), grid, block, 0, s, in, out, plane, g.Ny, g.Nx, iz, a)
  if (a.kind == CXM_H_TRANSFER) {
    if (a.phase_mode == CXM_PHASE_REF32) CXM_AH(CXM_H_TRANSFER, CXM_PHASE_REF32);
    else CXM_AH(CXM_H_TRANSFER, CXM_PHASE_ACCURATE);
  } else if (a.kind == CXM_H_ASM) {
    if (a.phase_mode == CXM_PHASE_REF32) CXM_AH(CXM_H_ASM, CXM_PHASE_REF32);
    else CXM_AH(CXM_H_ASM, CXM_PHASE_ACCURATE);
  } else if (a.kind == CXM_H_ARRAY) {
    CXM_AH(CXM_H_ARRAY, 0);
  } else {
    CXM_AH(CXM_H_NONE, 0);
  }
#undef CXM_AH
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

int generic_impl(void* stream, const void* in, void* out, void* workspace,
                 const cxm_prop2d_desc* d, const Geom& g) {
  hipStream_t s = (hipStream_t)stream;
  int rc;
  ColArgs ca;
  memset(&ca, 0, sizeof(ca));
  if (d->kind == CXM_H_TRANSFER || d->kind == CXM_H_ASM) {
    if ((rc = build_hparams(d, g.Ny, g.Nx, ca.hp, kMaxHP))) return rc;
    ca.nhp = (d->num_wv > 0 ? d->num_wv : 1) * d->num_z;
  }
  ca.H = (const float2*)d->propagator;
  ca.nz = d->num_z;
  ca.z_zip = d->z_zip;
  ca.nwv = d->num_wv > 0 ? d->num_wv : 1;
  ca.kind = d->kind;
  ca.phase_mode = d->phase_mode;
  ca.remove_evanescent = d->remove_evanescent;
  ca.bandlimit = d->bandlimit;
  ca.scale = (float)(1.0 / ((double)g.Ny * (double)g.Nx));
  const int nout = out_per_plane(d);
  const bool inten = d->out_kind == CXM_OUT_INTENSITY;
  const long long plane_f = (long long)g.Ny * g.Nx;
  const long long out_plane = (long long)g.Hout * g.Wout;
  float2* A1 = (float2*)workspace;
  float2* S = A1 + (size_t)g.chunk * g.Hin * g.Nx;
  float2* T = S + (size_t)g.chunk * plane_f;
  float2* U = nout > 1 ? T + (size_t)g.chunk * plane_f : T;
  float2* V = U + (size_t)g.chunk * g.Hout * g.Nx;
  void* cws = (void*)(A1 + (size_t)g.chunk * g.gen_plane);
  for (int p0 = 0; p0 < d->batch; p0 += g.chunk) {
    const int kc = std::min(g.chunk, d->batch - p0);
    GenCzt z;
    gen_czt(d, g, kc, z);
    const float2* src = (const float2*)in + (size_t)p0 * g.plane_in;
    if ((rc = cxm_czt_axis_c64(s, src, A1, cws, g.gen_czt_ws, &z.c[0]))) return rc;
    if ((rc = cxm_czt_axis_c64(s, A1, S, cws, g.gen_czt_ws, &z.c[1]))) return rc;
    ca.plane0 = p0;
    for (int iz = 0; iz < nout; ++iz) {
      float2* Z = nout > 1 ? T : S;
      if ((rc = launch_apply_H(S, Z, kc, g, iz, ca, s))) return rc;
      if ((rc = cxm_czt_axis_c64(s, Z, U, cws, g.gen_czt_ws, &z.c[2]))) return rc;
      const size_t o = ((size_t)p0 * nout + iz) * out_plane;
      if (inten) {
        if ((rc = cxm_czt_axis_c64(s, U, V, cws, g.gen_czt_ws, &z.c[3]))) return rc;
        if ((rc = cxm_intensity_c64(s, V, (float*)out + o, (int64_t)kc * out_plane, 1, nullptr)))
          return rc;
      } else if ((rc = cxm_czt_axis_c64(s, U, (float2*)out + o, cws, g.gen_czt_ws, &z.c[3]))) {
        return rc;
      }
    }
  }
  return CXM_OK;
}
}  // namespace

namespace {
// noalloc: the workspace is the library's static scratch array (workspace
// and workspace_bytes are ignored) and the chunk is sized to fit it.
int prop2d_impl(void* stream, const void* in, void* out, void* workspace, size_t workspace_bytes,
                const cxm_prop2d_desc* d, const HParam* hp_dev, bool noalloc = false) {
  Geom g;
  int rc = validate(d, g, noalloc ? kNoallocScratchBytes : 0);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const bool table = !hp_dev && use_h_table(d);
  const bool persist = !hp_dev && use_persist(d);
  const size_t need = ws_elems(d, g, table) * sizeof(float2);
  if (!noalloc && (!workspace || workspace_bytes < need))
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if (g.generic) {
    if (noalloc)
      return fail(CXM_ERR_UNSUPPORTED, "grid lengths without a fused FFT are not available "
                                       "without a workspace");
    if (hp_dev) return fail(CXM_ERR_UNSUPPORTED, "a parameter table needs a fused FFT grid");
    if (!workspace || workspace_bytes < need)
      return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
    if ((rc = ensure_init())) return rc;
    return generic_impl(stream, in, out, workspace, d, g);
  }
  if ((rc = ensure_init())) return rc;
  if (noalloc) {
    workspace = noalloc_scratch();
    if (!workspace || need > kNoallocScratchBytes)
      return fail(CXM_ERR_WORKSPACE, "static scratch unavailable or too small");
  }
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
  float2* ws2 = (nz == 1) ? ws1 : ws1 + (size_t)g.chunk * g.plane_ws;
  const long long tstride = tile_strip_stride(g.Rws);  // elements between tiled strips
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
    r1.out_plane_stride = g.plane_ws;
    r1.out_tiled = g.tiled;
    r1.out_strip_stride = tstride;
    r1.in_row_stride = g.Win;
    r1.out_row_stride = g.Nx;
    r1.n_rows = kc * g.Hin;
    r1.rows_per_plane = g.Hin;
    r1.in_off = d->pad_x;
    r1.in_len = g.Win;
    r1.out_off = 0;
    r1.out_len = g.Nx;
    r1.scale = 1.0f;
    r1.nt_in = (option(CXM_OPT_NT_IO) == 0 || option(CXM_OPT_NT_IO) == 3);
    if ((rc = launch_row(g.Nx, -1, r1, s))) return rc;
    // 2) columns: FFT * H * IFFT (in place when num_z == 1)
    ca.in = ws1;
    ca.out = ws2;
    ca.tw = twiddles_for(g.Ny);
    ca.H = (const float2*)d->propagator;
    ca.in_plane_stride = g.plane_ws;
    ca.out_plane_stride = g.plane_ws;
    ca.row_stride = g.tiled ? kTileW : g.Nx;
    ca.strip_stride = g.tiled ? tstride : 0;
    if (col_width(g.Ny, kc) != g.W) return fail(CXM_ERR_INVALID, "internal: column strip width changed");
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
    if (table) {
      // the scale is part of the tabulated H; the column pass multiplies by
      // the table and by 1 (exact)
      float2* htab = ws1 + chunk_elems(d, g);
      if (p0 == 0 && (rc = launch_h_table(ca, g.Ny, g.W, htab, s))) return rc;
      ca.kind = CXM_H_ARRAY;
      ca.H = htab;
      ca.h_strip_stride = (long long)g.Ny * g.W;
      ca.h_row_stride = g.W;
      ca.scale = 1.0f;
    }
    ca.strip_major = option(CXM_OPT_COL_ORDER) == 1;
    if (persist) {
      if ((rc = launch_col_persist(g.Ny, ca, s))) return rc;
    } else if ((rc = launch_col_prop(g.Ny, ca, s))) {
      return rc;
    }
    // 3) rows: inverse FFT, crop, complex or intensity
    RowArgs r2;
    memset(&r2, 0, sizeof(r2));
    r2.in = ws2;
    r2.tw = twiddles_for(g.Nx);
    r2.in_plane_stride = g.plane_ws;
    r2.in_tiled = g.tiled;
    r2.in_strip_stride = tstride;
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
    r2.nt_out = (option(CXM_OPT_NT_IO) == 0 || option(CXM_OPT_NT_IO) == 2);
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

// The same passes with the workspace in the library's static scratch array:
// no allocation by the caller or the library.  Calls that overlap on one GPU
// share the array and race.
extern "C" int cxm_prop2d_c64_noalloc(void* stream, const void* in, void* out,
                                      const cxm_prop2d_desc* d) {
  return prop2d_impl(stream, in, out, nullptr, 0, d, nullptr, true);
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
