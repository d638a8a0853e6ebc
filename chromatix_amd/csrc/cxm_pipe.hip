// Compositions of pipeline passes: the lens/far-field FFT family
// (optical_fft, ff_lens, transform_propagate, shifted fft2/ifft2) and the
// scalable angular spectrum propagation.  Host code only: it computes the
// per-plane operator parameters (in the reference's float32 order for ref32,
// in FP64 for the accurate mode) and chains row and column passes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "cxm_pipe_host.h"

namespace cxm {

// Per-wavelength sampling of a descriptor.
struct Wv {
  double lam, dy, dx, oy, ox;
};
static Wv wv_of(const cxm_optfft_desc* d, int w) {
  Wv r;
  r.lam = d->wavelengths ? d->wavelengths[w] : d->wavelength;
  r.dy = d->spacing ? d->spacing[2 * w] : d->dy;
  r.dx = d->spacing ? d->spacing[2 * w + 1] : d->dx;
  r.oy = d->origin ? d->origin[2 * w] : d->oy;
  r.ox = d->origin ? d->origin[2 * w + 1] : d->ox;
  return r;
}

// Optical-FFT constants of one wavelength on an (Ny, Nx) grid
// (convenience.py:33-54, propagation.py:71-81).
struct OptConsts {
  float2 norm;        // multiplies the (possibly conjugated) input
  float chirp_a;      // ref32: fl(pi / L^2)
  double chirp_turns; // accurate: 1 / (2 L^2)  (phase pi |x|^2 / L^2 in turns)
  double du_y, du_x;  // output spacing
  float du_y32, du_x32;
};

static OptConsts opt_consts(const Wv& w, double z, double n, int Ny, int Nx, int mode) {
  OptConsts c;
  const double L2 = w.lam * z / n;
  c.du_y = std::fabs(L2) / (Ny * w.dy);
  c.du_x = std::fabs(L2) / (Nx * w.dx);
  c.chirp_turns = 1.0 / (2.0 * L2);
  // float32 chain of the reference
  const float lam = (float)w.lam, zf = (float)z, nf = (float)n;
  const float L2f = (lam * zf) / nf;
  const float dyf = (float)w.dy, dxf = (float)w.dx;
  const float dfy = 1.0f / (dyf * (float)Ny), dfx = 1.0f / (dxf * (float)Nx);
  c.du_y32 = dfy * std::fabs(L2f);
  c.du_x32 = dfx * std::fabs(L2f);
  c.chirp_a = (float)M_PI / L2f;
  if (mode == CXM_PHASE_REF32) {
    if (z >= 0) {
      c.norm = make_float2(0.0f, -((dyf * dxf) / L2f));
    } else {
      const float pdu = c.du_y32 * c.du_x32;
      c.norm = make_float2(0.0f, -((L2f / pdu) / (float)((double)Ny * Nx)));
    }
  } else {
    if (z >= 0) c.norm = make_float2(0.0f, (float)(-(w.dy * w.dx) / L2));
    else c.norm = make_float2(0.0f, (float)(-(L2 / (c.du_y * c.du_x)) / ((double)Ny * Nx)));
  }
  return c;
}

// Fill a CHIRP operator slot: phase pi |g|^2 / L^2 on a grid of spacing (sy, sx).
static void chirp_slot(Ops& B, OpDesc* o, int slot, const OptConsts& c, double sy, double sx,
                       float sy32, float sx32, const Wv& w, int mode) {
  float4& p0 = B.p(o, slot, 0);
  float4& p1 = B.p(o, slot, 1);
  float4& p2 = B.p(o, slot, 2);
  if (mode == CXM_PHASE_REF32) {
    p0 = make_float4(c.chirp_a, 0.0f, 0.0f, 0.0f);
    p1 = make_float4(sy32, sx32, (float)w.oy, (float)w.ox);
  } else {
    float hi, lo;
    split(c.chirp_turns, hi, lo);
    p0 = make_float4(hi, 0.0f, lo, 0.0f);
    p1 = make_float4((float)sy, (float)sx, (float)w.oy, (float)w.ox);
  }
  p2 = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---- common validation -----------------------------------------------------------
static int check_common(const cxm_optfft_desc* d) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->batch < 0 || d->height <= 0 || d->width <= 0 || d->pad_y < 0 || d->pad_x < 0 ||
      d->crop_y < 0 || d->crop_x < 0)
    return fail(CXM_ERR_INVALID, "invalid shape parameters");
  const int nwv = d->num_wv > 0 ? d->num_wv : 1;
  if (nwv > 32) return fail(CXM_ERR_INVALID, "at most 32 wavelengths");
  if (d->phase_mode != CXM_PHASE_REF32 && d->phase_mode != CXM_PHASE_ACCURATE)
    return fail(CXM_ERR_INVALID, "invalid phase mode");
  if (d->out_kind != CXM_OUT_COMPLEX && d->out_kind != CXM_OUT_INTENSITY)
    return fail(CXM_ERR_INVALID, "invalid output kind");
  if (d->mode < CXM_OFFT_OPTICAL || d->mode > CXM_OFFT_RAW_INV)
    return fail(CXM_ERR_INVALID, "invalid optical-FFT mode");
  for (int w = 0; w < nwv; ++w) {
    const Wv v = wv_of(d, w);
    if (!(v.dy > 0 && v.dx > 0)) return fail(CXM_ERR_INVALID, "spacing must be positive");
    if (d->mode <= CXM_OFFT_TRANSFORM && !(v.lam > 0))
      return fail(CXM_ERR_INVALID, "wavelength must be positive");
  }
  if (d->mode <= CXM_OFFT_TRANSFORM && (d->z == 0 || !(d->n > 0)))
    return fail(CXM_ERR_INVALID, "z must be non-zero and n positive");
  return CXM_OK;
}

static long long chunk_planes(long long per_plane_elems, int batch) {
  const long long budget = (96ll << 20) / 8;
  return std::max(1ll, std::min<long long>(std::max(1, batch), budget / std::max(1ll, per_plane_elems)));
}

}  // namespace cxm

using namespace cxm;

// ---------------------------------------------------------------------------
// optical FFT family
// ---------------------------------------------------------------------------
namespace {
struct OptGeom { int Ny, Nx, Hout, Wout; long long chunk; };

int opt_geom(const cxm_optfft_desc* d, OptGeom& g) {
  int rc = check_common(d);
  if (rc) return rc;
  g.Ny = d->height + 2 * d->pad_y;
  g.Nx = d->width + 2 * d->pad_x;
  g.Hout = g.Ny - 2 * d->crop_y;
  g.Wout = g.Nx - 2 * d->crop_x;
  if (g.Hout <= 0 || g.Wout <= 0) return fail(CXM_ERR_INVALID, "crop removes the whole plane");
  if (!fft_len_supported(g.Ny) || !fft_len_supported(g.Nx))
    return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT grid " + std::to_string(g.Ny) + "x" +
                                         std::to_string(g.Nx));
  g.chunk = chunk_planes((long long)d->height * g.Nx, d->batch);
  return CXM_OK;
}
}  // namespace

// The pass geometry of a call: the descriptor itself, or for the adjoint
// (input = the forward output, output = the forward input) the descriptor
// with heights, widths, pads and crops exchanged.
cxm_optfft_desc pass_desc(const cxm_optfft_desc* d) {
  cxm_optfft_desc e = *d;
  if (d->adjoint) {
    e.height = d->height + 2 * d->pad_y - 2 * d->crop_y;
    e.width = d->width + 2 * d->pad_x - 2 * d->crop_x;
    e.pad_y = d->crop_y;
    e.pad_x = d->crop_x;
    e.crop_y = d->pad_y;
    e.crop_x = d->pad_x;
  }
  return e;
}

extern "C" size_t cxm_optical_fft_workspace_size(const cxm_optfft_desc* d) {
  OptGeom g;
  if (!d) return 0;
  const cxm_optfft_desc e = pass_desc(d);
  if (opt_geom(&e, g)) return 0;
  return (size_t)g.chunk * e.height * g.Nx * sizeof(float2);
}

extern "C" int cxm_optical_fft_c64(void* stream, const void* in, void* out, void* workspace,
                                   size_t workspace_bytes, const cxm_optfft_desc* d0) {
  OptGeom g;
  if (!d0) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d0->adjoint != 0 && d0->adjoint != 1) return fail(CXM_ERR_INVALID, "adjoint must be 0 or 1");
  if (d0->adjoint && d0->out_kind != CXM_OUT_COMPLEX)
    return fail(CXM_ERR_INVALID, "the adjoint has a complex output");
  const cxm_optfft_desc e = pass_desc(d0);
  const cxm_optfft_desc* d = &e;
  const bool adj = d0->adjoint != 0;
  int rc = opt_geom(d, g);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const size_t need = (size_t)g.chunk * d->height * g.Nx * sizeof(float2);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int nwv = d->num_wv > 0 ? d->num_wv : 1;
  const int mode = d->phase_mode;
  const bool optical = d->mode == CXM_OFFT_OPTICAL || d->mode == CXM_OFFT_TRANSFORM;
  const bool transform = d->mode == CXM_OFFT_TRANSFORM;
  const bool neg = optical && d->z < 0;
  const int shift = optical ? 1 : (d->shift ? 1 : 0);

  PipeArgs ra, ca;
  memset(&ra, 0, sizeof(ra));
  memset(&ca, 0, sizeof(ca));
  Ops rb(ra), cb(ca);
  if (optical && adj) {
    // adjoint chain: conj(Q2) and sigma on load, the DFT of the other
    // direction (unnormalised), conj(Q1) and the pupil on store
    OpDesc *q2 = nullptr, *sc = nullptr, *q1 = nullptr, *pup = nullptr;
    if (transform && !d->skip_final_phase) q2 = rb.add(ra.pre, OP_CHIRP, nwv, 3, mode);
    sc = rb.add(ra.pre, OP_SCALE, nwv, 1);
    if (transform && !d->skip_initial_phase) q1 = cb.add(ca.post, OP_CHIRP, nwv, 3, mode);
    if (d->mode == CXM_OFFT_OPTICAL && d->pupil) pup = cb.add(ca.post, OP_CIRC, nwv, 3);
    if (!sc || (transform && !d->skip_final_phase && !q2) ||
        (transform && !d->skip_initial_phase && !q1) || (d->mode == CXM_OFFT_OPTICAL && d->pupil && !pup))
      return fail(CXM_ERR_INVALID, "operator table overflow");
    for (int w = 0; w < nwv; ++w) {
      const Wv v = wv_of(d, w);
      const OptConsts c = opt_consts(v, d->z, d->n, g.Ny, g.Nx, mode);
      // sigma = conj(c) (z >= 0) or c (z < 0), times -i for the transform
      float2 sg = neg ? c.norm : make_float2(c.norm.x, -c.norm.y);
      if (transform) sg = make_float2(sg.y, -sg.x);
      rb.p(sc, w, 0) = make_float4(sg.x, sg.y, 0.f, 0.f);
      if (q2) {
        chirp_slot(rb, q2, w, c, c.du_y, c.du_x, c.du_y32, c.du_x32, v, mode);
        float4& p0 = rb.p(q2, w, 0);
        p0.x = -p0.x;
        p0.z = -p0.z;
      }
      if (q1) {
        chirp_slot(cb, q1, w, c, v.dy, v.dx, (float)v.dy, (float)v.dx, v, mode);
        float4& p0 = cb.p(q1, w, 0);
        p0.x = -p0.x;
        p0.z = -p0.z;
      }
      if (pup) {
        const float r = ((float)d->central_wavelength / (float)v.lam) * (float)d->pupil_d / 2.0f;
        cb.p(pup, w, 0) = make_float4(r * r, 0.f, 0.f, 0.f);
        cb.p(pup, w, 1) = make_float4((float)v.dy, (float)v.dx, (float)v.oy, (float)v.ox);
        cb.p(pup, w, 2) = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  } else if (optical) {
    OpDesc *pup = nullptr, *q1 = nullptr, *sc = nullptr;
    if (d->mode == CXM_OFFT_OPTICAL && d->pupil) pup = rb.add(ra.pre, OP_CIRC, nwv, 3);
    if (transform && !d->skip_initial_phase) q1 = rb.add(ra.pre, OP_CHIRP, nwv, 3, mode);
    if (neg) rb.add(ra.pre, OP_CONJ, 1, 0);
    sc = rb.add(ra.pre, OP_SCALE, nwv, 1);
    OpDesc *q2 = nullptr, *ij = nullptr;
    if (neg) cb.add(ca.post, OP_CONJ, 1, 0);
    if (transform) ij = cb.add(ca.post, OP_SCALE, 1, 1);
    if (transform && !d->skip_final_phase) q2 = cb.add(ca.post, OP_CHIRP, nwv, 3, mode);
    if (!sc || (transform && !ij) || (d->mode == CXM_OFFT_OPTICAL && d->pupil && !pup))
      return fail(CXM_ERR_INVALID, "operator table overflow");
    if (ij) cb.p(ij, 0, 0) = make_float4(0.f, 1.f, 0.f, 0.f);
    for (int w = 0; w < nwv; ++w) {
      const Wv v = wv_of(d, w);
      const OptConsts c = opt_consts(v, d->z, d->n, g.Ny, g.Nx, mode);
      rb.p(sc, w, 0) = make_float4(c.norm.x, c.norm.y, 0.f, 0.f);
      if (pup) {
        // circular_pupil (pupils.py:17-23): |g|^2 <= ((lam0/lam) * D / 2)^2 in float32
        const float r = ((float)d->central_wavelength / (float)v.lam) * (float)d->pupil_d / 2.0f;
        rb.p(pup, w, 0) = make_float4(r * r, 0.f, 0.f, 0.f);
        rb.p(pup, w, 1) = make_float4((float)v.dy, (float)v.dx, (float)v.oy, (float)v.ox);
        rb.p(pup, w, 2) = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      if (q1) chirp_slot(rb, q1, w, c, v.dy, v.dx, (float)v.dy, (float)v.dx, v, mode);
      if (q2) chirp_slot(cb, q2, w, c, c.du_y, c.du_x, c.du_y32, c.du_x32, v, mode);
    }
  }
  if (d->spec_mul) {
    // the spectrum multiply on the column pass's store
    if (optical || adj || d->out_kind != CXM_OUT_COMPLEX)
      return fail(CXM_ERR_INVALID, "spec_mul applies to the raw transforms with complex output");
    OpDesc* m = cb.add(ca.post, OP_CARRAY, 1, 0, shift);
    if (!m) return fail(CXM_ERR_INVALID, "operator table overflow");
    m->arr0 = d->spec_mul;
    m->arr_plane_stride = 0;
    m->arr_row_stride = g.Nx;
    m->arr_rows = g.Ny;
    m->arr_cols = g.Nx;
  }
  // transform direction: the forward DFT, or for the raw inverse and the
  // adjoint of the forward modes (z >= 0) the inverse one; the adjoint of
  // the raw inverse is the forward DFT scaled by 1/N, that of the raw
  // forward the inverse DFT without 1/N
  const bool inv = adj ? ((optical && !neg) || d->mode == CXM_OFFT_RAW_FWD)
                       : d->mode == CXM_OFFT_RAW_INV;
  const bool scale_n = adj ? d->mode == CXM_OFFT_RAW_INV : d->mode == CXM_OFFT_RAW_INV;
  // row pass
  ra.tw = twiddles_for(g.Nx);
  ra.in_plane_stride = (long long)d->height * d->width;
  ra.in_line_stride = d->width;
  ra.out_plane_stride = (long long)d->height * g.Nx;
  ra.out_line_stride = g.Nx;
  ra.n_lines = d->height;
  ra.grid_ny = g.Ny;
  ra.grid_nx = g.Nx;
  ra.line_off = d->pad_y;
  ra.in_off = d->pad_x;
  ra.in_len = d->width;
  ra.in_shift = shift;
  ra.out_off = 0;
  ra.out_len = g.Nx;
  ra.seq = inv ? SEQ_I : SEQ_F;
  ra.scale = 1.0f;
  ra.mid_scale = 1.0f;
  // column pass
  ca.tw = twiddles_for(g.Ny);
  ca.in_plane_stride = ra.out_plane_stride;
  ca.in_line_stride = g.Nx;
  ca.out_plane_stride = (long long)g.Hout * g.Wout;
  ca.out_line_stride = g.Wout;
  ca.n_lines = g.Nx;
  ca.grid_ny = g.Ny;
  ca.grid_nx = g.Nx;
  ca.line_off = 0;
  ca.in_off = d->pad_y;
  ca.in_len = d->height;
  ca.in_shift = shift;
  ca.out_shift = shift;
  ca.out_line_shift = shift;
  ca.out_off = d->crop_y;
  ca.out_len = g.Hout;
  ca.out_line_off = d->crop_x;
  ca.out_line_len = g.Wout;
  ca.seq = inv ? SEQ_I : SEQ_F;
  ca.intensity = d->out_kind == CXM_OUT_INTENSITY;
  ca.scale = scale_n ? (float)(1.0 / ((double)g.Ny * g.Nx)) : 1.0f;
  ca.mid_scale = 1.0f;
  // plane parameters cycle over wavelengths
  for (OpList* L : {&ra.pre, &ra.mid, &ra.post})
    for (int i = 0; i < L->n; ++i) if (L->op[i].pmod > 1) L->op[i].pmod = nwv;
  for (OpList* L : {&ca.pre, &ca.mid, &ca.post})
    for (int i = 0; i < L->n; ++i) if (L->op[i].pmod > 1) L->op[i].pmod = nwv;

  const size_t esz = ca.intensity ? sizeof(float) : sizeof(float2);
  for (long long p0 = 0; p0 < d->batch; p0 += g.chunk) {
    const int kc = (int)std::min<long long>(g.chunk, d->batch - p0);
    ra.in = (const float2*)in + p0 * ra.in_plane_stride;
    ra.out = workspace;
    ra.n_planes = kc;
    ra.plane0 = (int)p0;
    if ((rc = launch_pipe_row(g.Nx, ra, s))) return rc;
    ca.in = (const float2*)workspace;
    ca.out = (char*)out + (size_t)p0 * ca.out_plane_stride * esz;
    ca.n_planes = kc;
    ca.plane0 = (int)p0;
    if ((rc = launch_pipe_col(g.Ny, ca, s))) return rc;
  }
  return CXM_OK;
}

// ---------------------------------------------------------------------------
// Scalable angular spectrum
// ---------------------------------------------------------------------------
namespace {
struct SasGeom { int H, W, Ny, Nx; long long chunk; };

int sas_geom(const cxm_optfft_desc* d, SasGeom& g) {
  cxm_optfft_desc t = *d;
  t.mode = CXM_OFFT_TRANSFORM;
  int rc = check_common(&t);
  if (rc) return rc;
  g.H = d->height;
  g.W = d->width;
  g.Ny = g.H + 2 * (g.H / 2);
  g.Nx = g.W + 2 * (g.W / 2);
  if (!fft_len_supported(g.Ny) || !fft_len_supported(g.Nx))
    return fail(CXM_ERR_UNSUPPORTED, "unsupported SAS grid " + std::to_string(g.Ny) + "x" +
                                         std::to_string(g.Nx));
  g.chunk = chunk_planes((long long)g.Ny * g.Nx, d->batch);
  return CXM_OK;
}

// Fill an OP_SASDH operator for every wavelength: grid of spacing (dy, dx)
// and size (Ny, Nx) after padding (compute_sas_precompensation).
void sasdh_slots(Ops& B, OpDesc* o, const cxm_optfft_desc* d, int nwv, int Ny, int Nx,
                 const double* sy, const double* sx) {
  for (int w = 0; w < nwv; ++w) {
    const Wv v = wv_of(d, w);
    const float lam = (float)v.lam, nf = (float)d->n, zf = (float)d->z;
    // kz = 2 z pi n / lam with z float32 (traced under the reference's lax.cond)
    const float kz = (((2.0f * zf) * (float)M_PI) * nf) / lam;
    const float dyf = (float)sy[w], dxf = (float)sx[w];
    // t = L / pad_factor / |z| + |s| with L = 2 * extent
    const float ext_y = dyf * (float)Ny, ext_x = dxf * (float)Nx;
    const float ty = ((2.0f * ext_y) / 2.0f) / std::fabs(zf);
    const float tx = ((2.0f * ext_x) / 2.0f) / std::fabs(zf);
    if (d->phase_mode == CXM_PHASE_REF32) {
      B.p(o, w, 0) = make_float4(kz, 0.f, lam, nf);
      B.p(o, w, 1) = make_float4(1.0f / dyf, 1.0f / dxf, ty, tx);
    } else {
      float th, tl, yh, yl, xh, xl;
      split(d->z * d->n / v.lam, th, tl);
      split(v.lam / (d->n * sy[w] * Ny), yh, yl);
      split(v.lam / (d->n * sx[w] * Nx), xh, xl);
      const double z_abs = std::fabs(d->z);
      B.p(o, w, 0) = make_float4(th, tl, yh, yl);
      B.p(o, w, 1) = make_float4(xh, xl, (float)(sy[w] * Ny / z_abs), (float)(sx[w] * Nx / z_abs));
    }
  }
}
}  // namespace

// Two padded planes per chunk plane: a column pass that stores with a
// column fftshift writes into other workgroups' strips, so it never runs in
// place.
extern "C" size_t cxm_sas_workspace_size(const cxm_optfft_desc* d) {
  SasGeom g;
  if (sas_geom(d, g)) return 0;
  return 2 * (size_t)g.chunk * g.Ny * g.Nx * sizeof(float2);
}

extern "C" int cxm_sas_c64(void* stream, const void* in, void* out, void* workspace,
                           size_t workspace_bytes, const cxm_optfft_desc* d) {
  SasGeom g;
  int rc = sas_geom(d, g);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const size_t need = 2 * (size_t)g.chunk * g.Ny * g.Nx * sizeof(float2);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int nwv = d->num_wv > 0 ? d->num_wv : 1;
  const int mode = d->phase_mode;
  const bool neg = d->z < 0;
  const int py = g.H / 2, px = g.W / 2;
  const long long plane_ws = (long long)g.Ny * g.Nx;
  const long long plane_in = (long long)g.H * g.W;
  const float inv_n = (float)(1.0 / ((double)g.Ny * g.Nx));
  float2* ws = (float2*)workspace;
  float2* ws2 = ws + (size_t)g.chunk * plane_ws;

  // per-wavelength optical constants on the padded grid
  double sy[32], sx[32], duy[32], dux[32];
  OptConsts oc[32];
  for (int w = 0; w < nwv; ++w) {
    const Wv v = wv_of(d, w);
    oc[w] = opt_consts(v, d->z, d->n, g.Ny, g.Nx, mode);
    sy[w] = v.dy;
    sx[w] = v.dx;
    duy[w] = oc[w].du_y;
    dux[w] = oc[w].du_x;
  }

  // Pass builders ------------------------------------------------------------
  auto base_row = [&](PipeArgs& a) {
    a.tw = twiddles_for(g.Nx);
    a.grid_ny = g.Ny;
    a.grid_nx = g.Nx;
    a.scale = 1.0f;
    a.mid_scale = 1.0f;
    a.out_plane_stride = plane_ws;
    a.out_line_stride = g.Nx;
    a.out_len = g.Nx;
  };
  auto base_col = [&](PipeArgs& a) {
    a.tw = twiddles_for(g.Ny);
    a.grid_ny = g.Ny;
    a.grid_nx = g.Nx;
    a.scale = 1.0f;
    a.mid_scale = 1.0f;
    a.n_lines = g.Nx;
    a.in_plane_stride = plane_ws;
    a.in_line_stride = g.Nx;
    a.out_plane_stride = plane_ws;
    a.out_line_stride = g.Nx;
    a.in_len = g.Ny;
    a.out_len = g.Ny;
    a.out_line_len = g.Nx;
  };
  auto fix_pmod = [&](PipeArgs& a) {
    for (OpList* L : {&a.pre, &a.mid, &a.post})
      for (int i = 0; i < L->n; ++i) if (L->op[i].pmod > 1) L->op[i].pmod = nwv;
  };

  // transform-step operators (pre: chirp, conj, norm; post: conj, 1j, chirp)
  auto transform_pre = [&](Ops& B, OpList& L) -> int {
    OpDesc* q1 = d->skip_initial_phase ? nullptr : B.add(L, OP_CHIRP, nwv, 3, mode);
    if (neg) B.add(L, OP_CONJ, 1, 0);
    OpDesc* sc = B.add(L, OP_SCALE, nwv, 1);
    if (!sc) return fail(CXM_ERR_INVALID, "operator table overflow");
    for (int w = 0; w < nwv; ++w) {
      const Wv v = wv_of(d, w);
      if (q1) chirp_slot(B, q1, w, oc[w], v.dy, v.dx, (float)v.dy, (float)v.dx, v, mode);
      // the 1j of transform_propagate commutes with the FFT for z >= 0
      float2 nm = oc[w].norm;
      if (!neg) nm = make_float2(-nm.y, nm.x);
      B.p(sc, w, 0) = make_float4(nm.x, nm.y, 0.f, 0.f);
    }
    return CXM_OK;
  };
  auto transform_post = [&](Ops& B, OpList& L) -> int {
    if (neg) {
      B.add(L, OP_CONJ, 1, 0);
      OpDesc* ij = B.add(L, OP_SCALE, 1, 1);
      if (!ij) return fail(CXM_ERR_INVALID, "operator table overflow");
      B.p(ij, 0, 0) = make_float4(0.f, 1.f, 0.f, 0.f);
    }
    if (!d->skip_final_phase) {
      OpDesc* q2 = B.add(L, OP_CHIRP, nwv, 3, mode);
      if (!q2) return fail(CXM_ERR_INVALID, "operator table overflow");
      for (int w = 0; w < nwv; ++w) {
        const Wv v = wv_of(d, w);
        chirp_slot(B, q2, w, oc[w], oc[w].du_y, oc[w].du_x, oc[w].du_y32, oc[w].du_x32, v, mode);
      }
    }
    return CXM_OK;
  };

  const size_t esz = d->out_kind == CXM_OUT_INTENSITY ? sizeof(float) : sizeof(float2);
  const long long out_plane = plane_in;
  for (long long p0 = 0; p0 < d->batch; p0 += g.chunk) {
    const int kc = (int)std::min<long long>(g.chunk, d->batch - p0);
    const float2* src = (const float2*)in + p0 * plane_in;
    void* dst = (char*)out + (size_t)p0 * out_plane * esz;
    if (!neg) {
      // R1: zero-padded row FFT
      PipeArgs r1; memset(&r1, 0, sizeof(r1)); base_row(r1);
      r1.in = src; r1.out = ws; r1.in_plane_stride = plane_in; r1.in_line_stride = g.W;
      r1.n_lines = g.H; r1.line_off = py; r1.in_off = px; r1.in_len = g.W; r1.seq = SEQ_F;
      r1.n_planes = kc; r1.plane0 = (int)p0;
      // keep the rows in padded-grid position for the column pass
      r1.out = ws + (size_t)py * g.Nx;
      if ((rc = launch_pipe_row(g.Nx, r1, s))) return rc;
      // C1: column FFT * dH * IFFT (rows outside the input are zero)
      PipeArgs c1; memset(&c1, 0, sizeof(c1)); base_col(c1);
      Ops b1(c1);
      OpDesc* dh = b1.add(c1.mid, OP_SASDH, nwv, 2, mode);
      if (!dh) return fail(CXM_ERR_INVALID, "operator table overflow");
      sasdh_slots(b1, dh, d, nwv, g.Ny, g.Nx, sy, sx);
      fix_pmod(c1);
      c1.out = ws; c1.in_off = py; c1.in_len = g.H; c1.in = ws + (size_t)py * g.Nx;
      c1.seq = SEQ_FI; c1.mid_scale = inv_n; c1.n_planes = kc; c1.plane0 = (int)p0;
      if ((rc = launch_pipe_col(g.Ny, c1, s))) return rc;
      // R2: row IFFT -> input chirp, norm, 1j -> ifftshift -> row FFT
      PipeArgs r2; memset(&r2, 0, sizeof(r2)); base_row(r2);
      Ops b2(r2);
      if ((rc = transform_pre(b2, r2.mid))) return rc;
      fix_pmod(r2);
      r2.in = ws; r2.out = ws; r2.in_plane_stride = plane_ws; r2.in_line_stride = g.Nx;
      r2.n_lines = g.Ny; r2.line_off = 0; r2.in_len = g.Nx; r2.seq = SEQ_IF; r2.mid_shift = 1;
      r2.n_planes = kc; r2.plane0 = (int)p0;
      if ((rc = launch_pipe_row(g.Nx, r2, s))) return rc;
      // C2: ifftshift rows -> column FFT -> fftshift, output chirp, crop
      PipeArgs c2; memset(&c2, 0, sizeof(c2)); base_col(c2);
      Ops b3(c2);
      if ((rc = transform_post(b3, c2.post))) return rc;
      fix_pmod(c2);
      c2.in = ws; c2.out = dst; c2.in_shift = 1; c2.out_shift = 1; c2.out_line_shift = 1;
      c2.out_off = py; c2.out_len = g.H; c2.out_line_off = px; c2.out_line_len = g.W;
      c2.out_plane_stride = out_plane; c2.out_line_stride = g.W; c2.seq = SEQ_F;
      c2.intensity = d->out_kind == CXM_OUT_INTENSITY; c2.n_planes = kc; c2.plane0 = (int)p0;
      if ((rc = launch_pipe_col(g.Ny, c2, s))) return rc;
    } else {
      // R1: transform step rows (pad, ifftshift, chirp, conj, norm, row FFT)
      PipeArgs r1; memset(&r1, 0, sizeof(r1)); base_row(r1);
      Ops b1(r1);
      if ((rc = transform_pre(b1, r1.pre))) return rc;
      fix_pmod(r1);
      r1.in = src; r1.out = ws + (size_t)py * g.Nx; r1.in_plane_stride = plane_in;
      r1.in_line_stride = g.W; r1.n_lines = g.H; r1.line_off = py; r1.in_off = px;
      r1.in_len = g.W; r1.in_shift = 1; r1.seq = SEQ_F; r1.n_planes = kc; r1.plane0 = (int)p0;
      if ((rc = launch_pipe_row(g.Nx, r1, s))) return rc;
      // C1: ifftshift rows -> column FFT -> fftshift, conj, 1j, output chirp (full grid)
      PipeArgs c1; memset(&c1, 0, sizeof(c1)); base_col(c1);
      Ops b2(c1);
      if ((rc = transform_post(b2, c1.post))) return rc;
      fix_pmod(c1);
      c1.in = ws + (size_t)py * g.Nx; c1.out = ws2; c1.in_off = py; c1.in_len = g.H;
      c1.in_shift = 1; c1.out_shift = 1; c1.out_line_shift = 1; c1.seq = SEQ_F;
      c1.n_planes = kc; c1.plane0 = (int)p0;
      if ((rc = launch_pipe_col(g.Ny, c1, s))) return rc;
      // R2: row FFT of the transformed field
      PipeArgs r2; memset(&r2, 0, sizeof(r2)); base_row(r2);
      r2.in = ws2; r2.out = ws2; r2.in_plane_stride = plane_ws; r2.in_line_stride = g.Nx;
      r2.n_lines = g.Ny; r2.in_len = g.Nx; r2.seq = SEQ_F; r2.n_planes = kc; r2.plane0 = (int)p0;
      if ((rc = launch_pipe_row(g.Nx, r2, s))) return rc;
      // C2: column FFT * dH (on the transformed grid) * IFFT
      PipeArgs c2; memset(&c2, 0, sizeof(c2)); base_col(c2);
      Ops b3(c2);
      OpDesc* dh = b3.add(c2.mid, OP_SASDH, nwv, 2, mode);
      if (!dh) return fail(CXM_ERR_INVALID, "operator table overflow");
      sasdh_slots(b3, dh, d, nwv, g.Ny, g.Nx, duy, dux);
      fix_pmod(c2);
      c2.in = ws2; c2.out = ws2; c2.seq = SEQ_FI; c2.mid_scale = inv_n; c2.n_planes = kc;
      c2.plane0 = (int)p0;
      if ((rc = launch_pipe_col(g.Ny, c2, s))) return rc;
      // R3: row IFFT, crop
      PipeArgs r3; memset(&r3, 0, sizeof(r3)); base_row(r3);
      r3.in = ws2 + (size_t)py * g.Nx; r3.out = dst; r3.in_plane_stride = plane_ws;
      r3.in_line_stride = g.Nx; r3.n_lines = g.H; r3.line_off = py; r3.in_len = g.Nx;
      r3.seq = SEQ_I; r3.out_off = px; r3.out_len = g.W; r3.out_plane_stride = out_plane;
      r3.out_line_stride = g.W; r3.intensity = d->out_kind == CXM_OUT_INTENSITY;
      r3.n_planes = kc; r3.plane0 = (int)p0;
      if ((rc = launch_pipe_row(g.Nx, r3, s))) return rc;
    }
  }
  return CXM_OK;
}

// ---------------------------------------------------------------------------
// Materialised operators (compute_sas_precompensation)
// ---------------------------------------------------------------------------
namespace cxm {
__global__ void op_grid_kernel(float2* out, int ny, int nx, OpList L, OpTable T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int plane = blockIdx.y;
  if (i >= (long long)ny * nx) return;
  const int iy = (int)(i / nx), ix = (int)(i - (long long)iy * nx);
  out[(long long)plane * ny * nx + i] = apply_ops(L, T, make_float2(1.f, 0.f), plane, iy, ix, ny, nx);
}
}  // namespace cxm

extern "C" int cxm_sas_precompensation_c64(void* stream, void* out, const cxm_optfft_desc* d) {
  cxm_optfft_desc t = *d;
  t.mode = CXM_OFFT_TRANSFORM;
  int rc = check_common(&t);
  if (rc) return rc;
  if (!out) return fail(CXM_ERR_INVALID, "null buffer");
  if ((rc = ensure_init())) return rc;
  const int nwv = d->num_wv > 0 ? d->num_wv : 1;
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  Ops b(a);
  OpDesc* dh = b.add(a.pre, OP_SASDH, nwv, 2, d->phase_mode);
  if (!dh) return fail(CXM_ERR_INVALID, "operator table overflow");
  double sy[32], sx[32];
  for (int w = 0; w < nwv; ++w) { const Wv v = wv_of(d, w); sy[w] = v.dy; sx[w] = v.dx; }
  sasdh_slots(b, dh, d, nwv, d->height, d->width, sy, sx);
  if (nwv > 1) a.pre.op[0].pmod = nwv;
  const long long n = (long long)d->height * d->width;
  hipLaunchKernelGGL(op_grid_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nwv), dim3(256), 0,
                     (hipStream_t)stream, (float2*)out, d->height, d->width, a.pre, a.tab);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}
