// Multislice propagation through a thick sample and its checkpointed
// adjoint (multislice_thick_sample, src/chromatix/functional/samples.py:137-282;
// thin_sample :85-134), and the fluorescent Monte-Carlo variant
// (fluorescent_multislice_thick_sample, samples.py:285-441).
//
// The field stays in row-transformed form between slices, so each slice is
// two passes over one padded plane per field:
//   column pass : column FFT -> * H_t (on the fly) -> column IFFT   (in place)
//   row pass    : row IFFT -> * exp(i 2 pi (dn + i abs) t / lam) -> row FFT
// The first slice starts with a zero-padded row FFT of the input; the end is
// the back-propagation column pass (H_{-d} times the optional NA mask) and a
// cropping row IFFT.
//
// Adjoint (linear in the field for fixed dn): every pass has the same form
// with conjugated factors; conj(H_t) = H_{-t} for the angular-spectrum
// propagator as the reference defines it (z < 0 conjugates, the band limit
// depends on z^2), and the adjoint of the sample multiply is conj(T).  The
// dn gradient at slice i is k Im(g conj(u_{i+1})), where u_{i+1} is the
// forward field after the slice, kept for one segment of k slices that is
// recomputed from a stored state.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cxm_pipe_host.h"

namespace cxm {
namespace {

struct MsGeom {
  int h, w, Ny, Nx, D, k, nseg;
  long long plane;  // Ny * Nx
};

int ms_geom(const cxm_multislice_desc* d, MsGeom& g) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->batch < 0 || d->height <= 0 || d->width <= 0 || d->depth < 1 || d->pad_y < 0 ||
      d->pad_x < 0)
    return fail(CXM_ERR_INVALID, "invalid shape parameters");
  if (d->phase_mode != CXM_PHASE_REF32 && d->phase_mode != CXM_PHASE_ACCURATE)
    return fail(CXM_ERR_INVALID, "invalid phase mode");
  if (d->out_kind != CXM_OUT_COMPLEX && d->out_kind != CXM_OUT_INTENSITY)
    return fail(CXM_ERR_INVALID, "invalid output kind");
  if (!(d->dy > 0 && d->dx > 0 && d->wavelength > 0 && d->n > 0))
    return fail(CXM_ERR_INVALID, "spacing, wavelength and n must be positive");
  if (d->use_na && !(d->NA > 0)) return fail(CXM_ERR_INVALID, "NA must be positive");
  if (d->checkpoint_every < 1) return fail(CXM_ERR_INVALID, "checkpoint_every must be >= 1");
  g.h = d->height;
  g.w = d->width;
  g.Ny = d->height + 2 * d->pad_y;
  g.Nx = d->width + 2 * d->pad_x;
  g.D = d->depth;
  g.k = d->checkpoint_every;
  g.nseg = (g.D + g.k - 1) / g.k;
  g.plane = (long long)g.Ny * g.Nx;
  if (!fft_len_supported(g.Ny) || !fft_len_supported(g.Nx))
    return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT grid " + std::to_string(g.Ny) + "x" +
                                         std::to_string(g.Nx));
  return CXM_OK;
}

// ---- dedicated row pass of a slice -------------------------------------------
// In place on row-transformed padded planes: inverse row FFT, the thin-sample
// factor exp(i 2 pi (dn + i abs) t / lam) (forward) or its adjoint with the
// dn / absorption gradient k Im(g conj(u)), -k Re(g conj(u)) (ADJ, u the
// stored forward field after the sample, g the adjoint value before the
// factor), forward row FFT.  The arithmetic of the operator pipeline's
// OP_THIN / OP_THINADJ row pass (same transforms for the same E), with the
// factor applied to the values in registers instead of through the
// pipeline's operator stage.  One row per workgroup of N/E threads.
struct MsRowArgs {
  float2* ws;
  const float2* tw;
  long long plane;  // elements per padded plane
  int Ny;
  const float* dn;  // slice arrays, h x w, centred at (pad_y, pad_x)
  const float* ab;  // may be null
  int h, w, pad_y, pad_x;
  float t, lam, k;  // thickness, wavelength, 2 pi t / lam (ADJ)
  // forward: optional store of the field after the sample, grid position
  // (y, x) -> mid_out[p * mid_pstride + (y - mid_y0) * mid_rstride + x - mid_x0]
  float2* mid_out;
  long long mid_pstride;
  int mid_rstride, mid_y0, mid_x0, mid_h, mid_w;
  const float2* u_fwd;  // ADJ: field after the sample (one padded plane)
  float* g_dn;          // ADJ: gradients (h x w), added to when accum
  float* g_ab;
  int accum;
};

template <int N, int E, bool ADJ>
__global__ void __launch_bounds__(N / E) ms_row_kernel(MsRowArgs a) {
  constexpr int T = N / E;
  static_assert(T % 64 == 0, "one row per workgroup of whole wavefronts");
  __shared__ float2 lds[lds_span<N, 1>()];
  const int p = blockIdx.x / a.Ny;
  const int y = blockIdx.x - p * a.Ny;
  const int t = threadIdx.x;
  float2* row = a.ws + p * a.plane + (long long)y * N;
  float2 v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = row[t + k * T];
  fft<N, E, 1, 1, true>(v, lds, t, a.tw);
  const int ay = y - a.pad_y;
  if (ay >= 0 && ay < a.h) {
    const long long rb = (long long)ay * a.w;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int x = t + k * T;
      const int ax = x - a.pad_x;
      if (ax >= 0 && ax < a.w) {
        const float dn = a.dn[rb + ax];
        const float ab = a.ab ? a.ab[rb + ax] : 0.0f;
        float ang, m, sn, cs;
        thin_angle(dn, ab, a.t, a.lam, &ang, &m);
        sincos_exact(ang, &sn, &cs);
        if constexpr (ADJ) {
          const float2 u = a.u_fwd[(long long)y * N + x];
          const float gdn = a.k * (v[k].y * u.x - v[k].x * u.y);
          float* g0 = a.g_dn + rb + ax;
          *g0 = a.accum ? *g0 + gdn : gdn;
          if (a.g_ab) {
            const float gab = -a.k * (v[k].x * u.x + v[k].y * u.y);
            float* g1 = a.g_ab + rb + ax;
            *g1 = a.accum ? *g1 + gab : gab;
          }
          v[k] = cmul(v[k], make_float2(m * cs, -m * sn));
        } else {
          v[k] = cmul(v[k], make_float2(m * cs, m * sn));
        }
      }
    }
  }
  if constexpr (!ADJ) {
    if (a.mid_out) {
      const int yl = y - a.mid_y0;
      if (yl >= 0 && yl < a.mid_h) {
        float2* mo = a.mid_out + p * a.mid_pstride + (long long)yl * a.mid_rstride;
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const int xl = t + k * T - a.mid_x0;
          if (xl >= 0 && xl < a.mid_w) mo[xl] = v[k];
        }
      }
    }
  }
  fft<N, E, -1, 1, true>(v, lds, t, a.tw);
#pragma unroll
  for (int k = 0; k < E; ++k) row[t + k * T] = v[k];
}

// Row lengths with a dedicated pass (values per thread; 1024 as the
// pipeline's single-plane shape, so the two agree bit for bit)
#define CXM_MS_ROW_SIZES(X) X(512, 8) X(1024, 4) X(2048, 8)

template <int N, int E>
int launch_ms_row_t(bool adj, const MsRowArgs& a, unsigned blocks, hipStream_t s) {
  if (adj)
    hipLaunchKernelGGL((ms_row_kernel<N, E, true>), dim3(blocks), dim3(N / E), 0, s, a);
  else
    hipLaunchKernelGGL((ms_row_kernel<N, E, false>), dim3(blocks), dim3(N / E), 0, s, a);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

// -1 when N has no dedicated pass
int launch_ms_row(int N, bool adj, const MsRowArgs& a, int planes, hipStream_t s) {
  const unsigned blocks = (unsigned)((long long)planes * a.Ny);
  if (blocks == 0) return CXM_OK;
  if (N == 1024) {
    // values per thread of the 1024-point rows (CXM_OPT_MS_ROW_E)
    switch (option(CXM_OPT_MS_ROW_E)) {
      case 2: return launch_ms_row_t<1024, 2>(adj, a, blocks, s);
      case 8: return launch_ms_row_t<1024, 8>(adj, a, blocks, s);
      case 16: return launch_ms_row_t<1024, 16>(adj, a, blocks, s);
      default: break;
    }
  }
  switch (N) {
#define CXM_MSR(NN, EE)                                                                        \
  case NN:                                                                                     \
    if (adj)                                                                                   \
      hipLaunchKernelGGL((ms_row_kernel<NN, EE, true>), dim3(blocks), dim3(NN / EE), 0, s, a); \
    else                                                                                       \
      hipLaunchKernelGGL((ms_row_kernel<NN, EE, false>), dim3(blocks), dim3(NN / EE), 0, s, a);\
    CXM_LAUNCH_CHECK();                                                                        \
    return CXM_OK;
    CXM_MS_ROW_SIZES(CXM_MSR)
#undef CXM_MSR
    default:
      return -1;
  }
}

// The dedicated row passes and the H tables (CXM_OPT_MS_PASSES 0), or the
// operator-pipeline row passes with H evaluated per bin (1); 2: the
// dedicated row passes with H evaluated per bin.
bool ms_dedicated() { return option(CXM_OPT_MS_PASSES) != 1; }
bool ms_tables() { return option(CXM_OPT_MS_PASSES) == 0; }

// Angular-spectrum parameters for distance z on the padded grid
// (compute_asm_propagator, src/chromatix/functional/propagation.py:499-580).
int asm_hparam(const cxm_multislice_desc* d, const MsGeom& g, double z, bool per_slice,
               HParam& hp) {
  cxm_prop2d_desc p;
  memset(&p, 0, sizeof(p));
  p.batch = 1;
  p.height = g.Ny;
  p.width = g.Nx;
  p.num_z = 1;
  p.kind = CXM_H_ASM;
  p.phase_mode = d->phase_mode;
  p.remove_evanescent = per_slice ? d->remove_evanescent : 0;
  p.bandlimit = per_slice ? d->bandlimit : 0;
  p.num_wv = 1;
  p.dy = d->dy;
  p.dx = d->dx;
  p.wavelength = d->wavelength;
  p.n = d->n;
  p.ky = d->ky;
  p.kx = d->kx;
  p.z = &z;
  return build_hparams(&p, g.Ny, g.Nx, &hp, 1);
}

// H(z) of the per-slice propagation (the parameters asm_hparam gives) into
// one padded plane in FFT order, for the table-reading column pass.
int slice_table(const cxm_multislice_desc* d, const MsGeom& g, double z, float2* out, hipStream_t s) {
  cxm_prop2d_desc p;
  memset(&p, 0, sizeof(p));
  p.batch = 1;
  p.height = g.Ny;
  p.width = g.Nx;
  p.num_z = 1;
  p.kind = CXM_H_ASM;
  p.phase_mode = d->phase_mode;
  p.remove_evanescent = d->remove_evanescent;
  p.bandlimit = d->bandlimit;
  p.num_wv = 1;
  p.dy = d->dy;
  p.dx = d->dx;
  p.wavelength = d->wavelength;
  p.n = d->n;
  p.ky = d->ky;
  p.kx = d->kx;
  p.z = &z;
  return cxm_propagator_c64((void*)s, out, &p);
}

// Column pass over `planes` padded planes in place, H for distance z; rows
// [in_off, in_off + in_len) of the FFT grid are read from `in` rows
// [0, in_len) (the rest are zero).
int col_pass(const cxm_multislice_desc* d, const MsGeom& g, const float2* in, float2* out,
             int planes, double z, bool per_slice, bool na, int in_off, int in_len,
             hipStream_t s, const float2* H = nullptr) {
  ColArgs ca;
  memset(&ca, 0, sizeof(ca));
  int rc = asm_hparam(d, g, z, per_slice, ca.hp[0]);
  if (rc) return rc;
  if (H && !na) {
    // H(z) read from a table of the same values (one padded plane, FFT order)
    ca.kind = CXM_H_ARRAY;
    ca.H = H;
  }
  ca.nhp = 1;
  ca.in = in;
  ca.out = out;
  ca.tw = twiddles_for(g.Ny);
  ca.in_plane_stride = g.plane;
  ca.out_plane_stride = g.plane;
  ca.row_stride = g.Nx;
  ca.ncols = g.Nx;
  ca.n_planes = planes;
  ca.nz = 1;
  ca.nwv = 1;
  ca.in_off = in_off;
  ca.in_len = in_len;
  ca.out_off = 0;
  ca.out_len = g.Ny;
  if (ca.kind != CXM_H_ARRAY) ca.kind = CXM_H_ASM;
  ca.phase_mode = d->phase_mode;
  ca.remove_evanescent = per_slice ? d->remove_evanescent : 0;
  ca.bandlimit = per_slice ? d->bandlimit : 0;
  ca.scale = (float)(1.0 / ((double)g.Ny * (double)g.Nx));
  if (na) {
    const float r = (float)d->NA / (float)d->wavelength;
    ca.na_r2 = r * r;
  }
  return launch_col_prop(g.Ny, ca, s);
}

// Common row-pass arguments over full padded planes.
void row_common(PipeArgs& a, const MsGeom& g, int planes) {
  a.tw = twiddles_for(g.Nx);
  a.in_plane_stride = g.plane;
  a.in_line_stride = g.Nx;
  a.out_plane_stride = g.plane;
  a.out_line_stride = g.Nx;
  a.n_lines = g.Ny;
  a.n_planes = planes;
  a.grid_ny = g.Ny;
  a.grid_nx = g.Nx;
  a.in_len = g.Nx;
  a.out_len = g.Nx;
  a.scale = 1.0f;
  a.mid_scale = 1.0f;
}

void slice_arrays(OpDesc* o, const cxm_multislice_desc* d, const MsGeom& g, const float* dn,
                  const float* ab, int i) {
  const long long sl = (long long)g.h * g.w;
  o->arr0 = dn + i * sl;
  o->arr1 = ab ? ab + i * sl : nullptr;
  o->arr_plane_stride = 0;
  o->arr_row_stride = g.w;
  o->arr_rows = g.h;
  o->arr_cols = g.w;
  o->arr_row_off = d->pad_y;
  o->arr_col_off = d->pad_x;
  o->t = (float)d->thickness;
}

// Row pass of slice i: IFFT, thin sample, FFT, in place; optional store of
// the field after the sample (mid_*).
int sample_pass(const cxm_multislice_desc* d, const MsGeom& g, float2* ws, int planes,
                const float* dn, const float* ab, int i, float2* mid_out, long long mid_pstride,
                int mid_rstride, int y0, int x0, int mh, int mw, hipStream_t s) {
  if (ms_dedicated()) {
    const long long sl = (long long)g.h * g.w;
    MsRowArgs r;
    memset(&r, 0, sizeof(r));
    r.ws = ws;
    r.tw = twiddles_for(g.Nx);
    r.plane = g.plane;
    r.Ny = g.Ny;
    r.dn = dn + i * sl;
    r.ab = ab ? ab + i * sl : nullptr;
    r.h = g.h;
    r.w = g.w;
    r.pad_y = d->pad_y;
    r.pad_x = d->pad_x;
    r.t = (float)d->thickness;
    r.lam = (float)d->wavelength;
    r.mid_out = mid_out;
    r.mid_pstride = mid_pstride;
    r.mid_rstride = mid_rstride;
    r.mid_y0 = y0;
    r.mid_x0 = x0;
    r.mid_h = mh;
    r.mid_w = mw;
    const int rc = launch_ms_row(g.Nx, false, r, planes, s);
    if (rc >= 0) return rc;
  }
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  Ops b(a);
  OpDesc* th = b.add(a.mid, OP_THIN, 1, 1);
  if (!th) return fail(CXM_ERR_INVALID, "operator table overflow");
  slice_arrays(th, d, g, dn, ab, i);
  b.p(th, 0, 0) = make_float4((float)d->wavelength, 0.f, 0.f, 0.f);
  row_common(a, g, planes);
  a.in = ws;
  a.out = ws;
  a.seq = SEQ_IF;
  a.mid_out = mid_out;
  a.mid_plane_stride = mid_pstride;
  a.mid_row_stride = mid_rstride;
  a.mid_y0 = y0;
  a.mid_x0 = x0;
  a.mid_h = mh;
  a.mid_w = mw;
  return launch_pipe_row(g.Nx, a, s);
}

// Zero-padded row FFT of `planes` input fields into ws rows [pad_y, pad_y + h).
int input_pass(const cxm_multislice_desc* d, const MsGeom& g, const float2* in, float2* ws,
               int planes, hipStream_t s) {
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  row_common(a, g, planes);
  a.in = in;
  a.in_plane_stride = (long long)g.h * g.w;
  a.in_line_stride = g.w;
  a.n_lines = g.h;
  a.line_off = d->pad_y;
  a.in_off = d->pad_x;
  a.in_len = g.w;
  a.out = ws + (long long)d->pad_y * g.Nx;
  a.seq = SEQ_F;
  return launch_pipe_row(g.Nx, a, s);
}

double back_distance(const cxm_multislice_desc* d) { return -d->reverse_distance; }

int copy_planes(const float2* src, float2* dst, long long n, hipStream_t s) {
  return cxm_copy_c64((void*)s, src, dst, n);
}

}  // namespace
}  // namespace cxm

using namespace cxm;

// workspace: the batch's padded planes | H(t) (one plane)
extern "C" size_t cxm_multislice_workspace_size(const cxm_multislice_desc* d) {
  MsGeom g;
  if (ms_geom(d, g)) return 0;
  return (size_t)(std::max(1, d->batch) + 1) * g.plane * sizeof(float2);
}

extern "C" size_t cxm_multislice_checkpoint_size(const cxm_multislice_desc* d) {
  MsGeom g;
  if (ms_geom(d, g)) return 0;
  return (size_t)std::max(0, d->batch) * (g.nseg - 1) * g.plane * sizeof(float2);
}

extern "C" int cxm_multislice_fwd(void* stream, const void* in, const float* dn,
                                  const float* absorption, void* out, void* checkpoints,
                                  void* workspace, size_t workspace_bytes,
                                  const cxm_multislice_desc* d) {
  MsGeom g;
  int rc = ms_geom(d, g);
  if (rc) return rc;
  if (d->batch == 0) return CXM_OK;
  if (!in || !dn || !out) return fail(CXM_ERR_INVALID, "null buffer");
  const size_t need = cxm_multislice_workspace_size(d);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int B = d->batch;
  float2* ws = (float2*)workspace;
  const long long fld = (long long)g.h * g.w;
  float2* Ht = nullptr;
  if (ms_tables()) {
    Ht = ws + (long long)B * g.plane;
    if ((rc = slice_table(d, g, d->thickness, Ht, s))) return rc;
  }
  if ((rc = input_pass(d, g, (const float2*)in, ws, B, s))) return rc;
  for (int i = 0; i < g.D; ++i) {
    // the first column pass reads only the input rows (the rest is padding)
    const bool first = i == 0;
    if ((rc = col_pass(d, g, first ? ws + (long long)d->pad_y * g.Nx : ws, ws, B, d->thickness,
                       true, false, first ? d->pad_y : 0, first ? g.h : g.Ny, s, Ht)))
      return rc;
    float2* stack = d->return_stack ? (float2*)out + (long long)i * B * fld : nullptr;
    if ((rc = sample_pass(d, g, ws, B, dn, absorption, i, stack, fld, g.w, d->pad_y, d->pad_x,
                          g.h, g.w, s)))
      return rc;
    if (checkpoints && (i + 1) % g.k == 0 && i + 1 < g.D) {
      const int c = (i + 1) / g.k;  // state before slice c k
      float2* dst = (float2*)checkpoints + (long long)(c - 1) * B * g.plane;
      if ((rc = copy_planes(ws, dst, (long long)B * g.plane, s))) return rc;
    }
  }
  if (d->return_stack) return CXM_OK;
  if ((rc = col_pass(d, g, ws, ws, B, back_distance(d), false, d->use_na != 0, 0, g.Ny, s)))
    return rc;
  PipeArgs a;
  memset(&a, 0, sizeof(a));
  row_common(a, g, B);
  a.in = ws + (long long)d->pad_y * g.Nx;
  a.n_lines = g.h;
  a.line_off = d->pad_y;
  a.out = out;
  a.out_plane_stride = fld;
  a.out_line_stride = g.w;
  a.out_off = d->pad_x;
  a.out_len = g.w;
  a.seq = SEQ_I;
  a.intensity = d->out_kind == CXM_OUT_INTENSITY;
  return launch_pipe_row(g.Nx, a, s);
}

extern "C" size_t cxm_multislice_bwd_workspace_size(const cxm_multislice_desc* d) {
  MsGeom g;
  if (ms_geom(d, g)) return 0;
  // forward state, adjoint state, two segments of fields (double-buffered),
  // internal checkpoints, H(t) and H(-t)
  return (size_t)(2 + 2 * g.k + (g.nseg - 1) + 2) * g.plane * sizeof(float2);
}

namespace cxm {
namespace {
// A second stream per device for the adjoint's segment recomputation (default
// priority, non-blocking; created once and kept until the process ends).
hipStream_t ms_aux_stream() {
  static std::mutex mu;
  static hipStream_t st[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!st[dev]) {
    hipStream_t x;
    if (hipStreamCreateWithFlags(&x, hipStreamNonBlocking) != hipSuccess) return nullptr;
    st[dev] = x;
  }
  return st[dev];
}

// Events of one call (timing disabled); destroyed at scope exit, which HIP
// defers until the work they mark has completed.
struct EventSet {
  std::vector<hipEvent_t> ev;
  int make(int n) {
    ev.assign(n, nullptr);
    for (auto& e : ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
        return fail(CXM_ERR_HIP, "hipEventCreateWithFlags failed");
    return CXM_OK;
  }
  ~EventSet() {
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// `waiter` waits for the work enqueued on `src` so far (nothing when both are
// the same stream).
int ms_join(hipStream_t src, hipStream_t waiter, hipEvent_t e) {
  if (src == waiter) return CXM_OK;
  if (hipEventRecord(e, src) != hipSuccess || hipStreamWaitEvent(waiter, e, 0) != hipSuccess)
    return fail(CXM_ERR_HIP, "event record / wait failed");
  return CXM_OK;
}
}  // namespace
}  // namespace cxm

// The adjoint sweeps the segments of k slices from the last to the first.
// A segment's forward fields are recomputed from its checkpoint (stream F)
// into one of two segment buffers while the reverse sweep of the segment
// after it (the caller's stream R) reads the other: the two single-plane
// chains run side by side.  F recomputes segment sg once R has finished the
// segment sg + 2, which used the same buffer; R sweeps segment sg once F has
// recomputed it.  F's last work is waited for on R, so the caller's stream
// orders everything.
extern "C" int cxm_multislice_bwd(void* stream, const void* in, const float* dn,
                                  const float* absorption, const void* checkpoints,
                                  const void* cot, float* grad_dn, float* grad_absorption,
                                  void* workspace, size_t workspace_bytes,
                                  const cxm_multislice_desc* d) {
  MsGeom g;
  int rc = ms_geom(d, g);
  if (rc) return rc;
  if (d->return_stack) return fail(CXM_ERR_INVALID, "no adjoint for return_stack");
  if (d->grad_kind < CXM_GRAD_UPSTREAM || d->grad_kind > CXM_GRAD_COMPLEX)
    return fail(CXM_ERR_INVALID, "invalid gradient kind");
  if (d->batch == 0) return CXM_OK;
  if (!in || !dn || !cot || !grad_dn) return fail(CXM_ERR_INVALID, "null buffer");
  const size_t need = cxm_multislice_bwd_workspace_size(d);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t R = (hipStream_t)stream;
  hipStream_t F = R;
  if (g.nseg > 1) {
    F = ms_aux_stream();
    if (!F) F = R;
  }
  const int B = d->batch;
  const long long fld = (long long)g.h * g.w;
  const long long sl = fld;
  float2* wf = (float2*)workspace;  // forward state (F)
  float2* wg = wf + g.plane;        // adjoint state (R)
  float2* segs = wg + g.plane;      // two buffers of k fields after each slice
  float2* ick = segs + 2LL * g.k * g.plane;  // internal checkpoints (one field)
  float2* Hf = ick + (long long)(g.nseg - 1) * g.plane;  // H(t), H(-t) tables
  float2* Hb = Hf + g.plane;
  const float kk = (float)(2.0 * M_PI * d->thickness / d->wavelength);
  const bool ded = ms_dedicated();
  if (!ms_tables()) Hf = Hb = nullptr;
  EventSet es;
  // per field: start, one "recomputed" and one "swept" event per segment
  const int ne = 1 + 2 * g.nseg;
  if ((rc = es.make(B * ne))) return rc;
  if (Hf) {
    if ((rc = slice_table(d, g, d->thickness, Hf, R))) return rc;
    if ((rc = slice_table(d, g, -d->thickness, Hb, R))) return rc;
  }
  auto segbuf = [&](int sg) { return segs + (long long)(sg & 1) * g.k * g.plane; };
  for (int b = 0; b < B; ++b) {
    hipEvent_t* E = es.ev.data() + (long long)b * ne;
    hipEvent_t* e_rec = E + 1;             // F: segment sg recomputed
    hipEvent_t* e_swept = E + 1 + g.nseg;  // R: segment sg swept
    const float2* inb = (const float2*)in + b * fld;
    // F starts after everything enqueued on R so far (tables, previous field)
    if ((rc = ms_join(R, F, E[0]))) return rc;
    // states before slices c k, c = 1 .. nseg-1, of field b
    auto state = [&](int c) -> const float2* {
      if (checkpoints) return (const float2*)checkpoints + ((long long)(c - 1) * B + b) * g.plane;
      return ick + (long long)(c - 1) * g.plane;
    };
    if (!checkpoints && g.nseg > 1) {
      if ((rc = input_pass(d, g, inb, wf, 1, F))) return rc;
      for (int i = 0; i + 1 < (g.nseg - 1) * g.k + 1 && i < g.D; ++i) {
        const bool first = i == 0;
        if ((rc = col_pass(d, g, first ? wf + (long long)d->pad_y * g.Nx : wf, wf, 1,
                           d->thickness, true, false, first ? d->pad_y : 0, first ? g.h : g.Ny, F, Hf)))
          return rc;
        if ((rc = sample_pass(d, g, wf, 1, dn, absorption, i, nullptr, 0, 0, 0, 0, 0, 0, F)))
          return rc;
        if ((i + 1) % g.k == 0 && i + 1 < g.D &&
            (rc = copy_planes(wf, ick + (long long)((i + 1) / g.k - 1) * g.plane, g.plane, F)))
          return rc;
      }
    }
    for (int sg = g.nseg - 1; sg >= 0; --sg) {
      const int i0 = sg * g.k, i1 = std::min(g.D, i0 + g.k);
      float2* seg = segbuf(sg);
      // ---- F: recompute the segment, keeping the field after each slice
      if (sg + 2 <= g.nseg - 1 && F != R && hipStreamWaitEvent(F, e_swept[sg + 2], 0) != hipSuccess)
        return fail(CXM_ERR_HIP, "hipStreamWaitEvent failed");
      if (sg == 0) {
        if ((rc = input_pass(d, g, inb, wf, 1, F))) return rc;
      } else if ((rc = copy_planes(state(sg), wf, g.plane, F))) {
        return rc;
      }
      for (int i = i0; i < i1; ++i) {
        const bool first = i == 0;
        if ((rc = col_pass(d, g, first ? wf + (long long)d->pad_y * g.Nx : wf, wf, 1,
                           d->thickness, true, false, first ? d->pad_y : 0, first ? g.h : g.Ny, F, Hf)))
          return rc;
        if ((rc = sample_pass(d, g, wf, 1, dn, absorption, i, seg + (long long)(i - i0) * g.plane,
                              g.plane, g.Nx, 0, 0, g.Ny, g.Nx, F)))
          return rc;
      }
      if (sg == g.nseg - 1) {
        // output cotangent: back-propagate, inverse rows, cotangent of the
        // cropped output, forward rows -> adjoint state; then the adjoint of
        // the back-propagation
        if ((rc = col_pass(d, g, wf, wf, 1, back_distance(d), false, d->use_na != 0, 0, g.Ny, F)))
          return rc;
        PipeArgs a;
        memset(&a, 0, sizeof(a));
        Ops ob(a);
        OpDesc* ig = ob.add(a.mid, OP_IGRAD, 1, 1, d->grad_kind);
        if (!ig) return fail(CXM_ERR_INVALID, "operator table overflow");
        const size_t esz = d->grad_kind == CXM_GRAD_COMPLEX ? sizeof(float2) : sizeof(float);
        ig->arr0 = (const char*)cot + (size_t)b * fld * esz;
        ig->arr_plane_stride = 0;
        ig->arr_row_stride = g.w;
        ig->arr_rows = g.h;
        ig->arr_cols = g.w;
        ig->arr_row_off = d->pad_y;
        ig->arr_col_off = d->pad_x;
        ob.p(ig, 0, 0) = make_float4((float)(2.0 / ((double)B * fld)), 0.f, 0.f, 0.f);
        row_common(a, g, 1);
        a.in = wf;
        a.out = wg;
        a.seq = SEQ_IF;
        if ((rc = launch_pipe_row(g.Nx, a, F))) return rc;
        if ((rc = col_pass(d, g, wg, wg, 1, -back_distance(d), false, d->use_na != 0, 0, g.Ny, F)))
          return rc;
      }
      // ---- R: reverse through the segment once F has recomputed it
      if ((rc = ms_join(F, R, e_rec[sg]))) return rc;
      for (int i = i1 - 1; i >= i0; --i) {
        if (ded) {
          MsRowArgs r;
          memset(&r, 0, sizeof(r));
          r.ws = wg;
          r.tw = twiddles_for(g.Nx);
          r.plane = g.plane;
          r.Ny = g.Ny;
          r.dn = dn + i * sl;
          r.ab = absorption ? absorption + i * sl : nullptr;
          r.h = g.h;
          r.w = g.w;
          r.pad_y = d->pad_y;
          r.pad_x = d->pad_x;
          r.t = (float)d->thickness;
          r.lam = (float)d->wavelength;
          r.k = kk;
          r.u_fwd = seg + (long long)(i - i0) * g.plane;
          r.g_dn = grad_dn + (long long)i * fld;
          r.g_ab = grad_absorption ? grad_absorption + (long long)i * fld : nullptr;
          r.accum = b > 0;
          rc = launch_ms_row(g.Nx, true, r, 1, R);
          if (rc > 0) return rc;
          if (rc == 0) {
            if (i > 0 && (rc = col_pass(d, g, wg, wg, 1, -d->thickness, true, false, 0, g.Ny, R, Hb)))
              return rc;
            continue;
          }
        }
        PipeArgs a;
        memset(&a, 0, sizeof(a));
        Ops ob(a);
        OpDesc* ta = ob.add(a.mid, OP_THINADJ, 1, 1);
        if (!ta) return fail(CXM_ERR_INVALID, "operator table overflow");
        slice_arrays(ta, d, g, dn, absorption, i);
        ta->arr2 = seg + (long long)(i - i0) * g.plane;
        ta->arr2_plane_stride = 0;
        ta->arr2_row_stride = g.Nx;
        ta->out0 = grad_dn + (long long)i * fld;
        ta->out1 = grad_absorption ? grad_absorption + (long long)i * fld : nullptr;
        ta->accum = b > 0;
        ob.p(ta, 0, 0) = make_float4((float)d->wavelength, kk, 0.f, 0.f);
        row_common(a, g, 1);
        a.in = wg;
        a.out = wg;
        a.seq = SEQ_IF;
        if ((rc = launch_pipe_row(g.Nx, a, R))) return rc;
        if (i > 0 && (rc = col_pass(d, g, wg, wg, 1, -d->thickness, true, false, 0, g.Ny, R, Hb)))
          return rc;
      }
      if (F != R && hipEventRecord(e_swept[sg], R) != hipSuccess)
        return fail(CXM_ERR_HIP, "hipEventRecord failed");
    }
  }
  return CXM_OK;
}

// ---------------------------------------------------------------------------
// Fluorescent multislice, one Monte-Carlo sample (samples.py:285-441).
// Forward: the field stays row-transformed between slices as above; slice
// i's row pass applies the thin sample and adds f_i exp(i r_i) between the
// row IFFT and FFT, the column pass multiplies by P(+t); the last column
// pass stops after the column FFT and the multiply, which is fft2 of the
// field after the last slice.  Backward step i: U *= P(-t), row IFFT,
// column IFFT storing |.|^2 of the cropped window accumulated into
// intensity plane depth-1-i; the steps' multiplies form an elementwise
// chain, so the transforms of many steps run as one batched pass.
// ---------------------------------------------------------------------------
namespace cxm {
namespace {
struct FlGeom {
  int h, w, Ny, Nx, D;
  long long plane;
};

int fl_geom(const cxm_fluor_desc* d, FlGeom& g) {
  if (!d) return fail(CXM_ERR_INVALID, "null descriptor");
  if (d->height <= 0 || d->width <= 0 || d->depth < 1 || d->pad_y < 0 || d->pad_x < 0)
    return fail(CXM_ERR_INVALID, "invalid shape parameters");
  if (d->phase_mode != CXM_PHASE_REF32 && d->phase_mode != CXM_PHASE_ACCURATE)
    return fail(CXM_ERR_INVALID, "invalid phase mode");
  if (!(d->dy > 0 && d->dx > 0 && d->wavelength > 0 && d->n > 0))
    return fail(CXM_ERR_INVALID, "spacing, wavelength and n must be positive");
  g.h = d->height;
  g.w = d->width;
  g.Ny = d->height + 2 * d->pad_y;
  g.Nx = d->width + 2 * d->pad_x;
  g.D = d->depth;
  g.plane = (long long)g.Ny * g.Nx;
  if (!fft_len_supported(g.Ny) || !fft_len_supported(g.Nx))
    return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT grid " + std::to_string(g.Ny) + "x" +
                                         std::to_string(g.Nx));
  return CXM_OK;
}

// The slice arrays (unpadded, centred in the padded grid) for an operator.
void fl_arrays(OpDesc* o, const cxm_fluor_desc* d, const FlGeom& g, const float* a0,
               const float* a1, int i) {
  const long long sl = (long long)g.h * g.w;
  o->arr0 = a0 + i * sl;
  o->arr1 = a1 ? a1 + i * sl : nullptr;
  o->arr_plane_stride = 0;
  o->arr_row_stride = g.w;
  o->arr_rows = g.h;
  o->arr_cols = g.w;
  o->arr_row_off = d->pad_y;
  o->arr_col_off = d->pad_x;
  o->t = (float)d->thickness;
}

// ---- dedicated row passes of the fluorescent sample -------------------------
// The row passes of the sweep without the operator interpreter (the same
// arithmetic in the same order, so the bits of the operator pipeline):
//   row r of W: inverse row FFT; inside the slice window the thin sample
//   factor exp(i 2 pi dn t / lam) and the fluorescence f exp(i r) added;
//   forward row FFT, in place.  Slice 0 (FIRST) reads the spatial field
//   and skips the inverse transform.
// Rows shorter than a wavefront run 64 / T rows per wavefront, like the
// operator pipeline's row pass.
struct FlRowArgs {
  float2* W;
  const float2* src;  // FIRST: the spatial field of slice 0
  const float2* tw;
  int Ny;
  const float4* tf;   // the slice's factors, h x w at (pad_y, pad_x)
  int h, w, pad_y, pad_x;
};

// The slice factors of a block of slices, which do not depend on the field:
// tf = (T, F), T = exp(i 2 pi dn t / lam) (thin_sample's factor, its float32
// order), F = f exp(i r) (the fluorescence term) -- the values the row pass
// combines with the field as x = v T, v = F + x.  One full-chip pass per
// block, so that the sine/cosine work leaves the slice chain.
__global__ void fl_factors_kernel(const float* __restrict__ dn, const float* __restrict__ fl,
                                  const float* __restrict__ ph, float4* __restrict__ tf, long long n,
                                  float t, float lam) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float ang, m, sn, cs;
  thin_angle(dn[i], 0.0f, t, lam, &ang, &m);
  sincos_exact(ang, &sn, &cs);
  const float tx = m * cs, ty = m * sn;
  const float f = fl[i];
  sincos_exact(ph[i], &sn, &cs);
  tf[i] = make_float4(tx, ty, __fmul_rn(f, cs), __fmul_rn(f, sn));
}

template <int N, int E>
struct FlRowCfg {
  static constexpr int T = N / E;
  static constexpr int G = T >= 64 ? 1 : 64 / T;
};

template <int N, int E, bool FIRST>
__global__ void __launch_bounds__((FlRowCfg<N, E>::G * (N / E))) fl_row_kernel(FlRowArgs a) {
  constexpr int T = N / E;
  constexpr int G = FlRowCfg<N, E>::G;
  constexpr bool SYNC = T > 64;
  constexpr int SPAN = lds_span<N, 1>();
  __shared__ float2 lds[G * SPAN];
  const int g = threadIdx.x / T;
  const int t = threadIdx.x - g * T;
  const int r = blockIdx.x * G + g;
  const bool ok = r < a.Ny;
  const long long rb = (long long)(ok ? r : 0) * N;
  float2* st = lds + g * SPAN;
  float2 v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = (FIRST ? a.src : a.W)[rb + t + k * T];
  // the slice factors loaded before the first transform, which hides their
  // latency
  const int ay = r - a.pad_y;
  const bool rin = ok && ay >= 0 && ay < a.h;
  float4 f4[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int ax = t + k * T - a.pad_x;
    const bool in = rin && ax >= 0 && ax < a.w;
    f4[k] = in ? a.tf[(long long)ay * a.w + ax] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if constexpr (!FIRST) fft<N, E, 1, 1, SYNC>(v, st, t, a.tw);
  // the thread index laundered: the forward transform's twiddle loads are
  // not hoisted above the slice's operators
  int tl = t;
  asm volatile("" : "+v"(tl));
  if (rin) {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int ax = tl + k * T - a.pad_x;
      if (ax < 0 || ax >= a.w) continue;
      const float2 x = cmul(v[k], make_float2(f4[k].x, f4[k].y));
      v[k] = make_float2(__fadd_rn(f4[k].z, x.x), __fadd_rn(f4[k].w, x.y));
    }
  }
  fft<N, E, -1, 1, SYNC>(v, st, tl, a.tw);
  if (!ok) return;
#pragma unroll
  for (int k = 0; k < E; ++k) a.W[rb + tl + k * T] = v[k];
}

// (row length, values per thread): the operator pipeline's row shapes for
// one plane (1024: 256 threads per row), except 640 (10 values per thread,
// one 64-thread row per wavefront: twice the pipeline's waves, 97 -> 114
// Monte-Carlo stacks/s at 512^2 + pad 64; CXM_OPT_FL_ROW_E 20 gives the
// pipeline's shape and bits)
#define CXM_FL_ROW_SIZES(X) X(512, 16) X(640, 10) X(1024, 4) X(1280, 20) X(2048, 16)

bool fl_dedicated(int n) {
  if (!ms_dedicated()) return false;
  switch (n) {
#define CXM_FL_CASE(N, E) case N:
    CXM_FL_ROW_SIZES(CXM_FL_CASE)
#undef CXM_FL_CASE
    return true;
    default:
      return false;
  }
}

template <int N, int E>
int launch_fl_row_t(bool first, const FlRowArgs& a, hipStream_t s) {
  constexpr int G = FlRowCfg<N, E>::G;
  const dim3 grid((unsigned)((a.Ny + G - 1) / G)), block((unsigned)(G * (N / E)));
  if (first) hipLaunchKernelGGL((fl_row_kernel<N, E, true>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((fl_row_kernel<N, E, false>), grid, block, 0, s, a);
  CXM_LAUNCH_CHECK();
  return CXM_OK;
}

// first: slice 0 from the spatial field a.src (no inverse transform)
int launch_fl_row(int n, bool first, const FlRowArgs& a, hipStream_t s) {
  if (n == 640 && option(CXM_OPT_FL_ROW_E) == 20) return launch_fl_row_t<640, 20>(first, a, s);
  switch (n) {
#define CXM_FL_LAUNCH(N, E) \
  case N: return launch_fl_row_t<N, E>(first, a, s);
    CXM_FL_ROW_SIZES(CXM_FL_LAUNCH)
#undef CXM_FL_LAUNCH
    default:
      return fail(CXM_ERR_UNSUPPORTED, "no dedicated fluorescent row pass for " + std::to_string(n));
  }
}

// ---- backward sweep, batched over its steps --------------------------------
// U_{i+1} = U_i * Pb is an elementwise chain, so the steps' inverse
// transforms do not depend on one another.  Workgroup (row group, chunk c)
// loads its rows of U_{i0} and Pb, advances them through c * chunk
// multiplies in registers (the products of the step-by-step sweep, in its
// order, so the same bits), then for each of its steps multiplies once
// more, inverse-transforms a copy along the row and stores it scaled by
// 1 / Nx in the step's workspace plane (planes in reverse step order, so
// plane p belongs to intensity plane depth - i1 + p).  The last chunk
// stores U_{i1} in `dst` (a buffer other than `src`: every chunk of a row
// reads `src`).
struct FlBwdArgs {
  const float2* src;
  float2* dst;
  const float2* Pb;
  float2* W;
  const float2* tw;
  int Ny;
  int steps;          // i1 - i0
  int chunk;          // steps per workgroup
  long long plane;    // Ny * Nx
  float scale;        // 1 / Nx
};

template <int N, int E>
__global__ void __launch_bounds__((FlRowCfg<N, E>::G * (N / E))) fl_bwd_rows_kernel(FlBwdArgs a) {
  constexpr int T = N / E;
  constexpr int G = FlRowCfg<N, E>::G;
  constexpr bool SYNC = T > 64;
  constexpr int SPAN = lds_span<N, 1>();
  __shared__ float2 lds[G * SPAN];
  const int g = threadIdx.x / T;
  const int t = threadIdx.x - g * T;
  const int r = blockIdx.x * G + g;
  const bool ok = r < a.Ny;
  const long long rb = (long long)(ok ? r : 0) * N;
  float2* st = lds + g * SPAN;
  const int i0 = blockIdx.y * a.chunk;
  const int i1 = min(i0 + a.chunk, a.steps);
  float2 u[E], p[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    u[k] = a.src[rb + t + k * T];
    p[k] = a.Pb[rb + t + k * T];
  }
  for (int i = 0; i < i0; ++i) {
#pragma unroll
    for (int k = 0; k < E; ++k) u[k] = cmul(u[k], p[k]);
  }
  for (int i = i0; i < i1; ++i) {
    // the thread index laundered per step: twiddle loads stay in the loop
    int tl = t;
    asm volatile("" : "+v"(tl));
    float2 v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      u[k] = cmul(u[k], p[k]);
      v[k] = u[k];
    }
    fft<N, E, 1, 1, SYNC>(v, st, tl, a.tw);
    if (ok) {
      float2* w = a.W + (long long)(a.steps - 1 - i) * a.plane + rb;
#pragma unroll
      for (int k = 0; k < E; ++k) w[tl + k * T] = cscale(v[k], a.scale);
    }
  }
  if (ok && i1 == a.steps) {
#pragma unroll
    for (int k = 0; k < E; ++k) a.dst[rb + t + k * T] = u[k];
  }
}

// Steps per workgroup of the batched backward rows (CXM_OPT_FL_BWD_CHUNK).
int fl_bwd_chunk() {
  const int c = option(CXM_OPT_FL_BWD_CHUNK);
  return c > 0 ? c : 8;
}

int launch_fl_bwd_rows(int n, const FlBwdArgs& a, hipStream_t s) {
  const unsigned chunks = (unsigned)((a.steps + a.chunk - 1) / a.chunk);
  if (n == 640 && option(CXM_OPT_FL_ROW_E) == 20) {
    constexpr int G = FlRowCfg<640, 20>::G;
    const dim3 grid((unsigned)((a.Ny + G - 1) / G), chunks), block((unsigned)(G * 32));
    hipLaunchKernelGGL((fl_bwd_rows_kernel<640, 20>), grid, block, 0, s, a);
    CXM_LAUNCH_CHECK();
    return CXM_OK;
  }
  switch (n) {
#define CXM_FL_BWD_LAUNCH(N, E)                                                                  \
  case N: {                                                                                      \
    constexpr int G = FlRowCfg<N, E>::G;                                                         \
    const dim3 grid((unsigned)((a.Ny + G - 1) / G), chunks), block((unsigned)(G * (N / E)));    \
    hipLaunchKernelGGL((fl_bwd_rows_kernel<N, E>), grid, block, 0, s, a);                        \
    CXM_LAUNCH_CHECK();                                                                          \
    return CXM_OK;                                                                               \
  }
    CXM_FL_ROW_SIZES(CXM_FL_BWD_LAUNCH)
#undef CXM_FL_BWD_LAUNCH
    default:
      return fail(CXM_ERR_UNSUPPORTED, "no batched backward row pass for " + std::to_string(n));
  }
}

// Steps of the backward sweep per batched launch pair: the workspace holds
// that many planes, at most 512 MiB of them (at least one), or the number
// CXM_OPT_FL_BWD_BATCH sets.
int fl_bwd_batch(const FlGeom& g) {
  const int o = option(CXM_OPT_FL_BWD_BATCH);
  if (o > 0) return std::min(g.D, o);
  const long long cap = (512ll << 20) / (g.plane * (long long)sizeof(float2));
  return (int)std::max(1ll, std::min((long long)g.D, cap));
}

// Slices per block of precomputed slice factors (16 bytes per slice
// element): at most 512 MiB of them (at least one slice), or the number
// CXM_OPT_FL_FACTOR_BLOCK sets.
int fl_factor_block(const FlGeom& g) {
  const int o = option(CXM_OPT_FL_FACTOR_BLOCK);
  if (o > 0) return std::min(g.D, o);
  const long long cap = (512ll << 20) / ((long long)g.h * g.w * (long long)sizeof(float4));
  return (int)std::max(1ll, std::min((long long)g.D, cap));
}

// Byte offset of the factor block in the workspace (256-byte aligned) and
// the workspace size.
size_t fl_factor_offset(const FlGeom& g) {
  const size_t planes = (size_t)(3 + fl_bwd_batch(g)) * g.plane * sizeof(float2);
  return (planes + 255) / 256 * 256;
}
size_t fl_workspace(const FlGeom& g) {
  return fl_factor_offset(g) + (size_t)fl_factor_block(g) * g.h * g.w * sizeof(float4);
}

// Thin sample and fluorescence of slice i as operators of list L.
int fl_slice_ops(Ops& b, OpList& L, const cxm_fluor_desc* d, const FlGeom& g, const float* dn,
                 const float* fl, const float* ph, int i) {
  OpDesc* th = b.add(L, OP_THIN, 1, 1);
  OpDesc* fo = b.add(L, OP_FLUOR, 1, 0);
  if (!th || !fo) return fail(CXM_ERR_INVALID, "operator table overflow");
  fl_arrays(th, d, g, dn, nullptr, i);
  b.p(th, 0, 0) = make_float4((float)d->wavelength, 0.f, 0.f, 0.f);
  fl_arrays(fo, d, g, fl, ph, i);
  return CXM_OK;
}

void fl_row_common(PipeArgs& a, const FlGeom& g) {
  a.tw = twiddles_for(g.Nx);
  a.in_plane_stride = g.plane;
  a.in_line_stride = g.Nx;
  a.out_plane_stride = g.plane;
  a.out_line_stride = g.Nx;
  a.n_lines = g.Ny;
  a.n_planes = 1;
  a.grid_ny = g.Ny;
  a.grid_nx = g.Nx;
  a.in_len = g.Nx;
  a.out_len = g.Nx;
  a.scale = 1.0f;
  a.mid_scale = 1.0f;
}

void fl_col_common(PipeArgs& a, const FlGeom& g) {
  a.tw = twiddles_for(g.Ny);
  a.in_plane_stride = g.plane;
  a.in_line_stride = g.Nx;
  a.out_plane_stride = g.plane;
  a.out_line_stride = g.Nx;
  a.n_lines = g.Nx;
  a.n_planes = 1;
  a.grid_ny = g.Ny;
  a.grid_nx = g.Nx;
  a.in_len = g.Ny;
  a.out_len = g.Ny;
  a.out_line_len = g.Nx;
  a.scale = 1.0f;
  a.mid_scale = 1.0f;
}

int fl_propagator(const cxm_fluor_desc* d, const FlGeom& g, double z, void* out, hipStream_t s) {
  cxm_prop2d_desc p;
  memset(&p, 0, sizeof(p));
  p.batch = 1;
  p.height = g.Ny;
  p.width = g.Nx;
  p.num_z = 1;
  p.kind = CXM_H_ASM;
  p.phase_mode = d->phase_mode;
  p.remove_evanescent = d->remove_evanescent;
  p.bandlimit = d->bandlimit;
  p.num_wv = 1;
  p.dy = d->dy;
  p.dx = d->dx;
  p.wavelength = d->wavelength;
  p.n = d->n;
  p.ky = d->ky;
  p.kx = d->kx;
  p.z = &z;
  return cxm_propagator_c64((void*)s, out, &p);
}
}  // namespace
}  // namespace cxm

extern "C" size_t cxm_fluor_multislice_workspace_size(const cxm_fluor_desc* d) {
  FlGeom g;
  if (fl_geom(d, g)) return 0;
  // P(+t), P(-t), the carried spectrum's second buffer, the row-transformed
  // field (forward) or the backward steps' planes, then a block of slice
  // factors
  return fl_workspace(g);
}

extern "C" int cxm_fluor_multislice_sample_c64(void* stream, void* state, const float* fluorescence,
                                               const float* dn, const float* phases, float* intensity,
                                               void* workspace, size_t workspace_bytes,
                                               const cxm_fluor_desc* d) {
  FlGeom g;
  int rc = fl_geom(d, g);
  if (rc) return rc;
  if (!state || !fluorescence || !dn || !phases || !intensity)
    return fail(CXM_ERR_INVALID, "null buffer");
  const int B = fl_bwd_batch(g);
  const size_t need = fl_workspace(g);
  if (!workspace || workspace_bytes < need)
    return fail(CXM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  float2* u = (float2*)state;
  float2* Pf = (float2*)workspace;
  float2* Pb = Pf + g.plane;
  float2* U2 = Pb + g.plane;
  float2* W = U2 + g.plane;
  float4* TF = (float4*)((char*)workspace + fl_factor_offset(g));
  const float2* Pfw = (const float2*)d->prop_fwd;
  const float2* Pbw = (const float2*)d->prop_bwd;
  if (!Pfw) {
    if ((rc = fl_propagator(d, g, d->thickness, Pf, s))) return rc;
    Pfw = Pf;
  }
  if (!Pbw) {
    if ((rc = fl_propagator(d, g, -d->thickness, Pb, s))) return rc;
    Pbw = Pb;
  }
  const float inv_n = (float)(1.0 / ((double)g.Ny * (double)g.Nx));
  const bool ded = fl_dedicated(g.Nx);
  FlRowArgs fr;
  memset(&fr, 0, sizeof(fr));
  fr.W = W;
  fr.tw = twiddles_for(g.Nx);
  fr.Ny = g.Ny;
  fr.h = g.h;
  fr.w = g.w;
  fr.pad_y = d->pad_y;
  fr.pad_x = d->pad_x;
  const long long sl = (long long)g.h * g.w;
  // the factors of slices s0 .. s0 + NB - 1 (at TF + (slice % NB) * sl)
  const int NB = fl_factor_block(g);
  auto factors = [&](int s0) -> int {
    const long long n = (long long)std::min(NB, g.D - s0) * sl;
    hipLaunchKernelGGL(fl_factors_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       dn + s0 * sl, fluorescence + s0 * sl, phases + s0 * sl, TF, n,
                       (float)d->thickness, (float)d->wavelength);
    CXM_LAUNCH_CHECK();
    return CXM_OK;
  };
  // forward, slice 0: thin sample + fluorescence, row FFT (state -> W)
  if (ded) {
    if ((rc = factors(0))) return rc;
    fr.src = u;
    fr.tf = TF;
    if ((rc = launch_fl_row(g.Nx, true, fr, s))) return rc;
  } else {
    PipeArgs a;
    memset(&a, 0, sizeof(a));
    Ops b(a);
    if ((rc = fl_slice_ops(b, a.pre, d, g, dn, fluorescence, phases, 0))) return rc;
    fl_row_common(a, g);
    a.in = u;
    a.out = W;
    a.seq = SEQ_F;
    if ((rc = launch_pipe_row(g.Nx, a, s))) return rc;
  }
  for (int i = 0; i < g.D; ++i) {
    if (i + 1 < g.D) {
      // column FFT * P(+t) * column IFFT (normalised for the 2D IFFT), in place
      ColArgs ca;
      memset(&ca, 0, sizeof(ca));
      ca.in = W;
      ca.out = W;
      ca.tw = twiddles_for(g.Ny);
      ca.H = Pfw;
      ca.in_plane_stride = ca.out_plane_stride = g.plane;
      ca.row_stride = g.Nx;
      ca.ncols = g.Nx;
      ca.n_planes = 1;
      ca.nz = 1;
      ca.nwv = 1;
      ca.in_len = ca.out_len = g.Ny;
      ca.kind = CXM_H_ARRAY;
      ca.scale = inv_n;
      if ((rc = launch_col_prop(g.Ny, ca, s))) return rc;
      // row IFFT, slice i+1's thin sample + fluorescence, row FFT, in place
      if (ded) {
        if ((i + 1) % NB == 0 && (rc = factors(i + 1))) return rc;
        fr.tf = TF + (long long)((i + 1) % NB) * sl;
        if ((rc = launch_fl_row(g.Nx, false, fr, s))) return rc;
        continue;
      }
      PipeArgs a;
      memset(&a, 0, sizeof(a));
      Ops b(a);
      if ((rc = fl_slice_ops(b, a.mid, d, g, dn, fluorescence, phases, i + 1))) return rc;
      fl_row_common(a, g);
      a.in = W;
      a.out = W;
      a.seq = SEQ_IF;
      if ((rc = launch_pipe_row(g.Nx, a, s))) return rc;
    } else {
      // last slice: U = column FFT * P(+t) = fft2 of the field after the slice
      PipeArgs a;
      memset(&a, 0, sizeof(a));
      Ops b(a);
      OpDesc* op = b.add(a.post, OP_CARRAY, 1, 0);
      if (!op) return fail(CXM_ERR_INVALID, "operator table overflow");
      op->arr0 = Pfw;
      op->arr_row_stride = g.Nx;
      op->arr_rows = g.Ny;
      op->arr_cols = g.Nx;
      fl_col_common(a, g);
      a.in = W;
      a.out = u;
      a.seq = SEQ_F;
      if ((rc = launch_pipe_col(g.Ny, a, s))) return rc;
    }
  }
  // backward sweep: the inverse column FFT of `planes` workspace planes with
  // |.|^2 of the cropped window accumulated into intensity planes from `first`
  auto bwd_cols = [&](int first, int planes) {
    PipeArgs c;
    memset(&c, 0, sizeof(c));
    fl_col_common(c, g);
    c.n_planes = planes;
    c.in = W;
    c.out = intensity + (long long)first * sl;
    c.seq = SEQ_I;
    c.scale = (float)(1.0 / (double)g.Ny);
    c.intensity = 2;
    c.iscale = (float)d->intensity_scale;
    c.out_plane_stride = sl;
    c.out_line_stride = g.w;
    c.out_off = d->pad_y;
    c.out_len = g.h;
    c.out_line_off = d->pad_x;
    c.out_line_len = g.w;
    return launch_pipe_col(g.Ny, c, s);
  };
  if (ded) {
    // B steps per launch pair: every step's U *= Pb and inverse row FFT in
    // one batched pass, then the column pass over the B planes
    FlBwdArgs ba;
    memset(&ba, 0, sizeof(ba));
    ba.Pb = Pbw;
    ba.W = W;
    ba.tw = twiddles_for(g.Nx);
    ba.Ny = g.Ny;
    ba.chunk = fl_bwd_chunk();
    ba.plane = g.plane;
    ba.scale = (float)(1.0 / (double)g.Nx);
    float2* cur = u;  // holds the carried spectrum
    for (int i0 = 0; i0 < g.D; i0 += B) {
      float2* next = cur == u ? U2 : u;
      ba.src = cur;
      ba.dst = next;
      ba.steps = std::min(B, g.D - i0);
      if ((rc = launch_fl_bwd_rows(g.Nx, ba, s))) return rc;
      if ((rc = bwd_cols(g.D - i0 - ba.steps, ba.steps))) return rc;
      cur = next;
    }
    // the carried spectrum back in the caller's state
    if (cur != u && hipMemcpyAsync(u, cur, g.plane * sizeof(float2), hipMemcpyDeviceToDevice,
                                   s) != hipSuccess)
      return fail(CXM_ERR_HIP, "copy of the carried spectrum failed");
    return CXM_OK;
  }
  for (int i = 0; i < g.D; ++i) {
    if ((rc = cxm_cmul_c64((void*)s, u, Pbw, u, 1, g.Ny, g.Nx, 1, 0))) return rc;
    PipeArgs r;
    memset(&r, 0, sizeof(r));
    fl_row_common(r, g);
    r.in = u;
    r.out = W;
    r.seq = SEQ_I;
    r.scale = (float)(1.0 / (double)g.Nx);
    if ((rc = launch_pipe_row(g.Nx, r, s))) return rc;
    if ((rc = bwd_cols(g.D - 1 - i, 1))) return rc;
  }
  return CXM_OK;
}
