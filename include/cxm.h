/*
 * cxm.h -- C ABI of libcxm.so, the MI355X (gfx950) implementation of the
 * chromatix FFT-propagation hot path.
 *
 * Conventions for every entry point:
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *     calls only enqueue work on it and never synchronise the host, with
 *     one exception: cxm_enclosing_circle_c64 returns its result in host
 *     memory and so waits for `stream` (its host part runs on the pruned
 *     candidates).
 *   - library state is per process: cxm_set_option values and the timing
 *     switch (cxm_timing_*) apply to every later call on every thread and
 *     device; set them before concurrent work starts.
 *   - buffers are device pointers owned by the caller; complex data is
 *     interleaved complex64 (float re, float im), C order.
 *   - scratch comes from the caller: query *_workspace_size(), pass a device
 *     buffer at least that large.  NULL workspace with size 0 is allowed
 *     when the query returned 0.
 *   - return 0 on success, a CXM_ERR_* code otherwise; cxm_last_error()
 *     gives a thread-local message.  No C++ exception crosses the ABI.
 *   - host arrays inside descriptors (z lists) are read during the call only.
 *
 * Citations name the reference function (chromatix v0.5.0, paths relative
 * to the chromatix repository root) whose arithmetic an entry point
 * replaces.
 */
#ifndef CXM_H_
#define CXM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CXM_OK 0
#define CXM_ERR_INVALID 1     /* bad descriptor or argument */
#define CXM_ERR_UNSUPPORTED 2 /* FFT length or layout not implemented */
#define CXM_ERR_HIP 3         /* HIP runtime error */
#define CXM_ERR_WORKSPACE 4   /* workspace too small */

/* Transfer-function kinds for cxm_prop2d_c64. */
#define CXM_H_NONE 0     /* plain fft2 -> ifft2 round trip (identity up to rounding) */
#define CXM_H_TRANSFER 1 /* Fresnel: exp(-i pi lambda/n z |f-kykx|^2)
                            (src/chromatix/functional/propagation.py:466-496) */
#define CXM_H_ASM 2      /* exact angular spectrum incl. evanescent handling,
                            shift_yx ramp and optional band limit
                            (src/chromatix/functional/propagation.py:499-580) */
#define CXM_H_ARRAY 3    /* caller-provided propagator, (Ny, Nx) complex64 in
                            FFT (ifftshifted) order, shared by all planes
                            (src/chromatix/functional/propagation.py:391-394) */

/* Phase evaluation of the on-the-fly transfer function. */
#define CXM_PHASE_REF32 0    /* float32 in the reference's operation order */
#define CXM_PHASE_ACCURATE 1 /* double-float phase in turns, exact mod-1 reduction */

/* Output kinds. */
#define CXM_OUT_COMPLEX 0   /* complex64 field */
#define CXM_OUT_INTENSITY 1 /* float32 |u|^2 */

/*
 * 2D Fourier-convolution propagation of a batch of planes:
 *   out = crop( ifft2( fft2( pad(in) ) * H(z_j) ) )      for each plane, each z_j
 * Replaces kernel_propagate's default branch with the propagator built by
 * compute_transfer_propagator / compute_asm_propagator
 * (src/chromatix/functional/propagation.py:168-252, 255-352, 391-394,
 * 466-580) and, with out_kind = INTENSITY, the `.intensity` of the result
 * (src/chromatix/core/field.py:663-665).
 *
 * in : batch x height x width complex64 (unpadded)
 * out: batch x num_z x out_h x out_w (z_zip = 0) or batch x out_h x out_w
 *      (z_zip = 1), out_h = height + 2 pad_y - 2 crop_y, complex64 or float32
 *      by out_kind.  Plane order: input plane major, distance minor (the
 *      reference's z-batched propagator broadcast against a single field
 *      gives the same (z, h, w) order).
 */
typedef struct cxm_prop2d_desc {
  int32_t batch;
  int32_t height, width;
  int32_t pad_y, pad_x;   /* zero padding per side (cval = 0) */
  int32_t crop_y, crop_x; /* crop per side of the padded grid; "same" mode = pad */
  int32_t num_z;          /* >= 1 */
  int32_t z_zip;          /* 0: every plane x every z (out planes = batch*num_z);
                             1: plane b uses z[b % num_z] (out planes = batch),
                             the reference's broadcast of a (Z,1,1) propagator
                             against a field whose innermost batch is Z */
  int32_t kind;           /* CXM_H_* */
  int32_t phase_mode;     /* CXM_PHASE_* */
  int32_t out_kind;       /* CXM_OUT_* */
  int32_t remove_evanescent;
  int32_t bandlimit;
  int32_t num_wv;         /* wavelengths; input plane p uses wavelength p % num_wv
                             (chromatic fields passed wavelength-planar) */
  double dy, dx;          /* sample spacing (used when spacing == NULL) */
  double wavelength;      /* same units as dx (used when wavelengths == NULL) */
  double n;               /* refractive index of the medium */
  double ky, kx;          /* kykx, subtracted from the frequency grid */
  double shift_y, shift_x;/* output-plane shift (ASM linear phase, band limit) */
  const double* z;        /* num_z propagation distances, host memory */
  const double* wavelengths; /* num_wv entries or NULL, host memory */
  const double* spacing;  /* num_wv x (dy, dx) or NULL, host memory */
  const void* propagator; /* CXM_H_ARRAY only: device (Ny, Nx) complex64 */
} cxm_prop2d_desc;

size_t cxm_prop2d_workspace_size(const cxm_prop2d_desc* d);
int cxm_prop2d_c64(void* stream, const void* in, void* out, void* workspace,
                   size_t workspace_bytes, const cxm_prop2d_desc* d);

/*
 * cxm_prop2d_c64 without a workspace: the passes use the library's static
 * device scratch array of cxm_noalloc_scratch_bytes() (64 MiB) per GPU, and
 * neither the caller nor the library allocates memory during the call.  The
 * chunk of planes per pass sequence is sized to fit the array; a grid one
 * plane of which (with the shared H table) does not fit is rejected with
 * CXM_ERR_UNSUPPORTED (2048^2 fits, 4096^2 does not).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (two streams, two
 * host threads) race and must be serialised by the caller.  No counterpart
 * in the reference: for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).
 */
size_t cxm_noalloc_scratch_bytes(void);
int cxm_prop2d_c64_noalloc(void* stream, const void* in, void* out, const cxm_prop2d_desc* d);

/*
 * cxm_prop2d_c64 with a device table of per-output-plane transfer-function
 * parameters instead of the descriptor's z list: output plane (p, iz) of
 * input plane p uses params[p * num_z + iz].  The table is built on the host
 * by cxm_prop2d_build_params (cxm_prop2d_param_bytes() bytes per plane,
 * any distance per plane) and copied to the device by the caller, so a
 * batch of requests with different distances runs as one launch sequence
 * and a captured graph replays with new distances by refreshing the table.
 * z and z_zip of the descriptor are ignored / must be 0.  No counterpart in
 * the reference (serving support).
 */
size_t cxm_prop2d_param_bytes(void);
int cxm_prop2d_build_params(const cxm_prop2d_desc* d, const double* z_planes, int64_t n_planes,
                            void* host_out);
int cxm_prop2d_c64_params(void* stream, const void* in, void* out, void* workspace,
                          size_t workspace_bytes, const cxm_prop2d_desc* d, const void* params);

/*
 * The propagator itself, materialised: out is num_wv x num_z x Ny x Nx
 * complex64 in FFT (ifftshifted) order, Ny = height + 2 pad_y (pad/crop and
 * batch fields of the descriptor are otherwise ignored).  Same arithmetic as
 * the on-the-fly evaluation inside cxm_prop2d_c64.  Replaces
 * compute_transfer_propagator / compute_asm_propagator
 * (src/chromatix/functional/propagation.py:466-580).
 */
int cxm_propagator_c64(void* stream, void* out, const cxm_prop2d_desc* d);

/*
 * Lens / far-field FFT family on a batch of planes (two passes: rows, then
 * column strips; shifts, chirps, norms, pupils folded into the passes).
 *
 *   CXM_OFFT_OPTICAL   optical_fft  (src/chromatix/functional/convenience.py:9-59):
 *                      z >= 0: fftshift(fft2(ifftshift(-i dy dx / L^2 * u)))
 *                      z <  0: conj(fftshift(fft2(ifftshift(-i L^2/(duy dux NyNx) conj(u)))))
 *                      with L^2 = lam z / n; optional circular pupil of diameter
 *                      pupil_d scaled by central_wavelength/lam before it (ff_lens with
 *                      NA, src/chromatix/functional/lenses.py:64-95).
 *   CXM_OFFT_TRANSFORM transform_propagate (src/chromatix/functional/propagation.py:34-82):
 *                      zero pad, input chirp exp(i pi |x|^2/L^2), 1j * optical_fft,
 *                      output chirp on the new grid, crop.
 *   CXM_OFFT_RAW_FWD / CXM_OFFT_RAW_INV  fft2 / ifft2 (utils/fft.py:8-35), with
 *                      `shift` = 1 for fftshift(f(ifftshift(x))).
 *
 * in : batch x height x width complex64; plane p uses wavelength p % num_wv.
 * out: batch x (height + 2 pad_y - 2 crop_y) x (width + 2 pad_x - 2 crop_x),
 *      complex64 or float32 |u|^2.
 */
#define CXM_OFFT_OPTICAL 0
#define CXM_OFFT_TRANSFORM 1
#define CXM_OFFT_RAW_FWD 2
#define CXM_OFFT_RAW_INV 3

typedef struct cxm_optfft_desc {
  int32_t batch;
  int32_t height, width;
  int32_t pad_y, pad_x;
  int32_t crop_y, crop_x;
  int32_t num_wv;
  int32_t phase_mode;         /* CXM_PHASE_* for the chirps */
  int32_t out_kind;           /* CXM_OUT_* */
  int32_t mode;               /* CXM_OFFT_* */
  int32_t shift;              /* RAW modes only */
  int32_t skip_initial_phase; /* TRANSFORM */
  int32_t skip_final_phase;   /* TRANSFORM */
  int32_t pupil;              /* OPTICAL: apply the circular pupil */
  int32_t adjoint;            /* 1: the adjoint of the call the descriptor describes: input
                                 = its output shape, output = its input shape (complex) */
  double z, n, pupil_d, central_wavelength;
  double wavelength;          /* used when wavelengths == NULL */
  double dy, dx;              /* used when spacing == NULL */
  double oy, ox;              /* grid origin, used when origin == NULL */
  const double* wavelengths;  /* num_wv, host */
  const double* spacing;      /* num_wv x (dy, dx), host */
  const double* origin;       /* num_wv x (oy, ox), host */
  const void* spec_mul;       /* RAW_FWD / RAW_INV, not adjoint, complex output:
                                 optional height x width complex64 array (device, FFT
                                 order) multiplied onto every output plane, at the
                                 output's fftshifted positions when shift is set --
                                 fft2 (shifted) times fftshift(spec_mul) in one pass */
} cxm_optfft_desc;

size_t cxm_optical_fft_workspace_size(const cxm_optfft_desc* d);
int cxm_optical_fft_c64(void* stream, const void* in, void* out, void* workspace,
                        size_t workspace_bytes, const cxm_optfft_desc* d);

/*
 * Scalable angular spectrum propagation, transform_propagate_sas
 * (src/chromatix/functional/propagation.py:85-165): pad by half the size on
 * each side, z >= 0: kernel_propagate with the precompensation dH then
 * transform_propagate; z < 0: the reverse order; crop.  Uses the
 * cxm_optfft_desc fields batch, height, width, num_wv, wavelength(s),
 * spacing, origin, z, n, phase_mode, out_kind, skip_*_phase (pads/crops are
 * implied).  Four passes (z >= 0) or five (z < 0).
 */
size_t cxm_sas_workspace_size(const cxm_optfft_desc* d);
int cxm_sas_c64(void* stream, const void* in, void* out, void* workspace,
                size_t workspace_bytes, const cxm_optfft_desc* d);

/*
 * The SAS precompensation itself, materialised (num_wv x Ny x Nx complex64,
 * FFT order) for the grid described by d (height/width = the padded grid):
 * compute_sas_precompensation (src/chromatix/functional/propagation.py:85-103).
 */
int cxm_sas_precompensation_c64(void* stream, void* out, const cxm_optfft_desc* d);

/* Gradient kinds of the adjoint entry points (cxm_psf4f_bwd, cxm_multislice_bwd). */
#define CXM_GRAD_UPSTREAM 0 /* cot = dL/dI of the output intensity, float32 */
#define CXM_GRAD_MSE 1      /* cot = target intensity; L = mean((I - target)^2) over all elements */
#define CXM_GRAD_COMPLEX 2  /* cot = dL/dRe(u) + i dL/dIm(u) of the complex output, complex64
                               (multislice only) */

/*
 * 4f widefield PSF stack: a point source defocused by z_j in front of an
 * objective (focal length f, numerical aperture NA, immersion index n),
 * the phase mask exp(i phase) in its pupil plane, a tube lens of focal
 * length f_tube and its back focal plane.  Per plane j:
 *   u_j = objective_point_source(shape, (dy, dx), wavelength, z_j, f, n, NA,
 *                                power, amplitude, offset)
 *   out_j = ff_lens(phase_change(u_j, phase), f_tube, n)   (.intensity with
 *           out_kind = INTENSITY)
 * Replaces Optical4FSystemPSF.__call__ (src/chromatix/systems/microscopes.py:334-355)
 * with objective_point_source (src/chromatix/functional/sources.py:122-209),
 * phase_change (src/chromatix/functional/phase_masks.py:29-68), ff_lens
 * (src/chromatix/functional/lenses.py:64-95) and .intensity
 * (src/chromatix/core/field.py:663-665).  Monochromatic, scalar.
 *
 * phase: height x width float32 (device); out: num_z x height x width
 * float32 or complex64.  Output spacing: lam f_tube / (n N dx) per axis.
 */
typedef struct cxm_psf4f_desc {
  int32_t height, width;  /* pupil-plane grid */
  int32_t num_z;          /* defocus planes, 1..32 */
  int32_t phase_mode;     /* CXM_PHASE_* for the defocus phase */
  int32_t out_kind;       /* forward output: CXM_OUT_INTENSITY or CXM_OUT_COMPLEX */
  int32_t normalize;      /* 1: scale every plane to total power `power` */
  int32_t grad_kind;      /* CXM_GRAD_* (backward only) */
  int32_t reserved;
  double dy, dx;          /* pupil-plane spacing */
  double wavelength, n;   /* wavelength, immersion refractive index */
  double f, NA;           /* objective focal length, numerical aperture */
  double f_tube;          /* tube-lens focal length (> 0) */
  double power, amplitude;
  double offset_y, offset_x; /* lateral point-source offset */
  const double* z;        /* num_z defocus distances, host memory */
  double phase_scale;     /* the phase mask enters as fl(phase * phase_scale), the
                             spectral modulation lam_0 / lam of phase_change
                             (phase_masks.py:105-124) for one wavelength of a
                             chromatic field; 0 means 1 */
} cxm_psf4f_desc;

size_t cxm_psf4f_workspace_size(const cxm_psf4f_desc* d);
int cxm_psf4f_fwd(void* stream, const float* phase, void* out, void* workspace,
                  size_t workspace_bytes, const cxm_psf4f_desc* d);

/*
 * Gradient of a loss of the PSF intensity stack with respect to the phase
 * mask: grad_phase (height x width float32) = dL/dphase, from dL/dI
 * (CXM_GRAD_UPSTREAM) or from a target stack for the mean-squared error
 * (CXM_GRAD_MSE).  Recomputes the forward; the adjoint of what JAX's
 * reverse mode computes through the same composition.
 */
int cxm_psf4f_bwd(void* stream, const float* phase, const float* cot, float* grad_phase,
                  void* workspace, size_t workspace_bytes, const cxm_psf4f_desc* d);
/*
 * The loss value's forward and its gradient in one sweep (jax.value_and_grad
 * of a loss of Optical4FSystemPSF's intensity, microscopes.py:334-355): the
 * passes of cxm_psf4f_bwd, whose column pass also stores the intensity stack
 * it forms (num_z x height x width float32, as cxm_psf4f_fwd with
 * CXM_OUT_INTENSITY gives it).
 */
int cxm_psf4f_value_and_grad(void* stream, const float* phase, const float* cot, float* grad_phase,
                             float* intensity, void* workspace, size_t workspace_bytes,
                             const cxm_psf4f_desc* d);

/*
 * objective_point_source (src/chromatix/functional/sources.py:122-209),
 * materialised: out is num_z x height x width complex64 (uses the
 * descriptor's grid, optics, z, power, amplitude and offset fields).
 */
int cxm_point_source_c64(void* stream, void* out, const cxm_psf4f_desc* d);

/*
 * phase_change (src/chromatix/functional/phase_masks.py:29-68):
 *   out[p, y, x, w] = in[p, y, x, w] * exp(i phase[y, x] * modulation[w])
 * for n_planes planes of height x width (x num_wv innermost) complex64;
 * modulation = the spectral modulation lam0 / lam (host, NULL = 1).
 * in == out is allowed.
 */
int cxm_phase_change_c64(void* stream, const void* in, const float* phase, void* out,
                         int64_t n_planes, int32_t height, int32_t width, int32_t num_wv,
                         const double* modulation);

/*
 * Vector-Jacobian product of cxm_phase_change_c64 (kind 0) and
 * cxm_thin_sample_c64 (kind 1), the reverse-mode rules JAX derives for
 * phase_change / thin_sample (src/chromatix/functional/phase_masks.py:29-68,
 * src/chromatix/functional/samples.py:85-134).  g is the cotangent of the
 * output (dL/dRe + i dL/dIm), y the forward output:
 *   grad_in = g conj(T)
 *   grad_a[y, x] = sum over planes and wavelengths of k_w Im(g conj(y))
 *   grad_b[y, x] = - sum k_w Re(g conj(y))
 * a = phase (k_w = modulation[w]) or dn (k_w = 2 pi thickness / lam_w),
 * b = absorption (kind 1).  params: modulation (kind 0, NULL = 1) or the
 * wavelengths (kind 1).  Output pointers may be NULL (not computed).  The
 * per-pixel sums run in a fixed order (deterministic).
 */
int cxm_modulation_vjp_c64(void* stream, int32_t kind, const void* g, const void* y,
                           const float* a, const float* b, void* grad_in, float* grad_a,
                           float* grad_b, int64_t n_planes, int32_t height, int32_t width,
                           int32_t num_wv, const double* params, double thickness);

/*
 * Chirp-z transform along one axis of a batch of height x width complex64
 * planes (Bluestein), czt (src/chromatix/utils/czt.py:7-52):
 *   y_k = post_k * scale * sum_{j<n} x_j a^{-j} w^{j k},  k = 0 .. m-1
 * axis 1: along rows (n = width; out batch x height x m);
 * axis 0: along columns (n = height; out batch x m x width).
 * post_k = 1, or with mod_n = N >= 0 the modulation w^{-N k} a^{N} of the
 * CZT branch of kernel_propagate (src/chromatix/functional/propagation.py:427-441).
 * Plane p uses parameter set p % num_param (a, w: num_param complex values
 * as (re, im) pairs, host memory).  One pass with an FFT of the next
 * supported length >= n + m - 1.
 */
typedef struct cxm_czt_desc {
  int32_t batch;
  int32_t height, width;
  int32_t axis;
  int32_t m;
  int32_t num_param;      /* 1..64 */
  int32_t mod_n;          /* -1: post = 1 */
  int32_t reserved;
  const double* a;
  const double* w;
  double scale;
} cxm_czt_desc;

size_t cxm_czt_axis_workspace_size(const cxm_czt_desc* d);
int cxm_czt_axis_c64(void* stream, const void* in, void* out, void* workspace,
                     size_t workspace_bytes, const cxm_czt_desc* d);

/*
 * Matrix-DFT zoom on MFMA: the chirp-z transform of every height x width
 * plane along both axes, y[ky][kx] = sum_{jy,jx} x[jy][jx] A_y[ky][jy]
 * A_x[kx][jx] with A_d[k][j] = a_d^-j w_d^(j k), times w_d^(-N_d k) a_d^N_d
 * scale_d when mod_n_d = N_d >= 0 (mod_n_d < 0: no modulation, scale_d
 * still applies) -- the values of a cxm_czt_axis_c64 pass along each axis.
 * Computed as two complex GEMMs (X A_x^T, then A_y (X A_x^T)) on FP32
 * matrix cores (v_mfma_f32_16x16x4f32); the matrices are built in FP64.
 * For small output windows this contraction (8 (m n^2 + m^2 n) flops per
 * plane) is faster than Bluestein's three FFTs per line.  Replaces cztn /
 * zoomed_fft (src/chromatix/utils/czt.py:55-124) and the CZT pair of
 * kernel_propagate (src/chromatix/functional/propagation.py:427-441) for
 * such windows.
 * in : batch x height x width complex64;  out: batch x m_y x m_x complex64.
 */
typedef struct cxm_zoom_desc {
  int32_t batch, height, width, m_y, m_x;
  int32_t mod_n_y, mod_n_x;   /* output modulation index N per axis, < 0: none */
  int32_t reserved;
  double a_y[2], w_y[2];      /* complex (re, im) CZT parameters along the height */
  double a_x[2], w_x[2];      /* along the width */
  double scale_y, scale_x;    /* per-axis output scales */
} cxm_zoom_desc;

size_t cxm_zoom_dft_workspace_size(const cxm_zoom_desc* d);
int cxm_zoom_dft_mfma_c64(void* stream, const void* in, void* out, void* workspace,
                          size_t workspace_bytes, const cxm_zoom_desc* d);

/*
 * Batched complex64 GEMM on FP32 matrix cores, C[z] = A[z] B[z] (row-major,
 * leading dimensions in elements, batch strides in elements, 0 = shared),
 * the contraction of cxm_zoom_dft_mfma_c64.  No counterpart in the
 * reference (its matrix products are XLA dots); exposed for callers with
 * their own DFT matrices.
 */
int cxm_cgemm_mfma_c64(void* stream, const void* A, const void* B, void* C, int32_t M, int32_t N,
                       int32_t K, int64_t lda, int64_t ldb, int64_t ldc, int64_t stride_a,
                       int64_t stride_b, int64_t stride_c, int32_t batch);

/*
 * The scaled-convolution branch of kernel_propagate (Yu et al. 2012; used
 * when output_dx or output_shape is given and use_czt is False,
 * src/chromatix/functional/propagation.py:444-458), one axis of the
 * separable form, on batch planes of height x width complex64 (the shifted
 * spectrum fftshift(H * fft2(ifftshift(u)))):
 *   y_j = out_dx exp(i pi x_j^2 / alpha)
 *         sum_k u_k (1/alpha) exp(i pi alpha f_k^2) kern[j + s - k]
 *   kern[t] = exp(-i pi alpha f_t^2) for 0 <= t < n, 0 otherwise,
 *   f_k = (k - n/2 rounded down) df, x_j = x0 + j out_dx, s = (n - 1) / 2
 *   rounded down (the "same" window of the full linear convolution),
 * n the length along the axis (axis 1: rows, n = width; axis 0: columns,
 * n = height); the output has the input's shape.  The chirps are evaluated
 * in FP64 on the device (the reference builds them in float32); one pass
 * with an FFT of the next supported length >= 2n - 1.  in == out is not
 * allowed.
 */
typedef struct cxm_yu_desc {
  int32_t batch;
  int32_t height, width;
  int32_t axis;
  double alpha;           /* out_dx / df */
  double df;              /* input frequency spacing along the axis */
  double out_dx;          /* output sample spacing along the axis */
  double x0;              /* first output coordinate (grid origin included) */
} cxm_yu_desc;

size_t cxm_yu_axis_workspace_size(const cxm_yu_desc* d);
int cxm_yu_axis_c64(void* stream, const void* in, void* out, void* workspace,
                    size_t workspace_bytes, const cxm_yu_desc* d);

/*
 * Dense matrices of the one-axis passes, for their reverse-mode rules: a
 * pass along an axis is a product with its matrix, so its adjoint (the
 * vector-Jacobian product torch's autograd asks for) is the product with
 * the conjugate transpose, one cxm_cgemm_mfma_c64 call.  The reference
 * gets these rules from JAX's automatic differentiation of czt and of the
 * scaled convolution (src/chromatix/utils/czt.py:7-52,
 * src/chromatix/functional/propagation.py:427-458); here they are explicit.
 * Element (r, c) goes to out[r * ld_r + c * ld_c], conjugated when conj != 0
 * (ld_r = 1, ld_c = rows stores the conjugate transpose row-major).
 *   cxm_czt_matrix_c64: m x n, A[k][j] = post_k scale a^-j w^(j k), the
 *     values of cxm_czt_axis_c64 with parameter set `param` (FP64, rounded
 *     to complex64).
 *   cxm_yu_matrix_c64: n x n, A[j][k] = out_dx exp(i pi x_j^2 / alpha)
 *     (1/alpha) exp(i pi alpha f_k^2) kern[j + s - k], the values of
 *     cxm_yu_axis_c64 (above).
 */
int cxm_czt_matrix_c64(void* stream, const cxm_czt_desc* d, int32_t param, void* out,
                       int64_t ld_r, int64_t ld_c, int32_t conj);
int cxm_yu_matrix_c64(void* stream, const cxm_yu_desc* d, void* out, int64_t ld_r,
                      int64_t ld_c, int32_t conj);

/*
 * The scalar part of high_na_ff_lens (src/chromatix/functional/lenses.py:98-156):
 *   out = in * defocus,  defocus = exp(i k sz f) / sz where sz != 0, else 0,
 *   sz = sqrt(max(1 - NA^2 |f_grid lambda / n|^2, 0)), k = 2 pi n / lambda,
 * on the centred frequency grid of height x width planes (spacings dy, dx),
 * evaluated in FP64 (the phase k sz f reaches 1e5 rad).  The zoomed FFT
 * that follows is two cxm_czt_axis_c64 passes.  in == out is allowed.
 */
int cxm_high_na_defocus_c64(void* stream, const void* in, void* out, int64_t n_planes,
                            int32_t height, int32_t width, double dy, double dx,
                            double wavelength, double n, double NA, double f);

/*
 * Elementwise product of n_planes planes of height x width complex64 with a
 * stack of b_planes planes (plane p uses b[p % b_planes]), b optionally
 * read fftshifted: out = a * fftshift(b).  The propagator multiply of the
 * CZT branch of kernel_propagate, fftshift(H * fft2(...)) =
 * fftshift(H) * fftshift(fft2(...)) (src/chromatix/functional/propagation.py:413-417).
 * a == out is allowed.
 */
int cxm_cmul_c64(void* stream, const void* a, const void* b, void* out, int64_t n_planes,
                 int32_t height, int32_t width, int32_t b_planes, int32_t shift_b);

/*
 * thin_sample (src/chromatix/functional/samples.py:85-134):
 *   out[p, y, x, w] = in[p, y, x, w] * exp(i 2 pi (dn[y, x] + i absorption[y, x]) t / lam_w)
 * for n_planes planes of height x width (x num_wv innermost) complex64, in
 * the reference's complex64 operation order; absorption may be NULL (0);
 * wavelengths: num_wv host values.  in == out is allowed.
 */
int cxm_thin_sample_c64(void* stream, const void* in, const float* dn, const float* absorption,
                        void* out, int64_t n_planes, int32_t height, int32_t width,
                        int32_t num_wv, const double* wavelengths, double thickness);

/*
 * Multislice propagation through a thick sample (multislice_thick_sample,
 * src/chromatix/functional/samples.py:137-282, with thin_sample :85-134):
 * pad the field and the slices by (pad_y, pad_x); for each slice i,
 *   u <- exp(i 2 pi (dn_i + i absorption_i) t / lam) * ifft2(H_t * fft2(u))
 * with H_t the angular-spectrum propagator over the slice thickness t
 * (kykx, remove_evanescent, bandlimit as compute_asm_propagator); then,
 * unless return_stack, propagate by -reverse_distance (ASM, optional NA mask
 * |f|^2 <= (NA/lam)^2) and crop.  Each slice costs two passes: the column
 * pass (FFT, H, inverse FFT) and a row pass (inverse FFT, sample, FFT).
 *
 * in: batch x height x width complex64; dn, absorption: depth x height x
 * width float32 (absorption may be NULL = 0); out: batch x height x width
 * (complex64 or float32 by out_kind), or with return_stack depth x batch x
 * height x width complex64 (the field after each slice).
 * checkpoints (optional, for the adjoint): the row-transformed state before
 * slices k, 2k, ... (k = checkpoint_every); size from
 * cxm_multislice_checkpoint_size.
 */
typedef struct cxm_multislice_desc {
  int32_t batch;
  int32_t height, width;      /* unpadded field and slices */
  int32_t depth;              /* slices */
  int32_t pad_y, pad_x;       /* zero padding per side */
  int32_t phase_mode;         /* CXM_PHASE_* */
  int32_t remove_evanescent;  /* per-slice propagator */
  int32_t bandlimit;          /* per-slice propagator */
  int32_t use_na;             /* NA mask on the final back-propagation */
  int32_t return_stack;
  int32_t out_kind;           /* CXM_OUT_* of the final field */
  int32_t checkpoint_every;   /* k >= 1 */
  int32_t grad_kind;          /* CXM_GRAD_* (adjoint) */
  double dy, dx, wavelength, n, thickness, NA, ky, kx;
  double reverse_distance;    /* the reference's default is thickness * depth / 2 */
} cxm_multislice_desc;

size_t cxm_multislice_workspace_size(const cxm_multislice_desc* d);
size_t cxm_multislice_checkpoint_size(const cxm_multislice_desc* d);
int cxm_multislice_fwd(void* stream, const void* in, const float* dn, const float* absorption,
                       void* out, void* checkpoints, void* workspace, size_t workspace_bytes,
                       const cxm_multislice_desc* d);

/*
 * Fourier convolution over the last two or three axes (fourier_convolution,
 * src/chromatix/ops/ops.py:9-86): out = ifftn(fftn(image) * fftn(kernel))
 * with both zero-padded to at least (depth + k_depth - 1) x (height +
 * k_height - 1) x (width + k_width - 1) (the next implemented FFT lengths:
 * the linear convolution, which the reference's larger transform shapes
 * also give), then
 *   mode 0 ("same"): along each axis the window [(k - 1) / 2, + n) of the
 *                    linear convolution (k kernel, n image length; integer
 *                    halves rounded down);
 *   mode 1 ("full"): the window [0, out_n) of the reference's transform
 *                    shape; entries beyond the transform used here are not
 *                    written (the caller zeroes them).
 * depth == k_depth == 1 is the two-axis convolution.  Image volume b uses
 * kernel volume b % k_planes.  real_input: image, kernel and out are float32
 * and out is the real part (the reference's rfftn / irfftn); otherwise
 * complex64.  Layouts are C-contiguous (batch, depth, height, width).
 */
typedef struct cxm_fconv_desc {
  int32_t batch;                     /* image volumes */
  int32_t depth, height, width;      /* image volume */
  int32_t k_planes;                  /* kernel volumes */
  int32_t k_depth, k_height, k_width;
  int32_t real_input;
  int32_t mode;                      /* 0 same, 1 full */
  int32_t out_depth, out_height, out_width;  /* full mode output volume */
  int32_t sum_batch;                 /* 1: out is one volume, the sum over the batch of
                                        the convolutions (real two-axis input with an even
                                        row length only: the sensor's plane sum,
                                        src/chromatix/functional/sensors.py:79-80, fused
                                        into the transforms) */
} cxm_fconv_desc;

size_t cxm_fconv_workspace_size(const cxm_fconv_desc* d);
int cxm_fconv_c64(void* stream, const void* image, const void* kernel, void* out, void* workspace,
                  size_t workspace_bytes, const cxm_fconv_desc* d);

/*
 * Sum over the middle axis of an (outer, n_sum, inner) float32 array into
 * (outer, inner), in index order (the sensor's reduce_axis,
 * src/chromatix/functional/sensors.py:79-80).
 */
int cxm_sum_axis_f32(void* stream, const float* in, float* out, int64_t outer, int64_t n_sum,
                     int64_t inner);

/*
 * Approximate shot noise (approximate_shot_noise, src/chromatix/ops/noise.py:10-16):
 *   out = max(image + sqrt(image) * normal, 0)
 * with the standard-normal samples supplied by the caller.  in == out is
 * allowed.
 */
int cxm_shot_noise_approx_f32(void* stream, const float* image, const float* normal, float* out,
                              int64_t n);

/*
 * Fluorescent multislice (fluorescent_multislice_thick_sample,
 * src/chromatix/functional/samples.py:285-441), one Monte-Carlo sample on
 * one padded plane.  `state` (Ny x Nx complex64, Ny = height + 2 pad_y,
 * Nx = width + 2 pad_x) holds the field the sample starts from -- the padded
 * input field for the first sample -- and on return the field the
 * reference carries into the next sample (the spectrum after the backward
 * sweep).  Forward, slice i = 0 .. depth-1:
 *   u = u * exp(i 2 pi dn_i t / lam) + f_i exp(i r_i)   (f, r, dn: depth x
 *       height x width float32, zero outside the centred unpadded window)
 *   u = ifft2(fft2(u) * P(+t))
 * then U = fft2(u); backward, i = 0 .. depth-1:
 *   U = U * P(-t);  intensity[depth-1-i] += intensity_scale * |crop(ifft2(U))|^2
 * P(z): the angular-spectrum propagator of the padded grid (remove_evanescent,
 * bandlimit as given), or the caller's prop_fwd / prop_bwd (Ny x Nx, FFT
 * order, device) when non-null.  intensity: depth x height x width float32,
 * accumulated.  The workspace (cxm_fluor_multislice_workspace_size bytes)
 * holds 3 + B padded planes, B = min(depth, planes in 512 MiB): the backward
 * sweep transforms B steps per batched launch; and the precomputed factors
 * (16 bytes per slice element) of up to 512 MiB of slices.
 */
typedef struct cxm_fluor_desc {
  int32_t height, width;      /* unpadded field and slices */
  int32_t depth;
  int32_t pad_y, pad_x;
  int32_t phase_mode;         /* CXM_PHASE_* of the propagators */
  int32_t remove_evanescent;
  int32_t bandlimit;
  double dy, dx, wavelength, n, thickness, ky, kx;
  double intensity_scale;     /* 1 / num_samples */
  const void* prop_fwd;       /* optional caller propagators */
  const void* prop_bwd;
} cxm_fluor_desc;

size_t cxm_fluor_multislice_workspace_size(const cxm_fluor_desc* d);
int cxm_fluor_multislice_sample_c64(void* stream, void* state, const float* fluorescence,
                                    const float* dn, const float* phases, float* intensity,
                                    void* workspace, size_t workspace_bytes,
                                    const cxm_fluor_desc* d);

/*
 * Gradient of a loss of the multislice output with respect to dn (and
 * absorption): reverse sweep with checkpointing.  Segments of k slices are
 * recomputed from the stored states (or from a forward recomputation when
 * checkpoints is NULL), their fields kept, and the adjoint runs backwards:
 * the cotangent of the output (from cot by grad_kind), the adjoint of the
 * back-propagation, then per slice the adjoint row pass (gradient
 * k Im(g conj(u)) and -k Re(g conj(u)), k = 2 pi t / lam, then conj of the
 * sample) and the adjoint column pass (conj(H_t) = H_{-t}).  Fields of the
 * batch are processed one after another and their gradients summed.
 * grad_dn, grad_absorption: depth x height x width float32 (grad_absorption
 * may be NULL).  Not available with return_stack.
 */
size_t cxm_multislice_bwd_workspace_size(const cxm_multislice_desc* d);
int cxm_multislice_bwd(void* stream, const void* in, const float* dn, const float* absorption,
                       const void* checkpoints, const void* cot, float* grad_dn,
                       float* grad_absorption, void* workspace, size_t workspace_bytes,
                       const cxm_multislice_desc* d);

/*
 * Modified Born series solve (src/chromatix/experimental/modified_born_series/solver.py:42-146)
 * of (curl curl - k0^2 eps) E = -k0^2 S for a scalar permittivity eps
 * (nx x ny x nz complex64) or a 3x3 permittivity tensor per voxel (eps_tensor;
 * the tensorial permittivity the reference lists as future work,
 * modified_born_series/README.md) and the pre-scaled source S (Source.field,
 * sample.py:79-85; nx x ny x nz x 3 complex64), with the bias (shift) and
 * scale of get_shift_and_scale chosen by the caller:
 *   cxm_mbs_precondition: prec_y = prec(S) / num_scale, num_scale = max |prec(S)|
 *                         (prec_y: 3 x nx x ny x nz component-planar; num_scale:
 *                         one device double)
 *   cxm_mbs_iterate:      `iterations` steps x <- x + prec_y - prec_forward(x) on the
 *                         planar x; residual_sq (device double) = |r|^2 of the last
 *                         step, r = x_new - x
 *   cxm_mbs_iterate_stop: the same steps with the stopping test of
 *                         jaxopt.FixedPointIteration (run by solve,
 *                         solver.py:136-142: after every step, stop once the
 *                         update norm |r| <= tol, or after maxiter steps) on
 *                         the device: status (4 device doubles, zero-initialised
 *                         by the caller except [3]) = |r|^2 of the last step,
 *                         steps performed, stopped (0/1), |prec_y|^2 (set by the
 *                         caller, used when rtol >= 0); the tolerance is
 *                         rtol |prec_y| when rtol >= 0, else tol; steps enqueued
 *                         after the stop leave x and status unchanged and cost
 *                         one early return per launch
 *   cxm_mbs_output:       E = x num_scale, nx x ny x nz x 3
 * Workspace: cxm_mbs_workspace_size (one field plus reduction scratch).
 */
typedef struct cxm_mbs_desc {
  int32_t nx, ny, nz;     /* grid including absorbing boundaries */
  int32_t eps_tensor;     /* 0: eps nx x ny x nz; 1: eps 3 x 3 x nx x ny x nz (eps_ab
                             at (3 a + b) voxels + i), V(x) = (s eps - kb) x per voxel */
  double spacing, k0;     /* grid spacing, vacuum wavenumber (same length unit) */
  double bias_re, bias_im;
  double scale_re, scale_im;
} cxm_mbs_desc;

size_t cxm_mbs_workspace_size(const cxm_mbs_desc* d);
int cxm_mbs_precondition(void* stream, const void* eps, const void* source, void* prec_y,
                         double* num_scale, void* workspace, size_t workspace_bytes,
                         const cxm_mbs_desc* d);
int cxm_mbs_iterate(void* stream, const void* eps, const void* prec_y, void* x, int32_t iterations,
                    double* residual_sq, void* workspace, size_t workspace_bytes,
                    const cxm_mbs_desc* d);
int cxm_mbs_iterate_stop(void* stream, const void* eps, const void* prec_y, void* x,
                         int32_t iterations, int32_t maxiter, double tol, double rtol, double* status,
                         void* workspace, size_t workspace_bytes, const cxm_mbs_desc* d);
int cxm_mbs_output(void* stream, const void* x, const double* num_scale, void* e,
                   const cxm_mbs_desc* d);

/*
 * BiCGStab on the same preconditioned system (solve(..., use_bicgstab=True),
 * solver.py:141-143, which calls jax.scipy.sparse.linalg.bicgstab(prec_forward,
 * prec_y, x0, maxiter, tol) without a preconditioner): x solves
 * prec_forward(x) = prec_y; stop when |r|^2 <= max(tol^2 |prec_y|^2, atol^2) or on
 * breakdown (k = -10: <rh, r> = 0; k = -11: alpha or omega = 0).
 *   cxm_mbs_bicgstab_init:    r = prec_y - prec_forward(x), rh = p = q = r,
 *                             rho = alpha = omega = 1, k = 0
 *   cxm_mbs_bicgstab_iterate: `iterations` steps on the device; steps after the
 *                             stop leave x and the state unchanged
 * state: cxm_mbs_bicgstab_state_size bytes (four fields and the FP64 scalars),
 * owned by the caller between calls.  status: 4 device doubles written by every
 * call: |r|^2, k, stopped (0/1), |prec_y|^2.  x: 3 x nx x ny x nz planar.
 */
size_t cxm_mbs_bicgstab_state_size(const cxm_mbs_desc* d);
int cxm_mbs_bicgstab_init(void* stream, const void* eps, const void* prec_y, const void* x,
                          void* state, size_t state_bytes, double tol, double atol, double* status,
                          void* workspace, size_t workspace_bytes, const cxm_mbs_desc* d);
int cxm_mbs_bicgstab_iterate(void* stream, const void* eps, void* x, void* state, size_t state_bytes,
                             int32_t iterations, double* status, void* workspace,
                             size_t workspace_bytes, const cxm_mbs_desc* d);

/*
 * Reverse-mode rule of the solve: the implicit differentiation that
 * solve(..., implicit_diff=True) gets from jaxopt (solver.py:136-142).  With
 * A = prec_forward and b = prec(S) (E = A^-1 b x num_scale / num_scale), the
 * caller solves the adjoint system A^H lam = Ebar -- cxm_mbs_iterate or the
 * BiCGStab entries on conj(eps) with the desc's bias and scale conjugated,
 * right-hand side Ebar (3 x nx x ny x nz planar) -- and this entry returns
 *   grad_eps[i]       = conj(s sum_c (conj(lam_c) w_c + conj(mu_c) E_c)),
 *   grad_source[i][c] = -conj(s) mu_c,
 * mu = G^H V^H lam, w = G V E + E - G (s S), s the desc's scale, G the Green's
 * function of the desc, V = s eps - kb; E, S, grad_source nx x ny x nz x 3
 * interleaved, lam planar.  grad_eps and grad_source may be NULL.  Gradients
 * follow the conjugate-Wirtinger convention of torch: for a real loss L,
 * grad = dL/dRe + i dL/dIm.  Scalar permittivity only (eps_tensor = 0).
 * Workspace: cxm_mbs_vjp_workspace_size (three fields).
 */
size_t cxm_mbs_vjp_workspace_size(const cxm_mbs_desc* d);
int cxm_mbs_vjp(void* stream, const void* eps, const void* source, const void* e, const void* lam,
                void* grad_eps, void* grad_source, void* workspace, size_t workspace_bytes,
                const cxm_mbs_desc* d);

/*
 * Smallest circle enclosing n complex64 values (the shift of
 * get_shift_and_scale, solver.py:10-39, which minimises max |eps - shift|):
 * out[0], out[1] = centre (re, im), out[2] = max |value - centre|.  The device
 * prunes the values inside the octagon of the 8 directional extremes, the
 * host runs Welzl's algorithm on the rest.  Synchronises `stream`.
 */
size_t cxm_enclosing_circle_workspace_size(int64_t n);
int cxm_enclosing_circle_c64(void* stream, const void* values, int64_t n, double* out,
                             void* workspace, size_t workspace_bytes);

/*
 * Intensity |u|^2 of n_pixels complex64 values, or for an innermost
 * wavelength axis of length num_wv the density-weighted sum over it:
 *   out[p] = sum_w density[w] |u[p*num_wv + w]|^2   (float32 out)
 * density: num_wv host values, or NULL for weight 1.
 * Replaces ScalarField.intensity / ChromaticScalarField.intensity
 * (src/chromatix/core/field.py:663-665, 735-740).
 */
int cxm_intensity_c64(void* stream, const void* u, void* out, int64_t n_pixels,
                      int32_t num_wv, const double* density);

/*
 * Vector-Jacobian product of cxm_intensity_c64: out[p, w] = 2 density[w]
 * g[p] u[p, w] (complex64), g = dL/dI (float32, n_pixels values).  The
 * reverse-mode rule of .intensity (src/chromatix/core/field.py:663-665,
 * 735-740) that JAX derives automatically.
 */
int cxm_intensity_vjp_c64(void* stream, const void* u, const float* g, void* out,
                          int64_t n_pixels, int32_t num_wv, const double* density);

/*
 * Per-plane sums of |u|^2 (float64 out, one per plane), optionally
 * density-weighted over an innermost wavelength axis of length num_wv
 * (plane_elems counts complex values including that axis).  Fixed
 * reduction order: bitwise reproducible.  The reduction behind
 * Field.power (src/chromatix/core/field.py:656-661, 728-733).
 */
size_t cxm_plane_sumsq_workspace_size(int64_t n_planes);
int cxm_plane_sumsq_c64(void* stream, const void* u, void* out, int64_t n_planes,
                        int64_t plane_elems, int32_t num_wv, const double* density,
                        void* workspace, size_t workspace_bytes);

/*
 * Plain copy of n complex64 values (device to device), with the access
 * width of the FFT passes.  No counterpart in the reference: the bandwidth
 * reference and the known-size kernel for scaling HBM traffic counters.
 */
int cxm_copy_c64(void* stream, const void* in, void* out, int64_t n);

/*
 * Optional kernel timing: with timing enabled every kernel launch is
 * bracketed by HIP events on its stream.  cxm_timing_collect() waits for
 * them and returns the number of distinct kernels; cxm_timing_get(i, ...)
 * reads kernel i's name, summed milliseconds and launch count.  No
 * counterpart in the reference (measurement support).
 */
int cxm_timing_enable(int on);
int cxm_timing_collect(void);
int cxm_timing_get(int i, char* name, int name_len, double* total_ms, int64_t* launches);

/*
 * Launch-shape tuning (0 = the built-in choice, which for 1024-point
 * passes over at most a few planes picks a shape with more waves).  No
 * counterpart in the reference (performance tuning).
 *   CXM_OPT_COL_VARIANT   alternative column-pass shape for 640-, 1024- and
 *                         2048-point columns (640 and 1024: a narrower
 *                         default for up to 4 planes)
 *   CXM_OPT_CHUNK_PLANES  planes per chunk of the propagation passes
 *                         (default: a workspace of at most 128 MiB)
 *   CXM_OPT_ROW_VARIANT   alternative row-pass shape for 1024-point rows
 *   CXM_OPT_PIPE_ROW_VARIANT  alternative shape of the 1024-point rows of the
 *                         operator pipelines (multislice, optics, SAS, CZT)
 *   CXM_OPT_H_TABLE       cxm_prop2d_c64 with one parameter set (one
 *                         distance and wavelength) for the whole batch
 *                         evaluates H once into the workspace and the column
 *                         pass reads it: 0 = from CXM_H_TABLE_MIN_PLANES
 *                         planes on, 1 = never, 2 = always.  The table holds
 *                         exactly the values the column pass would evaluate,
 *                         so the results are bit-identical either way.
 *   CXM_OPT_COL_ORDER     block order of the column pass of cxm_prop2d_c64:
 *                         0 = plane by plane, 1 = strip by strip over all
 *                         planes of a chunk
 *   CXM_OPT_TILES         cxm_prop2d_c64 keeps its intermediate planes in
 *                         strips of 4 columns (each strip's rows contiguous)
 *                         where the pass shapes allow it: 0 = yes, 1 = never
 *                         (row-major intermediates).  Same results either way.
 *   CXM_OPT_ROW_LOOP      rows of one wavefront (1024, 1280, 1536 points):
 *                         0 = rows read from caller memory go through a pass
 *                         whose workgroups loop over groups of 4 rows and
 *                         load the next group during the current transform
 *                         (as many workgroups as stay resident), tiled
 *                         intermediates through one group per workgroup;
 *                         1 = one group per workgroup everywhere; k > 1 =
 *                         the looping pass everywhere with k workgroups per
 *                         CU.  Same results.
 *   CXM_OPT_NT_IO         cxm_prop2d_c64, cache hints for the planes read
 *                         and written once: 0 = non-temporal (streaming)
 *                         loads of the input and stores of the output, so
 *                         that the intermediates keep the Infinity Cache
 *                         (measured +11% on 64 x 1024^2); 1 = no hints;
 *                         2 = streaming stores only; 3 = streaming loads
 *                         only.  Same results.
 *   CXM_OPT_COL_PERSIST   cxm_prop2d_c64 with one parameter set and
 *                         512/1024/2048-point columns: 0 = the persistent
 *                         column pass (H of a strip evaluated once per
 *                         workgroup, no H table), 1 = the per-plane column
 *                         pass (with the H table), n > 1 = the persistent
 *                         pass with n workgroups per CU
 *   CXM_OPT_COL_PERSIST_E 8: the 1024-point persistent column pass with 8
 *                         values per thread (default 16)
 *   CXM_OPT_COL_PERSIST_MODE  the 1024-point persistent column pass: 0 =
 *                         twiddles in LDS and the next item's values loaded
 *                         during the current transforms, 1 = neither,
 *                         2 = LDS twiddles only, 3 = the early load only
 *   CXM_OPT_FCONV_PAIR    cxm_fconv_c64 with real input over two axes: 0 = the
 *                         image and kernel planes transformed together as one
 *                         complex plane (kernel scaled by a power of two to the
 *                         image's norm) and their spectra separated at the
 *                         product (even row lengths), 1 = separate kernel
 *                         spectra.  Equal to float32 rounding.
 *   CXM_OPT_MBS_XGREEN    modified Born series: 0 = the x-axis FFT, the
 *                         Green's projection and the inverse x-axis FFT run
 *                         as one pass where the x length has one; 2 = as
 *                         three launches.  Equal to float32 rounding.
 *   CXM_OPT_MBS_ZFUSE     modified Born series fixed-point iteration: 0 = the
 *                         first z-axis FFT pass fused with the V multiply and
 *                         the last with the update and its residual partial
 *                         sums, where the z length has the fused passes;
 *                         1 = separate V, z-pass and update launches.  Same
 *                         iterates; the residual sum's order differs.
 *   CXM_OPT_MS_PASSES     multislice (cxm_multislice_fwd / _bwd): 0 = the
 *                         per-slice column passes read H(+t) / H(-t) from
 *                         tables built once per call, and the slice's row
 *                         pass (thin sample or its adjoint) is a dedicated
 *                         kernel where the row length has one (512, 1024,
 *                         2048); 1 = the operator-pipeline row passes with H
 *                         evaluated per bin; 2 = the dedicated row passes with
 *                         H evaluated per bin.  Equal to float32 rounding.
 *                         cxm_fluor_multislice_sample_c64: 0 or 2 = the
 *                         slice's row pass (thin sample and fluorescence) and
 *                         the backward sweep batched over its steps (each
 *                         workgroup replays the propagator multiplies in
 *                         registers, then inverse-transforms its steps'
 *                         rows; one column pass over the steps' planes) as
 *                         dedicated kernels where the row length has them
 *                         (512, 640, 1024, 1280, 2048); 1 = the operator
 *                         pipeline step by step.  Bit-identical (640-point
 *                         rows: with CXM_OPT_FL_ROW_E 20).
 *   CXM_OPT_MS_ROW_E      values per thread of the dedicated 1024-point slice
 *                         row pass: 0 = 4, or 2, 8, 16
 *   CXM_OPT_PSF_PASSES    cxm_psf4f_fwd / _bwd / _value_and_grad: 0 = the
 *                         generated source from per-call tables (the
 *                         plane-invariant factor, and the defocus chirp as
 *                         row x column factors) in dedicated row and
 *                         phase-gradient passes, where the row length has them
 *                         (512, 1024, 1280, 1536, 2048); 1 = the operator
 *                         pipeline with the source regenerated per element.
 *                         Equal to float32 rounding.
 *   CXM_OPT_PSF_GRAD_WAVES  dedicated phase-gradient pass: row transforms
 *                         per workgroup (0 = 1, or 2, 4), each summing every
 *                         W-th plane; the W sums are added in LDS.  Equal
 *                         to float32 rounding.
 *   CXM_OPT_CZT_PASSES    cxm_czt_axis_c64 / cxm_yu_axis_c64: 0 = the
 *                         chirp multiplies in registers around the two
 *                         transforms in dedicated row and column passes,
 *                         where the FFT length has them (rows 1024-5120,
 *                         columns 1024-3072); 1 = the operator pipeline.
 *                         Bit-identical.
 *   CXM_OPT_FL_BWD_CHUNK  cxm_fluor_multislice_sample_c64's batched
 *                         backward rows: steps per workgroup (0 = 8).
 *                         Bit-identical.
 *   CXM_OPT_FL_ROW_E      the same, 640-point rows: values per thread of
 *                         the dedicated row passes (0 = 10, 20 = the operator
 *                         pipeline's shape and bits).  Equal to float32
 *                         rounding.
 *   CXM_OPT_FL_FACTOR_BLOCK  the same: slices per block of precomputed
 *                         slice factors (0 = the slices that fit 512 MiB);
 *                         sizes the workspace like CXM_OPT_FL_BWD_BATCH.
 *                         Bit-identical.
 *   CXM_OPT_FL_BWD_BATCH  the same: steps per batched launch pair (0 = the
 *                         planes that fit 512 MiB); sizes the workspace, so
 *                         set it before cxm_fluor_multislice_workspace_size.
 *                         Bit-identical.
 *   CXM_OPT_GRAD_GROUPS   cxm_psf4f_bwd: 0 = the default number of plane
 *                         groups of the phase-gradient pass (4), 1..12 =
 *                         that many (each group sums its planes in
 *                         registers and writes one complex partial plane).
 *                         Equal to float32 rounding.
 */
#define CXM_OPT_COL_VARIANT 1
#define CXM_OPT_CHUNK_PLANES 2
#define CXM_OPT_ROW_VARIANT 3
#define CXM_OPT_PIPE_ROW_VARIANT 4
#define CXM_OPT_H_TABLE 5
#define CXM_OPT_COL_ORDER 6
#define CXM_OPT_TILES 7
#define CXM_OPT_ROW_LOOP 8
#define CXM_OPT_MBS_XGREEN 9
#define CXM_OPT_NT_IO 10
#define CXM_OPT_FCONV_PAIR 11
#define CXM_OPT_COL_PERSIST 12
#define CXM_OPT_COL_PERSIST_E 13
#define CXM_OPT_COL_PERSIST_MODE 14
#define CXM_OPT_GRAD_GROUPS 15
#define CXM_OPT_MBS_ZFUSE 16
#define CXM_OPT_MS_PASSES 17
#define CXM_OPT_MS_ROW_E 18
#define CXM_OPT_PSF_PASSES 19
#define CXM_OPT_PSF_GRAD_WAVES 20
#define CXM_OPT_CZT_PASSES 21
#define CXM_OPT_FL_BWD_CHUNK 22
#define CXM_OPT_FL_BWD_BATCH 23
#define CXM_OPT_FL_ROW_E 24
#define CXM_OPT_FL_FACTOR_BLOCK 25
#define CXM_H_TABLE_MIN_PLANES 4
int cxm_set_option(int key, int value);

/*
 * Separable resampling of a batch of float32 planes (batch x in_h x in_w ->
 * batch x out_h x out_w) by per-output tap tables:
 *   out[b, i, j] = sum_u sum_v wy[i][u] wx[j][v] in[b, y0[i] + u, x0[j] + v]
 * (taps outside the input are skipped), with an optional mask (in_h x in_w,
 * multiplying the input; or out_h x out_w multiplying the output when
 * mask_on_output).  The caller builds the tables: sum pooling, or the
 * weight matrices of jax.image.scale_and_translate divided by the scale
 * product, which replaces the sensor's resampler (init_plane_resample,
 * src/chromatix/ops/resample.py:10-128, called by basic_sensor,
 * src/chromatix/functional/sensors.py:14-93, and Microscope._process_psf,
 * src/chromatix/systems/microscopes.py:229-269, whose PSF taper
 * utils.sigmoid_taper is the input mask); with the transposed tables and
 * the mask on the output it is the resampler's adjoint.
 * Workspace: cxm_resample_workspace_size (the row-resampled intermediate).
 */
typedef struct cxm_resample_desc {
  int64_t batch;
  int32_t in_h, in_w, out_h, out_w;
  int32_t ky, kx;           /* taps per output row / column (table widths) */
  int32_t mask_on_output;   /* 0: mask multiplies the input; 1: the output */
  int32_t reserved;
  const int32_t* y_start;   /* out_h first input rows (device) */
  const float* wy;          /* out_h x ky weights (device) */
  const int32_t* x_start;   /* out_w first input columns (device) */
  const float* wx;          /* out_w x kx weights (device) */
  const float* mask;        /* optional (device) */
} cxm_resample_desc;

size_t cxm_resample_workspace_size(const cxm_resample_desc* d);
int cxm_resample_f32(void* stream, const float* in, float* out, void* workspace,
                     size_t workspace_bytes, const cxm_resample_desc* d);

/* Library state. */
int cxm_init(int device);            /* build twiddle tables on `device` */
const char* cxm_last_error(void);    /* thread-local text of the last error */
int cxm_version(void);               /* major*10000 + minor*100 + patch */
int cxm_fft_supported(int64_t n);    /* 1 if an FFT of length n is implemented */

#ifdef __cplusplus
}
#endif
#endif /* CXM_H_ */
