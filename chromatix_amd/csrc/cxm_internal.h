// Host/device shared internals of libcxm: per-plane transfer-function
// parameters, kernel argument blocks, error plumbing and FFT-length dispatch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/cxm.h"
#include "cxm_common.h"
#include "cxm_params.h"

namespace cxm {

// ----- error handling -------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define CXM_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      return ::cxm::fail(CXM_ERR_HIP, std::string(#expr) + ": " +             \
                                          hipGetErrorString(_e));             \
  } while (0)
#define CXM_LAUNCH_CHECK()                                                    \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess)                                                     \
      return ::cxm::fail(CXM_ERR_HIP, std::string("kernel launch: ") +        \
                                          hipGetErrorString(_e));             \
  } while (0)

// ----- tuning options (cxm_set_option) ----------------------------------------
int option(int key);

// ----- optional kernel timing (HIP events on the launch stream) --------------
bool timing_on();
void* timing_start(hipStream_t s);
void timing_stop(void* start, hipStream_t s, const std::string& name);

// ----- twiddle tables (families: cxm_common.h) -------------------------------
// device pointer to the family table for length n on the current device
// (cxm_init must have run on it).
const float2* twiddles_for(int n);
int ensure_init();
// Static device workspace of the allocation-free entry points: its size, and
// its address on the current device (ensure_init must have run on it).
constexpr size_t kNoallocScratchBytes = 64ull << 20;
float2* noalloc_scratch();
int device_cus();  // compute units of the current device (cxm_init must have run on it)

constexpr int kMaxHP = 32;  // parameter sets carried by value per launch

// Tiled planes: strips of kTileW columns, each strip's rows contiguous, so
// that one 128-byte line holds 4 rows x 4 columns of complex64.
constexpr int kTileW = 4;

struct RowArgs {
  const float2* in;
  float2* out_c;
  float* out_i;
  const float2* tw;
  long long in_plane_stride, out_plane_stride;  // elements
  int in_row_stride, out_row_stride;            // elements
  int n_rows, rows_per_plane;
  int in_off, in_len;    // FFT index of input column 0, valid input columns
  int out_off, out_len;  // FFT indices [out_off, out_off+out_len) are stored
  float scale;
  int intensity;         // store: 0 complex, 1 |u|^2, 2 real part (float32 out_i)
  int in_real;           // input rows are float32 (imaginary part 0)
  // Tiled planes (strips of kTileW columns stored one after another, see
  // cxm_prop2d_kernels.h): element (y, x) at (x / 4) * strip_stride + y * 4 + x % 4.
  // Only with 4 rows per workgroup, rows_per_plane % 4 == 0 and the full
  // row range (in_off 0, in_len N, or out_off 0, out_len N).
  int in_tiled, out_tiled;
  long long in_strip_stride, out_strip_stride;  // elements
  int nt_in, nt_out;     // non-temporal loads of `in` / stores of the output
};

struct ColArgs {
  const float2* in;
  float2* out;
  const float2* tw;
  const float2* H;       // CXM_H_ARRAY
  long long in_plane_stride, out_plane_stride;
  int row_stride;        // elements between rows (Nx, or kTileW for tiled planes)
  long long strip_stride;    // elements between strips of C columns (0: C, untiled)
  long long h_strip_stride;  // CXM_H_ARRAY: elements between strips (0: C)
  int h_row_stride;          // CXM_H_ARRAY: elements between rows (0: ncols)
  int h_planes;              // CXM_H_ARRAY: planes of H (0 or 1: one for all input planes;
                             //   else input plane p uses H plane p % h_planes)
  long long h_plane_stride;  // CXM_H_ARRAY: elements between H planes
  int ncols;
  int n_planes;          // input planes in this launch
  int plane0;            // global index of the first input plane (param lookup)
  int nz;                // distances (parameter sets per wavelength)
  int z_zip;             // 1: plane b uses distance (b / nwv) % nz, one output plane
  int nwv;               // wavelengths cycling fastest over input planes
  int in_off, in_len;    // FFT rows [in_off, in_off+in_len) stored at rows [0, in_len)
  int out_off, out_len;  // FFT rows [out_off, out_off+out_len) written to rows [0, out_len)
  int kind, phase_mode, remove_evanescent, bandlimit;
  float scale;
  float na_r2;           // > 0: H also masked to |f|^2 <= na_r2 (fl((NA/lam)^2), samples.py:271-280)
  int nhp;
  int strip_major;       // block order: 0 plane-major, 1 strip-major (H shared by all planes)
  // optional device table, one HParam per output plane (global input plane
  // p, distance iz -> hp_dev[p * nout + iz]); replaces hp[] when non-null
  const HParam* hp_dev;
  HParam hp[kMaxHP];
};

static_assert(sizeof(ColArgs) <= 4096, "ColArgs exceeds the kernel-argument limit");

// Host builders.
int build_hparams(const cxm_prop2d_desc* d, int Ny, int Nx, HParam* out, int nmax);

// Kernel launchers (implemented per translation unit).
int launch_row(int N, int dir, const RowArgs& a, hipStream_t s);
int launch_col_prop(int N, const ColArgs& a, hipStream_t s);
// persistent column pass for one parameter set (cxm_prop2d_persist.hip);
// -1 when N has none
bool col_persist_available(int N);
int launch_col_persist(int N, const ColArgs& a, hipStream_t s);
bool fft_len_supported(long long n);

}  // namespace cxm
