// Row and column pass kernels of the batched 2D propagation (templates;
// instantiated per size group in cxm_inst_g*.hip).
//
// Row pass: G transforms of length N per workgroup, T = N/E threads each;
// zero-padded load (input column x lands at FFT index x + in_off), FFT in
// either direction, store of FFT indices [out_off, out_off + out_len) as
// complex64 or |u|^2.
//
// Column pass: C columns x T threads per workgroup; thread (ty, c) owns rows
// ty + k*T of column c and LDS holds the whole N x C strip.  Forward FFT,
// multiply by H (evaluated on the fly, or read from an array), inverse FFT,
// store of the requested rows.  All rows of the strip are read before any is
// written, so in == out is allowed when nz == 1.
#pragma once
#include <algorithm>
#include <string>

#include "cxm_fft.h"
#include "cxm_hfunc.h"
#include "cxm_internal.h"
#include "cxm_sizes.h"

namespace cxm {

template <int N, int E>
struct RowCfg {
  static constexpr int T = N / E;
  static constexpr int G = (T >= 256) ? 1 : 256 / T;
  static constexpr int THREADS = G * T;
  // LDS row pitch: the padded transform plus 8 slots, so that the four rows
  // of a tiled exchange fall on different banks
  static constexpr int PITCH = lds_span<N, 1>() + 8;
};

// Tiled planes (kTileW, cxm_internal.h): the column pass reads and writes
// its 4-column strips as whole 128-byte lines (a row-major plane gives it
// 32-byte pieces of lines, at four times the L2 requests), and a row pass
// moving 4 rows per workgroup exchanges whole lines through LDS.
// Lines of the 4 rows [r0, r0+4) of a tiled plane <-> the block's LDS rows.
// 16-byte pieces: piece c of the block = line c / 8 (strip), 8-byte pair
// (c % 8) -> row (c % 8) / 2, columns 4 * strip + 2 * (c % 2) + {0, 1}.
template <int N, int THREADS, int PITCH>
CXM_DEV void tiles_to_lds(const float2* base, long long strip_stride, float2* lds) {
  constexpr int PIECES = N / kTileW * 8;
  static_assert(PIECES % THREADS == 0, "tile exchange shape");
#pragma unroll
  for (int i = 0; i < PIECES / THREADS; ++i) {
    const int c = i * THREADS + threadIdx.x;
    const int q = c & 7, x = (c >> 3) * kTileW + 2 * (q & 1);
    const float4 w = *(const float4*)(base + (c >> 3) * strip_stride + 2 * q);
    float2* d = lds + (q >> 1) * PITCH + lds_at<N, 1>(x);  // x, x+1 share a 16-slot group
    d[0] = make_float2(w.x, w.y);
    d[1] = make_float2(w.z, w.w);
  }
}
template <int N, int THREADS, int PITCH>
CXM_DEV void lds_to_tiles(const float2* lds, float2* base, long long strip_stride) {
  constexpr int PIECES = N / kTileW * 8;
  static_assert(PIECES % THREADS == 0, "tile exchange shape");
#pragma unroll
  for (int i = 0; i < PIECES / THREADS; ++i) {
    const int c = i * THREADS + threadIdx.x;
    const int q = c & 7, x = (c >> 3) * kTileW + 2 * (q & 1);
    const float2* d = lds + (q >> 1) * PITCH + lds_at<N, 1>(x);
    *(float4*)(base + (c >> 3) * strip_stride + 2 * q) = make_float4(d[0].x, d[0].y, d[1].x, d[1].y);
  }
}

// Element xs of an input row: complex64, or float32 with imaginary part 0
// (streaming loads when nt_in).
CXM_DEV float2 row_load(const RowArgs& a, const float2* row, int xs) {
  if (a.in_real) {
    const float* q = (const float*)row + xs;
    return make_float2(a.nt_in ? __builtin_nontemporal_load(q) : *q, 0.f);
  }
  if (a.nt_in) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v v = __builtin_nontemporal_load((const f2v*)(row + xs));
    return make_float2(v.x, v.y);
  }
  return row[xs];
}

// Input row r of plane p (complex64 or float32 elements).
CXM_DEV const float2* row_ptr(const RowArgs& a, int p, int r) {
  const long long off = p * a.in_plane_stride + (long long)r * a.in_row_stride;
  return a.in_real ? (const float2*)((const float*)a.in + off) : a.in + off;
}

// Non-temporal (streaming) load and store of one complex64 value.
typedef float cxm_f2v __attribute__((ext_vector_type(2)));
CXM_DEV float2 nt_load2(const float2* p) {
  const cxm_f2v v = __builtin_nontemporal_load((const cxm_f2v*)p);
  return make_float2(v.x, v.y);
}
CXM_DEV void nt_store2(float2* p, float2 w) {
  cxm_f2v v;
  v.x = w.x;
  v.y = w.y;
  __builtin_nontemporal_store(v, (cxm_f2v*)p);
}

// Store of one output value: complex, |u|^2 or the real part.
CXM_DEV void row_store(const RowArgs& a, long long idx, float2 w) {
  if (a.intensity) {
    const float x = (a.intensity == 2) ? w.x : cabs2(w);
    if (a.nt_out) __builtin_nontemporal_store(x, a.out_i + idx);
    else a.out_i[idx] = x;
  } else if (a.nt_out) {
    nt_store2(a.out_c + idx, w);
  } else {
    a.out_c[idx] = w;
  }
}

template <int N, int E, int DIR>
__global__ void __launch_bounds__((RowCfg<N, E>::THREADS))
row_fft_kernel(RowArgs a) {
  using Cfg = RowCfg<N, E>;
  constexpr int T = Cfg::T;
  constexpr int G = Cfg::G;
  constexpr int PITCH = Cfg::PITCH;
  // transforms of at most one wave exchange through LDS without a block barrier
  constexpr bool SYNC = T > 64;
  __shared__ float2 lds[G * PITCH];
  const int g = threadIdx.x / T;
  const int t = threadIdx.x - g * T;
  const int row = blockIdx.x * G + g;
  const bool ok = row < a.n_rows;
  const int p = ok ? row / a.rows_per_plane : 0;
  const int r = ok ? row - p * a.rows_per_plane : 0;
  float2* st = lds + g * PITCH;
  float2 v[E];
  bool loaded = false;
  if constexpr (G == kTileW) {
    if (a.in_tiled) {  // block-uniform: n_rows and rows_per_plane are multiples of 4
      const int row0 = blockIdx.x * G;
      if (row0 < a.n_rows) {
        const int p0 = row0 / a.rows_per_plane;
        tiles_to_lds<N, Cfg::THREADS, PITCH>(
            a.in + p0 * a.in_plane_stride + (long long)(row0 - p0 * a.rows_per_plane) * kTileW,
            a.in_strip_stride, lds);
      }
      block_sync();
#pragma unroll
      for (int k = 0; k < E; ++k) v[k] = st[lds_at_split<N, 1>(t, k * T)];
      loaded = true;
    }
  }
  if (!loaded) {
    const float2* src = row_ptr(a, p, r);
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int xs = t + k * T - a.in_off;
      v[k] = (ok && xs >= 0 && xs < a.in_len) ? row_load(a, src, xs) : make_float2(0.f, 0.f);
    }
  }
  fft<N, E, DIR, 1, SYNC>(v, st, t, a.tw);
  if constexpr (G == kTileW) {
    if (a.out_tiled) {
      // each wave rewrites only its own row, which no other wave reads
#pragma unroll
      for (int k = 0; k < E; ++k) st[lds_at_split<N, 1>(t, k * T)] = cscale(v[k], a.scale);
      block_sync();
      const int row0 = blockIdx.x * G;
      if (row0 < a.n_rows) {
        const int p0 = row0 / a.rows_per_plane;
        lds_to_tiles<N, Cfg::THREADS, PITCH>(
            lds, a.out_c + p0 * a.out_plane_stride + (long long)(row0 - p0 * a.rows_per_plane) * kTileW,
            a.out_strip_stride);
      }
      return;
    }
  }
  if (!ok) return;
  const long long obase = p * a.out_plane_stride + (long long)r * a.out_row_stride;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int xo = t + k * T - a.out_off;
    if (xo >= 0 && xo < a.out_len) {
      row_store(a, obase + xo, cscale(v[k], a.scale));
    }
  }
}

// Row pass as a loop over groups of 4 rows (one row per wave) for rows of
// one wave (T == 64): the grid holds the workgroups the GPU keeps resident,
// and each workgroup loads its next group while it transforms the current
// one.  The twiddles of all stages are loaded once, before the first group:
// vector-memory counters complete in order, so twiddle loads issued after a
// prefetch would make the transform wait for the prefetched group.  Same
// results as row_fft_kernel.
template <int N, int E, int DIR>
__global__ void __launch_bounds__(256)
row_fft_loop_kernel(RowArgs a) {
  using Cfg = RowCfg<N, E>;
  static_assert(Cfg::G == 4 && Cfg::T == 64, "one row per wave, four rows per workgroup");
  static_assert(N / kTileW * 8 == 128 * E, "tile pieces per thread = E / 2");
  constexpr int T = 64, G = 4, PITCH = Cfg::PITCH;
  __shared__ float2 lds[G * PITCH];
  const int ngroups = (a.n_rows + G - 1) / G;
  AllTw<N, E, 1> twr;
  load_all_tw<N, E, DIR, 1>(twr, threadIdx.x & 63, a.tw);
  // next group's input: E values of this thread's row, or E/2 tile pieces
  float4 pf[E / 2];
  // Lane-dependent addresses are the same for every group; an opaque copy
  // of the thread index per group has them recomputed rather than held in
  // registers across the loop.
  auto fetch = [&](int grp, int tid) {
    const int g = tid >> 6, t = tid & 63;
    const int row0 = grp * G;
    if (a.in_tiled) {
      const int p0 = row0 / a.rows_per_plane;
      const float2* base =
          a.in + p0 * a.in_plane_stride + (long long)(row0 - p0 * a.rows_per_plane) * kTileW;
#pragma unroll
      for (int i = 0; i < E / 2; ++i) {
        const int c = i * 256 + tid;
        pf[i] = *(const float4*)(base + (c >> 3) * a.in_strip_stride + 2 * (c & 7));
      }
    } else {
      const int row = row0 + g;
      const bool ok = row < a.n_rows;
      const int p = ok ? row / a.rows_per_plane : 0;
      const int r = ok ? row - p * a.rows_per_plane : 0;
      const float2* src = a.in + p * a.in_plane_stride + (long long)r * a.in_row_stride;
#pragma unroll
      for (int k = 0; k < E; k += 2) {
        const int x0 = t + k * T - a.in_off, x1 = x0 + T;
        const bool in0 = ok && x0 >= 0 && x0 < a.in_len, in1 = ok && x1 >= 0 && x1 < a.in_len;
        float2 u0 = make_float2(0.f, 0.f), u1 = make_float2(0.f, 0.f);
        if (a.nt_in) {
          if (in0) u0 = nt_load2(src + x0);
          if (in1) u1 = nt_load2(src + x1);
        } else {
          if (in0) u0 = src[x0];
          if (in1) u1 = src[x1];
        }
        pf[k / 2] = make_float4(u0.x, u0.y, u1.x, u1.y);
      }
    }
  };
  int grp = blockIdx.x;
  if (grp < ngroups) fetch(grp, threadIdx.x);
  for (; grp < ngroups; grp += gridDim.x) {
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int g = tid >> 6, t = tid & 63;
    float2* st = lds + g * PITCH;
    float2 v[E];
    if (a.in_tiled) {
      block_sync();  // the previous group's LDS reads are done
#pragma unroll
      for (int i = 0; i < E / 2; ++i) {
        const int c = i * 256 + tid;
        const int q = c & 7, x = (c >> 3) * kTileW + 2 * (q & 1);
        float2* d = lds + (q >> 1) * PITCH + lds_at<N, 1>(x);
        d[0] = make_float2(pf[i].x, pf[i].y);
        d[1] = make_float2(pf[i].z, pf[i].w);
      }
      block_sync();
#pragma unroll
      for (int k = 0; k < E; ++k) v[k] = st[lds_at_split<N, 1>(t, k * T)];
    } else {
#pragma unroll
      for (int k = 0; k < E; k += 2) {
        v[k] = make_float2(pf[k / 2].x, pf[k / 2].y);
        v[k + 1] = make_float2(pf[k / 2].z, pf[k / 2].w);
      }
    }
    const int nxt = grp + gridDim.x;
    if (nxt < ngroups) fetch(nxt, tid);
    fft_pre<N, E, DIR, 1, false>(v, st, t, twr);
    const int row0 = grp * G;
    if (a.out_tiled) {
#pragma unroll
      for (int k = 0; k < E; ++k) st[lds_at_split<N, 1>(t, k * T)] = cscale(v[k], a.scale);
      block_sync();
      const int p0 = row0 / a.rows_per_plane;
      float2* base =
          a.out_c + p0 * a.out_plane_stride + (long long)(row0 - p0 * a.rows_per_plane) * kTileW;
#pragma unroll
      for (int i = 0; i < E / 2; ++i) {
        const int c = i * 256 + tid;
        const int q = c & 7, x = (c >> 3) * kTileW + 2 * (q & 1);
        const float2* d = lds + (q >> 1) * PITCH + lds_at<N, 1>(x);
        *(float4*)(base + (c >> 3) * a.out_strip_stride + 2 * q) =
            make_float4(d[0].x, d[0].y, d[1].x, d[1].y);
      }
      block_sync();  // before the next group's transforms reuse the rows
    } else {
      const int row = row0 + g;
      if (row < a.n_rows) {
        const int p = row / a.rows_per_plane;
        const int r = row - p * a.rows_per_plane;
        const long long obase = p * a.out_plane_stride + (long long)r * a.out_row_stride;
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const int xo = t + k * T - a.out_off;
          if (xo >= 0 && xo < a.out_len) {
            const float2 w = cscale(v[k], a.scale);
            if (a.intensity) a.out_i[obase + xo] = cabs2(w);
            else a.out_c[obase + xo] = w;
          }
        }
      }
    }
  }
}

template <int N, int E, int C>
struct ColCfg {
  static constexpr int T = N / E;
  static constexpr int THREADS = T * C;
};

// Block -> logical strip index, grouping consecutive strips on one XCD
// (workgroups are dealt round-robin over the 8 XCDs).
CXM_DEV int xcd_remap(int b, int nb) {
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// KIND: CXM_H_NONE, CXM_H_ARRAY, CXM_H_TRANSFER, CXM_H_ASM.  MODE: CXM_PHASE_*.
// MULTI: several output distances per input plane (runtime loop); the
// single-distance instantiation has no loop, so nothing loop-invariant is
// kept live across iterations (about 100 VGPRs instead of 170).
template <int N, int E, int C, int KIND, int MODE, bool MULTI>
__global__ void __launch_bounds__((ColCfg<N, E, C>::THREADS))
col_prop_kernel(ColArgs a) {
  constexpr int T = N / E;
  __shared__ float2 lds[lds_span<N, C>()];
  const int c = threadIdx.x % C;
  const int ty = threadIdx.x / C;
  const int nstrips = (a.ncols + C - 1) / C;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  // plane-major: each XCD works through whole planes; strip-major: each XCD
  // takes a range of strips of every plane (measured slower for 1024^2
  // batches, with H on the fly or from a table)
  const int p = a.strip_major ? L % a.n_planes : L / nstrips;
  const int col0 = (a.strip_major ? L / a.n_planes : L - p * nstrips) * C;
  const int col = col0 + c;
  const bool colok = col < a.ncols;
  // workgroup-uniform strip base + 32-bit lane offsets (SGPR-base addressing)
  const long long sstride = a.strip_stride ? a.strip_stride : C;
  const long long soff = (long long)(col0 / C) * sstride;
  const float2* src = a.in + p * a.in_plane_stride + soff;
  const int pg = a.plane0 + p;
  const int w = (a.nwv > 1) ? (pg % a.nwv) : 0;
  const int kx = colok ? col : 0;
  const int nout = MULTI ? (a.z_zip ? 1 : a.nz) : 1;
  // One register set per pass: with several distances the strip is
  // reloaded (from cache) and transformed again per distance, rather than
  // keeping the forward spectrum alive next to the working copy.
  for (int iz = 0; iz < nout; ++iz) {
    float2 u[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int ys = ty + k * T - a.in_off;
      u[k] = (colok && ys >= 0 && ys < a.in_len)
                 ? *(const float2*)((const char*)src + (unsigned)(ys * a.row_stride + c) * 8u)
                 : make_float2(0.f, 0.f);
    }
    fft<N, E, -1, C, true>(u, lds + c, ty, a.tw);
    if constexpr (KIND == CXM_H_TRANSFER || KIND == CXM_H_ASM) {
      const int zi = a.z_zip ? (pg / a.nwv) % a.nz : iz;
      const HParam* hpp = a.hp_dev ? a.hp_dev + ((long long)pg * nout + iz) : &a.hp[w * a.nz + zi];
      const HParam hp = *hpp;
      const XTerms xt = x_terms(hp, kx, a.ncols, MODE);
      const bool na = a.na_r2 > 0.0f;
      const float fx2 = na ? na_fx2(hp, kx, a.ncols) : 0.0f;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        float2 h = eval_H(hp, xt, ty + k * T, N, KIND, MODE, a.remove_evanescent,
                          a.bandlimit, a.scale);
        if (na && !na_keep(hp, ty + k * T, N, fx2, a.na_r2)) h = make_float2(0.f, 0.f);
        u[k] = cmul(u[k], h);
        // one H evaluation per scheduling region: the unrolled evaluations
        // are not interleaved into a register-spilling schedule
        __builtin_amdgcn_sched_barrier(0);
      }
    } else if constexpr (KIND == CXM_H_ARRAY) {
      const long long hss = a.h_strip_stride ? a.h_strip_stride : C;
      const int hrs = a.h_row_stride ? a.h_row_stride : a.ncols;
      const float2* hcol = a.H + (long long)(col0 / C) * hss + c +
                           (a.h_planes > 1 ? (long long)(pg % a.h_planes) * a.h_plane_stride : 0);
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const float2 h = colok ? hcol[(unsigned)((ty + k * T) * hrs)] : make_float2(0.f, 0.f);
        u[k] = cscale(cmul(u[k], h), a.scale);
      }
    } else {
#pragma unroll
      for (int k = 0; k < E; ++k) u[k] = cscale(u[k], a.scale);
    }
    fft<N, E, 1, C, true>(u, lds + c, ty, a.tw);
    if (colok) {
      float2* dst = a.out + ((long long)p * nout + iz) * a.out_plane_stride + soff;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const int yo = ty + k * T - a.out_off;
        if (yo >= 0 && yo < a.out_len)
          *(float2*)((char*)dst + (unsigned)(yo * a.row_stride + c) * 8u) = u[k];
      }
    }
  }
}

// ---- launchers --------------------------------------------------------------
template <int N, int E>
int launch_row_t(int dir, const RowArgs& a, hipStream_t s) {
  using Cfg = RowCfg<N, E>;
  const int blocks = (a.n_rows + Cfg::G - 1) / Cfg::G;
  if (blocks == 0) return CXM_OK;
  void* t0 = timing_start(s);
  if constexpr (Cfg::G == 4 && Cfg::T == 64) {
    // Looping pass for rows read from caller memory (HBM latency to hide);
    // tiled workspace input comes back faster than the loop gains.
    // (complex64 input and complex or |u|^2 output only: real input rows and
    // real-part stores take the plain pass)
    const int lo = option(CXM_OPT_ROW_LOOP);
    if (!a.in_real && a.intensity != 2 && (lo > 1 || (lo == 0 && !a.in_tiled))) {
      static int resident[2] = {0, 0};  // workgroups per CU, per direction
      int& rb = resident[dir < 0 ? 0 : 1];
      if (rb == 0) {
        int n = 0;
        if (dir < 0)
          (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, row_fft_loop_kernel<N, E, -1>, 256, 0);
        else
          (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, row_fft_loop_kernel<N, E, 1>, 256, 0);
        rb = n > 0 ? n : 1;
      }
      const int grid = std::min(blocks, (lo > 1 ? lo : rb) * device_cus());
      if (dir < 0)
        hipLaunchKernelGGL((row_fft_loop_kernel<N, E, -1>), dim3(grid), dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((row_fft_loop_kernel<N, E, 1>), dim3(grid), dim3(256), 0, s, a);
      CXM_LAUNCH_CHECK();
      if (t0) timing_stop(t0, s, std::string(dir < 0 ? "row_fft_fwd_" : "row_fft_inv_") + std::to_string(N));
      return CXM_OK;
    }
  }
  if (dir < 0)
    hipLaunchKernelGGL((row_fft_kernel<N, E, -1>), dim3(blocks), dim3(Cfg::THREADS), 0, s, a);
  else
    hipLaunchKernelGGL((row_fft_kernel<N, E, 1>), dim3(blocks), dim3(Cfg::THREADS), 0, s, a);
  CXM_LAUNCH_CHECK();
  if (t0) timing_stop(t0, s, std::string(dir < 0 ? "row_fft_fwd_" : "row_fft_inv_") + std::to_string(N));
  return CXM_OK;
}

template <int N, int E, int C>
int launch_col_t(const ColArgs& a, hipStream_t s) {
  using Cfg = ColCfg<N, E, C>;
  const long long nstrips = (a.ncols + C - 1) / C;
  const long long blocks = nstrips * a.n_planes;
  if (blocks == 0) return CXM_OK;
  if (blocks > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "column pass grid too large");
  const dim3 grid((unsigned)blocks), block(Cfg::THREADS);
  void* t0 = timing_start(s);
  const bool multi = !a.z_zip && a.nz > 1;
#define CXM_COL_LAUNCH(K, M)                                                                  \
  do {                                                                                        \
    if (multi)                                                                                \
      hipLaunchKernelGGL((col_prop_kernel<N, E, C, K, M, true>), grid, block, 0, s, a);      \
    else                                                                                      \
      hipLaunchKernelGGL((col_prop_kernel<N, E, C, K, M, false>), grid, block, 0, s, a);     \
  } while (0)
  if (a.kind == CXM_H_TRANSFER) {
    if (a.phase_mode == CXM_PHASE_REF32) CXM_COL_LAUNCH(CXM_H_TRANSFER, CXM_PHASE_REF32);
    else CXM_COL_LAUNCH(CXM_H_TRANSFER, CXM_PHASE_ACCURATE);
  } else if (a.kind == CXM_H_ASM) {
    if (a.phase_mode == CXM_PHASE_REF32) CXM_COL_LAUNCH(CXM_H_ASM, CXM_PHASE_REF32);
    else CXM_COL_LAUNCH(CXM_H_ASM, CXM_PHASE_ACCURATE);
  } else if (a.kind == CXM_H_ARRAY) {
    hipLaunchKernelGGL((col_prop_kernel<N, E, C, CXM_H_ARRAY, 0, false>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((col_prop_kernel<N, E, C, CXM_H_NONE, 0, false>), grid, block, 0, s, a);
  }
#undef CXM_COL_LAUNCH
  CXM_LAUNCH_CHECK();
  if (t0) timing_stop(t0, s, "col_prop_" + std::to_string(N));
  return CXM_OK;
}

}  // namespace cxm

// Defines launch_row_gK / launch_col_gK for one size group.
#define CXM_DEFINE_GROUP(K, SIZES)                                            \
  namespace cxm {                                                             \
  int launch_row_g##K(int N, int dir, const RowArgs& a, hipStream_t s) {      \
    switch (N) {                                                              \
      SIZES(CXM_ROW_CASE)                                                     \
      default: return -1;                                                     \
    }                                                                         \
  }                                                                           \
  int launch_col_g##K(int N, const ColArgs& a, hipStream_t s) {               \
    switch (N) {                                                              \
      SIZES(CXM_COL_CASE)                                                     \
      default: return -1;                                                     \
    }                                                                         \
  }                                                                           \
  }
#define CXM_ROW_CASE(NN, ER, EC, CC) case NN: return launch_row_t<NN, ER>(dir, a, s);
#define CXM_COL_CASE(NN, ER, EC, CC) case NN: return launch_col_t<NN, EC, CC>(a, s);
