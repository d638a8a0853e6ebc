// Column pass of cxm_prop2d_c64 for a call with one parameter set (one
// distance, one wavelength): a persistent grid over the chunk's
// (strip, plane) items.
//
// The items are ordered strip-major (all planes of strip 0, then strip 1,
// ...) and dealt in contiguous, equal ranges to the workgroups, so every
// workgroup meets one or two strips.  H of a strip -- the same for every
// plane of the call -- is evaluated into registers when the workgroup
// reaches the strip and reused for each of its planes: no H array in HBM,
// no table pass, and H costs one evaluation per strip and workgroup instead
// of one per plane.  The grid is as many workgroups as stay resident, so
// the chunk has no tail of partly filled workgroup waves.
//
// The values are those of the per-plane pass (same eval_H, same arguments,
// same scale), so results are bit-identical to it.
#include "cxm_prop2d_kernels.h"

namespace cxm {

// TWL: the length's twiddle table is copied to LDS once per workgroup and
// the transforms read it there (their waits then never wait for global
// loads).  PF: the next item's column values are loaded while the current
// item is transformed.
template <int N, int E, int C, int KIND, int MODE, bool TWL, bool PF>
__global__ void __launch_bounds__((ColCfg<N, E, C>::THREADS))
col_prop_persist_kernel(ColArgs a, int n_items) {
  constexpr int T = N / E;
  constexpr int THREADS = ColCfg<N, E, C>::THREADS;
  __shared__ float2 lds[lds_span<N, C>()];
  __shared__ float2 twl[TWL ? N : 1];
  const int c = threadIdx.x % C;
  const int ty0 = threadIdx.x / C;
  const int G = gridDim.x;
  const int i0 = (int)((long long)n_items * blockIdx.x / G);
  const int i1 = (int)((long long)n_items * (blockIdx.x + 1) / G);
  const long long sstride = a.strip_stride ? a.strip_stride : C;
  const HParam hp = a.hp[0];
  const bool na = a.na_r2 > 0.0f;
  if constexpr (TWL) {
    for (int i = threadIdx.x; i < N; i += THREADS) twl[i] = a.tw[i];
    block_sync();
  }
  // Strips whose rows all lie in the valid input (output) rows and whose
  // columns all exist -- every strip of an unpadded, uncropped call -- load
  // (store) without per-value guards, at one byte offset per thread plus a
  // scalar step per value.
  const bool rows_in = a.in_off == 0 && a.in_len >= N;
  const bool rows_out = a.out_off == 0 && a.out_len >= N;
  const unsigned rstep = (unsigned)(T * a.row_stride) * 8u;
  // column values of item `it` in the transform layout (zeros outside the
  // valid rows and columns)
  auto load = [&](float2 (&v)[E], int it, int ty) {
    const int strip = it / a.n_planes;
    const int p = it - strip * a.n_planes;
    const float2* src = a.in + p * a.in_plane_stride + (long long)strip * sstride;
    if (rows_in && strip * C + C <= a.ncols) {
      const unsigned off = (unsigned)(ty * a.row_stride + c) * 8u;
#pragma unroll
      for (int k = 0; k < E; ++k)
        v[k] = *(const float2*)((const char*)src + (off + (unsigned)k * rstep));
      return;
    }
    const bool ok = strip * C + c < a.ncols;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int ys = ty + k * T - a.in_off;
      v[k] = (ok && ys >= 0 && ys < a.in_len)
                 ? *(const float2*)((const char*)src + (unsigned)(ys * a.row_stride + c) * 8u)
                 : make_float2(0.f, 0.f);
    }
  };
  float2 h[E];
  float2 un[E];
  if constexpr (PF) {
    if (i0 < i1) load(un, i0, ty0);
  }
  int cur = -1;
  for (int it = i0; it < i1; ++it) {
    // Per-thread address terms of the transforms are the same in every
    // iteration; hoisted out of the loop they would all stay live across it.
    // Values the compiler cannot see through keep them in the loop body.
    int ty = ty0, tw_off = 0;
    asm volatile("" : "+v"(ty));
    asm volatile("" : "+s"(tw_off));
    const float2* tw = (TWL ? twl : a.tw) + tw_off;
    const int strip = it / a.n_planes;
    const int p = it - strip * a.n_planes;
    const int col = strip * C + c;
    const bool colok = col < a.ncols;
    float2 u[E];
    if constexpr (PF) {
#pragma unroll
      for (int k = 0; k < E; ++k) u[k] = un[k];
      if (it + 1 < i1) load(un, it + 1, ty);
    } else {
      load(u, it, ty);
    }
    if (strip != cur) {
      cur = strip;
      const int kx = colok ? col : 0;
      const XTerms xt = x_terms(hp, kx, a.ncols, MODE);
      const float fx2 = na ? na_fx2(hp, kx, a.ncols) : 0.0f;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        float2 v = eval_H(hp, xt, ty + k * T, N, KIND, MODE, a.remove_evanescent, a.bandlimit,
                          a.scale);
        if (na && !na_keep(hp, ty + k * T, N, fx2, a.na_r2)) v = make_float2(0.f, 0.f);
        h[k] = v;
        // one H evaluation per scheduling region (no register-spilling interleave)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    fft<N, E, -1, C, true>(u, lds + c, ty, tw);
#pragma unroll
    for (int k = 0; k < E; ++k) u[k] = cmul(u[k], h[k]);
    fft<N, E, 1, C, true>(u, lds + c, ty, tw);
    float2* dst = a.out + p * a.out_plane_stride + (long long)strip * sstride;
    if (rows_out && strip * C + C <= a.ncols) {
      const unsigned off = (unsigned)(ty * a.row_stride + c) * 8u;
#pragma unroll
      for (int k = 0; k < E; ++k) *(float2*)((char*)dst + (off + (unsigned)k * rstep)) = u[k];
    } else if (colok) {
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const int yo = ty + k * T - a.out_off;
        if (yo >= 0 && yo < a.out_len)
          *(float2*)((char*)dst + (unsigned)(yo * a.row_stride + c) * 8u) = u[k];
      }
    }
    // the next item's first exchange must not overwrite LDS a slower wave
    // still reads in this item's last one
    block_sync();
  }
}

// max_per_cu > 0 caps the resident workgroups per CU of the grid (the
// measured best for the instantiation); CXM_OPT_COL_PERSIST > 1 overrides.
template <int N, int E, int C, bool TWL, bool PF>
int launch_col_persist_t(const ColArgs& a, hipStream_t s, int max_per_cu = 0) {
  using Cfg = ColCfg<N, E, C>;
  const long long nstrips = (a.ncols + C - 1) / C;
  const long long items = nstrips * a.n_planes;
  if (items == 0) return CXM_OK;
  if (items > 0x7fffffffLL) return fail(CXM_ERR_INVALID, "column pass too large");
  // resident workgroups per CU of each instantiation, queried once
  static int resident[2][2] = {{0, 0}, {0, 0}};
  const int ki = a.kind == CXM_H_ASM ? 1 : 0, mi = a.phase_mode == CXM_PHASE_ACCURATE ? 1 : 0;
  int& rb = resident[ki][mi];
  const void* kern =
      a.kind == CXM_H_TRANSFER
          ? (mi ? (const void*)col_prop_persist_kernel<N, E, C, CXM_H_TRANSFER, CXM_PHASE_ACCURATE, TWL, PF>
                : (const void*)col_prop_persist_kernel<N, E, C, CXM_H_TRANSFER, CXM_PHASE_REF32, TWL, PF>)
          : (mi ? (const void*)col_prop_persist_kernel<N, E, C, CXM_H_ASM, CXM_PHASE_ACCURATE, TWL, PF>
                : (const void*)col_prop_persist_kernel<N, E, C, CXM_H_ASM, CXM_PHASE_REF32, TWL, PF>);
  if (rb == 0) {
    int n = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, Cfg::THREADS, 0);
    rb = n > 0 ? n : 1;
  }
  const int per_cu = option(CXM_OPT_COL_PERSIST) > 1 ? option(CXM_OPT_COL_PERSIST)
                     : max_per_cu > 0                 ? std::min(rb, max_per_cu)
                                                      : rb;
  const int grid = (int)std::min<long long>(items, (long long)per_cu * device_cus());
  void* t0 = timing_start(s);
  const int ni = (int)items;
#define CXM_PL(K, M)                                                                         \
  hipLaunchKernelGGL((col_prop_persist_kernel<N, E, C, K, M, TWL, PF>), dim3(grid),         \
                     dim3(Cfg::THREADS), 0, s, a, ni)
  if (a.kind == CXM_H_TRANSFER) {
    if (mi) CXM_PL(CXM_H_TRANSFER, CXM_PHASE_ACCURATE);
    else CXM_PL(CXM_H_TRANSFER, CXM_PHASE_REF32);
  } else {
    if (mi) CXM_PL(CXM_H_ASM, CXM_PHASE_ACCURATE);
    else CXM_PL(CXM_H_ASM, CXM_PHASE_REF32);
  }
#undef CXM_PL
  CXM_LAUNCH_CHECK();
  if (t0) timing_stop(t0, s, "col_prop_" + std::to_string(N));
  return CXM_OK;
}

// Lengths with a persistent column pass; its strip width is the default
// column-pass width of the length (the workspace layout depends on it).
bool col_persist_available(int N) { return N == 512 || N == 1024 || N == 2048; }

int launch_col_persist(int N, const ColArgs& a, hipStream_t s) {
  if (a.kind != CXM_H_TRANSFER && a.kind != CXM_H_ASM) return -1;
  const int m = option(CXM_OPT_COL_PERSIST_MODE);
  switch (N) {
    case 512: return launch_col_persist_t<512, 16, 16, true, false>(a, s);
    case 1024:
      if (option(CXM_OPT_COL_PERSIST_E) == 8) return launch_col_persist_t<1024, 8, 4, false, false>(a, s);
      if (m == 1) return launch_col_persist_t<1024, 16, 4, false, false>(a, s);
      if (m == 2) return launch_col_persist_t<1024, 16, 4, true, false>(a, s);
      if (m == 3) return launch_col_persist_t<1024, 16, 4, false, true>(a, s);
      // 16 planes of 1024^2: 60.6 us per launch at 2 workgroups per CU, 64.2
      // at 3 (LDS twiddles alone: 62.5 at 3; neither: 68.9; the per-plane
      // pass with the H table: 69.3 plus 8.2 for the table)
      return launch_col_persist_t<1024, 16, 4, true, true>(a, s, 2);
    case 2048: return launch_col_persist_t<2048, 16, 4, false, false>(a, s);
    default: return -1;
  }
}

}  // namespace cxm
