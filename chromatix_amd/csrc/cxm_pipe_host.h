// Host-side helpers shared by the pipeline entry points (cxm_pipe.hip,
// cxm_psf.hip): launch dispatch over the size groups and the operator-list
// builder.
#pragma once
#include <cstring>
#include <string>

#include "cxm_pipe_kernels.h"

namespace cxm {

#define CXM_PIPE_GROUP_DECL(K)                                            \
  int launch_pipe_row_g##K(int N, const PipeArgs& a, hipStream_t s);     \
  int launch_pipe_col_g##K(int N, const PipeArgs& a, hipStream_t s);     \
  int launch_pipe_phasegrad_g##K(int N, const PipeArgs& a, hipStream_t s, int groups);
CXM_PIPE_GROUP_DECL(0)
CXM_PIPE_GROUP_DECL(1)
CXM_PIPE_GROUP_DECL(2)
CXM_PIPE_GROUP_DECL(3)
CXM_PIPE_GROUP_DECL(4)
#undef CXM_PIPE_GROUP_DECL

// Row pass, column pass and phase-gradient row pass of length N (over the
// size groups); CXM_ERR_UNSUPPORTED for a length no group instantiates.
int launch_pipe_row_variant_1024(int v, const PipeArgs& a, hipStream_t s);

inline int launch_pipe_row(int N, const PipeArgs& a, hipStream_t s) {
  int rc;
  if (N == 1024) {
    // a few planes leave the default shape (4 rows of 64 threads per
    // workgroup) with ~1k waves for the whole GPU; 256 threads per row
    // quadruple that (measured 36 -> 23 us per plane in the multislice loop)
    int v = option(CXM_OPT_PIPE_ROW_VARIANT);
    if (v == 0 && (long long)a.n_planes * a.n_lines <= 4096) v = 2;
    if (v > 0 && (rc = launch_pipe_row_variant_1024(v, a, s)) >= 0) return rc;
  }
  if ((rc = launch_pipe_row_g0(N, a, s)) >= 0) return rc;
  if ((rc = launch_pipe_row_g1(N, a, s)) >= 0) return rc;
  if ((rc = launch_pipe_row_g2(N, a, s)) >= 0) return rc;
  if ((rc = launch_pipe_row_g3(N, a, s)) >= 0) return rc;
  if ((rc = launch_pipe_row_g4(N, a, s)) >= 0) return rc;
  return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(N));
}

inline int launch_pipe_col(int N, const PipeArgs& a, hipStream_t s) {
  int rc;
  if ((rc = launch_pipe_col_g0(N, a, s)) >= 0) return rc;
  if ((rc = launch_pipe_col_g1(N, a, s)) >= 0) return rc;
  if ((rc = launch_pipe_col_g2(N, a, s)) >= 0) return rc;
  if ((rc = launch_pipe_col_g3(N, a, s)) >= 0) return rc;
  if ((rc = launch_pipe_col_g4(N, a, s)) >= 0) return rc;
  return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(N));
}

// groups > 1: plane groups in parallel, partial sums at out + g * out_plane_stride
inline int launch_pipe_phasegrad(int N, const PipeArgs& a, hipStream_t s, int groups = 1) {
  int rc;
  if ((rc = launch_pipe_phasegrad_g0(N, a, s, groups)) >= 0) return rc;
  if ((rc = launch_pipe_phasegrad_g1(N, a, s, groups)) >= 0) return rc;
  if ((rc = launch_pipe_phasegrad_g2(N, a, s, groups)) >= 0) return rc;
  if ((rc = launch_pipe_phasegrad_g3(N, a, s, groups)) >= 0) return rc;
  if ((rc = launch_pipe_phasegrad_g4(N, a, s, groups)) >= 0) return rc;
  return fail(CXM_ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(N));
}

// ---- operator list builder ----------------------------------------------------
struct Ops {
  PipeArgs& a;
  int used = 0;
  explicit Ops(PipeArgs& args) : a(args) {}
  // Append an operator to `L` with `slots` parameter slots of `stride` float4s.
  OpDesc* add(OpList& L, int kind, int slots, int stride, int mode = CXM_PHASE_REF32) {
    if (L.n >= kMaxOps || used + slots * stride > kPTab) return nullptr;
    OpDesc& o = L.op[L.n++];
    memset(&o, 0, sizeof(o));
    o.kind = kind;
    o.mode = mode;
    o.pdiv = 1;
    o.pmod = slots;
    o.pbase = used;
    o.pstride = stride;
    used += slots * stride;
    return &o;
  }
  float4& p(const OpDesc* o, int slot, int j) { return a.tab.p[o->pbase + slot * o->pstride + j]; }
};

// v = hi + lo with hi = fl32(v)
inline void split(double v, float& hi, float& lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}

}  // namespace cxm
