// CPU check harness: evaluates the device transfer-function and pointwise
// operator arithmetic (cxm_hfunc.h, cxm_ops.h) and the host parameter
// builder (cxm_hparams.cpp) on the host, exposed through a C ABI for
// tests/test_cpu_check.py.  Test infrastructure only.
#include <string>

#include "cxm_hfunc.h"
#include "cxm_ops.h"

namespace cxm {
int build_hparams(const cxm_prop2d_desc* d, int Ny, int Nx, HParam* out, int nmax);
int fail(int code, const std::string&) { return code; }
}  // namespace cxm

using namespace cxm;

extern "C" int chk_hparam_size(void) { return (int)sizeof(HParam); }

extern "C" int chk_build_hparams(const cxm_prop2d_desc* d, int Ny, int Nx, HParam* out, int nmax) {
  return build_hparams(d, Ny, Nx, out, nmax);
}

// H over the whole grid in FFT order, interleaved complex (2 floats).
extern "C" void chk_eval_h(const HParam* hp, int Ny, int Nx, int kind, int mode, int rem_ev,
                           int bandlimit, float* out) {
  for (int kx = 0; kx < Nx; ++kx) {
    const XTerms xt = x_terms(*hp, kx, Nx, mode);
    for (int ky = 0; ky < Ny; ++ky) {
      const float2 h = eval_H(*hp, xt, ky, Ny, kind, mode, rem_ev, bandlimit, 1.0f);
      out[2 * ((long long)ky * Nx + kx)] = h.x;
      out[2 * ((long long)ky * Nx + kx) + 1] = h.y;
    }
  }
}

// One operator applied to ones over an (ny, nx) grid.
extern "C" void chk_apply_op(const OpDesc* op, const OpTable* tab, int plane, int ny, int nx,
                             float* out) {
  for (int iy = 0; iy < ny; ++iy)
    for (int ix = 0; ix < nx; ++ix) {
      const float2 v = apply_op(*op, *tab, make_float2(1.f, 0.f), plane, iy, ix, ny, nx);
      out[2 * ((long long)iy * nx + ix)] = v.x;
      out[2 * ((long long)iy * nx + ix) + 1] = v.y;
    }
}
