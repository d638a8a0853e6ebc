// Host construction of the per-plane transfer-function parameters (HParam):
// the reference's float32 intermediates computed in its operation order
// (compute_transfer_propagator / compute_asm_propagator,
// src/chromatix/functional/propagation.py:466-580) and the double-float
// constants of the accurate mode.  Plain C++ (no device code); compiled
// into libcxm and, for the CPU checks, on its own.
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/cxm.h"
#include "cxm_params.h"

namespace cxm {
int fail(int code, const std::string& msg);

static void split_df(double v, float& hi, float& lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}

// Band-limit rectangle of Matsushima & Shimobaba (2009), vectorised as in
// compute_asm_propagator (src/chromatix/functional/propagation.py:557-579),
// evaluated in float32 in the same order.
static void band_limit_axis(float shift, float df_axis, float extent, float z, float lam,
                            float& k0, float& kmax) {
  volatile float half = 1.0f / (2.0f * df_axis);
  float sp = shift + half, sn = shift - half;
  float z2 = z * z;
  float kp = (float)(1.0 / std::sqrt((double)(1.0f / (sp * sp) * z2 + 1.0f))) / lam;
  float kn = (float)(1.0 / std::sqrt((double)(1.0f / (sn * sn) * z2 + 1.0f))) / lam;
  float sgp = (shift + extent > 0.f) ? 1.f : (shift + extent < 0.f ? -1.f : 0.f);
  float sgn = (shift - extent > 0.f) ? 1.f : (shift - extent < 0.f ? -1.f : 0.f);
  k0 = 0.5f * (sgp * kp + sgn * kn);
  float kw = sgp * kp - sgn * kn;
  kmax = kw / 2.0f;
}

int build_hparams(const cxm_prop2d_desc* d, int Ny, int Nx, HParam* out, int nmax) {
  const int nwv = d->num_wv > 0 ? d->num_wv : 1;
  const int nz = d->num_z;
  if (nwv * nz > nmax) return fail(CXM_ERR_INVALID, "too many (wavelength, z) parameter sets");
  const float pi_f = (float)M_PI;
  for (int w = 0; w < nwv; ++w) {
    const double lam_d = d->wavelengths ? d->wavelengths[w] : d->wavelength;
    const double dy_d = d->spacing ? d->spacing[2 * w] : d->dy;
    const double dx_d = d->spacing ? d->spacing[2 * w + 1] : d->dx;
    for (int iz = 0; iz < nz; ++iz) {
      HParam h;
      memset(&h, 0, sizeof(h));
      const double z_d = d->z[iz];
      // float32 inputs exactly as the reference holds them
      const float lam = (float)lam_d, nf = (float)d->n, zf = (float)z_d;
      const float dyf = (float)dy_d, dxf = (float)dx_d;
      volatile float t;  // keep each rounding step observable (no contraction)
      h.inv_dy = 1.0f / dyf;
      h.inv_dx = 1.0f / dxf;
      h.ky0 = (float)d->ky;
      h.kx0 = (float)d->kx;
      h.sy = (float)d->shift_y;
      h.sx = (float)d->shift_x;
      if (d->kind == CXM_H_TRANSFER) {
        t = -pi_f * lam;
        t = t / nf;
        t = t * zf;
        h.c_l2 = t;
        split_df(-lam_d * z_d / (2.0 * d->n), h.q_hi, h.q_lo);
      } else {
        t = lam / nf;
        float ln = t;
        t = ln * ln;
        h.c_l2 = t;
        t = std::fabs(zf) * nf;
        t = t / lam;
        float an = t;
        t = (float)(2.0 * M_PI) * an;
        h.c_ph = t;
        double q = (lam_d / d->n);
        split_df(q * q, h.q_hi, h.q_lo);
        split_df(std::fabs(z_d) * d->n / lam_d, h.a_hi, h.a_lo);
        h.conj = z_d < 0 ? 1 : 0;
      }
      split_df(1.0 / (Ny * dy_d), h.ry_hi, h.ry_lo);
      split_df(1.0 / (Nx * dx_d), h.rx_hi, h.rx_lo);
      split_df(d->ky, h.kyc_hi, h.kyc_lo);
      split_df(d->kx, h.kxc_hi, h.kxc_lo);
      if (d->bandlimit) {
        const float ext_y = dyf * (float)Ny, ext_x = dxf * (float)Nx;
        const float dfy = 1.0f / ext_y, dfx = 1.0f / ext_x;
        band_limit_axis(h.sy, dfy, ext_y, zf, lam, h.by0, h.bym);
        band_limit_axis(h.sx, dfx, ext_x, zf, lam, h.bx0, h.bxm);
      }
      out[w * nz + iz] = h;
    }
  }
  return CXM_OK;
}

}  // namespace cxm
