// Plain parameter structs shared by host launch code and device kernels.
#pragma once

namespace cxm {

// ----- per-plane transfer function parameters ------------------------------
// Everything the column kernel needs to evaluate H(ky, kx) for one output
// plane.  f32 members hold the reference's float32 intermediates computed in
// its operation order on the host; df pairs hold the accurate-mode
// constants.
struct HParam {
  float inv_dy, inv_dx;  // fl(1/dy), fl(1/dx)                (field.py:40)
  float ky0, kx0;        // kykx as f32
  float c_l2;            // ASM: fl((lam/n)^2); transfer: fl(fl(fl(-pi*lam)/n)*z)
  float c_ph;            // ASM: fl(2pi * fl(fl(|z|*n)/lam))
  float sy, sx;          // shift_yx as f32
  float by0, bym, bx0, bxm;  // band-limit rectangle (propagation.py:557-579)
  // accurate mode: df constants (turns)
  float ry_hi, ry_lo, rx_hi, rx_lo;  // 1/(Ny dy), 1/(Nx dx)
  float kyc_hi, kyc_lo, kxc_hi, kxc_lo;  // kykx
  float q_hi, q_lo;      // ASM: (lam/n)^2;   transfer: -lam z/(2n)
  float a_hi, a_lo;      // ASM: |z| n / lam
  int conj;              // ASM with z < 0
  int pad_;
};
static_assert(sizeof(HParam) % 8 == 0, "HParam must stay 8-byte aligned");

}  // namespace cxm
