// Supported FFT lengths and their launch shapes, split into instantiation
// groups (one translation unit each, compiled in parallel).
//
// X(N, E_row, E_col, C_col): FFT length, values per thread in row and column
// transforms, columns per column-pass workgroup.  Column workgroups hold an
// N x C_col strip of complex64 in LDS, plus one pad row every 16 rows for
// power-of-two N (<= 136 KiB of the 160 KiB per CU).
#pragma once

#define CXM_SIZES_G0(X)                                                       \
  X(2, 2, 2, 64) X(4, 4, 4, 64) X(8, 8, 8, 64) X(16, 16, 16, 64)              \
  X(32, 16, 16, 64) X(64, 16, 16, 32) X(128, 16, 16, 32) X(256, 16, 16, 16)   \
  X(5, 5, 5, 64) X(10, 10, 10, 64) X(20, 20, 20, 64) X(40, 20, 20, 64)        \
  X(80, 20, 20, 32) X(160, 20, 20, 32)
#define CXM_SIZES_G1(X) X(512, 16, 16, 16) X(1024, 16, 16, 4)
#define CXM_SIZES_G2(X) X(2048, 16, 16, 4) X(4096, 16, 32, 4) X(8192, 32, 32, 2)
#define CXM_SIZES_G3(X)                                                       \
  X(320, 20, 20, 16) X(640, 20, 20, 8) X(1280, 20, 20, 4) X(2560, 20, 20, 4) \
  X(5120, 20, 20, 2) X(10240, 20, 20, 1)

#define CXM_SIZES_G4(X)                                                       \
  X(3, 3, 3, 64) X(6, 6, 6, 64) X(12, 12, 12, 64) X(24, 12, 12, 64)          \
  X(48, 12, 12, 32) X(96, 12, 12, 32) X(192, 12, 12, 16) X(384, 24, 24, 16)  \
  X(768, 24, 24, 8) X(1536, 24, 24, 4) X(3072, 24, 24, 4)

#define CXM_FOR_EACH_SIZE(X) \
  CXM_SIZES_G0(X) CXM_SIZES_G1(X) CXM_SIZES_G2(X) CXM_SIZES_G3(X) CXM_SIZES_G4(X)

namespace cxm {
struct RowArgs;
struct ColArgs;
// Per-group launchers; return -1 when N is not in the group.
int launch_row_g0(int N, int dir, const RowArgs& a, hipStream_t s);
int launch_row_g1(int N, int dir, const RowArgs& a, hipStream_t s);
int launch_row_g2(int N, int dir, const RowArgs& a, hipStream_t s);
int launch_row_g3(int N, int dir, const RowArgs& a, hipStream_t s);
int launch_row_g4(int N, int dir, const RowArgs& a, hipStream_t s);
int launch_col_g0(int N, const ColArgs& a, hipStream_t s);
int launch_col_g1(int N, const ColArgs& a, hipStream_t s);
int launch_col_g2(int N, const ColArgs& a, hipStream_t s);
int launch_col_g3(int N, const ColArgs& a, hipStream_t s);
int launch_col_g4(int N, const ColArgs& a, hipStream_t s);
}  // namespace cxm
