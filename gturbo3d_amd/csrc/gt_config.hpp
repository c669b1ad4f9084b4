// gt_config.hpp -- launch shapes of the pencil kernels per axis length.
//
// Pencil kernels (z and y passes) give a block TPB = T*P threads: T pencils
// (consecutive kx, so a row of T complex numbers is one contiguous HBM
// segment of T*16 bytes in FP64) times P threads per pencil.  P is the
// largest count that keeps every radix stage at a whole number of
// butterflies per thread where possible; guards handle the rest.
#pragma once

namespace gt {

__host__ __device__ constexpr int pencil_P(int n) {
  switch (n) {
    case 8: return 1;
    case 16: return 4;
    case 24: return 4;
    case 32: return 4;
    case 48: return 8;
    case 64: return 8;
    case 96: return 8;
    case 128: return 16;
    case 192: return 16;
    case 256: return 32;
    case 384: return 32;
    case 512: return 64;
    case 768: return 64;
    case 1024: return 128;
    default: return 0;
  }
}

#ifndef GT_PENCIL_T_LARGE
#define GT_PENCIL_T_LARGE 8
#endif
#ifndef GT_PENCIL_T_SMALL
#define GT_PENCIL_T_SMALL 16
#endif
__host__ __device__ constexpr int pencil_T(int n) { return n >= 384 ? GT_PENCIL_T_LARGE : GT_PENCIL_T_SMALL; }

// The other pencil count the autotuner may pick (0: none; the LDS tile
// T*N*16 bytes stays at or under 128 KB).
__host__ __device__ constexpr int alt_T(int n) {
  return n > 512 ? 0 : (pencil_T(n) == 16 ? 8 : 16);
}

// x passes: threads per row sequence (three sequences per block)
__host__ __device__ constexpr int row_P(int n) { return pencil_P(n); }

// lengths the library is compiled for (a build may restrict them with
// -D'GT_FOR_EACH_N(X)=X(512)' to time kernel variants at one size)
#ifndef GT_FOR_EACH_N
#define GT_FOR_EACH_N(X) \
  X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)
#endif

inline bool supported_n(int n) {
  switch (n) {
#define GT_CASE(N) case N:
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    return true;
    default:
      return false;
  }
}

}  // namespace gt
