// gt_internal.hpp -- shared definitions of the HIP library (kernels <-> host).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_core.hpp"

namespace gt {

// Per-launch description of the grid, the wavenumber tables and the
// truncation-driven extents, passed by value to every kernel.
//
// Storage: every spectral field is [NZ][NY][NXP] complex with the row pitch
// NXP = NX/2+1 rounded up to a multiple of 16, so every 16-wide tile of kx
// starts on a 256-byte boundary.  Pad columns kx > NX/2 are never read.
//
// Extents (pruned mode, when truncate = 1 and the state is truncated):
//   kxcut         largest kx index that any unmasked mode has
//   kyext[kx]     largest |ky| index with an unmasked mode in column kx
//   kzext[ky][kx] largest |kz| index unmasked at (kx, ky); -1 = none
// keep(n, N, K) <=> min(n, N - n) <= K.  In dense mode every keep() is true.
struct KP {
  int nx, ny, nz, nxh, nxp;
  long long fs;  // complex elements per field = nz*ny*nxp
  const double* fkx;
  const double* fky;
  const double* fkz;
  const double* fkxn;  // fkx^2 / fkxsmax (same expression as trunc.f90:14-16)
  const double* fkyn;
  const double* fkzn;
  double rtrunc2;
  int truncate;
  int pruned;
  int kxcut;
  int kycut, kzcut;  // largest kyext and kzext entries (ny/2, nz/2 in dense mode)
  const int16_t* kyext;
  const int16_t* kzext;
  double omega, dt, dt2, nu, mu, alpnu, alpmu;
  double scale;  // 1/(NX*NY*NZ), the fft_dir normalisation
  int visc_simple;  // alpnu == 1 and mu == 0: scra = nu*k2
  // with visc_simple, exp(-0.5*dt*nu*fk^2) per axis (nxh, ny, nz entries),
  // whose product is the integrating factor; nullptr otherwise
  const double* emx;
  const double* emy;
  const double* emz;
  int resident_blocks;  // grid size of the persistent kernels (a few blocks per CU)
  // batches: nmem independent members, member m's fields start m*bstride
  // (state arrays) or m*wstride (work array) elements further on
  int nmem;
  long long bstride, wstride;
  const double* omega_m;  // per-member omega, or nullptr for omega
  int tz, ty;  // pencils per block of the z-inverse and y passes (0: the default)
  int tze;     // pencils per block of the z-forward pass (0: the default)
  // planes z0 .. z0+zn-1 of the y and fused x passes (all of them by default)
  int z0, zn;
};

__device__ __forceinline__ bool keep_idx(int n, int N, int K) {
  const int a = (n <= (N >> 1)) ? n : N - n;
  return a <= K;
}

// One forcing increment (added to b at one spectral position).
struct FEnt {
  long long off;  // element offset within a field
  double d[6];    // (re, im) for bx, by, bz
};

// One step's forcing for the z-forward pass that adds it (the stage-2 pass
// fused with the next step's z-inverse pass): ent holds the step's entries,
// per for each member in turn; the entries of column key = (ky index) * nxp
// + (kx index) are idx[start[key] .. start[key + 1]), member-relative, in
// table order, so the columns of one kx tile are one contiguous range.
struct FAux {
  const FEnt* ent = nullptr;
  int per = 0;
  const int* start = nullptr;
  const int* idx = nullptr;
};

// pass identifiers for timing / byte accounting
enum PassId {
  PASS_A = 0,  // z inverse + curl/k^2 (3 -> 6 fields)
  PASS_B = 1,  // y inverse (6 fields)
  PASS_C = 2,  // x C2R, u x w, x R2C (6 -> 3 fields)
  PASS_D = 3,  // y forward (3 fields)
  PASS_E1 = 4, // z forward + Coriolis/invcurl/trunc + step1
  PASS_E2 = 5, // z forward + Coriolis/invcurl/trunc + step2
  PASS_F = 6,  // forcing scatter
  PASS_EA = 7, // E1 fused with pass A of stage 2 (bh kept in registers)
  PASS_EF = 8, // E2 + forcing fused with pass A of the next step
  NPASS = 9
};

}  // namespace gt
