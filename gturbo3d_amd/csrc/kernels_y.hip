// kernels_y.hip -- the y-axis pencil passes (the y stages of the 3D
// transforms behind fft_inv / fft_dir, src/fft_inv.f90:12, src/fft_dir.f90:12).
//
// Pass B (SIGN = +1) and pass D (SIGN = -1) transform every field of the
// work array in place, one block per (kx tile, z, field).  In pruned mode
// pass B reads only the ky rows that pass A wrote (the columns alive under
// the truncation) and pass D writes only those rows back.
#include "gt_config.hpp"
#include "gt_device.hpp"

namespace gt {

// FP64 twiddles loaded with the data (1) or by each stage (0).  Measured at
// 512^3 with 64-plane z chunks: 1 makes pass B 5% slower (the other passes'
// preloads pay; this one does not), so it is off.
#ifndef GT_Y_PRETW
#define GT_Y_PRETW 0
#endif
// FP32 work fields: the twiddle table copied into LDS behind the transform's
// rows at the block start, so every stage reads its R-1 twiddles with LDS
// latency instead of L2 latency.  Measured at 512^3: passes B and D -11%
// in FP32-bulk; with FP64 (one twiddle per stage) pass B +2%, so FP64 reads
// the global table.
template <typename W>
constexpr bool y_twlds() { return sizeof(W) == 4 && GT_TWLDS32; }
template <typename W, int N, int T>
constexpr size_t yfft_lds() {
  return (size_t)N * T * sizeof(cx<W>) + (y_twlds<W>() ? Plan<N>::twsize() * sizeof(cx<W>) : 0);
}

// PZ: the slots that mid_slot() prunes (kycut < 3N/8) are zeros known at
// compile time on input (SIGN > 0) or never stored on output (SIGN < 0).
template <typename W, int N, int T, int P, int SIGN, bool PZ = false>
__global__ __launch_bounds__(T* P) void k_yfft(KP kp, cx<W>* work, const cx<W>* __restrict__ tw, int nfields) {
  using LI = LayInterleaved<T, sizeof(cx<W>) == 8 && T == 8>;  // swizzled for 8 pencils of 8 bytes
  using F = Fft<W, N, P, 1, SIGN, LI>;
  constexpr int R0 = F::R0, RL = F::RL, M0 = F::M0, ML = F::ML;
  GT_DYN_LDS(cx<W>, lds);  // N T elements
  const int tid = threadIdx.x, p = tid % T, t = tid / T;
  const int kx = blockIdx.x * T + p;
  const int z = kp.z0 + blockIdx.y;
  const int f = blockIdx.z % nfields;
  work += (blockIdx.z / nfields) * kp.wstride;
  const bool on = kx <= kp.kxcut;
  // inverse: every load depends on Ky, so the extents come through scalar
  // loads; forward: Ky only guards the stores, and a vector load measured faster
  int ke = N >> 1;
  if (kp.pruned) {
    if constexpr (SIGN > 0) ke = tile_ext<T>(kp.kyext + blockIdx.x * T, p);
    else ke = on ? (int)kp.kyext[kx] : -1;
  }
  const int Ky = on ? ke : -1;
  const int ys = kp.nxp;
  // element n = t + c of the thread's pencil (c = P*m + r*N/R, a constant of
  // the slot) sits at tile + c*ys + lane: a start uniform over the block and
  // one 32-bit lane offset, instead of a 64-bit address per element
  cx<W>* tile = work + f * kp.fs + (long long)z * kp.ny * kp.nxp + blockIdx.x * T;
  const unsigned lane = p + t * ys;
  LI lay{p};

  cx<W> in[1][M0 * R0], out[1][ML * RL];
  // the stages after the first read their twiddles from this copy; the
  // first stage's barrier makes it visible
  cx<W>* twl = lds + N * T;
  if constexpr (y_twlds<W>()) tw_to_lds<W, N>(tw, twl, tid, T * P);
  static_for<0, M0 * R0>([&](auto qc) {
    constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
    cx<W> v = {W(0), W(0)};
    if constexpr (!(PZ && SIGN > 0 && mid_slot<N, P, P * m + r * F::NR0>())) {
      const int n = F::in_pos(t, m, r);
      bool ld = on && F::in_ok(t, m);
      if constexpr (SIGN > 0) ld = ld && keep_idx(n, N, Ky);
      if (ld) v = ld_row_p<(SIGN < 0 ? NT_D : NT_B)>(tile + (long long)(P * m + r * F::NR0) * ys, lane);
    }
    in[0][qs] = v;
  });
  if constexpr (sizeof(W) == 8 && GT_Y_PRETW) {
    // FP64: the twiddles are loaded with the data, so no stage waits for them
    typename F::Twr twr;
    F::load_tw(tw, t, twr);
    F::run(lay, lds, tw, twr, t, in, out);
  } else if constexpr (y_twlds<W>()) {
    F::run(lay, lds, twl, t, in, out);
  } else {
    F::run(lay, lds, tw, t, in, out);
  }
  if (!on) return;
  static_for<0, ML * RL>([&](auto qc) {
    constexpr int qs = decltype(qc)::value, m = qs / RL, r = qs % RL;
    if constexpr (!(PZ && SIGN < 0 && mid_slot<N, P, P * m + r * F::NRL>())) {
      if (!F::out_ok(t, m)) return;
      const int n = F::out_pos(t, m, r);
      if constexpr (SIGN < 0) {
        if (!keep_idx(n, N, Ky)) return;
      }
      st_row_p<(SIGN < 0 ? NT_D : NT_B)>(tile + (long long)(P * m + r * F::NRL) * ys, lane, out[0][qs]);
    }
  });
}

template <typename W, int N, int T>
static hipError_t yfft_nt(int sign, int nfields, const KP& kp, void* work, const void* tw, hipStream_t s) {
  constexpr int P = pencil_P(N);
  dim3 grid((kp.kxcut + T) / T, kp.zn, nfields * kp.nmem);
  cx<W>* w = (cx<W>*)work;
  const cx<W>* t = (const cx<W>*)tw;
  if constexpr (pz_len(N)) {
    if (kp.pruned && pz_cut(N, kp.kycut)) {
      if (sign > 0) k_yfft<W, N, T, P, +1, true><<<grid, T * P, yfft_lds<W, N, T>(), s>>>(kp, w, t, nfields);
      else k_yfft<W, N, T, P, -1, true><<<grid, T * P, yfft_lds<W, N, T>(), s>>>(kp, w, t, nfields);
      return hipGetLastError();
    }
  }
  if (sign > 0) k_yfft<W, N, T, P, +1><<<grid, T * P, yfft_lds<W, N, T>(), s>>>(kp, w, t, nfields);
  else k_yfft<W, N, T, P, -1><<<grid, T * P, yfft_lds<W, N, T>(), s>>>(kp, w, t, nfields);
  return hipGetLastError();
}

template <typename W, int N>
static hipError_t yfft_n(int sign, int nfields, const KP& kp, void* work, const void* tw, hipStream_t s) {
  if constexpr (alt_T(N) > 0) {
    if (kp.ty == alt_T(N)) return yfft_nt<W, N, alt_T(N)>(sign, nfields, kp, work, tw, s);
  }
  return yfft_nt<W, N, pencil_T(N)>(sign, nfields, kp, work, tw, s);
}

hipError_t launch_yfft(int prec, int sign, int nfields, const KP& kp, void* work, const void* tw, hipStream_t s) {
  switch (kp.ny) {
#define GT_CASE(N)                                                         \
  case N:                                                                  \
    return prec ? yfft_n<float, N>(sign, nfields, kp, work, tw, s)         \
                : yfft_n<double, N>(sign, nfields, kp, work, tw, s);
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace gt
