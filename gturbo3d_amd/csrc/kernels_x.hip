// kernels_x.hip -- x-axis (contiguous row) passes and the forcing scatter.
//
// Every x kernel works on a pair of rows (y0, y0+1) at one z and packs the
// same field of the two rows into one complex sequence, z = row0 + i*row1,
// so a real transform of length NX costs half a complex one:
//   C2R: Z[n] = X0[n] + i X1[n] for n <= NX/2 and conj(X0[NX-n]) + i conj(X1[NX-n])
//        above, with the imaginary parts of X[0] and X[NX/2] dropped (the
//        C2R of fft_inv, src/fft_inv.f90:12); row0 = Re z, row1 = Im z.
//   R2C: Y = FFT(row0 + i row1), X0[k] = (Y[k] + conj Y[NX-k])/2,
//        X1[k] = (Y[k] - conj Y[NX-k])/(2i) (the D2Z of src/fft_dir.f90:12).
//
// Pass C (k_xfused): the six inverse transforms of u and w, the cross
//   product u x w of src/produ.f90:9-28 in LDS, and the three forward
//   transforms, so the physical-space fields never reach HBM.
// k_xc2r / k_xr2c: plain transforms for the velocity output
//   (src/main.f90:196-201) and the initial transform (src/main.f90:91-93).
#include "gt_config.hpp"
#include "gt_device.hpp"

namespace gt {

// One block per row pair, six groups of P threads (one wavefront for
// N <= 512).  Group w inverse-transforms field w in its own LDS row; with one
// wave per row the stage exchanges need only wavefront synchronisation.
// The products are formed in registers at the forward transform's input
// positions, groups 0-2 run the three forward transforms, and all six
// groups unpack and store.  Only the transforms' own stage exchanges use the
// swizzled layout; the rows handed between phases are stored plainly at
// g*N + n, since every such access is at positions consecutive across lanes
// (conflict-free, with constant offsets instead of per-element address
// arithmetic).
//
// PZ (pruned, kxcut < 3N/8): a thread slot whose positions all lie in
// [3N/8, 5N/8) holds only modes above kxcut, so its inverse input is zero
// (no loads) and its forward output is never unpacked (not stored, and the
// butterfly outputs feeding it are dropped).
#ifndef GT_XFUSED_REGU
#define GT_XFUSED_REGU 1
#endif
#ifndef GT_XFUSED_WAVES_PER_EU
#define GT_XFUSED_WAVES_PER_EU 1
#endif
// FP64 twiddles of the three-wave kernel held in registers (loaded once)
#ifndef GT_X_PRETW
#define GT_X_PRETW 1
#endif
// FP32 work fields in the three-wave kernel (twiddles from an LDS copy of
// the table) instead of the six-wave one
#ifndef GT_X3_FP32
#define GT_X3_FP32 0
#endif
// FP32 six-wave kernel: twiddles from an LDS copy of the table (0: from the
// global table, measured 2.5% faster at 512^3)
#ifndef GT_X_TWLDS32
#define GT_X_TWLDS32 0
#endif
template <typename W, int N, int P, bool PZ>
__global__ __launch_bounds__(6 * P) __attribute__((amdgpu_waves_per_eu(GT_XFUSED_WAVES_PER_EU)))
void k_xfused(KP kp, cx<W>* work, const cx<W>* __restrict__ tw) {
  using L = LaySwz<N>;
  constexpr int SY = (P <= 64) ? SYNC_WAVE : SYNC_BLOCK;
  using FI = Fft<W, N, P, 1, +1, L, SY>;
  using FF = Fft<W, N, P, 1, -1, L, SY>;
  GT_DYN_LDS(cx<W>, lds);  // 6 N elements
  const int tid = threadIdx.x, t = tid % P;
  // the group of a thread is uniform over its wavefront when P is a multiple
  // of the wavefront size: a scalar, so row starts are scalar too
  const int w = (P % 64 == 0) ? __builtin_amdgcn_readfirstlane(tid / P) : tid / P;
  const int y0 = 2 * blockIdx.x, z = kp.z0 + blockIdx.y;
  work += blockIdx.z * kp.wstride;
  const int r0 = (z * kp.ny + y0) * kp.nxp;
  const int r1 = r0 + kp.nxp;
  const int kc = kp.kxcut;
  const L lay{w, 0};

  // With the inverse's output positions equal to the forward's input
  // positions (same first and last radix), wave w keeps its u_w in registers
  // and computes product c = w + 2 (mod 3), whose first term is u_w w_c2:
  // three reads from LDS per position instead of four.
  constexpr bool REGU = GT_XFUSED_REGU && FI::RL == FF::R0 && FI::ML == FF::M0;
  cx<W> uo[1][FI::EL];
  // FP32: the twiddles from a copy of the table in LDS behind the six rows
  // (tw_to_lds), made visible by a block barrier before the first transform
  // (GT_X_TWLDS32; measured 2.5% slower here, so off)
  constexpr bool TWL = sizeof(W) == 4 && GT_TWLDS32 && GT_X_TWLDS32;
  const cx<W>* twt = TWL ? lds + 6 * N : tw;
  // inverse transform of field w (u_x, u_y, u_z, w_x, w_y, w_z), rows y0 and y0+1
  {
    cx<W> in[1][FI::E0], out[1][FI::EL];
    const cx<W>* src = work + w * kp.fs;
    static_for<0, FI::M0 * FI::R0>([&](auto qc) {
      constexpr int q = decltype(qc)::value, m = q / FI::R0, r = q % FI::R0;
      constexpr int C = P * m + r * FI::NR0;  // in_pos(t, m, r) = t + C
      cx<W> v = {W(0), W(0)};
      if constexpr (!(PZ && mid_slot<N, P, C>())) {
        if (FI::in_ok(t, m)) v = c2r_in_at<W, N, P, C>(src + r0, src + r1, t, kc);
      }
      in[0][q] = v;
    });
    if constexpr (TWL) {
      tw_to_lds<W, N>(tw, lds + 6 * N, tid, 6 * P);
      __syncthreads();
    }
    FI::run(lay, lds, twt, t, in, out);
    stage_sync<SY>();
#pragma unroll
    for (int m = 0; m < FI::ML; m++)
      if (FI::out_ok(t, m))
#pragma unroll
        for (int r = 0; r < FI::RL; r++) lds[w * N + FI::out_pos(t, m, r)] = out[0][m * FI::RL + r];
    if constexpr (REGU) {
#pragma unroll
      for (int q = 0; q < FI::EL; q++) uo[0][q] = out[0][q];
    }
  }
  __syncthreads();
  // u x w (src/produ.f90:14-16 with a = u, b = w), component w, both rows
  cx<W> fin[1][FF::E0], fout[1][FF::EL];
  const int comp = (w < 3) ? w : w - 3;
  // the product component this group forms and transforms (groups 0-2)
  const int pc = REGU ? (comp + 2) % 3 : comp;
  const int c1 = (pc + 1) % 3, c2 = (pc + 2) % 3;
#pragma unroll
  for (int m = 0; m < FF::M0; m++)
#pragma unroll
    for (int r = 0; r < FF::R0; r++) {
      cx<W> v = {W(0), W(0)};
      if (w < 3 && FF::in_ok(t, m)) {
        const int o = FF::in_pos(t, m, r);
        // (u x w)_c = u_c1 w_c2 - u_c2 w_c1 with c1 = c+1, c2 = c+2 (mod 3):
        // one code path with computed rows (three branches on c would keep
        // registers for all of them)
        // u_c1 is this wave's own field when REGU (c1 = w)
        const cx<W> u1 = REGU ? uo[0][m * FF::R0 + r] : lds[c1 * N + o];
        const cx<W> u2 = lds[c2 * N + o];
        const cx<W> w1 = lds[(3 + c1) * N + o], w2 = lds[(3 + c2) * N + o];
        v = {u1.x * w2.x - u2.x * w1.x, u1.y * w2.y - u2.y * w1.y};
      }
      fin[0][m * FF::R0 + r] = v;
    }
  __syncthreads();  // every product read is done before any forward transform writes LDS
  if (SY == SYNC_BLOCK || w < 3) {
    FF::run(lay, lds, twt, t, fin, fout);
    stage_sync<SY>();
    static_for<0, FF::ML * FF::RL>([&](auto qc) {
      constexpr int q = decltype(qc)::value, m = q / FF::RL, r = q % FF::RL;
      if constexpr (!(PZ && mid_slot<N, P, P * m + r * FF::NRL>())) {
        if (FF::out_ok(t, m)) lds[w * N + FF::out_pos(t, m, r)] = fout[0][q];
      }
    });
  }
  __syncthreads();
  // unpack the two rows of component c = w % 3 (groups c and c+3 share it)
  {
    const int c = comp;
    // group ca transformed product c (c = ca + 2 mod 3 when REGU) and left
    // it in its own rows: the transforms of groups 0-2 run unsynchronised,
    // each exchanging through its own rows only
    const int ca = REGU ? (c + 1) % 3 : c;
    cx<W>* d0 = work + c * kp.fs + r0;
    cx<W>* d1 = d0 + kp.nxp;
    for (int k = t + P * (w / 3); k <= kc; k += 2 * P) {
      const int kk = (k == 0) ? 0 : N - k;
      const cx<W> Y = lds[ca * N + k], Yc = lds[ca * N + kk];
      st_row_p<NT_C>(d0, (unsigned)k, cx<W>{(Y.x + Yc.x) * W(0.5), (Y.y - Yc.y) * W(0.5)});
      st_row_p<NT_C>(d1, (unsigned)k, cx<W>{(Y.y + Yc.y) * W(0.5), (Yc.x - Y.x) * W(0.5)});
    }
  }
}

// Pass C with three wavefronts per row pair (FP64, one wavefront per
// transform, first and last radix equal).  Wave c inverse-transforms u_c and
// w_c (its component of both fields, both loads issued first) and keeps them
// in registers at the forward transform's input positions, then forms
//   P_{c+2} = u_c w_{c+1} - u_{c+1} w_c      (src/produ.f90:14-16)
// from its neighbour's fields in two LDS rounds of one row per wave (u_c
// out, Q = -u_{c+1} w_c; w_c out, P = Q + u_c w_{c+1}), forward-transforms P
// in its own row and unpacks it.  Against the six-wave kernel above no wave
// waits through another's transform, and a block needs 3 rows of LDS
// instead of 6: measured at the 512^3 chunk shape, 0.184 against 0.228 ms
// per 64 planes, with the same bits.  FP32 work fields keep the six-wave
// kernel, which is faster there (its passes are bound by VALU issue).
// Rows shorter than a wavefront (P < 64 threads per transform): each
// wavefront holds component c of G = 64 / P row pairs, one transform per P
// lanes, and the block's LDS row of (c, pair g) is c G + g.
template <typename W, int N, int P, bool PZ>
__global__ __launch_bounds__(3 * 64) void k_xfused3(KP kp, cx<W>* work, const cx<W>* __restrict__ tw) {
  using L = LaySwz<N>;
  constexpr int SY = SYNC_WAVE;
  static_assert(64 % P == 0, "whole transforms in each wavefront");
  constexpr int G = 64 / P;  // row pairs per block
  using FI = Fft<W, N, P, 1, +1, L, SY>;
  using FF = Fft<W, N, P, 1, -1, L, SY>;
  static_assert(FI::RL == FF::R0 && FI::ML == FF::M0, "inverse outputs sit at the forward inputs");
  GT_DYN_LDS(cx<W>, lds);  // 3 G N elements
  const int tid = threadIdx.x, t = tid % P;
  const int c = __builtin_amdgcn_readfirstlane(tid / 64);  // component of this wavefront
  const int g = (tid % 64) / P;                              // row pair of this transform
  const int cn = (c + 1) % 3;
  const int y0 = 2 * (blockIdx.x * G + g), z = kp.z0 + blockIdx.y;
  work += blockIdx.z * kp.wstride;
  const int r0 = (z * kp.ny + y0) * kp.nxp;
  const int r1 = r0 + kp.nxp;
  const int kc = kp.kxcut;
  const int row = c * G + g, rown = cn * G + g;  // LDS rows of this and the next component
  const L lay{row, 0};
  auto load = [&](const cx<W>* src, cx<W> (&in)[1][FI::E0]) {
    static_for<0, FI::M0 * FI::R0>([&](auto qc) {
      constexpr int q = decltype(qc)::value, m = q / FI::R0, r = q % FI::R0;
      constexpr int C = P * m + r * FI::NR0;  // in_pos(t, m, r) = t + C
      cx<W> v = {W(0), W(0)};
      if constexpr (!(PZ && mid_slot<N, P, C>())) {
        if (FI::in_ok(t, m)) v = c2r_in_at<W, N, P, C>(src + r0, src + r1, t, kc);
      }
      in[0][q] = v;
    });
  };
  // inverse transforms of u_c and w_c (rows y0 and y0+1 packed)
  cx<W> uo[1][FI::EL], wo[1][FI::EL];
  typename FI::Twr twr;  // the same entries serve the forward transform
  {
    cx<W> iu[1][FI::E0], iw[1][FI::E0];
    load(work + c * kp.fs, iu);
    load(work + (3 + c) * kp.fs, iw);
    if constexpr (sizeof(W) == 4) {
      // FP32 (GT_X3_FP32): the twiddles from an LDS copy of the table
      tw_to_lds<W, N>(tw, lds + 3 * G * N, tid, 3 * 64);
      __syncthreads();
      FI::run(lay, lds, lds + 3 * G * N, t, iu, uo);
      stage_sync<SY>();
      FI::run(lay, lds, lds + 3 * G * N, t, iw, wo);
    } else if constexpr (GT_X_PRETW) {
      // the twiddles of all three transforms, loaded with the data
      FI::load_tw(tw, t, twr);
      FI::run(lay, lds, tw, twr, t, iu, uo);
      stage_sync<SY>();  // row c is free for the second transform
      FI::run(lay, lds, tw, twr, t, iw, wo);
    } else {
      FI::run(lay, lds, tw, t, iu, uo);
      stage_sync<SY>();
      FI::run(lay, lds, tw, t, iw, wo);
    }
  }
  stage_sync<SY>();
  // exchange 1: u_c to row c; Q = -u_{c+1} w_c
#pragma unroll
  for (int m = 0; m < FI::ML; m++)
    if (FI::out_ok(t, m))
#pragma unroll
      for (int r = 0; r < FI::RL; r++) lds[row * N + FI::out_pos(t, m, r)] = uo[0][m * FI::RL + r];
  __syncthreads();
  cx<W> pr[1][FF::E0];
#pragma unroll
  for (int m = 0; m < FF::M0; m++)
#pragma unroll
    for (int r = 0; r < FF::R0; r++) {
      const int q = m * FF::R0 + r;
      cx<W> v = {W(0), W(0)};
      if (FF::in_ok(t, m)) {
        const cx<W> un = lds[rown * N + FF::in_pos(t, m, r)];
        v = {-(un.x * wo[0][q].x), -(un.y * wo[0][q].y)};
      }
      pr[0][q] = v;
    }
  __syncthreads();  // every u row has been read
  // exchange 2: w_c to row c; P = Q + u_c w_{c+1}
#pragma unroll
  for (int m = 0; m < FI::ML; m++)
    if (FI::out_ok(t, m))
#pragma unroll
      for (int r = 0; r < FI::RL; r++) lds[row * N + FI::out_pos(t, m, r)] = wo[0][m * FI::RL + r];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < FF::M0; m++)
#pragma unroll
    for (int r = 0; r < FF::R0; r++) {
      const int q = m * FF::R0 + r;
      if (FF::in_ok(t, m)) {
        const cx<W> wn = lds[rown * N + FF::in_pos(t, m, r)];
        pr[0][q] = {uo[0][q].x * wn.x + pr[0][q].x, uo[0][q].y * wn.y + pr[0][q].y};
      }
    }
  __syncthreads();  // every w row has been read before the rows are reused
  // forward transform of P_{c+2} in row c, then its two rows unpacked
  cx<W> fout[1][FF::EL];
  if constexpr (sizeof(W) == 4) FF::run(lay, lds, lds + 3 * G * N, t, pr, fout);
  else if constexpr (GT_X_PRETW) FF::run(lay, lds, tw, twr, t, pr, fout);
  else FF::run(lay, lds, tw, t, pr, fout);
  stage_sync<SY>();
  static_for<0, FF::ML * FF::RL>([&](auto qc) {
    constexpr int q = decltype(qc)::value, m = q / FF::RL, r = q % FF::RL;
    if constexpr (!(PZ && mid_slot<N, P, P * m + r * FF::NRL>())) {
      if (FF::out_ok(t, m)) lds[row * N + FF::out_pos(t, m, r)] = fout[0][q];
    }
  });
  stage_sync<SY>();
  const int pc = (c + 2) % 3;
  cx<W>* d0 = work + pc * kp.fs + r0;
  cx<W>* d1 = d0 + kp.nxp;
  for (int k = t; k <= kc; k += P) {
    const int kk = (k == 0) ? 0 : N - k;
    const cx<W> Y = lds[row * N + k], Yc = lds[row * N + kk];
    st_row_p<NT_C>(d0, (unsigned)k, cx<W>{(Y.x + Yc.x) * W(0.5), (Y.y - Yc.y) * W(0.5)});
    st_row_p<NT_C>(d1, (unsigned)k, cx<W>{(Y.y + Yc.y) * W(0.5), (Yc.x - Y.x) * W(0.5)});
  }
}

// velocity output: 3 half spectra -> 3 physical fields [nz][ny][nx]
template <int N, int P>
__global__ __launch_bounds__(3 * P) void k_xc2r(KP kp, const cx<double>* __restrict__ work, double* __restrict__ phys,
                                                const cx<double>* __restrict__ tw) {
  using L = LaySwz<N>;
  using FI = Fft<double, N, P, 1, +1, L>;
  GT_DYN_LDS(cx<double>, lds);  // 3 N elements
  const int tid = threadIdx.x, sq = tid / P, t = tid % P;
  const int y0 = 2 * blockIdx.x, z = blockIdx.y;
  const long long r0 = ((long long)z * kp.ny + y0) * kp.nxp;
  const long long r1 = r0 + kp.nxp;
  const cx<double>* src = work + sq * kp.fs;
  cx<double> in[1][FI::E0], out[1][FI::EL];
#pragma unroll
  for (int m = 0; m < FI::M0; m++)
#pragma unroll
    for (int r = 0; r < FI::R0; r++) {
      cx<double> v = {0.0, 0.0};
      if (FI::in_ok(t, m)) v = c2r_in<double, N>(src + r0, src + r1, FI::in_pos(t, m, r), kp.kxcut);
      in[0][m * FI::R0 + r] = v;
    }
  L lay{sq, 0};
  FI::run(lay, lds, tw, t, in, out);
  const long long plane = (long long)kp.nx * kp.ny * kp.nz;
  double* d0 = phys + sq * plane + ((long long)z * kp.ny + y0) * kp.nx;
  double* d1 = d0 + kp.nx;
#pragma unroll
  for (int m = 0; m < FI::ML; m++)
    if (FI::out_ok(t, m))
#pragma unroll
      for (int r = 0; r < FI::RL; r++) {
        const int n = FI::out_pos(t, m, r);
        d0[n] = out[0][m * FI::RL + r].x;
        d1[n] = out[0][m * FI::RL + r].y;
      }
}

// initial transform: 3 physical fields -> 3 half spectra (all kx <= NX/2)
template <int N, int P>
__global__ __launch_bounds__(3 * P) void k_xr2c(KP kp, const double* __restrict__ phys, cx<double>* __restrict__ work,
                                                const cx<double>* __restrict__ tw) {
  using L = LaySwz<N>;
  using FF = Fft<double, N, P, 1, -1, L>;
  GT_DYN_LDS(cx<double>, lds);  // 3 N elements
  const int tid = threadIdx.x, sq = tid / P, t = tid % P;
  const int y0 = 2 * blockIdx.x, z = blockIdx.y;
  const long long plane = (long long)kp.nx * kp.ny * kp.nz;
  const double* s0 = phys + sq * plane + ((long long)z * kp.ny + y0) * kp.nx;
  const double* s1 = s0 + kp.nx;
  cx<double> in[1][FF::E0], out[1][FF::EL];
#pragma unroll
  for (int m = 0; m < FF::M0; m++)
#pragma unroll
    for (int r = 0; r < FF::R0; r++) {
      cx<double> v = {0.0, 0.0};
      if (FF::in_ok(t, m)) {
        const int n = FF::in_pos(t, m, r);
        v = {s0[n], s1[n]};
      }
      in[0][m * FF::R0 + r] = v;
    }
  L lay{sq, 0};
  FF::run(lay, lds, tw, t, in, out);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < FF::ML; m++)
    if (FF::out_ok(t, m))
#pragma unroll
      for (int r = 0; r < FF::RL; r++) lds[sq * N + FF::out_pos(t, m, r)] = out[0][m * FF::RL + r];
  __syncthreads();
  const long long r0 = ((long long)z * kp.ny + y0) * kp.nxp;
  const long long r1 = r0 + kp.nxp;
  cx<double>* dst = work + sq * kp.fs;
  for (int k = t; k <= N / 2; k += P) {
    const int kk = (k == 0) ? 0 : N - k;
    const cx<double> Y = lds[sq * N + k], Yc = lds[sq * N + kk];
    dst[r0 + k] = {(Y.x + Yc.x) * 0.5, (Y.y - Yc.y) * 0.5};
    dst[r1 + k] = {(Y.y + Yc.y) * 0.5, (Yc.x - Y.x) * 0.5};
  }
}

// forcing: b[f][off] += d[f] for every entry (positions are distinct)
__global__ void k_forcing(cx<double>* b, long long fs, const FEnt* __restrict__ ent, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FEnt e = ent[i];
#pragma unroll
  for (int f = 0; f < 3; f++) {
    cx<double>* q = b + f * fs + e.off;
    cx<double> v = *q;
    v.x = v.x + e.d[2 * f];
    v.y = v.y + e.d[2 * f + 1];
    *q = v;
  }
}

// forcing in entry order by one thread, for tables whose positions repeat
__global__ void k_forcing_seq(cx<double>* b, long long fs, const FEnt* __restrict__ ent, int n) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int i = 0; i < n; i++) {
    const FEnt e = ent[i];
    for (int f = 0; f < 3; f++) {
      cx<double>* q = b + f * fs + e.off;
      cx<double> v = *q;
      v.x = v.x + e.d[2 * f];
      v.y = v.y + e.d[2 * f + 1];
      *q = v;
    }
  }
}

// ---------------------------------------------------------------------------
// the three-wave kernel applies: FP64, whole transforms in each wavefront
// (rows of 64 threads, or 16 or 32 with several row pairs per block), and
// the inverse's last-stage positions equal to the forward's first-stage ones
template <typename W, int N>
constexpr bool use_x3() {
  using PL = Plan<N>;
  constexpr int P = row_P(N);
  return (sizeof(W) == 8 || GT_X3_FP32) && (P == 64 || P == 32 || P == 16) &&
         PL::radix(0) == PL::radix(PL::S - 1) && cdiv(N / PL::radix(0), P) == cdiv(N / PL::radix(PL::S - 1), P);
}

template <typename W, int N>
static hipError_t xfused_n(const KP& kp, void* work, const void* tw, hipStream_t s) {
  constexpr int P = row_P(N);
  dim3 grid(kp.ny / 2, kp.zn, kp.nmem);
  if constexpr (use_x3<W, N>()) {
    constexpr int G = 64 / P;  // row pairs per block
    if ((kp.ny / 2) % G == 0) {
      const size_t lds3 = 3 * G * N * sizeof(cx<W>) + (sizeof(W) == 4 ? Plan<N>::twsize() * sizeof(cx<W>) : 0);
      dim3 grid3(kp.ny / 2 / G, kp.zn, kp.nmem);
      if (pz_cut(N, kp.kxcut))
        k_xfused3<W, N, P, true><<<grid3, 3 * 64, lds3, s>>>(kp, (cx<W>*)work, (const cx<W>*)tw);
      else
        k_xfused3<W, N, P, false><<<grid3, 3 * 64, lds3, s>>>(kp, (cx<W>*)work, (const cx<W>*)tw);
      return hipGetLastError();
    }
  }
  const size_t lds6 = 6 * N * sizeof(cx<W>) + ((sizeof(W) == 4 && GT_TWLDS32 && GT_X_TWLDS32) ? Plan<N>::twsize() * sizeof(cx<W>) : 0);
  if (pz_cut(N, kp.kxcut))
    k_xfused<W, N, P, true><<<grid, 6 * P, lds6, s>>>(kp, (cx<W>*)work, (const cx<W>*)tw);
  else
    k_xfused<W, N, P, false><<<grid, 6 * P, lds6, s>>>(kp, (cx<W>*)work, (const cx<W>*)tw);
  return hipGetLastError();
}

hipError_t launch_xfused(int prec, const KP& kp, void* work, const void* tw, hipStream_t s) {
  switch (kp.nx) {
#define GT_CASE(N) \
  case N:          \
    return prec ? xfused_n<float, N>(kp, work, tw, s) : xfused_n<double, N>(kp, work, tw, s);
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_xc2r(const KP& kp, const void* work, double* phys, const void* tw, hipStream_t s) {
  dim3 grid(kp.ny / 2, kp.nz);
  switch (kp.nx) {
#define GT_CASE(N)                                                                                             \
  case N:                                                                                                      \
    k_xc2r<N, row_P(N)><<<grid, 3 * row_P(N), 3 * N * sizeof(cx<double>), s>>>(kp, (const cx<double>*)work, phys, (const cx<double>*)tw); \
    return hipGetLastError();
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_xr2c(const KP& kp, const double* phys, void* work, const void* tw, hipStream_t s) {
  dim3 grid(kp.ny / 2, kp.nz);
  switch (kp.nx) {
#define GT_CASE(N)                                                                                       \
  case N:                                                                                                \
    k_xr2c<N, row_P(N)><<<grid, 3 * row_P(N), 3 * N * sizeof(cx<double>), s>>>(kp, phys, (cx<double>*)work, (const cx<double>*)tw); \
    return hipGetLastError();
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_forcing(void* b, long long fs, const FEnt* ent, int n, bool sequential, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (sequential)
    k_forcing_seq<<<1, 64, 0, s>>>((cx<double>*)b, fs, ent, n);
  else
    k_forcing<<<(n + 127) / 128, 128, 0, s>>>((cx<double>*)b, fs, ent, n);
  return hipGetLastError();
}

}  // namespace gt
