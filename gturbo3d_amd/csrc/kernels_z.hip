// kernels_z.hip -- the z-axis pencil passes of the step.
//
// Pass A (k_zinv): z-inverse transform of u = curl b and w = k^2 b, with the
//   curl (src/produ.f90:44-63) and the k^2 scale (src/nlt.f90:26-40) applied
//   to b as it is loaded, so b is read once and the six spectral fields of
//   src/nlt.f90:16-40 never exist in HBM.
// Pass E (k_zfwd): z-forward transform of the three products, then in
//   registers: the fft_dir normalisation (src/fft_dir.f90:17-25), the
//   Coriolis term (src/nlt.f90:67-81), invcurl (src/produ.f90:79-101),
//   trunc (src/trunc.f90:10-28) and the integrating-factor update of step1
//   (src/rkstep.f90:30-48) or step2 (src/rkstep.f90:86-104), with
//   emk = exp(-0.5*dt*(nu*k2**alpnu + mu*k2**alpmu)) (src/defk.f90:57-68)
//   evaluated on the fly instead of read from a table.
//
// Both passes touch only the (kx, ky, kz) the truncation leaves alive in
// pruned mode (see KP in gt_internal.hpp).
#include "gt_config.hpp"
#include "gt_device.hpp"

namespace gt {

enum { EPI_STEP1 = 0, EPI_STEP2 = 1, EPI_NONLIN = 2, EPI_INIT = 3 };

#ifndef GT_ZINV_PAIR
#define GT_ZINV_PAIR 1
#endif
#ifndef GT_ZINV_XCD
#define GT_ZINV_XCD 1
#endif
#ifndef GT_ZFWD_PAIR
#define GT_ZFWD_PAIR 1
#endif
// FP64 pass A: the transforms' twiddles in registers, loaded with the data
// (1), before it (2) or read from the table by each stage (0)
#ifndef GT_ZINV_PRETW
#define GT_ZINV_PRETW 1
#endif
// FP32 pass A with PAIR: the second transform's input kept in registers (1)
// or b read again for it (0)
#ifndef GT_ZINV_IN2
#define GT_ZINV_IN2 1
#endif
// FP64 pass E: each field's loads issued before the previous field's
// transform (1) or when its own transform starts (0)
#ifndef GT_ZFWD_PREF
#define GT_ZFWD_PREF 1
#endif
// output slots per load batch of the FP64 epilogue: all 8 (6 live at 512)
// spill; 4 does not
#ifndef GT_ZFWD_EPB64
#define GT_ZFWD_EPB64 4
#endif
// FA: the slots' kz wavenumbers kept in registers from the epilogue (1) or
// read from the table again for the z-inverse transforms (0), and the output
// slots per load batch of the epilogue, for FA = 1 and FA = 2.  Measured at
// 512^3 FP64: FA = 1 with 0 and 2 runs without spills, 5% faster than with
// 1 and 4 (11 spilled registers); FA = 2 is 4% slower with 0.
#ifndef GT_EA_FZREG
#define GT_EA_FZREG 0
#endif
#ifndef GT_EA_EPB
#define GT_EA_EPB 2
#endif
#ifndef GT_EF_FZREG
#define GT_EF_FZREG 1
#endif
#ifndef GT_EF_EPB
#define GT_EF_EPB 4
#endif

// Pass A: one output component q per block (0..2: u = curl b,
// 3..5: w = k^2 b), so a block holds only the one or two fields of b it
// needs in registers.
//   u_x = i(ky bz - kz by), u_y = i(kz bx - kx bz), u_z = i(kx by - ky bx)
//   (src/produ.f90:48-53);  w_f = (fkxs + fkys + fkzs) b_f (src/nlt.f90:30-36)
// EMK: transform a = Z emk b instead of b (the stage-1 state bh of a step in
// the reference dt2 mode, formed as it is loaded; Z sets k = 0 to zero)
// PZ: the slots that mid_slot() prunes (kzcut < 3N/8) are zeros, known at
// compile time.
// PAIR: nout = 3 blocks per column, block q computing u_q and then w_fb,
// the k^2 component of the second field u_q reads (u_0 + w_1, u_1 + w_2,
// u_2 + w_0): that field is loaded once for both transforms.
template <typename W, int N, int T, int P, bool EMK = false, bool PZ = false, bool PAIR = false>
__global__ __launch_bounds__(T* P) void k_zinv(KP kp, const cx<double>* __restrict__ b, cx<W>* __restrict__ work,
                                               const cx<W>* __restrict__ tw, int nout) {
  static_assert(!(PAIR && EMK) || sizeof(W) == 8, "the paired EMK blocks keep b in registers (FP64)");
  using LI = LayInterleaved<T, sizeof(cx<W>) == 8 && T == 8>;  // swizzled for 8 pencils of 8 bytes
  using F = Fft<W, N, P, 1, +1, LI>;
  constexpr int R0 = F::R0, RL = F::RL, M0 = F::M0, ML = F::ML;
  GT_DYN_LDS(cx<W>, lds);  // N T elements
  const int tid = threadIdx.x, p = tid % T, t = tid / T;
  // 1-D grid over (column, q).  Workgroups go to the 8 XCDs in turn, so the
  // nout blocks of one column are given the consecutive slots of one XCD
  // (ids i, i+8, i+16, ...): they read the same fields of b close in time
  // and all but the first find them in that XCD's L2.  When a tile row is
  // shorter than a 128-byte line, the PF adjacent tiles that share the line
  // follow on the same XCD too.  A tail of fewer than 8*nout*PF blocks keeps
  // the plain order.
  constexpr int PF = (T * (int)sizeof(cx<W>) >= 128) ? 1 : 128 / (T * (int)sizeof(cx<W>));
  const int L = blockIdx.x, grp = 8 * nout * PF, bulk = (gridDim.x / grp) * grp;
  int col, q;
  if (GT_ZINV_XCD && L < bulk) {
    const int s = L >> 3, u = s / nout;
    q = s % nout;
    col = PF * (8 * (u / PF) + (L & 7)) + u % PF;
  } else {
    q = L % nout;
    col = L / nout;
  }
  const int ntx = (kp.kxcut + T) / T;
  const int tx = col % ntx, j = (col / ntx) % kp.ny, mem = col / (ntx * kp.ny);
  const int kx = tx * T + p;
  b += mem * kp.bstride;
  work += mem * kp.wstride;
  int K = -1;
  if (kx <= kp.kxcut) K = kp.pruned ? (int)kp.kzext[j * kp.nxp + kx] : (N >> 1);
  static_assert(64 % T == 0, "every wavefront holds all T pencils");
  if (!wave_any(K >= 0)) return;  // uniform over the block

  const int zs = kp.ny * kp.nxp;
  const double fx = kp.fkx[kx < kp.nxh ? kx : kp.nxh - 1];
  const double fy = kp.fky[j];
  const bool curl = q < 3;
  // fields and wavenumber roles, uniform per block
  const int fa = curl ? (q == 0 ? 2 : q == 1 ? 0 : 1) : q - 3;
  const int fb = q == 0 ? 1 : q == 1 ? 2 : 0;
  // element n = t + c of the thread's pencil (c = P*m + r*N/R, a constant of
  // the slot) sits at tile + c*zs + lane: a start uniform over the block and
  // one 32-bit lane offset, instead of a 64-bit address per element
  const long long tile = (long long)j * kp.nxp + tx * T;
  const cx<double>* A = b + fa * kp.fs + tile;
  const cx<double>* B = b + fb * kp.fs + tile;
  const unsigned lane = p + t * zs;

  cx<W> in[1][M0 * R0], out[1][ML * RL];
  // FP64 work fields: all loads first, unconditionally from a valid element
  // (element 0 of the column for an unused slot), so every load is in flight
  // before the first wait; the selects come after.  FP32 work fields keep
  // predicated loads (below).
  constexpr bool BATCH = sizeof(W) == 8;
  cx<double> la[M0 * R0], lb[M0 * R0];
  // FP64: the kz wavenumbers of the thread's slots and the transforms'
  // twiddles are loaded with the data, so that nothing after the first wait
  // waits for memory again
  // (measured at 512; other lengths keep the table loads of each slot and
  // stage, whose registers would cost them resident blocks)
  constexpr bool PRELD = BATCH && N == 512;
  double fzv[M0 * R0];
  typename F::Twr twr;
  if constexpr (BATCH) {
    if constexpr (PRELD && GT_ZINV_PRETW == 2) F::load_tw(tw, t, twr);
    static_for<0, M0 * R0>([&](auto qc) {
      constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
      if constexpr (!(PZ && mid_slot<N, P, P * m + r * F::NR0>())) {
        const int n = F::in_pos(t, m, r);
        // an unused slot reads element c of the column instead
        const unsigned o = (F::in_ok(t, m) && keep_idx(n, N, K)) ? lane : (unsigned)p;
        const long long c = (long long)(P * m + r * F::NR0) * zs;
        la[qs] = ld_row(A + c, o);
        // curl is uniform per block: the k^2 blocks skip the second field
        lb[qs] = curl ? ld_row(B + c, o) : cx<double>{0.0, 0.0};
      }
    });
    if constexpr (PRELD) {
      static_for<0, M0 * R0>([&](auto qc) {
        constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
        if constexpr (!(PZ && mid_slot<N, P, P * m + r * F::NR0>())) fzv[qs] = kp.fkz[F::in_ok(t, m) ? F::in_pos(t, m, r) : 0];
      });
      if constexpr (GT_ZINV_PRETW == 1) F::load_tw(tw, t, twr);
    }
  }
  // kz wavenumber of slot qs (in_ok)
  auto fz_of = [&](auto qc) -> double {
    constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
    if constexpr (PRELD) return fzv[qs];
    else return kp.fkz[F::in_pos(t, m, r)];
  };
  // FP32: the twiddles from a copy of the table in LDS behind the pencils
  // (tw_to_lds; the first stage's barrier makes it visible)
  constexpr bool TWL = sizeof(W) == 4 && GT_TWLDS32;
  auto fft = [&](cx<W>(&i)[1][M0 * R0], cx<W>(&o)[1][ML * RL], const LI& ly) {
    if constexpr (PRELD && GT_ZINV_PRETW) F::run(ly, lds, tw, twr, t, i, o);
    else if constexpr (TWL) F::run(ly, lds, lds + N * T, t, i, o);
    else F::run(ly, lds, tw, t, i, o);
  };
  // the transform input of slot qs from the loaded b values (zero where the
  // slot is unused); FP32 work fields with PAIR also keep the second
  // transform's input, k^2 b of field fb, rounded to FP32 as it is formed,
  // so b is not read again (GT_ZINV_IN2)
  constexpr bool IN2 = PAIR && !BATCH && GT_ZINV_IN2;
  cx<W> in2[M0 * R0];
  auto form = [&](auto qc, cx<double> va, cx<double> vb, double fz) {
    constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
    const int n = F::in_pos(t, m, r);
    if constexpr (EMK) {
      const bool k0 = (kx == 0) && (j == 0) && (n == 0);
      const double e = k0 ? 0.0 : emk_at(kp, kx < kp.nxh ? kx : kp.nxh - 1, j, n, fx * fx + fy * fy + fz * fz);
      va = {e * va.x, e * va.y};
      vb = {e * vb.x, e * vb.y};
    }
    double re, im;
    if (curl) {
      const double ca = (q == 0) ? fy : (q == 1) ? fz : fx;
      const double cb = (q == 0) ? fz : (q == 1) ? fx : fy;
      re = -ca * va.y + cb * vb.y;
      im = ca * va.x - cb * vb.x;
    } else {
      const double k2 = fx * fx + fy * fy + fz * fz;
      re = k2 * va.x;
      im = k2 * va.y;
    }
    in[0][qs] = {(W)re, (W)im};
    if constexpr (IN2) {
      const double k2 = fx * fx + fy * fy + fz * fz;
      in2[qs] = {(W)(k2 * vb.x), (W)(k2 * vb.y)};
      // held as FP32 from here (the conversion is not sunk to its use, where
      // the double products would have to stay live through the transform)
      asm volatile("" : "+v"(in2[qs].x), "+v"(in2[qs].y));
    }
  };
  if constexpr (BATCH) {
    static_for<0, M0 * R0>([&](auto qc) {
      constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
      if constexpr (PZ && mid_slot<N, P, P * m + r * F::NR0>()) {
        in[0][qs] = {W(0), W(0)};
      } else {
        const int n = F::in_pos(t, m, r);
        const bool use = F::in_ok(t, m) && keep_idx(n, N, K);
        const double fz = F::in_ok(t, m) ? fz_of(qc) : 0.0;
        form(qc, use ? la[qs] : cx<double>{0.0, 0.0}, (use && curl) ? lb[qs] : cx<double>{0.0, 0.0}, fz);
      }
    });
  } else {
    // FP32 work fields: predicated loads, slot by slot.  Batching them (all
    // loads of 4 slots issued before the first wait), with the paired k^2
    // input kept in registers instead of b read again, takes 81 registers
    // and halves the resident blocks: 10% slower at 16 pencils.
    static_for<0, M0 * R0>([&](auto qc) {
      constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
      if constexpr (PZ && mid_slot<N, P, P * m + r * F::NR0>()) {
        in[0][qs] = {W(0), W(0)};
      } else {
        const int n = F::in_pos(t, m, r);
        const bool use = F::in_ok(t, m) && keep_idx(n, N, K);
        const double fz = F::in_ok(t, m) ? kp.fkz[n] : 0.0;
        cx<double> va = {0.0, 0.0}, vb = {0.0, 0.0};
        if (use) {
          const long long c = (long long)(P * m + r * F::NR0) * zs;
          va = ld_row(A + c, lane);
          if (curl) vb = ld_row(B + c, lane);
        }
        form(qc, va, vb, fz);
      }
    });
  }
  if constexpr (TWL) tw_to_lds<W, N>(tw, lds + N * T, tid, T * P);  // after the data loads
  LI lay{p};
  fft(in, out, lay);
  auto store = [&](int comp) {
    cx<W>* dst = work + comp * kp.fs + tile;
#pragma unroll
    for (int m = 0; m < ML; m++)
      if (F::out_ok(t, m))
#pragma unroll
        for (int r = 0; r < RL; r++) st_row_p<NT_A>(dst + (long long)(P * m + r * F::NRL) * zs, lane, out[0][m * RL + r]);
  };
  if constexpr (PAIR) {
    // every thread takes part in the second transform's exchanges, so the
    // idle pencils skip the stores instead of returning
    if (K >= 0) store(q);
    static_for<0, M0 * R0>([&](auto qc) {
      constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
      if constexpr (PZ && mid_slot<N, P, P * m + r * F::NR0>()) {
        in[0][qs] = {W(0), W(0)};
      } else if constexpr (BATCH) {
        const int n = F::in_pos(t, m, r);
        const bool use = F::in_ok(t, m) && keep_idx(n, N, K);
        const double fz = F::in_ok(t, m) ? fz_of(qc) : 0.0;
        cx<double> vb = use ? lb[qs] : cx<double>{0.0, 0.0};
        const double k2 = fx * fx + fy * fy + fz * fz;
        if constexpr (EMK) {  // as form() scales it for the first transform
          const bool k0 = (kx == 0) && (j == 0) && (n == 0);
          const double e = k0 ? 0.0 : emk_at(kp, kx < kp.nxh ? kx : kp.nxh - 1, j, n, k2);
          vb = {e * vb.x, e * vb.y};
        }
        in[0][qs] = {(W)(k2 * vb.x), (W)(k2 * vb.y)};
      } else if constexpr (IN2) {
        in[0][qs] = in2[qs];
      } else {
        const int n = F::in_pos(t, m, r);
        const bool use = F::in_ok(t, m) && keep_idx(n, N, K);
        const double fz = F::in_ok(t, m) ? kp.fkz[n] : 0.0;
        cx<double> vb = {0.0, 0.0};
        if (use) vb = ld_row(B + (long long)(P * m + r * F::NR0) * zs, lane);
        const double k2 = fx * fx + fy * fy + fz * fz;
        in[0][qs] = {(W)(k2 * vb.x), (W)(k2 * vb.y)};
      }
    });
    fft(in, out, lay);
    if (K < 0) return;
    store(3 + fb);
  } else {
    if (K < 0) return;
    store(q);
  }
}

// VS: simple viscosity, the integrating factor from the per-axis tables
// (emk_at) with no exponential in the kernel
// Waves per SIMD the register allocation targets.  FP32 work fields: 4
// (128 registers, a few spills), so two blocks of 8 pencils share a CU
// instead of one.  FP64: the three transforms' outputs alone take 96
// registers, and a bound costs more in spills than it gains.
#ifndef GT_ZFWD_WPE64
#define GT_ZFWD_WPE64 1
#endif
#ifndef GT_ZFWD_WPE32
#define GT_ZFWD_WPE32 4
#endif
// output slots per load batch of the FP32 epilogue: 1 (2 spills and makes
// the pass 25% slower)
#ifndef GT_ZFWD_EPB32
#define GT_ZFWD_EPB32 1
#endif
// PZ: the output slots that mid_slot() prunes (kzcut < 3N/8) are never
// stored, so the transform's last stage drops them.
// FA = 1 (with EPI_STEP1): the stage-1 state bh is not stored.  It stays
// in registers, and the block runs pass A of stage 2 on it at once: the six
// z-inverse transforms of u = curl bh and w = k^2 bh of its column, written
// to wout (the work fields whose products it has just read; every load of
// the column precedes the first store, and no other block touches the
// column).  bh then never goes through HBM.
// FA = 2 (with EPI_STEP2): the block adds this step's forcing increments of
// its column (src/main.f90:149) to the new b, stores it, and runs pass A of
// the next step's stage 1 on it in the same way, so the next step does not
// read b back.
// Both need a first radix equal to the last, so that the forward transform
// leaves every value where the inverse transform takes it.
template <typename W, int N, int T, int P, int EPI, bool VS = false, bool PZ = false, int FA = 0>
__global__ __launch_bounds__(T* P)
__attribute__((amdgpu_waves_per_eu(FA ? 1 : sizeof(W) == 8 ? GT_ZFWD_WPE64 : GT_ZFWD_WPE32))) void k_zfwd(KP kp, const cx<W>* __restrict__ work, const cx<double>* b,
                                               cx<double>* out, const cx<W>* __restrict__ tw, cx<W>* wout, FAux fx2) {
  // out may be b itself (step2 updates b in place), so neither is restrict
  using LI = LayInterleaved<T, sizeof(cx<W>) == 8 && T == 8>;  // swizzled for 8 pencils of 8 bytes
  using F = Fft<W, N, P, 1, -1, LI>;
  constexpr int R0 = F::R0, RL = F::RL, M0 = F::M0, ML = F::ML;
  static_assert(FA == 0 || ((FA == 1 ? EPI == EPI_STEP1 : EPI == EPI_STEP2) && R0 == RL && M0 == ML),
                "FA: stage 1 (1) or 2 (2), first radix equal to the last");
  GT_DYN_LDS(cx<W>, lds);  // N T elements
  const int tid = threadIdx.x, p = tid % T, t = tid / T;
  // 1-D grid over columns (kx tile, ky, member).  When a tile row is shorter
  // than a 128-byte line (FP32 work fields, 8 pencils), PF adjacent tiles
  // share each line: they are given consecutive slots of one XCD (workgroups
  // go to the 8 XCDs in turn), so the line is fetched once into that XCD's
  // L2.  A tail of fewer than 8*PF blocks keeps the plain order.
  constexpr int PF = (T * (int)sizeof(cx<W>) >= 128) ? 1 : 128 / (T * (int)sizeof(cx<W>));
  const int L = blockIdx.x, grp = 8 * PF, bulk = (gridDim.x / grp) * grp;
  int col = L;
  if (GT_ZFWD_PAIR && PF > 1 && L < bulk) {
    const int s = L >> 3;
    col = PF * (8 * (s / PF) + (L & 7)) + s % PF;
  }
  const int ntx = (kp.kxcut + T) / T;
  const int tx = col % ntx, j = (col / ntx) % kp.ny, mem = col / (ntx * kp.ny);
  const int kx = tx * T + p;
  work += mem * kp.wstride;
  if constexpr (FA != 0) wout += mem * kp.wstride;
  if (b) b += mem * kp.bstride;
  if constexpr (FA != 1) out += mem * kp.bstride;
  const double omega = kp.omega_m ? kp.omega_m[mem] : kp.omega;
  int K = -1;
  if (kx <= kp.kxcut) K = kp.pruned ? (int)kp.kzext[j * kp.nxp + kx] : (N >> 1);
  static_assert(64 % T == 0, "every wavefront holds all T pencils");
  if (!wave_any(K >= 0)) return;  // uniform over the block

  const int zs = kp.ny * kp.nxp;
  // element n = t + c of the thread's pencil at tile + c*zs + lane (as in pass A)
  const long long tile = (long long)j * kp.nxp + tx * T;
  const unsigned lane = p + t * zs;
  LI lay{p};

  // FP64 (PRE): the transforms' twiddles are loaded once, with the first
  // field, and each field's loads are issued before the previous field's
  // transform, so they are in flight while it runs (the LDS barriers of a
  // transform wait for LDS operations only).  FP32 loads each field when
  // its transform starts (registers set its occupancy).
  // (measured at 512; other lengths keep the loads of the original order,
  // whose registers would cost them resident blocks or spills)
  constexpr bool TUNED = sizeof(W) == 8 && N == 512;
  constexpr bool PRE = TUNED && GT_ZFWD_PREF;
  typename F::Twr twr;
  auto load_field = [&](int f, cx<W>(&v)[M0 * R0]) {
    const cx<W>* src = work + f * kp.fs + tile;
    static_for<0, M0 * R0>([&](auto qc) {
      constexpr int qs = decltype(qc)::value, m = qs / R0, r = qs % R0;
      const cx<W>* row = src + (long long)(P * m + r * F::NR0) * zs;
      if constexpr (PRE) {
        // unconditional, from a valid element (plane P*m + r*N/R0 < N); the
        // values of unused slots and of pencils past the x cut are dropped
        v[qs] = ld_row_p<NT_E>(row, F::in_ok(t, m) ? lane : (unsigned)p);
      } else {
        cx<W> x = {W(0), W(0)};
        if (K >= 0 && F::in_ok(t, m)) x = ld_row_p<NT_E>(row, lane);
        v[qs] = x;
      }
    });
  };
  cx<W> y[3][ML * RL];
  cx<W> nx[M0 * R0];
  if constexpr (PRE) {
    load_field(0, nx);
    F::load_tw(tw, t, twr);
  }
  // FP32: the twiddles from a copy of the table in LDS behind the pencils
  // (tw_to_lds; the first stage's barrier makes it visible)
  constexpr bool TWL = sizeof(W) == 4 && GT_TWLDS32;
  const cx<W>* twt = TWL ? lds + N * T : tw;
#pragma unroll
  for (int f = 0; f < 3; f++) {
    cx<W> in[1][M0 * R0], o[1][ML * RL];
    if constexpr (PRE) {
#pragma unroll
      for (int m = 0; m < M0; m++)
#pragma unroll
        for (int r = 0; r < R0; r++)
          in[0][m * R0 + r] = (K >= 0 && F::in_ok(t, m)) ? nx[m * R0 + r] : cx<W>{W(0), W(0)};
      if (f < 2) load_field(f + 1, nx);
      F::run(lay, lds, tw, twr, t, in, o);
    } else {
      load_field(f, nx);
      if constexpr (TWL) {
        if (f == 0) tw_to_lds<W, N>(tw, lds + N * T, tid, T * P);  // after the first field's loads
      }
#pragma unroll
      for (int q = 0; q < M0 * R0; q++) in[0][q] = nx[q];
      F::run(lay, lds, twt, t, in, o);
    }
#pragma unroll
    for (int q = 0; q < ML * RL; q++) y[f][q] = o[0][q];
  }
  // with FA every thread takes part in the transforms after the epilogue;
  // a pencil with K < 0 keeps no slot (keep_idx) and stores nothing
  if constexpr (FA == 0) {
    if (K < 0) return;
  }

  const int kxc = kx < kp.nxh ? kx : kp.nxh - 1;
  const double fx = kp.fkx[kxc], fy = kp.fky[j];
  const double fxs = fx * fx, fys = fy * fy;
  const double kxn = kp.fkxn[kxc], kyn = kp.fkyn[j];
  const double exy = VS ? kp.emx[kxc] * kp.emy[j] : 0.0;
  // The epilogue in batches of EPB output slots: every load of a batch (b
  // and the kz tables) is issued before the first is waited for.  A slot
  // that is not stored reads the column's kz = 0 element instead of its own
  // (a line the kept slots fetch anyway).
  constexpr int NS = ML * RL;
  // FA: the new state of every slot (zero where not kept) and the slot's
  // kz wavenumber
  constexpr bool FZR = FA == 1 ? GT_EA_FZREG : FA == 2 ? GT_EF_FZREG : 0;
  cx<double> bhr[FA ? 3 : 1][FA ? NS : 1];
  double fza[FZR ? NS : 1];
  // FA = 2: this step's forcing entries of the block's T columns,
  // fx2.idx[fa0 .. fa1) (member-relative entry numbers, in table order)
  int fa0 = 0, fa1 = 0;
  const FEnt* fent = nullptr;
  if constexpr (FA == 2) {
    const int key0 = j * kp.nxp + tx * T;
    fa0 = fx2.start[key0];
    fa1 = fx2.start[min(key0 + T, (j + 1) * kp.nxp)];
    fent = fx2.ent + (long long)mem * fx2.per;
  }
  if constexpr (FA) {
#pragma unroll
    for (int q = 0; q < NS; q++) {
      if constexpr (FZR) fza[q] = 0.0;
#pragma unroll
      for (int f = 0; f < 3; f++) bhr[f][q] = {0.0, 0.0};
    }
  }
  constexpr int EPB0 = TUNED ? (FA == 1 ? GT_EA_EPB : FA == 2 ? GT_EF_EPB : GT_ZFWD_EPB64)
                             : sizeof(W) == 8 ? 1 : GT_ZFWD_EPB32;
  constexpr int EPB = (NS % EPB0 != 0) ? NS : EPB0;
  static_for<0, NS / EPB>([&](auto bc) {
    constexpr int q0 = decltype(bc)::value * EPB;
    cx<double> bvs[EPB][3];
    double fzs[EPB], fzns[EPB], ezs[EPB];
    static_for<0, EPB>([&](auto ic) {
      constexpr int i = decltype(ic)::value, qs = q0 + i, m = qs / RL, r = qs % RL;
      if constexpr (!(PZ && mid_slot<N, P, P * m + r * F::NRL>())) {
        const int n = F::out_ok(t, m) ? F::out_pos(t, m, r) : 0;
        const bool ok = F::out_ok(t, m) && keep_idx(n, N, K);
        if constexpr (TUNED) {
          if constexpr (EPI != EPI_INIT) {
            const long long c = ok ? tile + (long long)(P * m + r * F::NRL) * zs : tile;
            const unsigned o = ok ? lane : (unsigned)p;
#pragma unroll
            for (int f = 0; f < 3; f++) bvs[i][f] = ld_row(b + f * kp.fs + c, o);
          }
          fzs[i] = kp.fkz[n];
          fzns[i] = kp.fkzn[n];
          if constexpr (VS) ezs[i] = kp.emz[n];
        } else if (ok) {
          // only the stored slots load, and the trunc and integrating-factor
          // tables are read where they are used (registers set the occupancy)
          if constexpr (EPI != EPI_INIT) {
            const long long c = tile + (long long)(P * m + r * F::NRL) * zs;
#pragma unroll
            for (int f = 0; f < 3; f++) bvs[i][f] = ld_row(b + f * kp.fs + c, lane);
          }
          fzs[i] = kp.fkz[n];
        }
      }
    });
    static_for<0, EPB>([&](auto ic) {
      constexpr int i = decltype(ic)::value, qs = q0 + i, m = qs / RL, r = qs % RL;
      if constexpr (!(PZ && mid_slot<N, P, P * m + r * F::NRL>())) {
        if (!F::out_ok(t, m)) return;
        const int n = F::out_pos(t, m, r);
        if (!keep_idx(n, N, K)) return;
        const long long c = tile + (long long)(P * m + r * F::NRL) * zs;  // + lane
        const double fz = fzs[i];
        const bool k0 = (kx == 0) && (j == 0) && (n == 0);
        cx<double> a[3], bv[3];
#pragma unroll
        for (int f = 0; f < 3; f++) a[f] = {(double)y[f][qs].x * kp.scale, (double)y[f][qs].y * kp.scale};
        const double k2 = fxs + fys + fz * fz;
        if constexpr (EPI == EPI_INIT) {
          // incompr (src/trunc.f90:47-63), k = 0 skipped
          if (!k0) {
            double sr = fx * a[0].x + fy * a[1].x;
            double si = fx * a[0].y + fy * a[1].y;
            sr = sr + fz * a[2].x;
            si = si + fz * a[2].y;
            a[0] = {a[0].x - sr * fx / k2, a[0].y - si * fx / k2};
            a[1] = {a[1].x - sr * fy / k2, a[1].y - si * fy / k2};
            a[2] = {a[2].x - sr * fz / k2, a[2].y - si * fz / k2};
          }
        } else {
          // Coriolis with the start-of-step b (src/nlt.f90:67-81); b is read
          // once and kept for the update below
          const double of = omega * fz;
#pragma unroll
          for (int f = 0; f < 3; f++) {
            bv[f] = bvs[i][f];
            a[f] = {a[f].x - of * bv[f].y, a[f].y + of * bv[f].x};
          }
        }
        // invcurl, k = 0 left untouched (src/produ.f90:83-99), with one
        // reciprocal of k2 instead of three divisions
        cx<double> nl[3] = {a[0], a[1], a[2]};
        if (!k0) {
          const double ik2 = 1.0 / k2;
          const double qx = fx * ik2, qy = fy * ik2, qz = fz * ik2;
          nl[0] = {-qy * a[2].y + qz * a[1].y, qy * a[2].x - qz * a[1].x};
          nl[1] = {-qz * a[0].y + qx * a[2].y, qz * a[0].x - qx * a[2].x};
          nl[2] = {-qx * a[1].y + qy * a[0].y, qx * a[1].x - qy * a[0].x};
        }
        if (kp.truncate == 1) {
          // trunc (src/trunc.f90:14-19,27)
          const double kn = ((TUNED ? fzns[i] : kp.fkzn[n]) + kyn) + kxn;
          if (kn > kp.rtrunc2 || k0) {
#pragma unroll
            for (int f = 0; f < 3; f++) nl[f] = {0.0, 0.0};
          }
        }
        if constexpr (EPI == EPI_NONLIN || EPI == EPI_INIT) {
#pragma unroll
          for (int f = 0; f < 3; f++) st_row_p<NT_E>(out + f * kp.fs + c, lane, nl[f]);
        } else {
          const double e = VS ? exy * (TUNED ? ezs[i] : kp.emz[n]) : emk_of(kp, k2);
#pragma unroll
          for (int f = 0; f < 3; f++) {
            const cx<double> bb = bv[f];
            cx<double> o;
            if constexpr (EPI == EPI_STEP1) o = {e * (bb.x + kp.dt2 * nl[f].x), e * (bb.y + kp.dt2 * nl[f].y)};
            else o = {e * (e * bb.x + kp.dt * nl[f].x), e * (e * bb.y + kp.dt * nl[f].y)};
            if (k0) o = {0.0, 0.0};
            if constexpr (FA == 2) {
              // forcing (src/main.f90:149): the increments at this mode, in
              // table order (a position may repeat)
              const long long pos = c + lane, mo = (long long)mem * 3 * kp.fs;
              for (int q = fa0; q < fa1; q++) {
                const FEnt& fe = fent[fx2.idx[q]];
                if (fe.off - mo == pos) {
                  o.x = o.x + fe.d[2 * f];
                  o.y = o.y + fe.d[2 * f + 1];
                }
              }
            }
            if constexpr (FA == 1) bhr[f][qs] = o;
            else st_row_p<NT_E>(out + f * kp.fs + c, lane, o);
            if constexpr (FA == 2) bhr[f][qs] = o;
          }
          if constexpr (FZR) fza[qs] = fz;
        }
      }
    });
  });
  if constexpr (FA) {
    // pass A on the new state (as k_zinv): output q < 3 is u_q = curl_q bh,
    // u_x = i(ky bz - kz by), u_y = i(kz bx - kx bz), u_z = i(kx by - ky bx)
    // (src/produ.f90:48-53); q >= 3 is w = k^2 bh of field q - 3
    // (src/nlt.f90:30-36).  Slot qs holds element t + P m + r N/R of the
    // forward output, which is the inverse transform's input slot qs.
    using FI = Fft<W, N, P, 1, +1, LI>;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      cx<W> in[1][M0 * R0], o[1][ML * RL];
      static_for<0, NS>([&](auto qc) {
        constexpr int qs = decltype(qc)::value, m = qs / RL, r = qs % RL;
        if constexpr (PZ && mid_slot<N, P, P * m + r * F::NRL>()) {
          in[0][qs] = {W(0), W(0)};
        } else {
          const double fz = FZR ? fza[qs] : kp.fkz[F::out_pos(t, m, r)];
          double re, im;
          if (q < 3) {
            const int fa = q == 0 ? 2 : q == 1 ? 0 : 1;
            const int fb = q == 0 ? 1 : q == 1 ? 2 : 0;
            const double ca = (q == 0) ? fy : (q == 1) ? fz : fx;
            const double cb = (q == 0) ? fz : (q == 1) ? fx : fy;
            const cx<double> va = bhr[fa][qs], vb = bhr[fb][qs];
            re = -ca * va.y + cb * vb.y;
            im = ca * va.x - cb * vb.x;
          } else {
            const double k2 = fx * fx + fy * fy + fz * fz;
            const cx<double> va = bhr[q - 3][qs];
            re = k2 * va.x;
            im = k2 * va.y;
          }
          in[0][qs] = {(W)re, (W)im};
        }
      });
      if constexpr (PRE) FI::run(lay, lds, tw, twr, t, in, o);
      else FI::run(lay, lds, twt, t, in, o);
      if (K >= 0) {
        cx<W>* dst = wout + q * kp.fs + tile;
#pragma unroll
        for (int m = 0; m < ML; m++)
          if (F::out_ok(t, m))
#pragma unroll
            for (int r = 0; r < RL; r++)
              st_row_p<NT_A>(dst + (long long)(P * m + r * F::NRL) * zs, lane, o[0][m * RL + r]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// dynamic LDS of the z passes: the transform's rows, and for FP32 the copy
// of the twiddle table behind them
template <typename W, int N, int T>
constexpr size_t zlds() {
  return (size_t)N * T * sizeof(cx<W>) + ((sizeof(W) == 4 && GT_TWLDS32) ? Plan<N>::twsize() * sizeof(cx<W>) : 0);
}
template <typename W, int N, int T>
static hipError_t zinv_nt(int nout, const KP& kp, const void* b, void* work, const void* tw, hipStream_t s) {
  constexpr int P = pencil_P(N);
  // all six outputs: three blocks per column, each a curl and a k^2 component
  const bool pair = GT_ZINV_PAIR && nout == 6;
  const int nb = pair ? 3 : nout;
  const int grid = (kp.kxcut + T) / T * kp.ny * nb * kp.nmem;
  const cx<double>* bb = (const cx<double>*)b;
  cx<W>* w = (cx<W>*)work;
  const cx<W>* t = (const cx<W>*)tw;
  if constexpr (pz_len(N)) {
    if (kp.pruned && pz_cut(N, kp.kzcut)) {
      if (pair) k_zinv<W, N, T, P, false, true, GT_ZINV_PAIR != 0><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, bb, w, t, nb);
      else k_zinv<W, N, T, P, false, true><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, bb, w, t, nb);
      return hipGetLastError();
    }
  }
  if (pair) k_zinv<W, N, T, P, false, false, GT_ZINV_PAIR != 0><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, bb, w, t, nb);
  else k_zinv<W, N, T, P><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, bb, w, t, nb);
  return hipGetLastError();
}

template <typename W, int N>
static hipError_t zinv_n(int nout, const KP& kp, const void* b, void* work, const void* tw, hipStream_t s) {
  if constexpr (alt_T(N) > 0) {
    if (kp.tz == alt_T(N)) return zinv_nt<W, N, alt_T(N)>(nout, kp, b, work, tw, s);
  }
  return zinv_nt<W, N, pencil_T(N)>(nout, kp, b, work, tw, s);
}

template <typename W, int N, int T>
static hipError_t zfwd_nt(int epi, const KP& kp, const void* work, const void* b, void* out, const void* tw,
                          hipStream_t s) {
  constexpr int P = pencil_P(N);
  const int grid = (kp.kxcut + T) / T * kp.ny * kp.nmem;
  const cx<W>* w = (const cx<W>*)work;
  const cx<double>* bb = (const cx<double>*)b;
  cx<double>* o = (cx<double>*)out;
  const cx<W>* t = (const cx<W>*)tw;
  const bool vs = kp.emx != nullptr;
  if constexpr (pz_len(N)) {
    // the step epilogues of the pruned step: dropped output slots
    if (kp.pruned && pz_cut(N, kp.kzcut) && (epi == EPI_STEP1 || epi == EPI_STEP2)) {
      if (epi == EPI_STEP1) {
        if (vs) k_zfwd<W, N, T, P, EPI_STEP1, true, true><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{});
        else k_zfwd<W, N, T, P, EPI_STEP1, false, true><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{});
      } else {
        if (vs) k_zfwd<W, N, T, P, EPI_STEP2, true, true><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{});
        else k_zfwd<W, N, T, P, EPI_STEP2, false, true><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{});
      }
      return hipGetLastError();
    }
  }
  switch (epi) {
    case EPI_STEP1:
      if (vs) k_zfwd<W, N, T, P, EPI_STEP1, true><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{});
      else k_zfwd<W, N, T, P, EPI_STEP1><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{});
      break;
    case EPI_STEP2:
      if (vs) k_zfwd<W, N, T, P, EPI_STEP2, true><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{});
      else k_zfwd<W, N, T, P, EPI_STEP2><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{});
      break;
    case EPI_NONLIN: k_zfwd<W, N, T, P, EPI_NONLIN><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{}); break;
    case EPI_INIT: k_zfwd<W, N, T, P, EPI_INIT><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, w, bb, o, t, nullptr, FAux{}); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename W, int N>
static hipError_t zfwd_n(int epi, const KP& kp, const void* work, const void* b, void* out, const void* tw,
                         hipStream_t s) {
  if constexpr (alt_T(N) > 0) {
    if (kp.tze == alt_T(N)) return zfwd_nt<W, N, alt_T(N)>(epi, kp, work, b, out, tw, s);
  }
  return zfwd_nt<W, N, pencil_T(N)>(epi, kp, work, b, out, tw, s);
}

hipError_t launch_zinv(int prec, int nout, const KP& kp, const void* b, void* work, const void* tw,
                       hipStream_t s) {
  switch (kp.nz) {
#define GT_CASE(N)                                                         \
  case N:                                                                  \
    return prec ? zinv_n<float, N>(nout, kp, b, work, tw, s)               \
                : zinv_n<double, N>(nout, kp, b, work, tw, s);
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

// pass A of a = Z emk b (FP64 work fields, default pencil width)
hipError_t launch_zinv_emk(const KP& kp, const void* b, void* work, const void* tw, hipStream_t s) {
  switch (kp.nz) {
#define GT_CASE(N)                                                                                     \
  case N: {                                                                                            \
    constexpr int T = pencil_T(N), P = pencil_P(N);                                                    \
    constexpr int NB = GT_ZINV_PAIR ? 3 : 6; /* blocks per column */                                   \
    const int grid = (kp.kxcut + T) / T * kp.ny * NB;                                                  \
    k_zinv<double, N, T, P, true, false, GT_ZINV_PAIR != 0><<<grid, T * P, N * T * sizeof(cx<double>), s>>>(       \
        kp, (const cx<double>*)b, (cx<double>*)work, (const cx<double>*)tw, NB);                         \
    return hipGetLastError();                                                                          \
  }
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_zfwd(int prec, int epi, const KP& kp, const void* work, const void* b, void* out, const void* tw,
                       hipStream_t s) {
  switch (kp.nz) {
#define GT_CASE(N)                                                          \
  case N:                                                                   \
    return prec ? zfwd_n<float, N>(epi, kp, work, b, out, tw, s)            \
                : zfwd_n<double, N>(epi, kp, work, b, out, tw, s);
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

// Pass E of stage 1 fused with pass A of stage 2, and pass E of stage 2
// fused with the forcing and pass A of the next step (k_zfwd with FA 1 and
// 2): they read the products and b and write the six z-inverse outputs into
// the work fields (FA 2 also the new b).  Compiled with simple viscosity for
// 512-point z transforms at the default width of 8 pencils (one block of 8
// wavefronts per CU at its register count; 4 pencils, two blocks per CU,
// measured slower: 1.19 against 1.03 ms per step in FP64, 1.63 against 0.78
// in FP32-bulk), where the z-forward width of the context must be the
// default one, which pass E of stage 2 uses; and for 128- and 256-point
// transforms at both widths.  gt_autotune decides whether a context uses
// them.
bool zfwd_zinv_supported(const KP& kp) {
  if (kp.emx == nullptr) return false;
  if (kp.nz == 128 || kp.nz == 256) return true;  // both widths
  return kp.nz == 512 && (alt_T(512) == 0 || kp.tze != alt_T(512));
}

template <typename W, int N, int T>
static hipError_t zfwd_zinv_w(int stage, const KP& kp, void* work, const void* b, void* out, const void* tw,
                              const FAux& fx, hipStream_t s) {
  constexpr int P = pencil_P(N);
  const int grid = (kp.kxcut + T) / T * kp.ny * kp.nmem;
  const cx<double>* bb = (const cx<double>*)b;
  cx<double>* o = (cx<double>*)out;
  cx<W>* wk = (cx<W>*)work;
  const cx<W>* t = (const cx<W>*)tw;
  // the pruned slots as in the unfused z-forward pass (lengths of 256 and up)
  const bool pz = pz_len(N) && kp.pruned && pz_cut(N, kp.kzcut);
  if (stage == 1) {
    // out is unused (bh stays in registers)
    if (pz) k_zfwd<W, N, T, P, EPI_STEP1, true, true, 1><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, wk, bb, nullptr, t, wk, fx);
    else k_zfwd<W, N, T, P, EPI_STEP1, true, false, 1><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, wk, bb, nullptr, t, wk, fx);
  } else {
    if (pz) k_zfwd<W, N, T, P, EPI_STEP2, true, true, 2><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, wk, bb, o, t, wk, fx);
    else k_zfwd<W, N, T, P, EPI_STEP2, true, false, 2><<<grid, T * P, zlds<W, N, T>(), s>>>(kp, wk, bb, o, t, wk, fx);
  }
  return hipGetLastError();
}

// stage 1: pass E of stage 1 with pass A of stage 2 (b read, out unused);
// stage 2: pass E of stage 2 into out (b in place), the forcing of fx, then
// pass A of the next step
hipError_t launch_zfwd_zinv(int prec, int stage, const KP& kp, void* work, const void* b, void* out, const void* tw,
                            const FAux& fx, hipStream_t s) {
  if (!zfwd_zinv_supported(kp) || (stage != 1 && stage != 2)) return hipErrorInvalidValue;
  if (kp.nz == 128) {
    if (kp.tze == alt_T(128))
      return prec ? zfwd_zinv_w<float, 128, alt_T(128)>(stage, kp, work, b, out, tw, fx, s)
                  : zfwd_zinv_w<double, 128, alt_T(128)>(stage, kp, work, b, out, tw, fx, s);
    return prec ? zfwd_zinv_w<float, 128, pencil_T(128)>(stage, kp, work, b, out, tw, fx, s)
                : zfwd_zinv_w<double, 128, pencil_T(128)>(stage, kp, work, b, out, tw, fx, s);
  }
  if (kp.nz == 256) {
    if (kp.tze == alt_T(256))
      return prec ? zfwd_zinv_w<float, 256, alt_T(256)>(stage, kp, work, b, out, tw, fx, s)
                  : zfwd_zinv_w<double, 256, alt_T(256)>(stage, kp, work, b, out, tw, fx, s);
    return prec ? zfwd_zinv_w<float, 256, pencil_T(256)>(stage, kp, work, b, out, tw, fx, s)
                : zfwd_zinv_w<double, 256, pencil_T(256)>(stage, kp, work, b, out, tw, fx, s);
  }
  return prec ? zfwd_zinv_w<float, 512, pencil_T(512)>(stage, kp, work, b, out, tw, fx, s)
              : zfwd_zinv_w<double, 512, pencil_T(512)>(stage, kp, work, b, out, tw, fx, s);
}

}  // namespace gt
