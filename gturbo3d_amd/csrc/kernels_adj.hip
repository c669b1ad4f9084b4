// kernels_adj.hip -- kernels of the adjoint (reverse-mode derivative) of the
// step, FP64.
//
// The forward nonlinear term is N(a; c) = T I (F[u x w] + i*omega*kz*c) with
// u = F^-1[curl a], w = F^-1[k^2 a] (src/nlt.f90:1-95), T = trunc, I = invcurl.
// With the real inner product sum Re(conj(x) y) over the stored half
// spectrum (the inner product of the (re, im) pairs):
//   curl, invcurl, k^2, trunc and emk are self-adjoint, (i omega kz)^T = -i omega kz,
//   (F^-1)^T = wgt * Fu with Fu the unnormalised R2C transform and wgt = 1 on
//     the kx = 0 and kx = NX/2 planes, 2 elsewhere (the C2R counts the other
//     modes twice and keeps only the real part of those two planes),
//   F^T = F^-1(wgt' .)/(NX NY NZ) with wgt' = 1 on those planes, 1/2 elsewhere.
// So for a cotangent lam of N, with mu = I T lam and p = F^T mu:
//   dN/da^T lam = wgt * (curl Fu[w x p] + k^2 Fu[p x u])
//   dN/dc^T lam = -i omega kz mu
// (from p.(du x w + u x dw) = du.(w x p) + dw.(p x u)).
// Passes of one adjoint evaluation (a = the state the forward step fed to nlt):
//   k_zinv (forward pass A) : curl a, k^2 a z-inverse     -> work fields 0-5
//   k_zinv_adj              : z stage of F^T mu            -> work fields 6-8
//   k_yfft, 9 fields, inverse
//   k_xadj                  : 9 C2R, w x p and p x u, 6 unnormalised R2C x wgt
//   k_yfft, 6 fields, forward
//   k_zfwd_adj              : z forward, curl A + k^2 B
// In pruned mode the state and every cotangent live on the modes trunc
// keeps (that set is invariant under the step), so outputs are projected on
// it; in dense mode nothing is projected.
#include "gt_config.hpp"
#include "gt_device.hpp"

namespace gt {

// z stage of F^T applied to mu = I T (s * nu): output component q = blockIdx.z.
// ZE: the input is lam and nu = Z emk lam is formed as it is loaded.
template <int N, int T, int P, bool ZE>
__global__ __launch_bounds__(T* P) void k_zinv_adj(KP kp, const cx<double>* __restrict__ nu, double s,
                                                   cx<double>* __restrict__ out, const cx<double>* __restrict__ tw) {
  using F = Fft<double, N, P, 1, +1, LayInterleaved<T>>;
  constexpr int R0 = F::R0, RL = F::RL, M0 = F::M0, ML = F::ML;
  GT_DYN_LDS(cx<double>, lds);  // N T elements
  const int tid = threadIdx.x, p = tid % T, t = tid / T;
  const int kx = blockIdx.x * T + p;
  const int j = blockIdx.y;
  const int q = blockIdx.z;
  int K = -1;
  if (kx <= kp.kxcut) K = kp.pruned ? (int)kp.kzext[j * kp.nxp + kx] : (N >> 1);
  static_assert(64 % T == 0, "every wavefront holds all T pencils");
  if (!wave_any(K >= 0)) return;  // uniform over the block
  const int col = j * kp.nxp + kx;
  const int zs = kp.ny * kp.nxp;
  const int kxc = kx < kp.nxh ? kx : kp.nxh - 1;
  const double fx = kp.fkx[kxc], fy = kp.fky[j];
  const double wp = (kxc == 0 || kxc == kp.nxh - 1) ? s * kp.scale : 0.5 * s * kp.scale;
  // (invcurl a)_q = i (ca a_fa - cb a_fb) / k^2 (src/produ.f90:83-99)
  const int fa = q == 0 ? 2 : q == 1 ? 0 : 1;
  const int fb = q == 0 ? 1 : q == 1 ? 2 : 0;
  const cx<double>* A = nu + fa * kp.fs + col;
  const cx<double>* B = nu + fb * kp.fs + col;
  const cx<double>* Q = nu + q * kp.fs + col;
  cx<double> in[1][M0 * R0], o[1][ML * RL];
  // loads first and unconditional (element 0 of the column for an unused
  // slot), so that they are all in flight before the first wait
  cx<double> la[M0 * R0], lb[M0 * R0];
#pragma unroll
  for (int m = 0; m < M0; m++)
#pragma unroll
    for (int r = 0; r < R0; r++) {
      const int n = F::in_pos(t, m, r);
      const int at = (F::in_ok(t, m) && K >= 0 && keep_idx(n, N, K)) ? n : 0;
      la[m * R0 + r] = A[at * zs];
      lb[m * R0 + r] = B[at * zs];
    }
  const cx<double> c0 = Q[0];
#pragma unroll
  for (int m = 0; m < M0; m++)
#pragma unroll
    for (int r = 0; r < R0; r++) {
      const int n = F::in_pos(t, m, r);
      cx<double> v = {0.0, 0.0};
      if (F::in_ok(t, m) && K >= 0 && keep_idx(n, N, K) && !trunc_zero(kp, kxc, j, n)) {
        const double fz = kp.fkz[n];
        const double k2 = fx * fx + fy * fy + fz * fz;
        const bool k0 = kxc == 0 && j == 0 && n == 0;
        const double e = ZE ? (k0 ? 0.0 : emk_at(kp, kxc, j, n, k2)) : 1.0;
        if (k0) {  // invcurl leaves k = 0 as it is
          v = {wp * e * c0.x, wp * e * c0.y};
        } else {
          const cx<double> a = la[m * R0 + r], bb = lb[m * R0 + r];
          const double ca = ((q == 0) ? fy : (q == 1) ? fz : fx) / k2;
          const double cb = ((q == 0) ? fz : (q == 1) ? fx : fy) / k2;
          const double we = wp * e;
          v = {we * (-ca * a.y + cb * bb.y), we * (ca * a.x - cb * bb.x)};
        }
      }
      in[0][m * R0 + r] = v;
    }
  LayInterleaved<T> lay{p};
  F::run(lay, lds, tw, t, in, o);
  if (K < 0) return;
  cx<double>* dst = out + q * kp.fs + col;
#pragma unroll
  for (int m = 0; m < ML; m++)
    if (F::out_ok(t, m))
#pragma unroll
      for (int r = 0; r < RL; r++) dst[F::out_pos(t, m, r) * zs] = o[0][m * RL + r];
}

// One block per row pair, nine groups of P threads: group g inverse-
// transforms field g (u 0-2, w 3-5, p 6-8) into LDS row g; groups 0-5 then
// form (w x p)_c (g = c) and (p x u)_c (g = 3 + c) at the forward
// transform's input positions, transform them and store wgt * Fu of both
// rows over the first six fields.
template <int N, int P>
__global__ __launch_bounds__(9 * P) void k_xadj(KP kp, cx<double>* work, const cx<double>* __restrict__ tw) {
  using L = LaySwz<N>;
  constexpr int SY = (P <= 64) ? SYNC_WAVE : SYNC_BLOCK;
  using FI = Fft<double, N, P, 1, +1, L, SY>;
  using FF = Fft<double, N, P, 1, -1, L, SY>;
  GT_DYN_LDS(cx<double>, lds);  // 9 N elements
  const int tid = threadIdx.x, g = tid / P, t = tid % P;
  const int y0 = 2 * blockIdx.x, z = blockIdx.y;
  const int r0 = (z * kp.ny + y0) * kp.nxp;
  const int r1 = r0 + kp.nxp;
  const int kc = kp.kxcut;
  const L lay{g, 0};
  // the twiddles of the inverse transform loaded with the data (the forward
  // transform of the same length reads the same entries), at the lengths
  // whose x passes do so too; the others read the table in every stage
  constexpr bool TWR = N == 128 || N == 256 || N == 512;
  typename FI::Twr twr;
  {
    cx<double> in[1][FI::E0], out[1][FI::EL];
    const cx<double>* src = work + g * kp.fs;
    static_for<0, FI::M0 * FI::R0>([&](auto qc) {
      constexpr int q = decltype(qc)::value, m = q / FI::R0, r = q % FI::R0;
      constexpr int C = P * m + r * FI::NR0;  // in_pos(t, m, r) = t + C
      cx<double> v = {0.0, 0.0};
      if (FI::in_ok(t, m)) v = c2r_in_at<double, N, P, C>(src + r0, src + r1, t, kc);
      in[0][q] = v;
    });
    if constexpr (TWR) {
      FI::load_tw(tw, t, twr);
      FI::run(lay, lds, tw, twr, t, in, out);
    } else {
      FI::run(lay, lds, tw, t, in, out);
    }
    stage_sync<SY>();
#pragma unroll
    for (int m = 0; m < FI::ML; m++)
      if (FI::out_ok(t, m))
#pragma unroll
        for (int r = 0; r < FI::RL; r++) lds[g * N + FI::out_pos(t, m, r)] = out[0][m * FI::RL + r];
  }
  __syncthreads();
  // (a x b)_c = a_{c+1} b_{c+2} - a_{c+2} b_{c+1}, rows in .x and .y
  cx<double> fin[1][FF::E0], fout[1][FF::EL];
  const int c = g % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
  const int fa = (g < 3) ? 3 : 6, fb = (g < 3) ? 6 : 0;
#pragma unroll
  for (int m = 0; m < FF::M0; m++)
#pragma unroll
    for (int r = 0; r < FF::R0; r++) {
      cx<double> v = {0.0, 0.0};
      if (g < 6 && FF::in_ok(t, m)) {
        const int o = FF::in_pos(t, m, r);
        const cx<double> a1 = lds[(fa + c1) * N + o], a2 = lds[(fa + c2) * N + o];
        const cx<double> b1 = lds[(fb + c1) * N + o], b2 = lds[(fb + c2) * N + o];
        v = {a1.x * b2.x - a2.x * b1.x, a1.y * b2.y - a2.y * b1.y};
      }
      fin[0][m * FF::R0 + r] = v;
    }
  __syncthreads();  // every product read is done before a forward transform writes LDS
  if (SY == SYNC_BLOCK || g < 6) {
    if constexpr (TWR) FF::run(lay, lds, tw, twr, t, fin, fout);
    else FF::run(lay, lds, tw, t, fin, fout);
    stage_sync<SY>();
#pragma unroll
    for (int m = 0; m < FF::ML; m++)
      if (FF::out_ok(t, m))
#pragma unroll
        for (int r = 0; r < FF::RL; r++) lds[g * N + FF::out_pos(t, m, r)] = fout[0][m * FF::RL + r];
  }
  __syncthreads();
  if (g < 6) {
    cx<double>* d0 = work + g * kp.fs + r0;
    cx<double>* d1 = work + g * kp.fs + r1;
    for (int k = t; k <= kc; k += P) {
      const int kk = (k == 0) ? 0 : N - k;
      const double h = (k == 0 || k == N / 2) ? 0.5 : 1.0;  // wgt / 2
      const cx<double> Y = lds[g * N + k], Yc = lds[g * N + kk];
      st_row_p<NT_C>(d0, (unsigned)k, cx<double>{h * (Y.x + Yc.x), h * (Y.y - Yc.y)});
      st_row_p<NT_C>(d1, (unsigned)k, cx<double>{h * (Y.y + Yc.y), h * (Yc.x - Y.x)});
    }
  }
}

// z forward of A = Fu[w x p] (fields 0-2) and B = Fu[p x u] (3-5), then
// g_q = (curl A)_q + k^2 B_q (wgt is already applied); component q = blockIdx.z.
// POST (reference dt2 mode): instead of g, write the cotangent of the step's
// input, out_q = emk nu_q - i omega kz (I T (dt nu))_q + Z emk g_q with
// nu = Z emk lam (lam = the cotangent of the step's output; out must not
// alias it: the three component blocks all read lam).
template <int N, int T, int P, bool POST>
__global__ __launch_bounds__(T* P) void k_zfwd_adj(KP kp, const cx<double>* __restrict__ work,
                                                   cx<double>* __restrict__ out, const cx<double>* __restrict__ tw,
                                                   const cx<double>* __restrict__ lam) {
  using F = Fft<double, N, P, 1, -1, LayInterleaved<T>>;
  constexpr int R0 = F::R0, RL = F::RL, M0 = F::M0, ML = F::ML;
  GT_DYN_LDS(cx<double>, lds);  // N T elements
  const int tid = threadIdx.x, p = tid % T, t = tid / T;
  const int kx = blockIdx.x * T + p;
  const int j = blockIdx.y;
  const int q = blockIdx.z;
  int K = -1;
  if (kx <= kp.kxcut) K = kp.pruned ? (int)kp.kzext[j * kp.nxp + kx] : (N >> 1);
  static_assert(64 % T == 0, "every wavefront holds all T pencils");
  if (!wave_any(K >= 0)) return;  // uniform over the block
  const int col = j * kp.nxp + kx;
  const int zs = kp.ny * kp.nxp;
  LayInterleaved<T> lay{p};
  const int fa = q == 0 ? 2 : q == 1 ? 0 : 1;
  const int fb = q == 0 ? 1 : q == 1 ? 2 : 0;
  cx<double> y[3][ML * RL];
#pragma unroll
  for (int h = 0; h < 3; h++) {
    cx<double> in[1][M0 * R0], o[1][ML * RL];
    const int f = (h == 0) ? fa : (h == 1) ? fb : 3 + q;
    const cx<double>* src = work + f * kp.fs + col;
#pragma unroll
    for (int m = 0; m < M0; m++)
#pragma unroll
      for (int r = 0; r < R0; r++) {
        cx<double> v = {0.0, 0.0};
        if (K >= 0 && F::in_ok(t, m)) v = src[F::in_pos(t, m, r) * zs];
        in[0][m * R0 + r] = v;
      }
    F::run(lay, lds, tw, t, in, o);
#pragma unroll
    for (int e = 0; e < ML * RL; e++) y[h][e] = o[0][e];
  }
  if (K < 0) return;
  const int kxc = kx < kp.nxh ? kx : kp.nxh - 1;
  const double fx = kp.fkx[kxc], fy = kp.fky[j];
  cx<double>* dst = out + q * kp.fs + col;
#pragma unroll
  for (int m = 0; m < ML; m++) {
    if (!F::out_ok(t, m)) continue;
#pragma unroll
    for (int r = 0; r < RL; r++) {
      const int n = F::out_pos(t, m, r);
      if (!keep_idx(n, N, K)) continue;
      const double fz = kp.fkz[n];
      const double k2 = fx * fx + fy * fy + fz * fz;
      // (curl A)_q = i (ca A_fa - cb A_fb) (src/produ.f90:48-53)
      const double ca = (q == 0) ? fy : (q == 1) ? fz : fx;
      const double cb = (q == 0) ? fz : (q == 1) ? fx : fy;
      const cx<double> a = y[0][m * RL + r], bb = y[1][m * RL + r], cc = y[2][m * RL + r];
      cx<double> v = {-ca * a.y + cb * bb.y + k2 * cc.x, ca * a.x - cb * bb.x + k2 * cc.y};
      if constexpr (POST) {
        const bool k0 = kxc == 0 && j == 0 && n == 0;
        const double e = emk_at(kp, kxc, j, n, k2);
        const double ez = k0 ? 0.0 : e;
        const int o = col + n * zs;
        const cx<double> lq = lam[q * kp.fs + o], la = lam[fa * kp.fs + o], lb = lam[fb * kp.fs + o];
        const cx<double> nq = {ez * lq.x, ez * lq.y};
        // Coriolis: -i omega kz (I T (dt nu))_q, I = invcurl (k = 0 untouched)
        cx<double> iv = {0.0, 0.0};
        if (!trunc_zero(kp, kxc, j, n)) {
          if (k0) {
            iv = {kp.dt * nq.x, kp.dt * nq.y};
          } else {
            const double sa = kp.dt * ez * ca / k2, sb = kp.dt * ez * cb / k2;
            iv = {-sa * la.y + sb * lb.y, sa * la.x - sb * lb.x};
          }
        }
        const double of = kp.omega * fz;
        v = {e * nq.x + of * iv.y + ez * v.x, e * nq.y - of * iv.x + ez * v.y};
      }
      if (kp.pruned && trunc_zero(kp, kxc, j, n)) v = {0.0, 0.0};
      dst[n * zs] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// pointwise parts of the step adjoint, one thread per stored mode
// ---------------------------------------------------------------------------
struct Mode {
  int i, j, k;
  long long q;
  double fx, fy, fz, k2;
};

__device__ __forceinline__ bool mode_of(const KP& kp, long long e, Mode& md) {
  const long long total = (long long)kp.nxh * kp.ny * kp.nz;
  if (e >= total) return false;
  md.i = (int)(e % kp.nxh);
  const long long rest = e / kp.nxh;
  md.j = (int)(rest % kp.ny);
  md.k = (int)(rest / kp.ny);
  md.q = ((long long)md.k * kp.ny + md.j) * kp.nxp + md.i;
  md.fx = kp.fkx[md.i];
  md.fy = kp.fky[md.j];
  md.fz = kp.fkz[md.k];
  md.k2 = md.fx * md.fx + md.fy * md.fy + md.fz * md.fz;
  return true;
}

__device__ __forceinline__ bool is_k0(const Mode& md) { return md.i == 0 && md.j == 0 && md.k == 0; }

// o = invcurl(a) (src/produ.f90:83-99), k = 0 left as it is
__device__ __forceinline__ void invcurl_at(const Mode& md, const cx<double> a[3], cx<double> o[3]) {
  if (is_k0(md)) {
    o[0] = a[0];
    o[1] = a[1];
    o[2] = a[2];
    return;
  }
  const double qx = md.fx / md.k2, qy = md.fy / md.k2, qz = md.fz / md.k2;
  o[0] = {-qy * a[2].y + qz * a[1].y, qy * a[2].x - qz * a[1].x};
  o[1] = {-qz * a[0].y + qx * a[2].y, qz * a[0].x - qx * a[2].x};
  o[2] = {-qx * a[1].y + qy * a[0].y, qx * a[1].x - qy * a[0].x};
}

// -i omega kz I T (s x): the Coriolis part of dN/dc^T applied to s x
__device__ __forceinline__ void coriolis_adj(const KP& kp, const Mode& md, const cx<double>* x, double s,
                                             cx<double> o[3]) {
  const bool tz = trunc_zero(kp, md.i, md.j, md.k);
  cx<double> tv[3], iv[3];
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const cx<double> v = x[f * kp.fs + md.q];
    tv[f] = tz ? cx<double>{0.0, 0.0} : cx<double>{s * v.x, s * v.y};
  }
  invcurl_at(md, tv, iv);
  const double of = kp.omega * md.fz;
#pragma unroll
  for (int f = 0; f < 3; f++) o[f] = {of * iv[f].y, -of * iv[f].x};
}

// out = Z emk in (Z: k = 0 set to zero, as step1/step2 do)
__global__ void k_adj_ze(KP kp, const cx<double>* __restrict__ in, cx<double>* __restrict__ out) {
  Mode md;
  if (!mode_of(kp, (long long)blockIdx.x * blockDim.x + threadIdx.x, md)) return;
  const double e = emk_at(kp, md.i, md.j, md.k, md.k2);
  const bool k0 = is_k0(md);
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const cx<double> v = in[f * kp.fs + md.q];
    out[f * kp.fs + md.q] = k0 ? cx<double>{0.0, 0.0} : cx<double>{e * v.x, e * v.y};
  }
}

// Cotangent of b through step2 and the stage-1 map bh = Z emk b (reference
// mode): lam_out = emk nu - i omega kz I T (dt nu) + Z emk g, with
// nu = Z emk lam+ and g = dN/da^T (dt nu) at a = bh.
__global__ void k_adj_post(KP kp, const cx<double>* __restrict__ nu, const cx<double>* __restrict__ g,
                           cx<double>* __restrict__ lam_out) {
  Mode md;
  if (!mode_of(kp, (long long)blockIdx.x * blockDim.x + threadIdx.x, md)) return;
  const double e = emk_at(kp, md.i, md.j, md.k, md.k2);
  const bool k0 = is_k0(md);
  const bool proj = kp.pruned && trunc_zero(kp, md.i, md.j, md.k);
  cx<double> co[3];
  coriolis_adj(kp, md, nu, kp.dt, co);
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const cx<double> v = nu[f * kp.fs + md.q];
    const cx<double> gg = g[f * kp.fs + md.q];
    cx<double> o = {e * v.x + co[f].x, e * v.y + co[f].y};
    if (!k0) o = {o.x + e * gg.x, o.y + e * gg.y};
    if (proj) o = {0.0, 0.0};
    lam_out[f * kp.fs + md.q] = o;
  }
}

// Midpoint mode, stage 1 (bh = Z emk (b + dt2 N(b; b))): with h = Z emk g,
// acc += g1 - i omega kz I T (dt2 h), g1 = dN/da^T (dt2 h) at a = b.
__global__ void k_adj_stage1(KP kp, const cx<double>* __restrict__ h, const cx<double>* __restrict__ g1,
                             cx<double>* __restrict__ acc) {
  Mode md;
  if (!mode_of(kp, (long long)blockIdx.x * blockDim.x + threadIdx.x, md)) return;
  const bool proj = kp.pruned && trunc_zero(kp, md.i, md.j, md.k);
  cx<double> co[3];
  coriolis_adj(kp, md, h, kp.dt2, co);
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const cx<double> a = acc[f * kp.fs + md.q];
    const cx<double> gg = g1[f * kp.fs + md.q];
    cx<double> o = {a.x + gg.x + co[f].x, a.y + gg.y + co[f].y};
    if (proj) o = {0.0, 0.0};
    acc[f * kp.fs + md.q] = o;
  }
}

// x <- x with the modes trunc zeroes set to zero (pruned mode: the cotangent
// of the final state restricted to the kept modes)
__global__ void k_adj_project(KP kp, cx<double>* __restrict__ x) {
  Mode md;
  if (!mode_of(kp, (long long)blockIdx.x * blockDim.x + threadIdx.x, md)) return;
  if (!trunc_zero(kp, md.i, md.j, md.k)) return;
#pragma unroll
  for (int f = 0; f < 3; f++) x[f * kp.fs + md.q] = {0.0, 0.0};
}

// dots[e] = sum_f Re(conj(lam_f[off_e]) d_f), the cotangent against one
// forcing increment of unit amplitude
__global__ void k_adj_fdot(const cx<double>* __restrict__ lam, long long fs, const FEnt* __restrict__ ent, int n,
                           double* __restrict__ dots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FEnt e = ent[i];
  double s = 0.0;
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const cx<double> l = lam[f * fs + e.off];
    s += l.x * e.d[2 * f] + l.y * e.d[2 * f + 1];
  }
  dots[i] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int N>
static hipError_t zinv_adj_n(const KP& kp, const void* nu, double s, bool ze, void* out, const void* tw,
                             hipStream_t st) {
  constexpr int T = pencil_T(N), P = pencil_P(N);
  dim3 grid((kp.kxcut + T) / T, kp.ny, 3);
  const cx<double>* x = (const cx<double>*)nu;
  if (ze)
    k_zinv_adj<N, T, P, true><<<grid, T * P, N * T * sizeof(cx<double>), st>>>(kp, x, s, (cx<double>*)out, (const cx<double>*)tw);
  else
    k_zinv_adj<N, T, P, false><<<grid, T * P, N * T * sizeof(cx<double>), st>>>(kp, x, s, (cx<double>*)out, (const cx<double>*)tw);
  return hipGetLastError();
}

template <int N>
static hipError_t zfwd_adj_n(const KP& kp, const void* work, void* out, const void* tw, const void* lam,
                             hipStream_t st) {
  constexpr int T = pencil_T(N), P = pencil_P(N);
  dim3 grid((kp.kxcut + T) / T, kp.ny, 3);
  const cx<double>* w = (const cx<double>*)work;
  const cx<double>* t = (const cx<double>*)tw;
  if (lam)
    k_zfwd_adj<N, T, P, true><<<grid, T * P, N * T * sizeof(cx<double>), st>>>(kp, w, (cx<double>*)out, t, (const cx<double>*)lam);
  else
    k_zfwd_adj<N, T, P, false><<<grid, T * P, N * T * sizeof(cx<double>), st>>>(kp, w, (cx<double>*)out, t, nullptr);
  return hipGetLastError();
}

// ze: the input is lam and the kernel forms nu = Z emk lam
hipError_t launch_zinv_adj(const KP& kp, const void* nu, double s, bool ze, void* out, const void* tw,
                           hipStream_t st) {
  switch (kp.nz) {
#define GT_CASE(N) \
  case N:          \
    return zinv_adj_n<N>(kp, nu, s, ze, out, tw, st);
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

// lam == nullptr: out = g; else out = the step cotangent formed from lam and g
hipError_t launch_zfwd_adj(const KP& kp, const void* work, void* out, const void* tw, const void* lam,
                           hipStream_t st) {
  switch (kp.nz) {
#define GT_CASE(N) \
  case N:          \
    return zfwd_adj_n<N>(kp, work, out, tw, lam, st);
    GT_FOR_EACH_N(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

// x lengths of the adjoint x pass: nine row groups per block must fit 1024
// threads and the 9 N complex LDS rows
#define GT_FOR_EACH_NX_ADJ(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

bool adjoint_nx_supported(int n) {
  switch (n) {
#define GT_CASE(N) case N:
    GT_FOR_EACH_NX_ADJ(GT_CASE)
#undef GT_CASE
    return true;
    default:
      return false;
  }
}

hipError_t launch_xadj(const KP& kp, void* work, const void* tw, hipStream_t st) {
  dim3 grid(kp.ny / 2, kp.nz);
  switch (kp.nx) {
#define GT_CASE(N)                                                                                    \
  case N:                                                                                             \
    k_xadj<N, row_P(N)><<<grid, 9 * row_P(N), 9 * N * sizeof(cx<double>), st>>>(kp, (cx<double>*)work, (const cx<double>*)tw); \
    return hipGetLastError();
    GT_FOR_EACH_NX_ADJ(GT_CASE)
#undef GT_CASE
    default:
      return hipErrorInvalidValue;
  }
}

static unsigned mode_blocks(const KP& kp) {
  const long long total = (long long)kp.nxh * kp.ny * kp.nz;
  return (unsigned)((total + 255) / 256);
}

hipError_t launch_adj_ze(const KP& kp, const void* in, void* out, hipStream_t st) {
  k_adj_ze<<<mode_blocks(kp), 256, 0, st>>>(kp, (const cx<double>*)in, (cx<double>*)out);
  return hipGetLastError();
}

hipError_t launch_adj_post(const KP& kp, const void* nu, const void* g, void* lam_out, hipStream_t st) {
  k_adj_post<<<mode_blocks(kp), 256, 0, st>>>(kp, (const cx<double>*)nu, (const cx<double>*)g, (cx<double>*)lam_out);
  return hipGetLastError();
}

hipError_t launch_adj_stage1(const KP& kp, const void* h, const void* g1, void* acc, hipStream_t st) {
  k_adj_stage1<<<mode_blocks(kp), 256, 0, st>>>(kp, (const cx<double>*)h, (const cx<double>*)g1, (cx<double>*)acc);
  return hipGetLastError();
}

hipError_t launch_adj_project(const KP& kp, void* x, hipStream_t st) {
  k_adj_project<<<mode_blocks(kp), 256, 0, st>>>(kp, (cx<double>*)x);
  return hipGetLastError();
}

hipError_t launch_adj_fdot(const void* lam, long long fs, const FEnt* ent, int n, double* dots, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_adj_fdot<<<(n + 127) / 128, 128, 0, st>>>((const cx<double>*)lam, fs, ent, n, dots);
  return hipGetLastError();
}

}  // namespace gt
