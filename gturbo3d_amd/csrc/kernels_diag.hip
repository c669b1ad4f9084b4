// kernels_diag.hip -- device diagnostics of the state (src/diagnostics.f90).
//
// One sweep over the half spectrum accumulates, per block, in LDS:
//   sp[ib-1]   += kk*kk*fac*(|bx|^2+|by|^2+|bz|^2), ib = int(kk)+1, 0 < ib < NBIN
//                 (spectrum, src/diagnostics.f90:22-35)
//   sp2d[ib-1] += kk*kk*fac*|bz|^2 on the kz = 0 plane (:37-48)
//   hel[ib-1]  += fac*Re(conj(u).w), u = curl b, w = curl u (helicity spectrum)
//   glob[0..8] the nine sums of glob over u = curl b, glob[9..17] over
//              w = curl u (:126-153, :164-191), k^2 > 0 only
//   flux[ib]   += fac2*kk*kk*Re(b.conj(N)), ib = int(kk) < NBIN (fluxes, :77-91)
// with fac = 1/2 (fac2 = 1) on the kx = 0 and kx = NX/2 planes, 1 (2)
// elsewhere, |z| the hypot of cdabs and kk the correctly rounded sqrt, so
// int(kk) bins like the reference's host sqrt.  A second kernel sums the
// per-block partials in block order.
#include "gt_config.hpp"
#include "gt_internal.hpp"

namespace gt {

constexpr int DIAG_NBIN_MAX = 512;

__device__ __forceinline__ double abs2_ref(cx<double> z) {
  const double h = hypot(z.x, z.y);
  return h * h;
}

__device__ __forceinline__ void curl_at(double fx, double fy, double fz, const cx<double> a[3], cx<double> u[3]) {
  // src/produ.f90:48-53
  u[0] = {-fy * a[2].y + fz * a[1].y, fy * a[2].x - fz * a[1].x};
  u[1] = {-fz * a[0].y + fx * a[2].y, fz * a[0].x - fx * a[2].x};
  u[2] = {-fx * a[1].y + fy * a[0].y, fx * a[1].x - fy * a[0].x};
}

// layout of one block's partial: sp[nbin] sp2d[nbin] hel[nbin] flux[nbin] glob[18]
__global__ __launch_bounds__(256) void k_diag(KP kp, const cx<double>* __restrict__ b, const cx<double>* __restrict__ nl,
                                              double* __restrict__ partial, int nbin) {
  __shared__ double acc[4 * DIAG_NBIN_MAX + 18];
  const int len = 4 * nbin + 18;
  for (int q = threadIdx.x; q < len; q += blockDim.x) acc[q] = 0.0;
  __syncthreads();
  double* sp = acc;
  double* sp2d = acc + nbin;
  double* hel = acc + 2 * nbin;
  double* flux = acc + 3 * nbin;
  double* glob = acc + 4 * nbin;
  double g[18];
#pragma unroll
  for (int q = 0; q < 18; q++) g[q] = 0.0;

  const long long total = (long long)kp.nxh * kp.ny * kp.nz;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e % kp.nxh);
    const long long rest = e / kp.nxh;
    const int j = (int)(rest % kp.ny);
    const int k = (int)(rest / kp.ny);
    const long long q = ((long long)k * kp.ny + j) * kp.nxp + i;
    const double fx = kp.fkx[i], fy = kp.fky[j], fz = kp.fkz[k];
    const double k2 = fx * fx + fy * fy + fz * fz;
    const double kk = __dsqrt_rn(k2);
    const bool edge = (i == 0) || (i == kp.nxh - 1);
    const double fac = edge ? 0.5 : 1.0;
    cx<double> a[3] = {b[q], b[kp.fs + q], b[2 * kp.fs + q]};
    const int ib = (int)kk + 1;
    if (ib > 0 && ib < nbin) {
      const double s = abs2_ref(a[0]) + abs2_ref(a[1]) + abs2_ref(a[2]);
      atomicAdd(&sp[ib - 1], kk * kk * fac * s);
      if (k == 0) atomicAdd(&sp2d[ib - 1], kk * kk * fac * abs2_ref(a[2]));
    }
    cx<double> u[3], w[3];
    curl_at(fx, fy, fz, a, u);
    curl_at(fx, fy, fz, u, w);
    if (ib > 0 && ib < nbin) {
      const double h = (u[0].x * w[0].x + u[0].y * w[0].y) + (u[1].x * w[1].x + u[1].y * w[1].y) +
                       (u[2].x * w[2].x + u[2].y * w[2].y);
      atomicAdd(&hel[ib - 1], fac * h);
    }
    if (k2 > 0.0) {
      const double pn = kp.visc_simple ? k2 : pow(k2, kp.alpnu);
      const double pm = pow(k2, kp.alpmu);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double su = fac * abs2_ref(u[c]);
        const double sw = fac * abs2_ref(w[c]);
        g[c] += su;
        g[3 + c] += 2.0 * kp.nu * su * pn;
        g[6 + c] += 2.0 * kp.mu * su * pm;
        g[9 + c] += sw;
        g[12 + c] += 2.0 * kp.nu * sw * pn;
        g[15 + c] += 2.0 * kp.mu * sw * pm;
      }
    }
    if (nl) {
      const int ibf = (int)kk;
      if (ibf < nbin) {
        const cx<double> n0 = nl[q], n1 = nl[kp.fs + q], n2 = nl[2 * kp.fs + q];
        const double s = (a[0].x * n0.x + a[0].y * n0.y) + (a[1].x * n1.x + a[1].y * n1.y) +
                         (a[2].x * n2.x + a[2].y * n2.y);
        const double fac2 = edge ? 1.0 : 2.0;
        atomicAdd(&flux[ibf], fac2 * kk * kk * s);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 18; q++) atomicAdd(&glob[q], g[q]);
  __syncthreads();
  double* out = partial + (long long)blockIdx.x * len;
  for (int q = threadIdx.x; q < len; q += blockDim.x) out[q] = acc[q];
}

__global__ void k_diag_sum(const double* __restrict__ partial, int nblocks, int len, double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= len) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; b++) s += partial[(long long)b * len + q];
  out[q] = s;
}

int diag_len(int nbin) { return 4 * nbin + 18; }

hipError_t launch_diag(const KP& kp, const void* b, const void* nl, double* partial, int nblocks, double* out,
                       int nbin, hipStream_t s) {
  if (nbin > DIAG_NBIN_MAX) return hipErrorInvalidValue;
  k_diag<<<nblocks, 256, 0, s>>>(kp, (const cx<double>*)b, (const cx<double>*)nl, partial, nbin);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int len = diag_len(nbin);
  k_diag_sum<<<(len + 255) / 256, 256, 0, s>>>(partial, nblocks, len, out);
  return hipGetLastError();
}

}  // namespace gt
