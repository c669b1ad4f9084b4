/*
 * gt_oracle.c -- CPU restatement of Gturbo3D's pseudo-spectral RK2 /
 * integrating-factor time step (triply periodic rotating Navier-Stokes in
 * vector-potential form).
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the parity checker for the HIP
 * library in gturbo3d_amd/ and the timed CPU baseline of bench.py.  The
 * product never links, loads or calls it.
 *
 * Every routine restates the reference routine cited next to it
 * (file:line relative to the Gturbo3D source tree), keeping the
 * reference's expression order so that the arithmetic matches to the last
 * bit wherever the reference's own arithmetic is deterministic.  The
 * reference's 3D FFTs are cuFFT calls (fft_dir.f90:12, fft_inv.f90:12); they
 * are restated here from the DFT definition with a mixed-radix Stockham
 * transform (radices 2, 3, 4, 5 and a generic O(R^2) DFT for other primes),
 * twiddles rounded from long-double sinl/cosl.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   or_rann / LCG           closed form mod 2^31 (Python big integers)
 *   defk tables, emk        closed form; jc/kc by definition
 *   iniforcing (both)       brute-force enumeration in the reference loop order
 *   trunc mask              brute force in numpy, same expression order
 *   3D R2C / C2R            numpy.fft.rfftn / irfftn (pocketfft), direct DFT
 *   curl / invcurl / cross  analytic fields transformed by the oracle FFT
 *   incompr                 divergence-free output, idempotence
 *   nlt                     closed forms: two shear modes, Coriolis on the
 *                           second argument (test_oracle_nlt_closed_form.py);
 *                           Beltrami (ABC) fields, energy / helicity invariants
 *   step1 / step2           linear decay closed form exp(-nu k^2 dt)
 *   forcing                 orthogonality to k, Hermitian pairs, amplitude
 *   step (composition)      rotating ABC closed form (both stage-1 modes);
 *                           composition of the separately pinned pieces
 *   diagnostics             Parseval against physical-space sums
 *
 * Layout: a complex field is complex128 (NX/2+1, NY, NZ) column-major, i.e.
 * C order [NZ][NY][NX/2+1] of (re, im) pairs (modules.f90 / main.f90:14);
 * the same memory is the padded real array [NZ][NY][NX+2] used by
 * fft_dir/fft_inv (fft_dir.f90:7) and cross (produ.f90:5).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_NFORC 2048 /* modules.f90:39 */
#define OR_MAXFAC 40

typedef struct {
  int nx, ny, nz;
  double xlx, xly, xlz;
  double nu, alpnu, mu, alpmu;
  double dt, omega;
  double rtrunc; /* as read from params.dat (not squared) */
  int truncate;
  double famp, k1f, k2f;
  int variant;  /* 3 = 3D3C forcing, 2 = 2D2C forcing */
  int dt2_mode; /* 0: reference (dt2 never assigned = 0), 1: midpoint dt2 = dt/2 */
} or_params;

typedef struct {
  int n;
  int nfac;
  int fac[OR_MAXFAC];
  double *tw; /* tw[2m], tw[2m+1] = cos, -sin of 2 pi m / n */
} or_plan;

typedef struct {
  or_params p;
  int nx, ny, nz, nxh, nxp2;
  double dkx, dky, dkz, dt2, sdt;
  double *fkx, *fky, *fkz, *fkxs, *fkys, *fkzs;
  int *jc, *kc; /* 1-based like the reference */
  double rtrunc2, fkxsmax, fkysmax, fkzsmax;
  double *emk; /* [nz][ny][nxh] */
  int nf;
  int ikxf[OR_NFORC], ikyf[OR_NFORC], ikzf[OR_NFORC]; /* 1-based */
  double ff[OR_NFORC];
  double inpe, inpz;
  or_plan px, py, pz;
  int err;
} or_ctx;

/* ------------------------------------------------------------------ */
/* FFT (restates the cuFFT D2Z/Z2D plans of main.f90:27-34)           */
/* ------------------------------------------------------------------ */

static int plan_init(or_plan *pl, int n) {
  int m = n, k = 0;
  pl->n = n;
  while (m % 4 == 0) { pl->fac[k++] = 4; m /= 4; }
  while (m % 2 == 0) { pl->fac[k++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { pl->fac[k++] = f; m /= f; }
  }
  pl->nfac = k;
  pl->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  if (!pl->tw) return -1;
  for (int j = 0; j < n; j++) {
    long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)n;
    pl->tw[2 * j] = (double)cosl(a);
    pl->tw[2 * j + 1] = (double)(-sinl(a));
  }
  return 0;
}

/* exp(sign * 2 pi i * q / n) for integer q, from the table */
static inline void tw_get(const or_plan *pl, long q, int sign, double *c, double *s) {
  long n = pl->n;
  q %= n;
  if (q < 0) q += n;
  *c = pl->tw[2 * q];
  *s = (sign < 0) ? pl->tw[2 * q + 1] : -pl->tw[2 * q + 1];
}

/*
 * Stockham auto-sort transform of B interleaved sequences of length n:
 * element m of sequence b sits at x[2*(m*B + b)].  sign = -1 forward,
 * +1 inverse (unnormalised).  work has the same size as x.
 */
static void fft_batch(const or_plan *pl, double *x, double *work, int B, int sign) {
  const int n = pl->n;
  double *src = x, *dst = work;
  int Ns = 1;
  double vr[64], vi[64], yr[64], yi[64];
  for (int s = 0; s < pl->nfac; s++) {
    const int R = pl->fac[s];
    const int nr = n / R;
    for (int j = 0; j < nr; j++) {
      const int k = j % Ns;
      const int idxD = (j / Ns) * Ns * R + k;
      double twc[64], tws[64];
      for (int r = 0; r < R; r++) {
        /* twiddle exp(sign 2 pi i r k / (Ns R)) */
        tw_get(pl, (long)r * k * (n / (Ns * R)), sign, &twc[r], &tws[r]);
      }
      for (int b = 0; b < B; b++) {
        for (int r = 0; r < R; r++) {
          double a = src[2 * ((size_t)(j + r * nr) * B + b)];
          double c = src[2 * ((size_t)(j + r * nr) * B + b) + 1];
          vr[r] = a * twc[r] - c * tws[r];
          vi[r] = a * tws[r] + c * twc[r];
        }
        if (R == 2) {
          yr[0] = vr[0] + vr[1]; yi[0] = vi[0] + vi[1];
          yr[1] = vr[0] - vr[1]; yi[1] = vi[0] - vi[1];
        } else if (R == 4) {
          double ar = vr[0] + vr[2], ai = vi[0] + vi[2];
          double br = vr[0] - vr[2], bi = vi[0] - vi[2];
          double cr = vr[1] + vr[3], ci = vi[1] + vi[3];
          double dr = vr[1] - vr[3], di = vi[1] - vi[3];
          /* (sign i) * d */
          double er = -sign * di, ei = sign * dr;
          yr[0] = ar + cr; yi[0] = ai + ci;
          yr[2] = ar - cr; yi[2] = ai - ci;
          yr[1] = br + er; yi[1] = bi + ei;
          yr[3] = br - er; yi[3] = bi - ei;
        } else {
          for (int q = 0; q < R; q++) {
            double sr = 0.0, si = 0.0;
            for (int r = 0; r < R; r++) {
              double c, sn;
              tw_get(pl, (long)((r * q) % R) * (n / R), sign, &c, &sn);
              sr += vr[r] * c - vi[r] * sn;
              si += vr[r] * sn + vi[r] * c;
            }
            yr[q] = sr;
            yi[q] = si;
          }
        }
        for (int r = 0; r < R; r++) {
          dst[2 * ((size_t)(idxD + r * Ns) * B + b)] = yr[r];
          dst[2 * ((size_t)(idxD + r * Ns) * B + b) + 1] = yi[r];
        }
      }
    }
    double *t = src; src = dst; dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n * B);
}

#define TILE 8

/* complex C2C along y of every (kx, z) column, in the [nz][ny][nxh] layout */
static void fft_axis_y(or_ctx *c, double *z, int sign) {
  const int nx = c->nxh, ny = c->ny, nz = c->nz;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 2 * ny * TILE);
    double *wrk = (double *)malloc(sizeof(double) * 2 * ny * TILE);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
      for (int i0 = 0; i0 < nx; i0 += TILE) {
        int B = (nx - i0 < TILE) ? nx - i0 : TILE;
        for (int j = 0; j < ny; j++)
          for (int b = 0; b < B; b++) {
            size_t o = 2 * (((size_t)k * ny + j) * nx + i0 + b);
            buf[2 * (j * B + b)] = z[o];
            buf[2 * (j * B + b) + 1] = z[o + 1];
          }
        fft_batch(&c->py, buf, wrk, B, sign);
        for (int j = 0; j < ny; j++)
          for (int b = 0; b < B; b++) {
            size_t o = 2 * (((size_t)k * ny + j) * nx + i0 + b);
            z[o] = buf[2 * (j * B + b)];
            z[o + 1] = buf[2 * (j * B + b) + 1];
          }
      }
    free(buf);
    free(wrk);
  }
}

static void fft_axis_z(or_ctx *c, double *z, int sign) {
  const int nx = c->nxh, ny = c->ny, nz = c->nz;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 2 * nz * TILE);
    double *wrk = (double *)malloc(sizeof(double) * 2 * nz * TILE);
#pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < ny; j++)
      for (int i0 = 0; i0 < nx; i0 += TILE) {
        int B = (nx - i0 < TILE) ? nx - i0 : TILE;
        for (int k = 0; k < nz; k++)
          for (int b = 0; b < B; b++) {
            size_t o = 2 * (((size_t)k * ny + j) * nx + i0 + b);
            buf[2 * (k * B + b)] = z[o];
            buf[2 * (k * B + b) + 1] = z[o + 1];
          }
        fft_batch(&c->pz, buf, wrk, B, sign);
        for (int k = 0; k < nz; k++)
          for (int b = 0; b < B; b++) {
            size_t o = 2 * (((size_t)k * ny + j) * nx + i0 + b);
            z[o] = buf[2 * (k * B + b)];
            z[o + 1] = buf[2 * (k * B + b) + 1];
          }
      }
    free(buf);
    free(wrk);
  }
}

/*
 * fft_dir (fft_dir.f90:1-28): in-place 3D R2C of the padded real array
 * [nz][ny][nx+2], then z/NY/NX/NZ over every real of the array
 * (fft_dir.f90:17-25).  The x transform is a complex FFT of the real row.
 */
void or_fft_dir(or_ctx *c, double *z) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nxh = c->nxh, nxp2 = c->nxp2;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 2 * nx);
    double *wrk = (double *)malloc(sizeof(double) * 2 * nx);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        double *row = z + ((size_t)k * ny + j) * nxp2;
        for (int i = 0; i < nx; i++) { buf[2 * i] = row[i]; buf[2 * i + 1] = 0.0; }
        fft_batch(&c->px, buf, wrk, 1, -1);
        for (int i = 0; i < nxh; i++) { row[2 * i] = buf[2 * i]; row[2 * i + 1] = buf[2 * i + 1]; }
      }
    free(buf);
    free(wrk);
  }
  fft_axis_y(c, z, -1);
  fft_axis_z(c, z, -1);
  const size_t tot = (size_t)nxp2 * ny * nz;
  const double dny = (double)ny, dnx = (double)nx, dnz = (double)nz;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < tot; q++) z[q] = z[q] / dny / dnx / dnz;
}

/*
 * fft_inv (fft_inv.f90:1-16): in-place unnormalised 3D C2R.  C2C inverse
 * along z and y, then per x row a Hermitian C2R that uses only the real
 * parts of the kx = 0 and kx = NX/2 coefficients (standard C2R semantics).
 * The two padding reals of each row are set to zero.
 */
void or_fft_inv(or_ctx *c, double *z) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nxh = c->nxh, nxp2 = c->nxp2;
  fft_axis_z(c, z, +1);
  fft_axis_y(c, z, +1);
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 2 * nx);
    double *wrk = (double *)malloc(sizeof(double) * 2 * nx);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        double *row = z + ((size_t)k * ny + j) * nxp2;
        for (int i = 0; i < nxh; i++) { buf[2 * i] = row[2 * i]; buf[2 * i + 1] = row[2 * i + 1]; }
        buf[1] = 0.0;
        if (nx % 2 == 0) buf[2 * (nx / 2) + 1] = 0.0;
        for (int i = nxh; i < nx; i++) {
          buf[2 * i] = buf[2 * (nx - i)];
          buf[2 * i + 1] = -buf[2 * (nx - i) + 1];
        }
        fft_batch(&c->px, buf, wrk, 1, +1);
        for (int i = 0; i < nx; i++) row[i] = buf[2 * i];
        for (int i = nx; i < nxp2; i++) row[i] = 0.0;
      }
    free(buf);
    free(wrk);
  }
}

/* ------------------------------------------------------------------ */
/* Host tables                                                         */
/* ------------------------------------------------------------------ */

/* rann (forcing_3d3c.f90:82-93, forcing_2d2c.f90:52-63) */
float or_rann(int32_t *irand) {
  const uint32_t mask = 0x7FFFFFFFu;
  uint32_t s = (uint32_t)(*irand);
  s = (69069u * s + 1u) & mask;
  *irand = (int32_t)s;
  return (float)((double)s / (double)mask);
}

/* defk (defk.f90:1-94) */
static void defk(or_ctx *c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nxh = c->nxh;
  const int ny2p1 = ny / 2 + 1, ny2p2 = ny / 2 + 2, nz2p1 = nz / 2 + 1, nz2p2 = nz / 2 + 2;
  (void)nx;
  c->jc[0] = 1;
  for (int j = 2; j <= ny; j++) c->jc[j - 1] = ny + 2 - j;
  c->kc[0] = 1;
  for (int k = 2; k <= nz; k++) c->kc[k - 1] = nz + 2 - k;
  c->dkx = 2.0 * 3.14159265358979323846 / c->p.xlx;
  c->dky = 2.0 * 3.14159265358979323846 / c->p.xly;
  c->dkz = 2.0 * 3.14159265358979323846 / c->p.xlz;
  for (int i = 1; i <= nxh; i++) c->fkx[i - 1] = (double)(i - 1) * c->dkx;
  for (int j = 1; j <= ny2p1; j++) c->fky[j - 1] = (double)(j - 1) * c->dky;
  for (int j = ny2p2; j <= ny; j++) c->fky[j - 1] = (double)(j - ny - 1) * c->dky;
  for (int k = 1; k <= nz2p1; k++) c->fkz[k - 1] = (double)(k - 1) * c->dkz;
  for (int k = nz2p2; k <= nz; k++) c->fkz[k - 1] = (double)(k - nz - 1) * c->dkz;
  for (int i = 0; i < nxh; i++) c->fkxs[i] = c->fkx[i] * c->fkx[i];
  for (int j = 0; j < ny; j++) c->fkys[j] = c->fky[j] * c->fky[j];
  for (int k = 0; k < nz; k++) c->fkzs[k] = c->fkz[k] * c->fkz[k];
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nxh; i++) {
        double fk2 = c->fkxs[i] + c->fkys[j] + c->fkzs[k];
        double scra = 0.0;
        if (fk2 != 0.0) scra = c->p.nu * pow(fk2, c->p.alpnu) + c->p.mu * pow(fk2, c->p.alpmu);
        double e = exp(-0.5 * c->p.dt * scra);
        if (fk2 == 0.0) e = 1.0;
        c->emk[((size_t)k * ny + j) * nxh + i] = e;
      }
  c->fkxsmax = c->fkxs[nxh - 1];
  c->fkysmax = c->fkys[ny2p1 - 1];
  c->fkzsmax = c->fkzs[nz2p1 - 1];
  c->rtrunc2 = c->p.rtrunc * c->p.rtrunc;
}

static int forced_add(or_ctx *c, int i, int j, int k, int variant) {
  double k2 = c->fkxs[i - 1] + c->fkys[j - 1] + c->fkzs[k - 1];
  double kk = sqrt(k2);
  double k2n = c->fkxs[i - 1] / c->fkxsmax + c->fkys[j - 1] / c->fkysmax + c->fkzs[k - 1] / c->fkzsmax;
  if (k2n <= c->rtrunc2) {
    if ((k2 <= c->p.k2f * c->p.k2f) && (k2 > c->p.k1f * c->p.k1f)) {
      if (c->nf >= OR_NFORC) return -1; /* 'NFORC too small !' */
      int n = c->nf++;
      c->ikxf[n] = i;
      c->ikyf[n] = j;
      c->ikzf[n] = k;
      c->ff[n] = 1.0 / kk;
      if (variant == 3) {
        double a = c->p.famp * c->ff[n];
        c->inpz = c->inpz + k2 * k2 * (a * a);
        c->inpe = c->inpe + k2 * (a * a);
      }
    }
  }
  return 0;
}

/* iniforcing, 3D3C (iniforcing_3d3c.f90:4-195): fixed enumeration order */
static int iniforcing_3d3c(or_ctx *c) {
  const int nx2 = c->nx / 2, nx2p1 = nx2 + 1, ny = c->ny, nz = c->nz;
  const int ny2 = ny / 2, ny2p1 = ny2 + 1, nz2 = nz / 2;
  int i, j, k;
  c->nf = 0;
  c->inpe = 0.0;
  c->inpz = 0.0;
  i = 1;
  j = 1;
  for (k = 2; k <= nz2; k++) if (forced_add(c, i, j, k, 3)) return -1;
  for (j = 2; j <= ny2; j++)
    for (k = 1; k <= nz; k++) if (forced_add(c, i, j, k, 3)) return -1;
  j = ny2p1;
  for (k = 2; k <= nz2; k++) if (forced_add(c, i, j, k, 3)) return -1;
  for (i = 2; i <= nx2; i++)
    for (j = 1; j <= ny; j++)
      for (k = 1; k <= nz; k++) if (forced_add(c, i, j, k, 3)) return -1;
  i = nx2p1;
  j = 1;
  for (k = 2; k <= nz2; k++) if (forced_add(c, i, j, k, 3)) return -1;
  for (j = 2; j <= ny2; j++)
    for (k = 1; k <= nz; k++) if (forced_add(c, i, j, k, 3)) return -1;
  j = ny2p1;
  for (k = 2; k <= nz2; k++) if (forced_add(c, i, j, k, 3)) return -1;
  return 0;
}

/* iniforcing, 2D2C (iniforcing_2d2c.f90:4-88): kz = 0 plane only */
static int iniforcing_2d2c(or_ctx *c) {
  const int nx2 = c->nx / 2, nx2p1 = nx2 + 1, ny = c->ny, ny2 = ny / 2;
  int i, j, k = 1;
  c->nf = 0;
  c->inpe = 0.0;
  c->inpz = 0.0;
  i = 1;
  for (j = 2; j <= ny2; j++) if (forced_add(c, i, j, k, 2)) return -1;
  for (i = 2; i <= nx2; i++)
    for (j = 1; j <= ny; j++) if (forced_add(c, i, j, k, 2)) return -1;
  i = nx2p1;
  for (j = 2; j <= ny2; j++) if (forced_add(c, i, j, k, 2)) return -1;
  for (int n = 0; n < c->nf; n++) {
    double k2 = c->fkxs[c->ikxf[n] - 1] + c->fkys[c->ikyf[n] - 1] + c->fkzs[c->ikzf[n] - 1];
    double a = c->p.famp * c->ff[n];
    c->inpe = c->inpe + k2 * (a * a);
    c->inpz = c->inpz + k2 * k2 * (a * a);
  }
  return 0;
}

or_ctx *or_create(const or_params *p) {
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  if (!c) return NULL;
  c->p = *p;
  c->nx = p->nx;
  c->ny = p->ny;
  c->nz = p->nz;
  c->nxh = p->nx / 2 + 1;
  c->nxp2 = p->nx + 2;
  c->fkx = (double *)malloc(sizeof(double) * c->nxh);
  c->fkxs = (double *)malloc(sizeof(double) * c->nxh);
  c->fky = (double *)malloc(sizeof(double) * c->ny);
  c->fkys = (double *)malloc(sizeof(double) * c->ny);
  c->fkz = (double *)malloc(sizeof(double) * c->nz);
  c->fkzs = (double *)malloc(sizeof(double) * c->nz);
  c->jc = (int *)malloc(sizeof(int) * c->ny);
  c->kc = (int *)malloc(sizeof(int) * c->nz);
  c->emk = (double *)malloc(sizeof(double) * (size_t)c->nxh * c->ny * c->nz);
  if (!c->fkx || !c->fkxs || !c->fky || !c->fkys || !c->fkz || !c->fkzs || !c->jc || !c->kc || !c->emk) {
    c->err = 1;
    return c;
  }
  /* inieuler.f90:67 */
  c->sdt = sqrt(p->dt);
  c->dt2 = (p->dt2_mode == 1) ? 0.5 * p->dt : 0.0;
  defk(c);
  if (p->variant == 2) {
    if (iniforcing_2d2c(c)) c->err = 2;
  } else {
    if (iniforcing_3d3c(c)) c->err = 2;
  }
  if (plan_init(&c->px, c->nx) || plan_init(&c->py, c->ny) || plan_init(&c->pz, c->nz)) c->err = 3;
  return c;
}

int or_error(const or_ctx *c) { return c->err; }

void or_destroy(or_ctx *c) {
  if (!c) return;
  free(c->fkx); free(c->fkxs); free(c->fky); free(c->fkys); free(c->fkz); free(c->fkzs);
  free(c->jc); free(c->kc); free(c->emk);
  free(c->px.tw); free(c->py.tw); free(c->pz.tw);
  free(c);
}

int or_nf(const or_ctx *c) { return c->nf; }

void or_forcing_table(const or_ctx *c, int *ikxf, int *ikyf, int *ikzf, double *ff, double *inp) {
  for (int n = 0; n < c->nf; n++) {
    ikxf[n] = c->ikxf[n];
    ikyf[n] = c->ikyf[n];
    ikzf[n] = c->ikzf[n];
    ff[n] = c->ff[n];
  }
  inp[0] = c->inpe;
  inp[1] = c->inpz;
}

/* wavenumber tables: out = fkx(nxh) fky(ny) fkz(nz) then fkxsmax fkysmax fkzsmax rtrunc2 */
void or_tables(const or_ctx *c, double *out, double *emk) {
  size_t o = 0;
  for (int i = 0; i < c->nxh; i++) out[o++] = c->fkx[i];
  for (int j = 0; j < c->ny; j++) out[o++] = c->fky[j];
  for (int k = 0; k < c->nz; k++) out[o++] = c->fkz[k];
  out[o++] = c->fkxsmax;
  out[o++] = c->fkysmax;
  out[o++] = c->fkzsmax;
  out[o++] = c->rtrunc2;
  if (emk) memcpy(emk, c->emk, sizeof(double) * (size_t)c->nxh * c->ny * c->nz);
}

/* ------------------------------------------------------------------ */
/* Point operations                                                    */
/* ------------------------------------------------------------------ */

#define RE(z, q) (z)[2 * (q)]
#define IM(z, q) (z)[2 * (q) + 1]
#define IDX(c, i, j, k) ((((size_t)(k)) * (c)->ny + (j)) * (c)->nxh + (i))

/* curl (produ.f90:34-66): a <- i k x a */
void or_curl(or_ctx *c, double *ax, double *ay, double *az) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++) {
        size_t q = IDX(c, i, j, k);
        double fx = c->fkx[i], fy = c->fky[j], fz = c->fkz[k];
        double x1 = -fy * IM(az, q) + fz * IM(ay, q);
        double x2 = fy * RE(az, q) - fz * RE(ay, q);
        double y1 = -fz * IM(ax, q) + fx * IM(az, q);
        double y2 = fz * RE(ax, q) - fx * RE(az, q);
        double z1 = -fx * IM(ay, q) + fy * IM(ax, q);
        double z2 = fx * RE(ay, q) - fy * RE(ax, q);
        RE(ax, q) = x1; IM(ax, q) = x2;
        RE(ay, q) = y1; IM(ay, q) = y2;
        RE(az, q) = z1; IM(az, q) = z2;
      }
}

/* invcurl (produ.f90:69-104): a <- i (k/k^2) x a, k = 0 left untouched */
void or_invcurl(or_ctx *c, double *ax, double *ay, double *az) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++) {
        if (i == 0 && j == 0 && k == 0) continue;
        size_t q = IDX(c, i, j, k);
        double k2 = c->fkxs[i] + c->fkys[j] + c->fkzs[k];
        double kx = c->fkx[i] / k2, ky = c->fky[j] / k2, kz = c->fkz[k] / k2;
        double x1 = -ky * IM(az, q) + kz * IM(ay, q);
        double x2 = ky * RE(az, q) - kz * RE(ay, q);
        double y1 = -kz * IM(ax, q) + kx * IM(az, q);
        double y2 = kz * RE(ax, q) - kx * RE(az, q);
        double z1 = -kx * IM(ay, q) + ky * IM(ax, q);
        double z2 = kx * RE(ay, q) - ky * RE(ax, q);
        RE(ax, q) = x1; IM(ax, q) = x2;
        RE(ay, q) = y1; IM(ay, q) = y2;
        RE(az, q) = z1; IM(az, q) = z2;
      }
}

/* the truncation test of trunc.f90:14-17 (rtrunc already squared, defk.f90:91) */
static inline int masked(const or_ctx *c, int i, int j, int k) {
  double k3 = c->fkzs[k] / c->fkzsmax;
  double k2 = k3 + c->fkys[j] / c->fkysmax;
  double k1 = k2 + c->fkxs[i] / c->fkxsmax;
  return k1 > c->rtrunc2;
}

/* trunc (trunc.f90:1-31) */
void or_trunc(or_ctx *c, double *zf) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++)
        if (masked(c, i, j, k)) {
          size_t q = IDX(c, i, j, k);
          RE(zf, q) = 0.0;
          IM(zf, q) = 0.0;
        }
  RE(zf, 0) = 0.0;
  IM(zf, 0) = 0.0;
}

/* mask as bytes (1 = zeroed by trunc), for tests */
void or_mask(const or_ctx *c, unsigned char *out) {
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++) out[IDX(c, i, j, k)] = (unsigned char)masked(c, i, j, k);
}

/* cross (produ.f90:1-31): a <- a x b on the padded real arrays, pads zeroed */
void or_cross(or_ctx *c, double *a1, double *a2, double *a3, const double *b1, const double *b2,
              const double *b3) {
  const int nxp2 = c->nxp2, nx = c->nx;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < nxp2; i++) {
        size_t q = ((size_t)k * c->ny + j) * nxp2 + i;
        if (i < nx) {
          double s1 = a2[q] * b3[q] - a3[q] * b2[q];
          double s2 = a3[q] * b1[q] - a1[q] * b3[q];
          double s3 = a1[q] * b2[q] - a2[q] * b1[q];
          a1[q] = s1;
          a2[q] = s2;
          a3[q] = s3;
        } else {
          a1[q] = 0.0;
          a2[q] = 0.0;
          a3[q] = 0.0;
        }
      }
}

/* incompr (trunc.f90:35-68) without the racy diagnostic sum */
void or_incompr(or_ctx *c, double *ax, double *ay, double *az) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++) {
        if (i == 0 && j == 0 && k == 0) continue;
        size_t q = IDX(c, i, j, k);
        double k2 = c->fkxs[i] + c->fkys[j] + c->fkzs[k];
        double fx = c->fkx[i], fy = c->fky[j], fz = c->fkz[k];
        /* scra = fkx*ax + fky*ay + fkz*az (complex, left to right) */
        double sr = fx * RE(ax, q) + fy * RE(ay, q);
        double si = fx * IM(ax, q) + fy * IM(ay, q);
        sr = sr + fz * RE(az, q);
        si = si + fz * IM(az, q);
        /* a - scra*fk/k2: (scra*fk)/k2 */
        RE(ax, q) = RE(ax, q) - sr * fx / k2; IM(ax, q) = IM(ax, q) - si * fx / k2;
        RE(ay, q) = RE(ay, q) - sr * fy / k2; IM(ay, q) = IM(ay, q) - si * fy / k2;
        RE(az, q) = RE(az, q) - sr * fz / k2; IM(az, q) = IM(az, q) - si * fz / k2;
      }
}

/* ------------------------------------------------------------------ */
/* nlt, step1, step2, forcing, the step body                           */
/* ------------------------------------------------------------------ */

/*
 * Mutation switches (OR_MUTANT, 0 = none, the only value the Makefile uses):
 * each plants one plausible misreading of the reference in nlt or the step,
 * so that tests/test_oracle_mutants.py can build the oracle once per value
 * and show that the closed-form tests reject every one of them.
 *   1  vorticity |k| a instead of k^2 a        (nlt.f90:26-40)
 *   2  w x u instead of u x w                  (nlt.f90:55, produ.f90:14-16)
 *   3  vorticity -k^2 a                        (nlt.f90:26-40)
 *   4  Coriolis -i omega kz b                  (nlt.f90:72-77)
 *   5  Coriolis on the first argument a        (nlt.f90:72-77)
 *   6  stage-2 Coriolis on bh, not the old b   (main.f90:144)
 *   7  step2 as emk^2 a + dt c                 (rkstep.f90:86-96)
 */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif

/*
 * nlt(a, v, b) (nlt.f90:1-95): on return a = T[invcurl(F[u x w] + i omega kz b)]
 * and v = F[u x w], with u = F^-1[curl a], w = F^-1[k^2 a].
 */
void or_nlt(or_ctx *c, double *ax, double *ay, double *az, double *vx, double *vy, double *vz,
            const double *bx, const double *by, const double *bz) {
  const size_t m = (size_t)c->nxh * c->ny * c->nz;
#if OR_MUTANT == 5
  double *a0 = malloc(sizeof(double) * 6 * m);
  memcpy(a0, ax, sizeof(double) * 2 * m);
  memcpy(a0 + 2 * m, ay, sizeof(double) * 2 * m);
  memcpy(a0 + 4 * m, az, sizeof(double) * 2 * m);
  bx = a0;
  by = a0 + 2 * m;
  bz = a0 + 4 * m;
#endif
  memcpy(vx, ax, sizeof(double) * 2 * m);
  memcpy(vy, ay, sizeof(double) * 2 * m);
  memcpy(vz, az, sizeof(double) * 2 * m);
  or_curl(c, vx, vy, vz);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++) {
        size_t q = IDX(c, i, j, k);
        double k2 = c->fkxs[i] + c->fkys[j] + c->fkzs[k];
#if OR_MUTANT == 1
        k2 = sqrt(k2);
#elif OR_MUTANT == 3
        k2 = -k2;
#endif
        RE(ax, q) = k2 * RE(ax, q); IM(ax, q) = k2 * IM(ax, q);
        RE(ay, q) = k2 * RE(ay, q); IM(ay, q) = k2 * IM(ay, q);
        RE(az, q) = k2 * RE(az, q); IM(az, q) = k2 * IM(az, q);
      }
  or_fft_inv(c, vx);
  or_fft_inv(c, vy);
  or_fft_inv(c, vz);
  or_fft_inv(c, ax);
  or_fft_inv(c, ay);
  or_fft_inv(c, az);
#if OR_MUTANT == 2
  or_cross(c, ax, ay, az, vx, vy, vz);
  memcpy(vx, ax, sizeof(double) * 2 * m);
  memcpy(vy, ay, sizeof(double) * 2 * m);
  memcpy(vz, az, sizeof(double) * 2 * m);
#else
  or_cross(c, vx, vy, vz, ax, ay, az);
#endif
  or_fft_dir(c, vx);
  or_fft_dir(c, vy);
  or_fft_dir(c, vz);
#if OR_MUTANT == 4
  const double om = -c->p.omega;
#else
  const double om = c->p.omega;
#endif
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++) {
        size_t q = IDX(c, i, j, k);
        double k2 = c->fkz[k];
        RE(ax, q) = RE(vx, q) - om * k2 * IM(bx, q);
        IM(ax, q) = IM(vx, q) + om * k2 * RE(bx, q);
        RE(ay, q) = RE(vy, q) - om * k2 * IM(by, q);
        IM(ay, q) = IM(vy, q) + om * k2 * RE(by, q);
        RE(az, q) = RE(vz, q) - om * k2 * IM(bz, q);
        IM(az, q) = IM(vz, q) + om * k2 * RE(bz, q);
      }
  or_invcurl(c, ax, ay, az);
  if (c->p.truncate == 1) {
    or_trunc(c, ax);
    or_trunc(c, ay);
    or_trunc(c, az);
  }
#if OR_MUTANT == 5
  free(a0);
#endif
}

/* step1, live branch (rkstep.f90:30-48): c <- emk (a + dt2 c), c(0) = 0 */
void or_step1(or_ctx *c, const double *ax, const double *ay, const double *az, double *cx, double *cy,
              double *cz) {
  const size_t m = (size_t)c->nxh * c->ny * c->nz;
  const double dt2 = c->dt2;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < m; q++) {
    double e = c->emk[q];
    RE(cx, q) = e * (RE(ax, q) + dt2 * RE(cx, q)); IM(cx, q) = e * (IM(ax, q) + dt2 * IM(cx, q));
    RE(cy, q) = e * (RE(ay, q) + dt2 * RE(cy, q)); IM(cy, q) = e * (IM(ay, q) + dt2 * IM(cy, q));
    RE(cz, q) = e * (RE(az, q) + dt2 * RE(cz, q)); IM(cz, q) = e * (IM(az, q) + dt2 * IM(cz, q));
  }
  RE(cx, 0) = IM(cx, 0) = 0.0;
  RE(cy, 0) = IM(cy, 0) = 0.0;
  RE(cz, 0) = IM(cz, 0) = 0.0;
}

/* step2, live branch (rkstep.f90:86-104): a <- emk (emk a + dt c), a(0) = 0 */
void or_step2(or_ctx *c, double *ax, double *ay, double *az, const double *cx, const double *cy,
              const double *cz) {
  const size_t m = (size_t)c->nxh * c->ny * c->nz;
  const double dt = c->p.dt;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < m; q++) {
    double e = c->emk[q];
#if OR_MUTANT == 7
    RE(ax, q) = e * e * RE(ax, q) + dt * RE(cx, q); IM(ax, q) = e * e * IM(ax, q) + dt * IM(cx, q);
    RE(ay, q) = e * e * RE(ay, q) + dt * RE(cy, q); IM(ay, q) = e * e * IM(ay, q) + dt * IM(cy, q);
    RE(az, q) = e * e * RE(az, q) + dt * RE(cz, q); IM(az, q) = e * e * IM(az, q) + dt * IM(cz, q);
#else
    RE(ax, q) = e * (e * RE(ax, q) + dt * RE(cx, q)); IM(ax, q) = e * (e * IM(ax, q) + dt * IM(cx, q));
    RE(ay, q) = e * (e * RE(ay, q) + dt * RE(cy, q)); IM(ay, q) = e * (e * IM(ay, q) + dt * IM(cy, q));
    RE(az, q) = e * (e * RE(az, q) + dt * RE(cz, q)); IM(az, q) = e * (e * IM(az, q) + dt * IM(cz, q));
#endif
  }
  RE(ax, 0) = IM(ax, 0) = 0.0;
  RE(ay, 0) = IM(ay, 0) = 0.0;
  RE(az, 0) = IM(az, 0) = 0.0;
}

/* forcing, 3D3C (forcing_3d3c.f90:11-78) */
static void forcing_3d3c(or_ctx *c, double *bx, double *by, double *bz, int32_t *seed) {
  const double pi2 = 2.0 * 3.14159265358979323846;
  const int nx2p1 = c->nxh;
  for (int ik = 0; ik < c->nf; ik++) {
    int i = c->ikxf[ik], j = c->ikyf[ik], k = c->ikzf[ik];
    double te = pi2 * (double)or_rann(seed);
    double sf = c->sdt * c->p.famp;
    double er = sf * cos(te), ei = sf * sin(te);
    te = pi2 * (double)or_rann(seed);
    double cte = cos(te), ste = sin(te);
    double kr = er * c->ff[ik], ki = ei * c->ff[ik];
    size_t q = IDX(c, i - 1, j - 1, k - 1);
    size_t qc = IDX(c, i - 1, c->jc[j - 1] - 1, c->kc[k - 1] - 1);
    if (i == 1 && j == 1) {
      RE(bx, q) = RE(bx, q) + kr * cte; IM(bx, q) = IM(bx, q) + ki * cte;
      RE(by, q) = RE(by, q) + kr * ste; IM(by, q) = IM(by, q) + ki * ste;
      RE(bx, qc) = RE(bx, qc) + kr * cte; IM(bx, qc) = IM(bx, qc) + (-ki) * cte;
      RE(by, qc) = RE(by, qc) + kr * ste; IM(by, qc) = IM(by, qc) + (-ki) * ste;
    } else {
      double k2 = c->fkxs[i - 1] + c->fkys[j - 1];
      double kxy = sqrt(k2);
      k2 = k2 + c->fkzs[k - 1];
      double kk = sqrt(k2);
      double c1 = c->fkz[k - 1] / kk, s1 = kxy / kk;
      double c2 = c->fkx[i - 1] / kxy, s2 = c->fky[j - 1] / kxy;
      double f1 = cte * c2 * c1 - ste * s2;
      double f2 = cte * s2 * c1 + ste * c2;
      RE(bx, q) = RE(bx, q) + kr * f1; IM(bx, q) = IM(bx, q) + ki * f1;
      RE(by, q) = RE(by, q) + kr * f2; IM(by, q) = IM(by, q) + ki * f2;
      RE(bz, q) = RE(bz, q) - kr * cte * s1; IM(bz, q) = IM(bz, q) - ki * cte * s1;
      if (i == 1 || i == nx2p1) {
        RE(bx, qc) = RE(bx, qc) + kr * f1; IM(bx, qc) = IM(bx, qc) + (-ki) * f1;
        RE(by, qc) = RE(by, qc) + kr * f2; IM(by, qc) = IM(by, qc) + (-ki) * f2;
        RE(bz, qc) = RE(bz, qc) - kr * cte * s1; IM(bz, qc) = IM(bz, qc) - (-ki) * cte * s1;
      }
    }
  }
}

/* forcing, 2D2C (forcing_2d2c.f90:11-48) */
static void forcing_2d2c(or_ctx *c, double *bz, int32_t *seed) {
  const double pi2 = 2.0 * 3.14159265358979;
  const int nx2p1 = c->nxh;
  for (int ik = 0; ik < c->nf; ik++) {
    int i = c->ikxf[ik], j = c->ikyf[ik], k = c->ikzf[ik];
    double te = pi2 * (double)or_rann(seed);
    double sf = c->sdt * c->p.famp;
    double er = sf * cos(te), ei = sf * sin(te);
    double kr = er * c->ff[ik], ki = ei * c->ff[ik];
    size_t q = IDX(c, i - 1, j - 1, k - 1);
    RE(bz, q) = RE(bz, q) + kr;
    IM(bz, q) = IM(bz, q) + ki;
    if (i == 1 || i == nx2p1) {
      size_t qc = IDX(c, i - 1, c->jc[j - 1] - 1, c->kc[k - 1] - 1);
      RE(bz, qc) = RE(bz, qc) + kr;
      IM(bz, qc) = IM(bz, qc) + (-ki);
    }
  }
}

void or_forcing(or_ctx *c, double *bx, double *by, double *bz, int32_t *seed) {
  if (c->p.variant == 2)
    forcing_2d2c(c, bz, seed);
  else
    forcing_3d3c(c, bx, by, bz, seed);
}

/*
 * nsteps iterations of the step body of main.f90:124-150 (bnl <- b; nlt;
 * step1; nlt with the start-of-step b; step2; forcing).  b is 3 fields,
 * bnl and u are 3 work fields each.  If nl1 is not NULL it receives the
 * first-stage nonlinear term of the last step (used by fluxes at
 * main.f90:134-138).
 */
void or_step(or_ctx *c, double *bx, double *by, double *bz, double *bnx, double *bny, double *bnz,
             double *ux, double *uy, double *uz, int32_t *seed, int nsteps) {
  const size_t m = (size_t)c->nxh * c->ny * c->nz;
  for (int s = 0; s < nsteps; s++) {
    memcpy(bnx, bx, sizeof(double) * 2 * m);
    memcpy(bny, by, sizeof(double) * 2 * m);
    memcpy(bnz, bz, sizeof(double) * 2 * m);
    or_nlt(c, bnx, bny, bnz, ux, uy, uz, bx, by, bz);
    or_step1(c, bx, by, bz, bnx, bny, bnz);
#if OR_MUTANT == 6
    double *bh = malloc(sizeof(double) * 6 * m);
    memcpy(bh, bnx, sizeof(double) * 2 * m);
    memcpy(bh + 2 * m, bny, sizeof(double) * 2 * m);
    memcpy(bh + 4 * m, bnz, sizeof(double) * 2 * m);
    or_nlt(c, bnx, bny, bnz, ux, uy, uz, bh, bh + 2 * m, bh + 4 * m);
    free(bh);
#else
    or_nlt(c, bnx, bny, bnz, ux, uy, uz, bx, by, bz);
#endif
    or_step2(c, bx, by, bz, bnx, bny, bnz);
    or_forcing(c, bx, by, bz, seed);
  }
}

/*
 * Initial transform of a physical velocity (main.f90:91-106): fft_dir x3,
 * incompr, invcurl, trunc x3.  In: padded real velocity; out: spectral b.
 */
void or_init_from_velocity(or_ctx *c, double *bx, double *by, double *bz) {
  or_fft_dir(c, bx);
  or_fft_dir(c, by);
  or_fft_dir(c, bz);
  or_incompr(c, bx, by, bz);
  or_invcurl(c, bx, by, bz);
  if (c->p.truncate == 1) {
    or_trunc(c, bx);
    or_trunc(c, by);
    or_trunc(c, bz);
  }
}

/* velocity output (main.f90:190-201): u = fft_inv(curl b), on copies */
void or_velocity(or_ctx *c, const double *bx, const double *by, const double *bz, double *ux, double *uy,
                 double *uz) {
  const size_t m = (size_t)c->nxh * c->ny * c->nz;
  memcpy(ux, bx, sizeof(double) * 2 * m);
  memcpy(uy, by, sizeof(double) * 2 * m);
  memcpy(uz, bz, sizeof(double) * 2 * m);
  or_curl(c, ux, uy, uz);
  or_fft_inv(c, ux);
  or_fft_inv(c, uy);
  or_fft_inv(c, uz);
}

/* ------------------------------------------------------------------ */
/* Diagnostics (diagnostics.f90)                                       */
/* ------------------------------------------------------------------ */

/* spectrum (diagnostics.f90:5-56): sp and sp2d, NBIN = nx/2 bins each */
void or_spectrum(const or_ctx *c, const double *bx, const double *by, const double *bz, double *sp,
                 double *sp2d) {
  const int nbin = c->nx / 2;
  for (int b = 0; b < nbin; b++) sp[b] = sp2d[b] = 0.0;
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++) {
        double fac = (i == 0 || i == c->nxh - 1) ? 0.5 : 1.0;
        double kk = sqrt(c->fkxs[i] + c->fkys[j] + c->fkzs[k]);
        int ib = (int)kk + 1;
        if (ib > 0 && ib < nbin) {
          size_t q = IDX(c, i, j, k);
          double ax = hypot(RE(bx, q), IM(bx, q)), ay = hypot(RE(by, q), IM(by, q)),
                 az = hypot(RE(bz, q), IM(bz, q));
          double scra = ax * ax + ay * ay + az * az;
          sp[ib - 1] = sp[ib - 1] + kk * kk * fac * scra;
        }
      }
  for (int j = 0; j < c->ny; j++)
    for (int i = 0; i < c->nxh; i++) {
      double fac = (i == 0 || i == c->nxh - 1) ? 0.5 : 1.0;
      double kk = sqrt(c->fkxs[i] + c->fkys[j] + c->fkzs[0]);
      int ib = (int)kk + 1;
      if (ib > 0 && ib < nbin) {
        size_t q = IDX(c, i, j, 0);
        double az = hypot(RE(bz, q), IM(bz, q));
        sp2d[ib - 1] = sp2d[ib - 1] + kk * kk * fac * az * az;
      }
    }
}

/*
 * fluxes (diagnostics.f90:62-103): sp(ib) with ib = int(kk) (the reference
 * writes sp(0) for kk < 1, out of bounds; it is kept here as bin 0 of an
 * (nbin+1)-long array), then spf(i) = sum(sp(i:NBIN)) for i = 1..NBIN.
 * spf has nbin entries.
 */
void or_fluxes(const or_ctx *c, const double *bx, const double *by, const double *bz, const double *ax,
               const double *ay, const double *az, double *spf) {
  const int nbin = c->nx / 2;
  double *sp = (double *)calloc((size_t)nbin + 1, sizeof(double));
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nxh; i++) {
        double fac = (i == 0 || i == c->nxh - 1) ? 1.0 : 2.0;
        double kk = sqrt(c->fkxs[i] + c->fkys[j] + c->fkzs[k]);
        int ib = (int)kk;
        size_t q = IDX(c, i, j, k);
        /* real(b * conjg(a)) summed over components */
        double scra = (RE(bx, q) * RE(ax, q) + IM(bx, q) * IM(ax, q)) +
                      (RE(by, q) * RE(ay, q) + IM(by, q) * IM(ay, q)) +
                      (RE(bz, q) * RE(az, q) + IM(bz, q) * IM(az, q));
        if (ib < nbin) sp[ib] = sp[ib] + fac * kk * kk * scra;
      }
  for (int i = 1; i <= nbin; i++) {
    double s = 0.0;
    for (int q = i; q <= nbin; q++) s += (q < nbin) ? sp[q] : 0.0;
    spf[i - 1] = s;
  }
  free(sp);
}

/*
 * glob (diagnostics.f90:114-199): out[0..8] energy quantities of curl b,
 * out[9..17] enstrophy quantities of curl curl b (x, y, z components of
 * energy, viscous and friction dissipation).
 */
void or_glob(or_ctx *c, const double *bx, const double *by, const double *bz, double *out) {
  const size_t m = (size_t)c->nxh * c->ny * c->nz;
  double *a[3];
  for (int f = 0; f < 3; f++) a[f] = (double *)malloc(sizeof(double) * 2 * m);
  memcpy(a[0], bx, sizeof(double) * 2 * m);
  memcpy(a[1], by, sizeof(double) * 2 * m);
  memcpy(a[2], bz, sizeof(double) * 2 * m);
  for (int pass = 0; pass < 2; pass++) {
    or_curl(c, a[0], a[1], a[2]);
    double t[9] = {0};
    for (int k = 0; k < c->nz; k++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nxh; i++) {
          double fac = (i == 0 || i == c->nxh - 1) ? 0.5 : 1.0;
          double k2 = c->fkxs[i] + c->fkys[j] + c->fkzs[k];
          if (k2 > 0.0) {
            size_t q = IDX(c, i, j, k);
            double s[3];
            for (int f = 0; f < 3; f++) {
              double h = hypot(RE(a[f], q), IM(a[f], q));
              s[f] = fac * (h * h);
            }
            for (int f = 0; f < 3; f++) {
              t[f] = t[f] + s[f];
              t[3 + f] = t[3 + f] + 2.0 * c->p.nu * s[f] * pow(k2, c->p.alpnu);
              t[6 + f] = t[6 + f] + 2.0 * c->p.mu * s[f] * pow(k2, c->p.alpmu);
            }
          }
        }
    for (int q = 0; q < 9; q++) out[9 * pass + q] = t[q];
  }
  for (int f = 0; f < 3; f++) free(a[f]);
}

int or_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
