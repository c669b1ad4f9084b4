// gt_api.cpp -- host side of the library: the context, the reference's
// host tables (defk, iniforcing), the forcing LCG, the truncation extents
// and the C ABI declared in include/gturbo3d.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/gturbo3d.h"
#include "gt_config.hpp"
#include "gt_internal.hpp"

namespace gt {
hipError_t launch_zinv(int prec, int nout, const KP& kp, const void* b, void* work, const void* tw, hipStream_t s);
hipError_t launch_zfwd(int prec, int epi, const KP& kp, const void* work, const void* b, void* out, const void* tw,
                       hipStream_t s);
hipError_t launch_yfft(int prec, int sign, int nfields, const KP& kp, void* work, const void* tw, hipStream_t s);
hipError_t launch_xfused(int prec, const KP& kp, void* work, const void* tw, hipStream_t s);
hipError_t launch_xc2r(const KP& kp, const void* work, double* phys, const void* tw, hipStream_t s);
hipError_t launch_xr2c(const KP& kp, const double* phys, void* work, const void* tw, hipStream_t s);
hipError_t launch_forcing(void* b, long long fs, const FEnt* ent, int n, bool sequential, hipStream_t s);
hipError_t launch_diag(const KP& kp, const void* b, const void* nl, double* partial, int nblocks, double* out,
                       int nbin, hipStream_t s);
int diag_len(int nbin);
hipError_t launch_zinv_emk(const KP& kp, const void* b, void* work, const void* tw, hipStream_t s);
bool zfwd_zinv_supported(const KP& kp);
hipError_t launch_zfwd_zinv(int prec, int stage, const KP& kp, void* work, const void* b, void* out, const void* tw,
                            const FAux& fx, hipStream_t s);
hipError_t launch_zinv_adj(const KP& kp, const void* nu, double s, bool ze, void* out, const void* tw,
                           hipStream_t st);
hipError_t launch_zfwd_adj(const KP& kp, const void* work, void* out, const void* tw, const void* lam,
                           hipStream_t st);
hipError_t launch_xadj(const KP& kp, void* work, const void* tw, hipStream_t st);
hipError_t launch_adj_ze(const KP& kp, const void* in, void* out, hipStream_t st);
hipError_t launch_adj_post(const KP& kp, const void* nu, const void* g, void* lam_out, hipStream_t st);
hipError_t launch_adj_stage1(const KP& kp, const void* h, const void* g1, void* acc, hipStream_t st);
hipError_t launch_adj_project(const KP& kp, void* x, hipStream_t st);
hipError_t launch_adj_fdot(const void* lam, long long fs, const FEnt* ent, int n, double* dots, hipStream_t st);
bool adjoint_nx_supported(int n);
enum { EPI_STEP1 = 0, EPI_STEP2 = 1, EPI_NONLIN = 2, EPI_INIT = 3 };
}  // namespace gt

using namespace gt;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(GT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int NFORC = 2048;  // src/modules.f90:39

struct Forced {
  int i, j, k;  // 1-based (kx, ky, kz) indices as in the reference
  double ff;
};

}  // namespace

struct gt_ctx {
  gt_params p{};
  int nx = 0, ny = 0, nz = 0, nxh = 0, nxp = 0;
  long long fs = 0;
  int prec = 0;
  int resident_blocks = 1024;
  int dev = 0;  // the device current at creation, which holds the context's memory
  int zchunk = 0;   // z planes per chunk of passes B, C (and D) (0: all at once)
  bool zchunk_d = true;  // pass D runs in the chunks too (else once, after them)
  hipStream_t ext_stream = nullptr;  // an ensemble's stream the batch also runs on
  // bumped whenever something an ensemble's captured graphs bake in changes:
  // the pruned/dense mode (extents of every launch), the tile widths, the z chunk
  unsigned gen = 0;
  // z chunks of passes B-D on one stream or alternating between two
  // (gt_set_zstreams); the second stream and its fork/join events
  int zstreams = 1;
  hipStream_t cstream = nullptr;
  hipEvent_t cev[2] = {nullptr, nullptr};
  // the z-forward passes fused with the z-inverse passes that follow them
  // where compiled (gt_set_fusion; gt_autotune times both); on by default
  // in FP64, off in FP32-bulk
  bool fuse_ea = true;
  int tz = 0, ty = 0, tze = 0;       // pencils per block of the z-inverse / y / z-forward passes (0: default)
  double tune_ms = 0.0;              // the autotuner's time per nonlinear evaluation
  // batches: nmem members with their own omega (and forcing seeds)
  int nmem = 1;
  std::vector<double> omegas;
  double* domega = nullptr;
  std::vector<char> mem_inside;  // member vanishes outside the kept extents
  hipStream_t stream = nullptr;

  // host tables (src/defk.f90)
  std::vector<double> fkx, fky, fkz, fkxs, fkys, fkzs, fkxn, fkyn, fkzn;
  std::vector<int> jc, kc;
  double fkxsmax = 0, fkysmax = 0, fkzsmax = 0, rtrunc2 = 0, sdt = 0, dt2 = 0;
  std::vector<Forced> forced;
  double inpe = 0, inpz = 0;
  bool force_seq = false;
  // forcing shells: the distinct |k| of the forced modes, ascending; a
  // per-shell amplitude replaces famp for the modes of that shell
  std::vector<double> shell_k2;
  std::vector<int> mode_shell;   // shell of each forced mode
  std::vector<int> entry_shell;  // shell of each forcing entry of one step
  std::vector<double> mode_amp;  // empty: famp for every mode

  // truncation extents
  int kxcut = 0;
  int kycut = 0, kzcut = 0;  // largest kyext / kzext entries
  std::vector<int16_t> kyext, kzext;
  bool pruned = false;

  // device memory
  cx<double>* b = nullptr;   // 3 fields
  cx<double>* bh = nullptr;  // 3 fields
  void* work = nullptr;      // 6 fields of the work precision
  double* dtab = nullptr;    // fkx fky fkz fkxn fkyn fkzn
  int16_t* dkyext = nullptr;
  int16_t* dkzext = nullptr;
  // the forcing entries of each (ky, kx) column (FAux): CSR over ky index *
  // nxp + kx index
  std::vector<int> fstart, fidx;
  int* dfstart = nullptr;
  int* dfidx = nullptr;
  void* tw64[3] = {nullptr, nullptr, nullptr};  // x, y, z
  void* tw32[3] = {nullptr, nullptr, nullptr};
  FEnt* dforce = nullptr;
  size_t dforce_cap = 0;
  FEnt* hforce = nullptr;  // pinned staging
  size_t hforce_cap = 0;
  hipEvent_t hforce_done = nullptr;
  bool hforce_pending = false;

  // diagnostics scratch (allocated on first use)
  double* ddiag_part = nullptr;
  double* ddiag_out = nullptr;
  int diag_blocks = 0;

  // adjoint: checkpoints of the forward run and the backward scratch
  struct Adjoint {
    int nsteps = 0, every = 0, nseg = 0;
    bool ready = false;
    // the forward left the states of its last segment in the segment buffer
    // (its half (nseg - 1) & 1), so the backward need not recompute them
    bool last_kept = false;
    std::vector<int32_t> seeds;     // forcing seed at the start of each segment
    cx<double>* ckpt = nullptr;     // nseg + 1 states (the last one: the final state)
    cx<double>* seg = nullptr;      // the states of one segment, every states
    cx<double>* work = nullptr;     // 9 fields
    cx<double>* vec = nullptr;      // 7 x 3 fields: lam (two, alternating) nu g bh h g1
    FEnt* dunit = nullptr;          // unit-amplitude forcing entries of every step
    double* ddots = nullptr;        // their dots with the cotangents
    size_t cap_ckpt = 0, cap_seg = 0, cap_unit = 0;
    hipStream_t rstream = nullptr;  // the segment recompute, beside the sweep
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // recompute done x2, sweep done x2
  } adj;

  // profiling
  bool prof = false;
  struct Rec {
    int pass;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double prof_ms[NPASS] = {0};
  long long prof_n[NPASS] = {0};
};

namespace {

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------

// defk (src/defk.f90:13-54, 88-91)
void build_wavenumbers(gt_ctx* c) {
  const gt_params& p = c->p;
  const int nx = c->nx, ny = c->ny, nz = c->nz, nxh = c->nxh;
  c->jc.assign(ny, 1);
  for (int j = 2; j <= ny; j++) c->jc[j - 1] = ny + 2 - j;
  c->kc.assign(nz, 1);
  for (int k = 2; k <= nz; k++) c->kc[k - 1] = nz + 2 - k;
  const double dkx = 2.0 * 3.14159265358979323846 / p.xlx;
  const double dky = 2.0 * 3.14159265358979323846 / p.xly;
  const double dkz = 2.0 * 3.14159265358979323846 / p.xlz;
  c->fkx.resize(nxh);
  c->fky.resize(ny);
  c->fkz.resize(nz);
  for (int i = 1; i <= nxh; i++) c->fkx[i - 1] = (double)(i - 1) * dkx;
  for (int j = 1; j <= ny / 2 + 1; j++) c->fky[j - 1] = (double)(j - 1) * dky;
  for (int j = ny / 2 + 2; j <= ny; j++) c->fky[j - 1] = (double)(j - ny - 1) * dky;
  for (int k = 1; k <= nz / 2 + 1; k++) c->fkz[k - 1] = (double)(k - 1) * dkz;
  for (int k = nz / 2 + 2; k <= nz; k++) c->fkz[k - 1] = (double)(k - nz - 1) * dkz;
  c->fkxs.resize(nxh);
  c->fkys.resize(ny);
  c->fkzs.resize(nz);
  for (int i = 0; i < nxh; i++) c->fkxs[i] = c->fkx[i] * c->fkx[i];
  for (int j = 0; j < ny; j++) c->fkys[j] = c->fky[j] * c->fky[j];
  for (int k = 0; k < nz; k++) c->fkzs[k] = c->fkz[k] * c->fkz[k];
  c->fkxsmax = c->fkxs[nxh - 1];
  c->fkysmax = c->fkys[ny / 2];
  c->fkzsmax = c->fkzs[nz / 2];
  c->rtrunc2 = p.rtrunc * p.rtrunc;
  // the three ratios of the trunc test (src/trunc.f90:14-16)
  c->fkxn.resize(nxh);
  c->fkyn.resize(ny);
  c->fkzn.resize(nz);
  for (int i = 0; i < nxh; i++) c->fkxn[i] = c->fkxs[i] / c->fkxsmax;
  for (int j = 0; j < ny; j++) c->fkyn[j] = c->fkys[j] / c->fkysmax;
  for (int k = 0; k < nz; k++) c->fkzn[k] = c->fkzs[k] / c->fkzsmax;
  (void)nx;
}

bool trunc_masked(const gt_ctx* c, int i, int j, int k) {
  // src/trunc.f90:14-17 (rtrunc squared at src/defk.f90:91)
  const double k3 = c->fkzn[k];
  const double k2 = k3 + c->fkyn[j];
  const double k1 = k2 + c->fkxn[i];
  return k1 > c->rtrunc2;
}

// iniforcing: one candidate mode (src/iniforcing_3d3c.f90:26-40)
int forced_try(gt_ctx* c, int i, int j, int k) {
  const double k2 = c->fkxs[i - 1] + c->fkys[j - 1] + c->fkzs[k - 1];
  const double kk = std::sqrt(k2);
  const double k2n = c->fkxs[i - 1] / c->fkxsmax + c->fkys[j - 1] / c->fkysmax + c->fkzs[k - 1] / c->fkzsmax;
  if (k2n <= c->rtrunc2 && k2 <= c->p.k2f * c->p.k2f && k2 > c->p.k1f * c->p.k1f) {
    if ((int)c->forced.size() >= NFORC) return -1;
    const double ff = 1.0 / kk;
    c->forced.push_back({i, j, k, ff});
    if (c->p.variant == GT_FORCING_3D3C) {
      const double a = c->p.famp * ff;
      c->inpz = c->inpz + k2 * k2 * (a * a);
      c->inpe = c->inpe + k2 * (a * a);
    }
  }
  return 0;
}

// iniforcing in the reference enumeration order, which fixes the order of
// the random draws (src/iniforcing_3d3c.f90:22-184, src/iniforcing_2d2c.f90:12-68)
int build_forcing(gt_ctx* c) {
  c->forced.clear();
  c->inpe = c->inpz = 0.0;
  const int nx2 = c->nx / 2, nx2p1 = nx2 + 1, ny = c->ny, nz = c->nz, ny2 = ny / 2, nz2 = nz / 2;
#define TRY(i, j, k) \
  if (forced_try(c, i, j, k)) return -1
  if (c->p.variant == GT_FORCING_3D3C) {
    for (int pl = 0; pl < 2; pl++) {  // kx = 0 plane, then kx = Nyquist plane
      if (pl == 1) {
        for (int i = 2; i <= nx2; i++)
          for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++) TRY(i, j, k);
      }
      const int i = (pl == 0) ? 1 : nx2p1;
      for (int k = 2; k <= nz2; k++) TRY(i, 1, k);
      for (int j = 2; j <= ny2; j++)
        for (int k = 1; k <= nz; k++) TRY(i, j, k);
      for (int k = 2; k <= nz2; k++) TRY(i, ny2 + 1, k);
    }
  } else {
    const int k = 1;
    for (int j = 2; j <= ny2; j++) TRY(1, j, k);
    for (int i = 2; i <= nx2; i++)
      for (int j = 1; j <= ny; j++) TRY(i, j, k);
    for (int j = 2; j <= ny2; j++) TRY(nx2p1, j, k);
    for (const Forced& f : c->forced) {
      const double k2 = c->fkxs[f.i - 1] + c->fkys[f.j - 1] + c->fkzs[f.k - 1];
      const double a = c->p.famp * f.ff;
      c->inpe = c->inpe + k2 * (a * a);
      c->inpz = c->inpz + k2 * k2 * (a * a);
    }
  }
#undef TRY
  return 0;
}

long long elem_off(const gt_ctx* c, int i0, int j0, int k0) {  // 0-based (kx, ky, kz)
  return ((long long)k0 * c->ny + j0) * c->nxp + i0;
}

// Extents of the modes that survive the truncation, widened to cover the
// forced modes and their conjugates.
void build_extents(gt_ctx* c) {
  const int nxh = c->nxh, nxp = c->nxp, ny = c->ny, nz = c->nz;
  c->kzext.assign((size_t)ny * nxp, (int16_t)-1);
  c->kyext.assign(nxp, (int16_t)-1);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nxh; i++) {
      int K = -1;
      for (int m = 0; m <= nz / 2; m++) {
        if (trunc_masked(c, i, j, m)) break;
        K = m;
      }
      c->kzext[(size_t)j * nxp + i] = (int16_t)K;
    }
  auto widen = [&](int i0, int j0, int k0) {
    const int mz = std::min(k0, nz - k0);
    int16_t& K = c->kzext[(size_t)j0 * nxp + i0];
    if (K < mz) K = (int16_t)mz;
  };
  for (const Forced& f : c->forced) {
    widen(f.i - 1, f.j - 1, f.k - 1);
    widen(f.i - 1, c->jc[f.j - 1] - 1, c->kc[f.k - 1] - 1);
  }
  // a column is alive iff |ky| <= kyext[kx]; every column inside is kept
  for (int i = 0; i < nxh; i++) {
    int Ky = -1;
    for (int j = 0; j < ny; j++)
      if (c->kzext[(size_t)j * nxp + i] >= 0) Ky = std::max(Ky, std::min(j, ny - j));
    c->kyext[i] = (int16_t)Ky;
    for (int j = 0; j < ny; j++)
      if (std::min(j, ny - j) <= Ky && c->kzext[(size_t)j * nxp + i] < 0) c->kzext[(size_t)j * nxp + i] = 0;
  }
  c->kxcut = -1;
  for (int i = 0; i < nxh; i++)
    if (c->kyext[i] >= 0) c->kxcut = i;
  if (c->kxcut < 0) c->kxcut = 0;
  c->kycut = 0;
  c->kzcut = 0;
  for (int i = 0; i < nxh; i++) c->kycut = std::max(c->kycut, (int)c->kyext[i]);
  for (int16_t K : c->kzext) c->kzcut = std::max(c->kzcut, (int)K);
}

// Stockham twiddles exp(-2 pi i r k/(Ns R)) per stage (fft_core.hpp)
template <typename T>
std::vector<T> twiddles(int n) {
  const PlanInfo pi = plan_of(n);
  std::vector<T> tw;
  int Ns = 1;
  for (int s = 0; s < pi.nst; s++) {
    const int R = pi.r[s];
    for (int k = 0; k < Ns; k++)
      for (int r = 1; r < R; r++) {
        const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)(r * k) /
                              (long double)(Ns * R);
        tw.push_back((T)cosl(a));
        tw.push_back((T)sinl(a));
      }
    Ns *= R;
  }
  if (tw.empty()) tw.push_back(T(0));
  return tw;
}

KP make_kp(const gt_ctx* c, bool dense) {
  KP kp{};
  kp.nx = c->nx;
  kp.ny = c->ny;
  kp.nz = c->nz;
  kp.nxh = c->nxh;
  kp.nxp = c->nxp;
  kp.fs = c->fs;
  kp.fkx = c->dtab;
  kp.fky = c->dtab + c->nxh;
  kp.fkz = c->dtab + c->nxh + c->ny;
  kp.fkxn = c->dtab + c->nxh + c->ny + c->nz;
  kp.fkyn = kp.fkxn + c->nxh;
  kp.fkzn = kp.fkyn + c->ny;
  kp.rtrunc2 = c->rtrunc2;
  kp.truncate = c->p.truncate;
  const bool pr = c->pruned && !dense;
  kp.pruned = pr ? 1 : 0;
  kp.kxcut = pr ? c->kxcut : c->nxh - 1;
  kp.kycut = pr ? c->kycut : c->ny / 2;
  kp.kzcut = pr ? c->kzcut : c->nz / 2;
  kp.kyext = c->dkyext;
  kp.kzext = c->dkzext;
  kp.omega = c->p.omega;
  kp.dt = c->p.dt;
  kp.dt2 = c->dt2;
  kp.nu = c->p.nu;
  kp.mu = c->p.mu;
  kp.alpnu = c->p.alpnu;
  kp.alpmu = c->p.alpmu;
  kp.scale = 1.0 / ((double)c->nx * (double)c->ny * (double)c->nz);
  kp.visc_simple = (c->p.alpnu == 1.0 && c->p.mu == 0.0) ? 1 : 0;
  kp.emx = kp.visc_simple ? kp.fkzn + c->nz : nullptr;
  kp.emy = kp.visc_simple ? kp.emx + c->nxh : nullptr;
  kp.emz = kp.visc_simple ? kp.emy + c->ny : nullptr;
  kp.resident_blocks = c->resident_blocks;
  kp.nmem = c->nmem;
  kp.bstride = 3 * c->fs;
  kp.wstride = 6 * c->fs;
  kp.omega_m = c->nmem > 1 ? c->domega : nullptr;
  kp.tz = c->tz;
  kp.ty = c->ty;
  kp.tze = c->tze;
  kp.z0 = 0;
  kp.zn = c->nz;
  return kp;
}

// launch description of one member (velocity, init and diagnostics work on
// one member at a time; the caller offsets the pointers)
KP make_kp1(const gt_ctx* c, bool dense, int member) {
  KP kp = make_kp(c, dense);
  kp.nmem = 1;
  kp.omega_m = nullptr;
  if (c->nmem > 1) kp.omega = c->omegas[member];
  return kp;
}

const void* twp(const gt_ctx* c, int axis) { return c->prec ? c->tw32[axis] : c->tw64[axis]; }

hipEvent_t pool_get(gt_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct ProfScope {
  gt_ctx* c;
  int pass;
  hipEvent_t a = nullptr;
  hipStream_t s;
  // events on the stream the pass is launched on (the context stream by default)
  ProfScope(gt_ctx* c_, int pass_, hipStream_t s_ = nullptr) : c(c_), pass(pass_), s(s_ ? s_ : c_->stream) {
    if (c->prof) {
      a = pool_get(c);
      if (a) hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (c->prof && a) {
      hipEvent_t b = pool_get(c);
      if (b) {
        hipEventRecord(b, s);
        c->recs.push_back({pass, a, b});
      }
    }
  }
};

#define LAUNCH_ON(pass, stream, expr)                                                            \
  do {                                                                                          \
    ProfScope ps_(c, pass, stream);                                                             \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return fail(GT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH(pass, expr)                                                                      \
  do {                                                                                          \
    ProfScope ps_(c, pass);                                                                     \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return fail(GT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Passes B-D of one evaluation: work fields 0-5 (pass A's output) -> 0-2.
int nonlinear_mid(gt_ctx* c, const KP& kp, hipStream_t st) {
  const int pr = c->prec;
  // B, C and D in chunks of zchunk planes along z: a chunk's intermediates
  // are read by the next pass while they are still in the device's
  // memory-side cache (256 MB on MI355X).  The chunk size is a launch shape,
  // picked by gt_autotune (0: all planes at once); results do not depend on it.
  const int zc = (c->zchunk > 0 && c->zchunk < c->nz) ? c->zchunk : c->nz;
  const bool dch = c->zchunk_d || zc == c->nz;
  // zstreams = 2: odd chunks on a second stream, so that two chunks overlap
  // (gt_set_zstreams; the autotuner picks it with the chunk size).  Per-pass
  // timing records each launch's events on its own stream.
  const bool two = c->zstreams == 2 && zc < c->nz && dch;
  if (two && !c->cstream) {
    // default priority, like the context stream
    HIPCHK(hipStreamCreate(&c->cstream));
    for (auto& e : c->cev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (two) {
    HIPCHK(hipEventRecord(c->cev[0], st));
    HIPCHK(hipStreamWaitEvent(c->cstream, c->cev[0], 0));
  }
  int k = 0;
  for (int z0 = 0; z0 < c->nz; z0 += zc, k++) {
    KP kc = kp;
    kc.z0 = z0;
    kc.zn = std::min(zc, c->nz - z0);
    hipStream_t cs = (two && (k & 1)) ? c->cstream : st;
    LAUNCH_ON(PASS_B, cs, launch_yfft(pr, +1, 6, kc, c->work, twp(c, 1), cs));
    LAUNCH_ON(PASS_C, cs, launch_xfused(pr, kc, c->work, twp(c, 0), cs));
    if (dch) LAUNCH_ON(PASS_D, cs, launch_yfft(pr, -1, 3, kc, c->work, twp(c, 1), cs));
  }
  if (two) {
    HIPCHK(hipEventRecord(c->cev[1], c->cstream));
    HIPCHK(hipStreamWaitEvent(st, c->cev[1], 0));
  }
  if (!dch) LAUNCH(PASS_D, launch_yfft(pr, -1, 3, kp, c->work, twp(c, 1), st));
  return GT_OK;
}

// One evaluation of N(a; b) through passes A-D, then the pass-E epilogue.
// bold: the start-of-step state read by the Coriolis term and the RK update
// of pass E (c->b when null)
int nonlinear_stage(gt_ctx* c, const KP& kp, const cx<double>* a, int epi, cx<double>* out, hipStream_t st,
                    const cx<double>* bold = nullptr) {
  const int pr = c->prec;
  LAUNCH(PASS_A, launch_zinv(pr, 6, kp, a, c->work, twp(c, 2), st));
  if (int rc = nonlinear_mid(c, kp, st)) return rc;
  const int pid = (epi == EPI_STEP2) ? PASS_E2 : PASS_E1;
  LAUNCH(pid, launch_zfwd(pr, epi, kp, c->work, bold ? bold : c->b, out, twp(c, 2), st));
  return GT_OK;
}

// Both stages of one step from bin into bout (which may be bin).  With
// fusion, stage 1's pass E leaves bh in registers and runs stage 2's pass A
// on it, so bh is neither stored nor read back.
int step_stages(gt_ctx* c, const KP& kp, const cx<double>* bin, cx<double>* bout, hipStream_t st) {
  if (c->fuse_ea && zfwd_zinv_supported(kp)) {
    const int pr = c->prec;
    // stage 1: passes A-D of N1 = nlt(b; b)                 (src/main.f90:124-131)
    LAUNCH(PASS_A, launch_zinv(pr, 6, kp, bin, c->work, twp(c, 2), st));
    if (int rc = nonlinear_mid(c, kp, st)) return rc;
    // bh = emk (b + dt2 N1) (src/main.f90:141), then pass A of nlt(bh; b)
    LAUNCH(PASS_EA, launch_zfwd_zinv(pr, 1, kp, c->work, bin, nullptr, twp(c, 2), FAux{}, st));
    // stage 2: passes B-E, b = emk (emk b + dt N2)          (src/main.f90:144-146)
    if (int rc = nonlinear_mid(c, kp, st)) return rc;
    LAUNCH(PASS_E2, launch_zfwd(pr, EPI_STEP2, kp, c->work, bin, bout, twp(c, 2), st));
    return GT_OK;
  }
  // stage 1: N1 = nlt(b; b), bh = emk (b + dt2 N1)      (src/main.f90:124-141)
  if (int rc = nonlinear_stage(c, kp, bin, EPI_STEP1, c->bh, st, bin)) return rc;
  // stage 2: N2 = nlt(bh; b), b = emk (emk b + dt N2)   (src/main.f90:144-146)
  return nonlinear_stage(c, kp, c->bh, EPI_STEP2, bout, st, bin);
}

// forcing increments of one step (src/forcing_3d3c.f90:26-72,
// src/forcing_2d2c.f90:11-48); appends entries and advances the seed
// unit: every amplitude 1 (the derivative of the increments with respect to
// their shell's amplitude), else famp or the per-shell amplitudes
void forcing_entries(const gt_ctx* c, int32_t* seed, FEnt* out, int member = 0, bool unit = false) {
  auto rann = [](int32_t* s) -> float {  // src/forcing_3d3c.f90:82-93
    const uint32_t mask = 0x7FFFFFFFu;
    uint32_t v = (69069u * (uint32_t)(*s) + 1u) & mask;
    *s = (int32_t)v;
    return (float)((double)v / (double)mask);
  };
  const int nxh = c->nxh;
  size_t o = 0;
  for (size_t im = 0; im < c->forced.size(); im++) {
    const Forced& fm = c->forced[im];
    const double amp = unit ? 1.0 : (c->mode_amp.empty() ? c->p.famp : c->mode_amp[im]);
    const double sf = c->sdt * amp;
    const int i = fm.i, j = fm.j, k = fm.k;
    const long long mo = (long long)member * 3 * c->fs;
    const long long q = mo + elem_off(c, i - 1, j - 1, k - 1);
    const long long qc = mo + elem_off(c, i - 1, c->jc[j - 1] - 1, c->kc[k - 1] - 1);
    if (!out) {  // draws only: advance the seed as the reference does
      rann(seed);
      if (c->p.variant != GT_FORCING_2D2C) rann(seed);
      continue;
    }
    if (c->p.variant == GT_FORCING_2D2C) {
      const double pi2 = 2.0 * 3.14159265358979;
      const double te = pi2 * (double)rann(seed);
      const double er = sf * std::cos(te), ei = sf * std::sin(te);
      const double kr = er * fm.ff, ki = ei * fm.ff;
      out[o++] = {q, {0.0, 0.0, 0.0, 0.0, kr, ki}};
      if (i == 1 || i == nxh) out[o++] = {qc, {0.0, 0.0, 0.0, 0.0, kr, -ki}};
      continue;
    }
    const double pi2 = 2.0 * 3.14159265358979323846;
    double te = pi2 * (double)rann(seed);
    const double er = sf * std::cos(te), ei = sf * std::sin(te);
    te = pi2 * (double)rann(seed);
    const double cte = std::cos(te), ste = std::sin(te);
    const double kr = er * fm.ff, ki = ei * fm.ff;
    if (i == 1 && j == 1) {
      out[o++] = {q, {kr * cte, ki * cte, kr * ste, ki * ste, 0.0, 0.0}};
      out[o++] = {qc, {kr * cte, (-ki) * cte, kr * ste, (-ki) * ste, 0.0, 0.0}};
    } else {
      double k2 = c->fkxs[i - 1] + c->fkys[j - 1];
      const double kxy = std::sqrt(k2);
      k2 = k2 + c->fkzs[k - 1];
      const double kk = std::sqrt(k2);
      const double c1 = c->fkz[k - 1] / kk, s1 = kxy / kk;
      const double c2 = c->fkx[i - 1] / kxy, s2 = c->fky[j - 1] / kxy;
      const double f1 = cte * c2 * c1 - ste * s2;
      const double f2 = cte * s2 * c1 + ste * c2;
      // bz - effek*cte*s1 == bz + (-(effek*cte*s1)) exactly
      out[o++] = {q, {kr * f1, ki * f1, kr * f2, ki * f2, -(kr * cte * s1), -(ki * cte * s1)}};
      if (i == 1 || i == nxh)
        out[o++] = {qc, {kr * f1, (-ki) * f1, kr * f2, (-ki) * f2, -(kr * cte * s1), -((-ki) * cte * s1)}};
    }
  }
}

size_t entries_per_step(const gt_ctx* c) {
  size_t n = 0;
  for (const Forced& f : c->forced) {
    n += 1;
    const bool conj = (c->p.variant == GT_FORCING_2D2C) ? (f.i == 1 || f.i == c->nxh)
                                                          : ((f.i == 1 && f.j == 1) || f.i == 1 || f.i == c->nxh);
    if (conj) n += 1;
  }
  return n;
}

// Forcing shells: forced modes grouped by k^2 (equal up to rounding), and
// the shell of each entry forcing_entries emits for one step.
void build_shells(gt_ctx* c) {
  std::vector<double> k2s;
  for (const Forced& f : c->forced) k2s.push_back(c->fkxs[f.i - 1] + c->fkys[f.j - 1] + c->fkzs[f.k - 1]);
  std::vector<double> sorted = k2s;
  std::sort(sorted.begin(), sorted.end());
  c->shell_k2.clear();
  for (double v : sorted)
    if (c->shell_k2.empty() || v > c->shell_k2.back() * (1.0 + 1e-9)) c->shell_k2.push_back(v);
  c->mode_shell.assign(c->forced.size(), 0);
  c->entry_shell.clear();
  for (size_t im = 0; im < c->forced.size(); im++) {
    int sh = 0;
    while (sh + 1 < (int)c->shell_k2.size() && k2s[im] > c->shell_k2[sh] * (1.0 + 1e-9)) sh++;
    c->mode_shell[im] = sh;
    const Forced& f = c->forced[im];
    const bool conj = (c->p.variant == GT_FORCING_2D2C) ? (f.i == 1 || f.i == c->nxh)
                                                          : ((f.i == 1 && f.j == 1) || f.i == 1 || f.i == c->nxh);
    c->entry_shell.push_back(sh);
    if (conj) c->entry_shell.push_back(sh);
  }
}

void free_ctx(gt_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->adj.rstream) hipStreamSynchronize(c->adj.rstream);
  if (c->cstream) {
    hipStreamSynchronize(c->cstream);
    hipStreamDestroy(c->cstream);
  }
  for (auto e : c->cev)
    if (e) hipEventDestroy(e);
  for (auto& r : c->recs) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto e : c->pool) hipEventDestroy(e);

  if (c->hforce_done) hipEventDestroy(c->hforce_done);
  hipFree(c->b);
  hipFree(c->bh);
  hipFree(c->domega);
  hipFree(c->work);
  hipFree(c->dtab);
  hipFree(c->dkyext);
  hipFree(c->dkzext);
  hipFree(c->dfstart);
  hipFree(c->dfidx);
  for (int a = 0; a < 3; a++) {
    hipFree(c->tw64[a]);
    hipFree(c->tw32[a]);
  }
  hipFree(c->dforce);
  hipFree(c->adj.ckpt);
  hipFree(c->adj.seg);
  hipFree(c->adj.work);
  hipFree(c->adj.vec);
  hipFree(c->adj.dunit);
  hipFree(c->adj.ddots);
  for (auto e : c->adj.ev)
    if (e) hipEventDestroy(e);
  if (c->adj.rstream) hipStreamDestroy(c->adj.rstream);
  hipFree(c->ddiag_part);
  hipFree(c->ddiag_out);
  if (c->hforce) hipHostFree(c->hforce);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int ensure_force_cap(gt_ctx* c, size_t n) {
  if (n <= c->dforce_cap && n <= c->hforce_cap) return GT_OK;
  if (c->hforce_pending) {
    HIPCHK(hipEventSynchronize(c->hforce_done));
    c->hforce_pending = false;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  size_t cap = std::max<size_t>(n, 2 * std::max(c->dforce_cap, (size_t)64));
  hipFree(c->dforce);
  c->dforce = nullptr;
  if (c->hforce) hipHostFree(c->hforce);
  c->hforce = nullptr;
  c->dforce_cap = c->hforce_cap = 0;
  HIPCHK(hipMalloc(&c->dforce, cap * sizeof(FEnt)));
  HIPCHK(hipHostMalloc((void**)&c->hforce, cap * sizeof(FEnt), hipHostMallocDefault));
  c->dforce_cap = c->hforce_cap = cap;
  return GT_OK;
}

// Set pruned mode and make bh vanish outside the kept extents.
int set_mode(gt_ctx* c, bool pruned) {
  if (c->pruned != pruned) c->gen++;
  c->pruned = pruned;
  HIPCHK(hipMemsetAsync(c->bh, 0, sizeof(cx<double>) * 3 * c->fs * c->nmem, c->stream));
  return GT_OK;
}

// Every entry point on a context starts here: a context bound to an
// ensemble first waits for the ensemble's stream, which runs on the same
// device fields.
int check_ctx(gt_ctx* c) {
  if (!c) return fail(GT_EINVAL, "null context");
  if (c->ext_stream) HIPCHK(hipStreamSynchronize(c->ext_stream));
  return GT_OK;
}

}  // namespace

extern "C" {

const char* gt_last_error(void) { return g_err.c_str(); }

static int create_ctx(const gt_params* p, int nmem, const double* omegas, gt_ctx** out) {
  if (!p || !out) return fail(GT_EINVAL, "gt_create: null argument");
  if (nmem < 1) return fail(GT_EINVAL, "gt_create: need at least one member");
  *out = nullptr;
  for (int n : {p->nx, p->ny, p->nz})
    if (!supported_n(n))
      return fail(GT_ESIZE, "gt_create: grid size " + std::to_string(n) +
                                " not compiled (supported: 16 24 32 48 64 96 128 192 256 384 512 768 1024)");
  if (p->variant != GT_FORCING_3D3C && p->variant != GT_FORCING_2D2C)
    return fail(GT_EINVAL, "gt_create: variant must be 3 (3D3C) or 2 (2D2C)");
  if (p->dt2_mode != GT_DT2_REFERENCE && p->dt2_mode != GT_DT2_MIDPOINT)
    return fail(GT_EINVAL, "gt_create: dt2_mode must be 0 or 1");
  if (p->precision != GT_PREC_FP64 && p->precision != GT_PREC_FP32_BULK)
    return fail(GT_EINVAL, "gt_create: precision must be 0 or 1");
  gt_ctx* c = new (std::nothrow) gt_ctx();
  if (!c) return fail(GT_ENOMEM, "gt_create: out of host memory");
  c->p = *p;
  c->nx = p->nx;
  c->ny = p->ny;
  c->nz = p->nz;
  c->nxh = p->nx / 2 + 1;
  c->nxp = (c->nxh + 15) / 16 * 16;
  c->fs = (long long)c->nz * c->ny * c->nxp;
  c->prec = p->precision;
  c->fuse_ea = c->prec == GT_PREC_FP64;
  c->nmem = nmem;
  c->omegas.assign(nmem, p->omega);
  if (omegas)
    for (int m = 0; m < nmem; m++) c->omegas[m] = omegas[m];
  c->mem_inside.assign(nmem, 1);
  c->sdt = std::sqrt(p->dt);                              // src/inieuler.f90:67
  c->dt2 = (p->dt2_mode == GT_DT2_MIDPOINT) ? 0.5 * p->dt : 0.0;
  build_wavenumbers(c);
  if (build_forcing(c)) {
    delete c;
    return fail(GT_ENFORC, "gt_create: NFORC too small ! (more than 2048 forced modes)");
  }
  build_extents(c);
  build_shells(c);
  {
    std::vector<long long> offs;
    for (const Forced& f : c->forced) {
      offs.push_back(elem_off(c, f.i - 1, f.j - 1, f.k - 1));
      const bool conj = (p->variant == GT_FORCING_2D2C) ? (f.i == 1 || f.i == c->nxh)
                                                        : ((f.i == 1 && f.j == 1) || f.i == 1 || f.i == c->nxh);
      if (conj) offs.push_back(elem_off(c, f.i - 1, c->jc[f.j - 1] - 1, c->kc[f.k - 1] - 1));
    }
    // entries by (ky, kx) column, key = ky index * nxp + kx index (the
    // offset within a z plane), in table order within a column (the order
    // forcing_entries emits them)
    const long long zs = (long long)c->ny * c->nxp;
    std::vector<int> key(offs.size());
    c->fstart.assign((size_t)zs + 1, 0);
    for (size_t e = 0; e < offs.size(); e++) {
      key[e] = (int)(offs[e] % zs);
      c->fstart[key[e] + 1]++;
    }
    for (size_t k = 1; k < c->fstart.size(); k++) c->fstart[k] += c->fstart[k - 1];
    c->fidx.assign(offs.size(), 0);
    std::vector<int> fill(c->fstart.begin(), c->fstart.end() - 1);
    for (size_t e = 0; e < offs.size(); e++) c->fidx[fill[key[e]]++] = (int)e;
    std::sort(offs.begin(), offs.end());
    c->force_seq = std::adjacent_find(offs.begin(), offs.end()) != offs.end();
  }

  int rc = GT_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == GT_OK) rc = fail(e == hipErrorOutOfMemory ? GT_ENOMEM : GT_EHIP,
                                                  std::string("gt_create: ") + what + ": " + hipGetErrorString(e));
  };
  chk(hipStreamCreate(&c->stream), "hipStreamCreate");
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess) c->dev = dev;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev) == hipSuccess && ncu > 0)
      c->resident_blocks = 4 * ncu;
  }
  const size_t cw = c->prec ? sizeof(cx<float>) : sizeof(cx<double>);
  const size_t nm = (size_t)nmem;
  if (rc == GT_OK) chk(hipMalloc(&c->b, sizeof(cx<double>) * 3 * c->fs * nm), "state");
  if (rc == GT_OK) chk(hipMalloc(&c->bh, sizeof(cx<double>) * 3 * c->fs * nm), "half-step state");
  if (rc == GT_OK) chk(hipMalloc(&c->work, cw * 6 * c->fs * nm), "work fields");
  if (rc == GT_OK) chk(hipMalloc(&c->domega, sizeof(double) * nm), "member omegas");
  if (rc == GT_OK)
    chk(hipMemcpy(c->domega, c->omegas.data(), sizeof(double) * nm, hipMemcpyHostToDevice), "member omegas");
  std::vector<double> tab;
  for (auto* v : {&c->fkx, &c->fky, &c->fkz, &c->fkxn, &c->fkyn, &c->fkzn}) tab.insert(tab.end(), v->begin(), v->end());
  // per-axis integrating factors exp(-0.5*dt*nu*fk^2) (simple viscosity)
  for (auto* v : {&c->fkxs, &c->fkys, &c->fkzs})
    for (double k2 : *v) tab.push_back(std::exp(-0.5 * c->p.dt * (c->p.nu * k2)));
  if (rc == GT_OK) chk(hipMalloc(&c->dtab, sizeof(double) * tab.size()), "tables");
  if (rc == GT_OK) chk(hipMemcpy(c->dtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice), "tables");
  if (rc == GT_OK) chk(hipMalloc(&c->dkyext, sizeof(int16_t) * c->kyext.size()), "extents");
  if (rc == GT_OK)
    chk(hipMemcpy(c->dkyext, c->kyext.data(), sizeof(int16_t) * c->kyext.size(), hipMemcpyHostToDevice), "extents");
  if (rc == GT_OK) chk(hipMalloc(&c->dfstart, sizeof(int) * c->fstart.size()), "forcing tiles");
  if (rc == GT_OK)
    chk(hipMemcpy(c->dfstart, c->fstart.data(), sizeof(int) * c->fstart.size(), hipMemcpyHostToDevice), "forcing tiles");
  if (rc == GT_OK) chk(hipMalloc(&c->dfidx, sizeof(int) * std::max<size_t>(c->fidx.size(), 1)), "forcing tiles");
  if (rc == GT_OK && !c->fidx.empty())
    chk(hipMemcpy(c->dfidx, c->fidx.data(), sizeof(int) * c->fidx.size(), hipMemcpyHostToDevice), "forcing tiles");
  if (rc == GT_OK) chk(hipMalloc(&c->dkzext, sizeof(int16_t) * c->kzext.size()), "extents");
  if (rc == GT_OK)
    chk(hipMemcpy(c->dkzext, c->kzext.data(), sizeof(int16_t) * c->kzext.size(), hipMemcpyHostToDevice), "extents");
  const int ns[3] = {c->nx, c->ny, c->nz};
  for (int a = 0; a < 3 && rc == GT_OK; a++) {
    std::vector<double> t64 = twiddles<double>(ns[a]);
    chk(hipMalloc(&c->tw64[a], sizeof(double) * t64.size()), "twiddles");
    if (rc == GT_OK) chk(hipMemcpy(c->tw64[a], t64.data(), sizeof(double) * t64.size(), hipMemcpyHostToDevice), "twiddles");
    if (c->prec && rc == GT_OK) {
      std::vector<float> t32 = twiddles<float>(ns[a]);
      chk(hipMalloc(&c->tw32[a], sizeof(float) * t32.size()), "twiddles");
      if (rc == GT_OK)
        chk(hipMemcpy(c->tw32[a], t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice), "twiddles");
    }
  }
  if (rc == GT_OK) chk(hipEventCreateWithFlags(&c->hforce_done, hipEventDisableTiming), "event");
  if (rc == GT_OK) chk(hipMemsetAsync(c->b, 0, sizeof(cx<double>) * 3 * c->fs * nm, c->stream), "memset");
  if (rc == GT_OK) chk(hipMemsetAsync(c->work, 0, cw * 6 * c->fs * nm, c->stream), "memset");
  if (rc == GT_OK) rc = set_mode(c, c->p.truncate == 1);
  if (rc == GT_OK) chk(hipStreamSynchronize(c->stream), "sync");
  if (rc != GT_OK) {
    free_ctx(c);
    return rc;
  }
  *out = c;
  return GT_OK;
}

int gt_create(const gt_params* p, gt_ctx** out) { return create_ctx(p, 1, nullptr, out); }

int gt_destroy(gt_ctx* c) {
  if (!c) return GT_OK;
  free_ctx(c);
  return GT_OK;
}

// Three host fields [nz][ny][nxh] <-> three pitched device fields: one linear
// copy per field through a contiguous device staging area (3 nz ny nxh
// complex) and one device-side 2D copy, instead of a 2D copy straight from
// host memory, which moves one short row at a time.
static int fields_to_device(gt_ctx* c, const double* const src[3], cx<double>* dst, void* staging, hipStream_t st) {
  const size_t plane = (size_t)c->nz * c->ny * c->nxh;
  cx<double>* stg = (cx<double>*)staging;
  for (int f = 0; f < 3; f++)
    HIPCHK(hipMemcpyAsync(stg + f * plane, src[f], sizeof(cx<double>) * plane, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpy2DAsync(dst, sizeof(cx<double>) * c->nxp, stg, sizeof(cx<double>) * c->nxh,
                          sizeof(cx<double>) * c->nxh, (size_t)3 * c->ny * c->nz, hipMemcpyDeviceToDevice, st));
  return GT_OK;
}

static int fields_to_host(gt_ctx* c, const cx<double>* src, double* const dst[3], void* staging, hipStream_t st) {
  const size_t plane = (size_t)c->nz * c->ny * c->nxh;
  cx<double>* stg = (cx<double>*)staging;
  HIPCHK(hipMemcpy2DAsync(stg, sizeof(cx<double>) * c->nxh, src, sizeof(cx<double>) * c->nxp,
                          sizeof(cx<double>) * c->nxh, (size_t)3 * c->ny * c->nz, hipMemcpyDeviceToDevice, st));
  for (int f = 0; f < 3; f++)
    HIPCHK(hipMemcpyAsync(dst[f], stg + f * plane, sizeof(cx<double>) * plane, hipMemcpyDeviceToHost, st));
  return GT_OK;
}

static int set_state_m(gt_ctx* c, int member, const double* bx, const double* by, const double* bz) {
  if (int rc = check_ctx(c)) return rc;
  if (c->ext_stream) HIPCHK(hipStreamSynchronize(c->ext_stream));
  if (!bx || !by || !bz) return fail(GT_EINVAL, "set_state: null field");
  if (member < 0 || member >= c->nmem) return fail(GT_EINVAL, "set_state: member out of range");
  const double* src[3] = {bx, by, bz};
  // pruned only if every member vanishes outside the kept extents
  bool inside = (c->p.truncate == 1);
  for (int f = 0; f < 3 && inside; f++) {
    const double* sf = src[f];
    for (int k = 0; k < c->nz && inside; k++) {
      const int mk = std::min(k, c->nz - k);
      for (int j = 0; j < c->ny && inside; j++)
        for (int i = 0; i < c->nxh; i++) {
          if (mk <= c->kzext[(size_t)j * c->nxp + i]) continue;
          const size_t q = (((size_t)k * c->ny + j) * c->nxh + i) * 2;
          if (sf[q] != 0.0 || sf[q + 1] != 0.0) {
            inside = false;
            break;
          }
        }
    }
  }
  cx<double>* dst = c->b + (size_t)member * 3 * c->fs;
  HIPCHK(hipMemsetAsync(dst, 0, sizeof(cx<double>) * 3 * c->fs, c->stream));
  // the work fields (at least 3 fs complex doubles in either precision) stage the copy
  if (int rc = fields_to_device(c, src, dst, c->work, c->stream)) return rc;
  c->mem_inside[member] = inside ? 1 : 0;
  c->adj.ready = false;
  bool all = true;
  for (char v : c->mem_inside) all = all && v;
  if (int rc = set_mode(c, all)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}

static int get_state_m(gt_ctx* c, int member, double* bx, double* by, double* bz) {
  if (int rc = check_ctx(c)) return rc;
  if (c->ext_stream) HIPCHK(hipStreamSynchronize(c->ext_stream));
  if (!bx || !by || !bz) return fail(GT_EINVAL, "get_state: null field");
  if (member < 0 || member >= c->nmem) return fail(GT_EINVAL, "get_state: member out of range");
  double* dst[3] = {bx, by, bz};
  const cx<double>* src = c->b + (size_t)member * 3 * c->fs;
  if (int rc = fields_to_host(c, src, dst, c->work, c->stream)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}

int gt_set_state(gt_ctx* c, const double* bx, const double* by, const double* bz) {
  return set_state_m(c, 0, bx, by, bz);
}

int gt_get_state(gt_ctx* c, double* bx, double* by, double* bz) { return get_state_m(c, 0, bx, by, bz); }

// Enqueue nsteps steps of every member on stream st; the forcing entries of
// step s are at dforce + s * per * nmem.
static int enqueue_steps(gt_ctx* c, hipStream_t st, int nsteps, const FEnt* dforce) {
  const size_t per = entries_per_step(c);
  const KP kp = make_kp(c, false);
  const bool fuse = c->fuse_ea && zfwd_zinv_supported(kp);
  const int pr = c->prec;
  for (int s = 0; s < nsteps; s++) {
    const FEnt* fstep = dforce + per * c->nmem * s;
    if (!fuse) {
      if (int rc = step_stages(c, kp, c->b, c->b, st)) return rc;
      // forcing (src/main.f90:149)
      if (per > 0) LAUNCH(PASS_F, launch_forcing(c->b, c->fs, fstep, (int)(per * c->nmem), c->force_seq, st));
      continue;
    }
    // fused: pass A of stage 1 was run by the previous step's last kernel,
    // except in the first step of the call
    if (s == 0) LAUNCH(PASS_A, launch_zinv(pr, 6, kp, c->b, c->work, twp(c, 2), st));
    // stage 1: passes B-D of N1 = nlt(b; b); bh = emk (b + dt2 N1) with
    // pass A of nlt(bh; b)                       (src/main.f90:124-141)
    if (int rc = nonlinear_mid(c, kp, st)) return rc;
    LAUNCH(PASS_EA, launch_zfwd_zinv(pr, 1, kp, c->work, c->b, nullptr, twp(c, 2), FAux{}, st));
    // stage 2: passes B-E, b = emk (emk b + dt N2) (src/main.f90:144-146)
    if (int rc = nonlinear_mid(c, kp, st)) return rc;
    if (s + 1 < nsteps) {
      // with the forcing (src/main.f90:149) and pass A of the next step
      FAux fx;
      fx.ent = fstep;
      fx.per = (int)per;
      fx.start = c->dfstart;
      fx.idx = c->dfidx;
      LAUNCH(PASS_EF, launch_zfwd_zinv(pr, 2, kp, c->work, c->b, c->b, twp(c, 2), fx, st));
    } else {
      LAUNCH(PASS_E2, launch_zfwd(pr, EPI_STEP2, kp, c->work, c->b, c->b, twp(c, 2), st));
      if (per > 0) LAUNCH(PASS_F, launch_forcing(c->b, c->fs, fstep, (int)(per * c->nmem), c->force_seq, st));
    }
  }
  return GT_OK;
}

// One step from bin into bout (single member, bout != bin): the same kernels
// as enqueue_steps, with the start-of-step state read from bin.  In pruned
// mode pass E stores only the kept modes, so bout must already vanish
// outside them.
static int enqueue_step_to(gt_ctx* c, hipStream_t st, const cx<double>* bin, cx<double>* bout, const FEnt* dforce) {
  const size_t per = entries_per_step(c);
  const KP kp = make_kp(c, false);
  if (int rc = step_stages(c, kp, bin, bout, st)) return rc;
  if (per > 0) LAUNCH(PASS_F, launch_forcing(bout, c->fs, dforce, (int)per, c->force_seq, st));
  return GT_OK;
}

// Host side of the forcing for nsteps steps of every member into buf
// (laid out [step][member][entry]), advancing seeds[member].
static void fill_forcing(const gt_ctx* c, int nsteps, int32_t* seeds, FEnt* buf) {
  const size_t per = entries_per_step(c);
  for (int s = 0; s < nsteps; s++)
    for (int m = 0; m < c->nmem; m++) forcing_entries(c, &seeds[m], per ? buf + per * (c->nmem * s + m) : nullptr, m);
}

// Compute the forcing of nsteps steps on the host and queue its upload to
// dforce on the context stream (advancing seeds).
static int stage_forcing(gt_ctx* c, int nsteps, int32_t* seeds, hipStream_t st = nullptr) {
  const size_t per = entries_per_step(c);
  const size_t need = per * (size_t)c->nmem * (size_t)nsteps;
  if (!st) st = c->stream;
  if (need > 0) {
    if (int rc = ensure_force_cap(c, need)) return rc;
    if (c->hforce_pending) {
      HIPCHK(hipEventSynchronize(c->hforce_done));
      c->hforce_pending = false;
    }
    fill_forcing(c, nsteps, seeds, c->hforce);
    HIPCHK(hipMemcpyAsync(c->dforce, c->hforce, sizeof(FEnt) * need, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(c->hforce_done, st));
    c->hforce_pending = true;
  } else {
    fill_forcing(c, nsteps, seeds, nullptr);
  }
  return GT_OK;
}

static int step_m(gt_ctx* c, int nsteps, int32_t* seeds) {
  if (int rc = check_ctx(c)) return rc;
  if (c->ext_stream) HIPCHK(hipStreamSynchronize(c->ext_stream));
  if (!seeds) return fail(GT_EINVAL, "step: null seed");
  if (nsteps < 0) return fail(GT_EINVAL, "step: negative step count");
  if (nsteps == 0) return GT_OK;
  if (int rc = stage_forcing(c, nsteps, seeds)) return rc;
  return enqueue_steps(c, c->stream, nsteps, c->dforce);
}

int gt_step(gt_ctx* c, int nsteps, int32_t* seed) {
  if (c && c->nmem != 1) return fail(GT_EINVAL, "gt_step: batch contexts step with gt_batch_step");
  return step_m(c, nsteps, seed);
}

int gt_nonlinear(gt_ctx* c, double* nx, double* ny, double* nz) {
  if (int rc = check_ctx(c)) return rc;
  if (c->nmem != 1) return fail(GT_EINVAL, "gt_nonlinear: single-member contexts only");
  if (!nx || !ny || !nz) return fail(GT_EINVAL, "gt_nonlinear: null field");
  const KP kp = make_kp(c, false);
  if (int rc = nonlinear_stage(c, kp, c->b, EPI_NONLIN, c->bh, c->stream)) return rc;
  double* dst[3] = {nx, ny, nz};
  if (int rc = fields_to_host(c, c->bh, dst, c->work, c->stream)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  // bh must vanish outside the kept extents before the next step
  if (int rc = set_mode(c, c->pruned)) return rc;
  return GT_OK;
}

int gt_velocity(gt_ctx* c, double* ux, double* uy, double* uz) {
  if (int rc = check_ctx(c)) return rc;
  if (c->nmem != 1) return fail(GT_EINVAL, "gt_velocity: single-member contexts only");
  if (!ux || !uy || !uz) return fail(GT_EINVAL, "gt_velocity: null field");
  void* w = nullptr;
  double* phys = nullptr;
  const size_t plane = (size_t)c->nx * c->ny * c->nz;
  HIPCHK(hipMalloc(&w, sizeof(cx<double>) * 3 * c->fs));
  if (hipMalloc((void**)&phys, sizeof(double) * 3 * plane) != hipSuccess) {
    hipFree(w);
    return fail(GT_ENOMEM, "gt_velocity: out of device memory");
  }
  int rc = GT_OK;
  const KP kp = make_kp1(c, false, 0);
  hipError_t e = launch_zinv(0, 3, kp, c->b, w, c->tw64[2], c->stream);
  if (e == hipSuccess) e = launch_yfft(0, +1, 3, kp, w, c->tw64[1], c->stream);
  if (e == hipSuccess) e = launch_xc2r(kp, w, phys, c->tw64[0], c->stream);
  double* dst[3] = {ux, uy, uz};
  for (int f = 0; f < 3 && e == hipSuccess; f++)
    e = hipMemcpyAsync(dst[f], phys + f * plane, sizeof(double) * plane, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) rc = fail(GT_EHIP, std::string("gt_velocity: ") + hipGetErrorString(e));
  hipFree(w);
  hipFree(phys);
  return rc;
}

int gt_init_velocity(gt_ctx* c, const double* ux, const double* uy, const double* uz) {
  if (int rc = check_ctx(c)) return rc;
  if (c->nmem != 1) return fail(GT_EINVAL, "gt_init_velocity: single-member contexts only");
  if (!ux || !uy || !uz) return fail(GT_EINVAL, "gt_init_velocity: null field");
  void* w = nullptr;
  double* phys = nullptr;
  const size_t plane = (size_t)c->nx * c->ny * c->nz;
  HIPCHK(hipMalloc(&w, sizeof(cx<double>) * 3 * c->fs));
  if (hipMalloc((void**)&phys, sizeof(double) * 3 * plane) != hipSuccess) {
    hipFree(w);
    return fail(GT_ENOMEM, "gt_init_velocity: out of device memory");
  }
  int rc = GT_OK;
  const double* src[3] = {ux, uy, uz};
  hipError_t e = hipSuccess;
  for (int f = 0; f < 3 && e == hipSuccess; f++)
    e = hipMemcpyAsync(phys + f * plane, src[f], sizeof(double) * plane, hipMemcpyHostToDevice, c->stream);
  const KP kp = make_kp1(c, true, 0);
  if (e == hipSuccess) e = hipMemsetAsync(w, 0, sizeof(cx<double>) * 3 * c->fs, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->b, 0, sizeof(cx<double>) * 3 * c->fs, c->stream);
  if (e == hipSuccess) e = launch_xr2c(kp, phys, w, c->tw64[0], c->stream);
  if (e == hipSuccess) e = launch_yfft(0, -1, 3, kp, w, c->tw64[1], c->stream);
  if (e == hipSuccess) e = launch_zfwd(0, EPI_INIT, kp, w, nullptr, c->b, c->tw64[2], c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) rc = fail(GT_EHIP, std::string("gt_init_velocity: ") + hipGetErrorString(e));
  hipFree(w);
  hipFree(phys);
  if (rc == GT_OK) rc = set_mode(c, c->p.truncate == 1);
  if (rc == GT_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(GT_EHIP, "gt_init_velocity: sync");
  return rc;
}

namespace {
// one diagnostics sweep; host receives sp, sp2d, hel, flux bins and glob
int run_diag(gt_ctx* c, int member, const void* nl, std::vector<double>& host) {
  const int nbin = c->nx / 2;
  if (nbin > 512) return fail(GT_ESIZE, "diagnostics: NX/2 above 512 bins");
  const int len = diag_len(nbin);
  if (!c->ddiag_part) {
    c->diag_blocks = std::max(64, c->resident_blocks);
    HIPCHK(hipMalloc(&c->ddiag_part, sizeof(double) * (size_t)len * c->diag_blocks));
    HIPCHK(hipMalloc(&c->ddiag_out, sizeof(double) * 4 * 512 + sizeof(double) * 18));
  }
  const KP kp = make_kp1(c, false, member);
  HIPCHK(launch_diag(kp, c->b + (size_t)member * 3 * c->fs, nl, c->ddiag_part, c->diag_blocks, c->ddiag_out, nbin,
                     c->stream));
  host.assign(len, 0.0);
  HIPCHK(hipMemcpyAsync(host.data(), c->ddiag_out, sizeof(double) * len, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// Diagnostics without allocation (gt_spectra_noalloc).
//
// The per-block partial sums of the diagnostics sweep and their total live in
// `scratch`, a file-scope device array of fixed size instead of the buffers
// gt_spectra allocates on its first call.  Every GPU the library runs on holds
// its own copy of the array until the process ends.  Host code must not use
// the array's name as a device address: scratch_ptr() takes the address from
// hipGetSymbolAddress once per device (on the entry point's first call with
// that device current) and every launch gets that pointer.  The kernels write
// every partial and every total they later read, so nothing is cleared
// between calls.  Calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
namespace {
constexpr int NA_BLOCKS = 1024;           // partial-sum blocks (gt_spectra uses max(64, 4 CUs) = 1024 on MI355X)
constexpr int NA_LEN = 4 * 512 + 18;      // diag_len of the largest binning (NX/2 = 512)
constexpr size_t NA_DOUBLES = (size_t)NA_BLOCKS * NA_LEN + NA_LEN;  // partials, then the totals
constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
void* na_ptr[NA_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) double scratch[NA_DOUBLES];

namespace {
// The device address of `scratch` on device dev, which must be current.
int scratch_ptr(int dev, double** out) {
  if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(GT_EINVAL, "spectra_noalloc: device id out of range");
  std::lock_guard<std::mutex> lk(na_mutex);
  if (!na_ptr[dev]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev)
      return fail(GT_EINVAL, "spectra_noalloc: make the context's device " + std::to_string(dev) +
                                 " current (device " + std::to_string(cur) + " is)");
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(GT_EHIP, "spectra_noalloc: no device address for the scratch array");
    na_ptr[dev] = p;
  }
  *out = (double*)na_ptr[dev];
  return GT_OK;
}
}  // namespace

extern "C" int gt_spectra_noalloc(gt_ctx* c, double* sp, double* sp2d, double* hel, double* glob) {
  if (int rc = check_ctx(c)) return rc;
  if (c->nmem != 1) return fail(GT_EINVAL, "gt_spectra_noalloc: single-member contexts only");
  const int nbin = c->nx / 2;
  const int len = diag_len(nbin);
  if (nbin > 512 || len > NA_LEN)
    return fail(GT_ESIZE, "gt_spectra_noalloc: NX/2 = " + std::to_string(nbin) + " above the 512 bins of the scratch");
  // the same block count as gt_spectra, so that the sums are added in the same order
  const int nblocks = std::max(64, c->resident_blocks);
  if (nblocks > NA_BLOCKS)
    return fail(GT_ESIZE, "gt_spectra_noalloc: " + std::to_string(nblocks) + " partial-sum blocks above the " +
                              std::to_string(NA_BLOCKS) + " of the scratch");
  double* base = nullptr;
  if (int rc = scratch_ptr(c->dev, &base)) return rc;
  double* partial = base;
  double* total = base + (size_t)NA_BLOCKS * NA_LEN;
  const KP kp = make_kp1(c, false, 0);
  HIPCHK(launch_diag(kp, c->b, nullptr, partial, nblocks, total, nbin, c->stream));
  double h[NA_LEN];
  HIPCHK(hipMemcpyAsync(h, total, sizeof(double) * len, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q = 0; q < nbin; q++) {
    if (sp) sp[q] = h[q];
    if (sp2d) sp2d[q] = h[nbin + q];
    if (hel) hel[q] = h[2 * nbin + q];
  }
  if (glob)
    for (int q = 0; q < 18; q++) glob[q] = h[4 * nbin + q];
  if (gt_check_finite(h + 4 * nbin, 18) != GT_OK)
    return fail(GT_ENONFINITE, "spectra_noalloc: the state is not finite (" + g_err + ")");
  return GT_OK;
}

extern "C" int gt_check_finite(const double* v, int n) {
  if (!v && n > 0) return fail(GT_EINVAL, "gt_check_finite: null values");
  for (int q = 0; q < n; q++)
    if (!std::isfinite(v[q]))
      return fail(GT_ENONFINITE, "non-finite value " + std::to_string(v[q]) + " at index " + std::to_string(q));
  return GT_OK;
}

static int spectra_m(gt_ctx* c, int member, double* sp, double* sp2d, double* hel, double* glob) {
  if (int rc = check_ctx(c)) return rc;
  if (member < 0 || member >= c->nmem) return fail(GT_EINVAL, "spectra: member out of range");
  std::vector<double> h;
  if (int rc = run_diag(c, member, nullptr, h)) return rc;
  const int nbin = c->nx / 2;
  for (int q = 0; q < nbin; q++) {
    if (sp) sp[q] = h[q];
    if (sp2d) sp2d[q] = h[nbin + q];
    if (hel) hel[q] = h[2 * nbin + q];
  }
  if (glob)
    for (int q = 0; q < 18; q++) glob[q] = h[4 * nbin + q];
  // the glob sums cover every mode with k != 0: a NaN or an infinity anywhere
  // in the state makes one of them non-finite
  if (gt_check_finite(h.data() + 4 * nbin, 18) != GT_OK)
    return fail(GT_ENONFINITE, "spectra: the state is not finite (" + g_err + ")");
  return GT_OK;
}

extern "C" int gt_spectra(gt_ctx* c, double* sp, double* sp2d, double* hel, double* glob) {
  return spectra_m(c, 0, sp, sp2d, hel, glob);
}

extern "C" int gt_fluxes(gt_ctx* c, double* spf) {
  if (int rc = check_ctx(c)) return rc;
  if (c->nmem != 1) return fail(GT_EINVAL, "gt_fluxes: single-member contexts only");
  if (!spf) return fail(GT_EINVAL, "gt_fluxes: null output");
  const KP kp = make_kp(c, false);
  if (int rc = nonlinear_stage(c, kp, c->b, EPI_NONLIN, c->bh, c->stream)) return rc;
  std::vector<double> h;
  int rc = run_diag(c, 0, c->bh, h);
  if (int r2 = set_mode(c, c->pruned)) return r2;
  if (rc) return rc;
  const int nbin = c->nx / 2;
  const double* flux = h.data() + 3 * nbin;
  for (int i = 1; i <= nbin; i++) {
    double s = 0.0;
    for (int q = i; q < nbin; q++) s += flux[q];
    spf[i - 1] = s;
  }
  return GT_OK;
}

extern "C" int gt_forcing_table(gt_ctx* c, int* nf, int* ikxf, int* ikyf, int* ikzf, double* ff, double* inpe, double* inpz,
                     int cap) {
  if (int rc = check_ctx(c)) return rc;
  if (!nf) return fail(GT_EINVAL, "gt_forcing_table: null nf");
  *nf = (int)c->forced.size();
  if (inpe) *inpe = c->inpe;
  if (inpz) *inpz = c->inpz;
  const int n = std::min(cap, *nf);
  for (int q = 0; q < n; q++) {
    if (ikxf) ikxf[q] = c->forced[q].i;
    if (ikyf) ikyf[q] = c->forced[q].j;
    if (ikzf) ikzf[q] = c->forced[q].k;
    if (ff) ff[q] = c->forced[q].ff;
  }
  return GT_OK;
}

int gt_forcing_count(const gt_params* p, int* nf, double* inpe, double* inpz) {
  if (!p || !nf) return fail(GT_EINVAL, "gt_forcing_count: null argument");
  if (p->nx < 2 || p->ny < 2 || p->nz < 2 || (p->nx | p->ny | p->nz) & 1)
    return fail(GT_EINVAL, "gt_forcing_count: grid sizes must be even");
  gt_ctx c;
  c.p = *p;
  c.nx = p->nx;
  c.ny = p->ny;
  c.nz = p->nz;
  c.nxh = p->nx / 2 + 1;
  c.nxp = (c.nxh + 15) / 16 * 16;
  build_wavenumbers(&c);
  if (build_forcing(&c)) return fail(GT_ENFORC, "gt_forcing_count: NFORC too small !");
  *nf = (int)c.forced.size();
  if (inpe) *inpe = c.inpe;
  if (inpz) *inpz = c.inpz;
  return GT_OK;
}

int gt_sync(gt_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}

int gt_stream(gt_ctx* c, void** stream) {
  if (int rc = check_ctx(c)) return rc;
  if (!stream) return fail(GT_EINVAL, "gt_stream: null");
  *stream = (void*)c->stream;
  return GT_OK;
}

int gt_pruned(gt_ctx* c, int* pruned) {
  if (int rc = check_ctx(c)) return rc;
  if (!pruned) return fail(GT_EINVAL, "gt_pruned: null");
  *pruned = c->pruned ? 1 : 0;
  return GT_OK;
}

int gt_prof_enable(gt_ctx* c, int on) {
  if (int rc = check_ctx(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& r : c->recs) {
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->recs.clear();
  for (int q = 0; q < NPASS; q++) {
    c->prof_ms[q] = 0;
    c->prof_n[q] = 0;
  }
  c->prof = on != 0;
  return GT_OK;
}

int gt_prof_read(gt_ctx* c, double* ms, int64_t* launches, double* bytes, int npass) {
  if (int rc = check_ctx(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& r : c->recs) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
    c->prof_ms[r.pass] += t;
    c->prof_n[r.pass] += 1;
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->recs.clear();
  // algorithmic bytes of one launch of each pass
  const double cw = c->prec ? 8.0 : 16.0, cb = 16.0;
  const bool pr = c->pruned;
  const int kxc = pr ? c->kxcut : c->nxh - 1;
  auto keepcount = [](int N, int K) -> double { return K < 0 ? 0.0 : (K >= N / 2 ? (double)N : 2.0 * K + 1.0); };
  double colsA = 0, elemsB = 0, keepE = 0, colsY = 0;
  for (int i = 0; i <= kxc; i++) {
    const int Ky = pr ? c->kyext[i] : c->ny / 2;
    colsY += keepcount(c->ny, Ky);
    for (int j = 0; j < c->ny; j++) {
      const int K = pr ? c->kzext[(size_t)j * c->nxp + i] : c->nz / 2;
      if (K >= 0) colsA += 1;
      keepE += keepcount(c->nz, K);
    }
  }
  (void)elemsB;
  const double full = (double)(kxc + 1) * c->ny * c->nz;
  double by[NPASS] = {0};
  by[PASS_A] = keepE * 3 * cb + colsA * c->nz * 6 * cw;  // b once per field (the six outputs re-read it, mostly from L2)
  by[PASS_B] = colsY * c->nz * 6 * cw + full * 6 * cw;
  by[PASS_C] = full * 6 * cw + full * 3 * cw;
  by[PASS_D] = full * 3 * cw + colsY * c->nz * 3 * cw;
  by[PASS_E1] = colsA * c->nz * 3 * cw + keepE * 3 * cb + keepE * 3 * cb;
  by[PASS_E2] = by[PASS_E1];
  by[PASS_EA] = colsA * c->nz * 3 * cw + keepE * 3 * cb + colsA * c->nz * 6 * cw;
  by[PASS_F] = (double)entries_per_step(c) * (sizeof(FEnt) + 6 * cb);
  // the new b stored too, and the step's forcing entries read
  by[PASS_EF] = by[PASS_EA] + keepE * 3 * cb + (double)entries_per_step(c) * sizeof(FEnt);
  // passes B-D launch once per z chunk: the mean bytes of one launch
  if (c->zchunk > 0 && c->zchunk < c->nz) {
    const double nch = (double)((c->nz + c->zchunk - 1) / c->zchunk);
    for (int q : {PASS_B, PASS_C}) by[q] /= nch;
    if (c->zchunk_d) by[PASS_D] /= nch;
  }
  for (int q = 0; q < npass && q < NPASS; q++) {
    if (ms) ms[q] = c->prof_ms[q];
    if (launches) launches[q] = c->prof_n[q];
    if (bytes) bytes[q] = by[q];
  }
  return GT_OK;
}

// ---------------------------------------------------------------------------
// Batches and ensembles
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Launch-shape autotuner
// ---------------------------------------------------------------------------

int gt_set_tiles(gt_ctx* c, int tz, int tze, int ty) {
  if (int rc = check_ctx(c)) return rc;
  auto ok = [](int n, int t) { return t == 0 || t == pencil_T(n) || (alt_T(n) > 0 && t == alt_T(n)); };
  if (!ok(c->nz, tz) || !ok(c->nz, tze) || !ok(c->ny, ty))
    return fail(GT_EINVAL, "gt_set_tiles: no such tile width for this grid");
  c->tz = (tz == pencil_T(c->nz)) ? 0 : tz;
  c->tze = (tze == pencil_T(c->nz)) ? 0 : tze;
  c->ty = (ty == pencil_T(c->ny)) ? 0 : ty;
  c->gen++;
  return GT_OK;
}

int gt_set_zchunk(gt_ctx* c, int planes, int with_d) {
  if (int rc = check_ctx(c)) return rc;
  if (planes < 0) return fail(GT_EINVAL, "gt_set_zchunk: planes must be >= 0");
  c->zchunk = (planes >= c->nz) ? 0 : planes;
  c->zchunk_d = with_d != 0;
  c->gen++;
  return GT_OK;
}

int gt_set_zstreams(gt_ctx* c, int nstreams) {
  if (int rc = check_ctx(c)) return rc;
  if (nstreams != 1 && nstreams != 2) return fail(GT_EINVAL, "gt_set_zstreams: 1 or 2 streams");
  if (c->zstreams != nstreams) c->gen++;
  c->zstreams = nstreams;
  return GT_OK;
}

int gt_get_zstreams(gt_ctx* c, int* nstreams) {
  if (int rc = check_ctx(c)) return rc;
  if (!nstreams) return fail(GT_EINVAL, "gt_get_zstreams: null");
  *nstreams = c->zstreams;
  return GT_OK;
}

int gt_set_fusion(gt_ctx* c, int on) {
  if (int rc = check_ctx(c)) return rc;
  if (c->fuse_ea != (on != 0)) c->gen++;
  c->fuse_ea = on != 0;
  return GT_OK;
}

int gt_get_fusion(gt_ctx* c, int* on, int* active) {
  if (int rc = check_ctx(c)) return rc;
  if (!on) return fail(GT_EINVAL, "gt_get_fusion: null");
  *on = c->fuse_ea ? 1 : 0;
  if (active) *active = (c->fuse_ea && zfwd_zinv_supported(make_kp(c, false))) ? 1 : 0;
  return GT_OK;
}

int gt_get_zchunk(gt_ctx* c, int* planes, int* with_d) {
  if (int rc = check_ctx(c)) return rc;
  if (!planes) return fail(GT_EINVAL, "gt_get_zchunk: null");
  *planes = c->zchunk;
  if (with_d) *with_d = c->zchunk_d ? 1 : 0;
  return GT_OK;
}

int gt_get_tiles(gt_ctx* c, int* tz, int* tze, int* ty, double* ms) {
  if (int rc = check_ctx(c)) return rc;
  if (tz) *tz = c->tz ? c->tz : pencil_T(c->nz);
  if (tze) *tze = c->tze ? c->tze : pencil_T(c->nz);
  if (ty) *ty = c->ty ? c->ty : pencil_T(c->ny);
  if (ms) *ms = c->tune_ms;
  return GT_OK;
}

// Each width acts on its own passes only (tz: pass A, tze: pass E, ty:
// passes B and D), so the tuner times nrep nonlinear evaluations of the
// current state (all members) with every width at each candidate, with an
// event pair around every pass launch, and picks each width from the time
// of its own passes.  The state is left unchanged; the widths do not change
// any result bit.
int gt_autotune(gt_ctx* c, int nrep) {
  if (int rc = check_ctx(c)) return rc;
  if (nrep < 1) return fail(GT_EINVAL, "gt_autotune: nrep must be >= 1");
  if (c->ext_stream) HIPCHK(hipStreamSynchronize(c->ext_stream));
  const std::vector<int> cand = {8, 16};
  const int NC = (int)cand.size();
  const bool prof = c->prof;
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<gt_ctx::Rec> saved;
  saved.swap(c->recs);
  std::vector<double> tA(NC, 1e30), tE(NC, 1e30);
  int rc = GT_OK;
  auto ok = [](int n, int t) { return t == pencil_T(n) || (alt_T(n) > 0 && t == alt_T(n)); };
  auto enc = [](int n, int t) { return t == pencil_T(n) ? 0 : t; };
  // two interleaved rounds over the candidates, the best time of each kept,
  // so that no candidate is measured only while the clocks ramp up
  for (int round = 0; round < 2 && rc == GT_OK; round++)
  for (int ci = 0; ci < NC && rc == GT_OK; ci++) {
    const int w = cand[ci];
    if (!ok(c->nz, w) && !ok(c->ny, w)) continue;  // no pass has this width
    c->tz = ok(c->nz, w) ? enc(c->nz, w) : 0;
    c->tze = ok(c->nz, w) ? enc(c->nz, w) : 0;
    c->ty = ok(c->ny, w) ? enc(c->ny, w) : 0;
    const KP kp = make_kp(c, false);
    c->prof = false;
    rc = nonlinear_stage(c, kp, c->b, EPI_STEP1, c->bh, c->stream);  // warm-up
    c->prof = true;
    for (int r = 0; r < nrep && rc == GT_OK; r++) rc = nonlinear_stage(c, kp, c->b, EPI_STEP1, c->bh, c->stream);
    c->prof = false;
    if (rc == GT_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(GT_EHIP, "autotune sync");
    double ms[NPASS] = {0};
    for (auto& q : c->recs) {
      float t = 0.f;
      if (rc == GT_OK && hipEventElapsedTime(&t, q.a, q.b) != hipSuccess) rc = fail(GT_EHIP, "autotune timing");
      ms[q.pass] += t;
      c->pool.push_back(q.a);
      c->pool.push_back(q.b);
    }
    c->recs.clear();
    if (ok(c->nz, w)) tA[ci] = std::min(tA[ci], ms[PASS_A] / nrep);
    if (ok(c->nz, w)) tE[ci] = std::min(tE[ci], ms[PASS_E1] / nrep);
  }
  c->recs.swap(saved);
  c->prof = prof;
  auto pick = [&](const std::vector<double>& t, int n) {
    int bi = -1;
    for (int i = 0; i < NC; i++)
      if (t[i] < 1e30 && (bi < 0 || t[i] < t[bi])) bi = i;
    return bi < 0 ? 0 : enc(n, cand[bi]);
  };
  c->tz = pick(tA, c->nz);
  c->tze = pick(tE, c->nz);
  // the fused z-forward passes (gt_set_fusion) are compiled at the default
  // width only for some lengths: with fusion on, a tuned width that they
  // lack gives way to the default one
  if (c->fuse_ea && !zfwd_zinv_supported(make_kp(c, false))) {
    const int te = c->tze;
    c->tze = 0;
    if (!zfwd_zinv_supported(make_kp(c, false))) c->tze = te;
  }
  // width of the y passes and z chunk of passes B-D together, by the wall
  // time of whole evaluations: the y passes share the memory-side cache with
  // the x pass chunk by chunk, so their own event times (taken unchunked)
  // picked widths that were slower in the chunked step, and the chunks add
  // launches that per-pass sums would not count
  if (rc == GT_OK) {
    // (y width, planes, pass D in the chunks, streams)
    std::vector<int> ycand;
    for (int w : cand)
      if (ok(c->ny, w)) ycand.push_back(enc(c->ny, w));
    std::vector<std::tuple<int, int, bool, int>> zcand;
    // planes per chunk: NZ/16 .. NZ/2, and 3, 5 and 6 NZ/64 (a ragged last
    // chunk; at 512^3 FP64, 40 planes measured 0.7% faster than 32)
    std::vector<int> planes;
    for (int d : {16, 8, 4, 2})
      if (c->nz % d == 0 && c->nz / d >= 8) planes.push_back(c->nz / d);
    for (int m : {3, 5, 6})
      if ((m * c->nz) % 64 == 0 && m * c->nz / 64 >= 8) planes.push_back(m * c->nz / 64);
    // (one stream: the two-stream chunks of gt_set_zstreams overlap each
    // pass with the other stream's, which per-pass timing cannot separate;
    // the autotuner keeps the configuration whose passes it can account for)
    for (int yw : ycand) {
      zcand.push_back({yw, 0, true, 1});
      for (int pl : planes)
        for (bool wd : {true, false}) zcand.push_back({yw, pl, wd, 1});
    }
    hipEvent_t e0 = pool_get(c), e1 = pool_get(c);
    if (!e0 || !e1) rc = fail(GT_EHIP, "autotune events");
    // two interleaved rounds over the candidates, the best time of each kept
    std::vector<double> tz(zcand.size(), 1e30);
    for (int round = 0; round < 2; round++)
      for (size_t i = 0; i < zcand.size() && rc == GT_OK; i++) {
        c->ty = std::get<0>(zcand[i]);
        c->zchunk = std::get<1>(zcand[i]);
        c->zchunk_d = std::get<2>(zcand[i]);
        c->zstreams = std::get<3>(zcand[i]);
        const KP kp = make_kp(c, false);
        rc = nonlinear_stage(c, kp, c->b, EPI_STEP1, c->bh, c->stream);  // warm-up
        if (rc == GT_OK && hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(GT_EHIP, "autotune record");
        for (int r = 0; r < nrep && rc == GT_OK; r++) rc = nonlinear_stage(c, kp, c->b, EPI_STEP1, c->bh, c->stream);
        float t = 0.f;
        if (rc == GT_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&t, e0, e1) != hipSuccess))
          rc = fail(GT_EHIP, "autotune timing");
        if (rc == GT_OK) tz[i] = std::min(tz[i], (double)t / nrep);
      }
    size_t bi = 0;
    for (size_t i = 1; i < zcand.size(); i++)
      if (tz[i] < tz[bi]) bi = i;
    c->ty = std::get<0>(zcand[bi]);
    c->zchunk = std::get<1>(zcand[bi]);
    c->zchunk_d = std::get<2>(zcand[bi]);
    c->zstreams = std::get<3>(zcand[bi]);
    if (rc == GT_OK) c->tune_ms = tz[bi];  // the tuned time per evaluation
    // fusion of the z passes, where the fused kernels exist: one step's
    // passes with and without it, timed whole in two interleaved rounds
    // (the best of each kept).  The state is only read: the new b goes to
    // bh, and the forcing adds zeros.  Fused: passes B-D, E1 with stage 2's
    // A, B-D, E2 with the next step's A; unfused: A, B-D, E1, A, B-D, E2.
    if (rc == GT_OK && zfwd_zinv_supported(make_kp(c, false))) {
      const KP kp = make_kp(c, false);
      const size_t per = entries_per_step(c);
      if (per > 0) {
        rc = ensure_force_cap(c, per * c->nmem);
        if (rc == GT_OK && hipMemsetAsync(c->dforce, 0, sizeof(FEnt) * per * c->nmem, c->stream) != hipSuccess)
          rc = fail(GT_EHIP, "autotune forcing buffer");
      }
      FAux fx;
      fx.ent = c->dforce;
      fx.per = (int)per;
      fx.start = c->dfstart;
      fx.idx = c->dfidx;
      auto unfused = [&]() -> int {
        if (int r = nonlinear_stage(c, kp, c->b, EPI_STEP1, c->bh, c->stream, c->b)) return r;
        return nonlinear_stage(c, kp, c->bh, EPI_STEP2, c->bh, c->stream, c->b);
      };
      auto fused = [&]() -> int {
        if (int r = nonlinear_mid(c, kp, c->stream)) return r;
        LAUNCH(PASS_EA, launch_zfwd_zinv(c->prec, 1, kp, c->work, c->b, nullptr, twp(c, 2), FAux{}, c->stream));
        if (int r = nonlinear_mid(c, kp, c->stream)) return r;
        LAUNCH(PASS_EF, launch_zfwd_zinv(c->prec, 2, kp, c->work, c->b, c->bh, twp(c, 2), fx, c->stream));
        return GT_OK;
      };
      double tf[2] = {1e30, 1e30};  // unfused, fused
      for (int round = 0; round < 2 && rc == GT_OK; round++)
        for (int f = 0; f < 2 && rc == GT_OK; f++) {
          if (f == 1) rc = launch_zinv(c->prec, 6, kp, c->b, c->work, twp(c, 2), c->stream) == hipSuccess
                               ? GT_OK : fail(GT_EHIP, "autotune pass A");
          if (rc == GT_OK) rc = f ? fused() : unfused();  // warm-up
          if (rc == GT_OK && hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(GT_EHIP, "autotune record");
          for (int r = 0; r < nrep && rc == GT_OK; r++) rc = f ? fused() : unfused();
          float t = 0.f;
          if (rc == GT_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            rc = fail(GT_EHIP, "autotune timing");
          if (rc == GT_OK) tf[f] = std::min(tf[f], (double)t / nrep);
        }
      if (rc == GT_OK) c->fuse_ea = tf[1] < tf[0];
    }
    if (e0) c->pool.push_back(e0);
    if (e1) c->pool.push_back(e1);
  }
  c->gen++;  // widths and z chunk may have changed
  if (rc != GT_OK) return rc;
  return set_mode(c, c->pruned);  // bh vanishes outside the kept extents again
}

int gt_batch_create(const gt_params* p, int nmembers, const double* omegas, gt_ctx** out) {
  return create_ctx(p, nmembers, omegas, out);
}

int gt_batch_size(gt_ctx* c, int* nmembers) {
  if (int rc = check_ctx(c)) return rc;
  if (!nmembers) return fail(GT_EINVAL, "gt_batch_size: null");
  *nmembers = c->nmem;
  return GT_OK;
}

int gt_batch_set_state(gt_ctx* c, int member, const double* bx, const double* by, const double* bz) {
  return set_state_m(c, member, bx, by, bz);
}

int gt_batch_get_state(gt_ctx* c, int member, double* bx, double* by, double* bz) {
  return get_state_m(c, member, bx, by, bz);
}

int gt_batch_step(gt_ctx* c, int nsteps, int32_t* seeds) { return step_m(c, nsteps, seeds); }

int gt_batch_spectra(gt_ctx* c, int member, double* sp, double* sp2d, double* hel, double* glob) {
  return spectra_m(c, member, sp, sp2d, hel, glob);
}

}  // extern "C"

struct gt_ensemble {
  std::vector<gt_ctx*> batches;
  std::vector<std::vector<int32_t>> seeds;  // per batch, per member
  int chunk = 0;
  hipStream_t stream = nullptr;
  // two captured graphs, each reading its own pinned forcing buffers
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<FEnt*> hbuf[2];  // per batch
  // per batch: the device forcing buffer the graphs copy into (the ensemble's
  // own, so that nothing a batch call reallocates is baked into a graph)
  std::vector<FEnt*> dbuf;
  // per batch: the generation of the context when the graphs were captured;
  // a batch whose generation moved on gets the graphs recaptured
  std::vector<unsigned> gen;
  hipEvent_t done[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  int next = 0;
};

namespace {

void free_ensemble(gt_ensemble* e) {
  if (!e) return;
  if (e->stream) hipStreamSynchronize(e->stream);
  for (gt_ctx* c : e->batches)
    if (c && c->ext_stream == e->stream) c->ext_stream = nullptr;
  for (int g = 0; g < 2; g++) {
    if (e->exec[g]) hipGraphExecDestroy(e->exec[g]);
    if (e->graph[g]) hipGraphDestroy(e->graph[g]);
    for (FEnt* h : e->hbuf[g])
      if (h) hipHostFree(h);
    if (e->done[g]) hipEventDestroy(e->done[g]);
  }
  for (FEnt* d : e->dbuf)
    if (d) hipFree(d);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

int ensemble_capture(gt_ensemble* e, int g) {
  if (e->exec[g]) {
    hipGraphExecDestroy(e->exec[g]);
    e->exec[g] = nullptr;
  }
  if (e->graph[g]) {
    hipGraphDestroy(e->graph[g]);
    e->graph[g] = nullptr;
  }
  // no timing events inside a captured graph
  std::vector<char> prof(e->batches.size());
  for (size_t q = 0; q < e->batches.size(); q++) {
    prof[q] = e->batches[q]->prof;
    e->batches[q]->prof = false;
  }
  int rc = GT_OK;
  hipError_t berr = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
  if (berr != hipSuccess) rc = fail(GT_EHIP, std::string("ensemble capture: ") + hipGetErrorString(berr));
  for (size_t q = 0; q < e->batches.size() && rc == GT_OK; q++) {
    gt_ctx* c = e->batches[q];
    const size_t need = entries_per_step(c) * (size_t)c->nmem * (size_t)e->chunk;
    if (need > 0 && hipMemcpyAsync(e->dbuf[q], e->hbuf[g][q], sizeof(FEnt) * need, hipMemcpyHostToDevice,
                                   e->stream) != hipSuccess)
      rc = fail(GT_EHIP, "ensemble capture: forcing copy");
    if (rc == GT_OK) rc = enqueue_steps(c, e->stream, e->chunk, e->dbuf[q]);
  }
  for (size_t q = 0; q < e->batches.size(); q++) e->batches[q]->prof = prof[q] != 0;
  if (berr != hipSuccess) return rc;
  hipGraph_t graph = nullptr;
  hipError_t err = hipStreamEndCapture(e->stream, &graph);
  if (rc != GT_OK) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (err != hipSuccess) return fail(GT_EHIP, std::string("ensemble capture: ") + hipGetErrorString(err));
  e->graph[g] = graph;
  HIPCHK(hipGraphInstantiate(&e->exec[g], graph, nullptr, nullptr, 0));
  return GT_OK;
}

// Recapture both graphs when a batch changed what they bake in (its mode,
// tile widths or z chunk) since the last capture.
int ensemble_refresh(gt_ensemble* e) {
  bool stale = false;
  for (size_t q = 0; q < e->batches.size(); q++) stale = stale || e->batches[q]->gen != e->gen[q];
  if (!stale) return GT_OK;
  HIPCHK(hipStreamSynchronize(e->stream));  // no replay of the old graphs is in flight
  for (int g = 0; g < 2; g++)
    if (int rc = ensemble_capture(e, g)) return rc;
  for (size_t q = 0; q < e->batches.size(); q++) e->gen[q] = e->batches[q]->gen;
  return GT_OK;
}

}  // namespace

extern "C" {

int gt_ensemble_create(gt_ctx** batches, int nbatches, int chunk, gt_ensemble** out) {
  if (!batches || nbatches < 1 || !out) return fail(GT_EINVAL, "gt_ensemble_create: bad arguments");
  if (chunk < 1) return fail(GT_EINVAL, "gt_ensemble_create: chunk must be >= 1");
  *out = nullptr;
  gt_ensemble* e = new (std::nothrow) gt_ensemble();
  if (!e) return fail(GT_ENOMEM, "gt_ensemble_create: out of host memory");
  e->chunk = chunk;
  int rc = GT_OK;
  for (int q = 0; q < nbatches; q++) {
    if (!batches[q]) rc = fail(GT_EINVAL, "gt_ensemble_create: null batch");
    e->batches.push_back(batches[q]);
    e->seeds.emplace_back(batches[q] ? batches[q]->nmem : 0, 0);
  }
  for (int q = 0; q < nbatches && rc == GT_OK; q++)
    for (int r = 0; r < q; r++)
      if (batches[q] == batches[r]) rc = fail(GT_EINVAL, "gt_ensemble_create: a batch is listed twice");
  for (int q = 0; q < nbatches && rc == GT_OK; q++)
    if (batches[q]->ext_stream) rc = fail(GT_EINVAL, "gt_ensemble_create: a batch already belongs to an ensemble");
  if (rc == GT_OK && hipStreamCreate(&e->stream) != hipSuccess) rc = fail(GT_EHIP, "gt_ensemble_create: stream");
  for (gt_ctx* c : e->batches) {
    if (rc != GT_OK) break;
    const size_t need = entries_per_step(c) * (size_t)c->nmem * (size_t)chunk;
    FEnt* d = nullptr;
    if (need > 0 && hipMalloc((void**)&d, sizeof(FEnt) * need) != hipSuccess)
      rc = fail(GT_ENOMEM, "gt_ensemble_create: device forcing buffer");
    e->dbuf.push_back(d);
  }
  for (int g = 0; g < 2 && rc == GT_OK; g++) {
    if (hipEventCreateWithFlags(&e->done[g], hipEventDisableTiming) != hipSuccess)
      rc = fail(GT_EHIP, "gt_ensemble_create: event");
    for (gt_ctx* c : e->batches) {
      if (rc != GT_OK) break;
      const size_t need = entries_per_step(c) * (size_t)c->nmem * (size_t)chunk;
      FEnt* h = nullptr;
      if (need > 0 && hipHostMalloc((void**)&h, sizeof(FEnt) * need, hipHostMallocDefault) != hipSuccess)
        rc = fail(GT_ENOMEM, "gt_ensemble_create: pinned forcing buffer");
      e->hbuf[g].push_back(h);
    }
  }
  for (gt_ctx* c : e->batches) {
    if (rc == GT_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(GT_EHIP, "gt_ensemble_create: sync");
    if (rc == GT_OK) c->ext_stream = e->stream;
    e->gen.push_back(c ? c->gen : 0);
  }
  for (int g = 0; g < 2 && rc == GT_OK; g++) rc = ensemble_capture(e, g);
  if (rc != GT_OK) {
    free_ensemble(e);
    return rc;
  }
  *out = e;
  return GT_OK;
}

int gt_ensemble_destroy(gt_ensemble* e) {
  free_ensemble(e);
  return GT_OK;
}

int gt_ensemble_set_seeds(gt_ensemble* e, int batch, const int32_t* seeds) {
  if (!e || batch < 0 || batch >= (int)e->batches.size() || !seeds)
    return fail(GT_EINVAL, "gt_ensemble_set_seeds: bad arguments");
  for (int m = 0; m < e->batches[batch]->nmem; m++) e->seeds[batch][m] = seeds[m];
  return GT_OK;
}

int gt_ensemble_get_seeds(gt_ensemble* e, int batch, int32_t* seeds) {
  if (!e || batch < 0 || batch >= (int)e->batches.size() || !seeds)
    return fail(GT_EINVAL, "gt_ensemble_get_seeds: bad arguments");
  for (int m = 0; m < e->batches[batch]->nmem; m++) seeds[m] = e->seeds[batch][m];
  return GT_OK;
}

// Advance every member of every batch by nsteps: whole chunks replay the
// captured graphs (alternating, so the host fills one pinned buffer while the
// other graph runs); a remainder runs eagerly on the ensemble stream.
int gt_ensemble_step(gt_ensemble* e, int nsteps) {
  if (!e) return fail(GT_EINVAL, "gt_ensemble_step: null ensemble");
  if (nsteps < 0) return fail(GT_EINVAL, "gt_ensemble_step: negative step count");
  // work a batch call queued on its own stream (a set_state copy, a
  // diagnostic) touches the same device fields: it finishes first
  for (gt_ctx* c : e->batches) HIPCHK(hipStreamSynchronize(c->stream));
  if (int rc = ensemble_refresh(e)) return rc;
  int left = nsteps;
  while (left >= e->chunk) {
    const int g = e->next;
    if (e->used[g]) HIPCHK(hipEventSynchronize(e->done[g]));
    for (size_t q = 0; q < e->batches.size(); q++)
      fill_forcing(e->batches[q], e->chunk, e->seeds[q].data(), e->hbuf[g][q]);
    HIPCHK(hipGraphLaunch(e->exec[g], e->stream));
    HIPCHK(hipEventRecord(e->done[g], e->stream));
    e->used[g] = true;
    e->next ^= 1;
    left -= e->chunk;
  }
  if (left > 0) {
    // the device forcing buffers are shared with the graphs, and the pinned
    // buffers of graph `next` are free once the stream is idle
    HIPCHK(hipStreamSynchronize(e->stream));
    const int g = e->next;
    for (size_t q = 0; q < e->batches.size(); q++) {
      gt_ctx* c = e->batches[q];
      const size_t need = entries_per_step(c) * (size_t)c->nmem * (size_t)left;
      if (need > 0) {
        fill_forcing(c, left, e->seeds[q].data(), e->hbuf[g][q]);
        HIPCHK(hipMemcpyAsync(e->dbuf[q], e->hbuf[g][q], sizeof(FEnt) * need, hipMemcpyHostToDevice, e->stream));
      } else {
        fill_forcing(c, left, e->seeds[q].data(), nullptr);
      }
      if (int rc = enqueue_steps(c, e->stream, left, e->dbuf[q])) return rc;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  return GT_OK;
}

int gt_ensemble_sync(gt_ensemble* e) {
  if (!e) return fail(GT_EINVAL, "gt_ensemble_sync: null ensemble");
  HIPCHK(hipStreamSynchronize(e->stream));
  return GT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Forcing shells and the adjoint of the step (FP64, single member).
//
// Forward: nsteps steps from the current state, the state and the forcing
// seed kept every `every` steps.  Backward: for each segment, last first,
// the states of the segment are recomputed from its checkpoint, then the
// cotangent is carried back through the steps one at a time:
//   dJ/dalpha_s += <lam_{n+1}, unit forcing increments of shell s at step n>
//   lam_n = (d b_{n+1} / d b_n)^T lam_{n+1}
// ---------------------------------------------------------------------------
namespace {

int adj_malloc(void** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? GT_ENOMEM : GT_EHIP,
                std::string("adjoint: ") + what + ": " + hipGetErrorString(e));
  }
  return GT_OK;
}

int adj_reserve(gt_ctx* c, size_t nckpt, size_t nseg_states, size_t nunit) {
  gt_ctx::Adjoint& A = c->adj;
  const size_t st = sizeof(cx<double>) * 3 * c->fs;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (A.rstream) HIPCHK(hipStreamSynchronize(A.rstream));  // nothing of an earlier backward in flight
  if (!A.work) {
    if (int rc = adj_malloc((void**)&A.work, sizeof(cx<double>) * 9 * c->fs, "work fields")) return rc;
    if (int rc = adj_malloc((void**)&A.vec, 7 * st, "cotangent fields")) return rc;
  }
  if (!A.rstream) {
    // default priority, like the context stream
    if (hipStreamCreate(&A.rstream) != hipSuccess) return fail(GT_EHIP, "adjoint: recompute stream");
    for (auto& e : A.ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(GT_EHIP, "adjoint: events");
  }
  if (nckpt > A.cap_ckpt) {
    hipFree(A.ckpt);
    A.cap_ckpt = 0;
    if (int rc = adj_malloc((void**)&A.ckpt, nckpt * st, "checkpoints")) return rc;
    A.cap_ckpt = nckpt;
  }
  if (nseg_states > A.cap_seg) {
    hipFree(A.seg);
    A.cap_seg = 0;
    if (int rc = adj_malloc((void**)&A.seg, nseg_states * st, "segment states")) return rc;
    A.cap_seg = nseg_states;
  }
  if (nunit > A.cap_unit) {
    hipFree(A.dunit);
    hipFree(A.ddots);
    A.ddots = nullptr;
    A.cap_unit = 0;
    if (int rc = adj_malloc((void**)&A.dunit, nunit * sizeof(FEnt), "forcing entries")) return rc;
    if (int rc = adj_malloc((void**)&A.ddots, nunit * sizeof(double), "forcing dots")) return rc;
    A.cap_unit = nunit;
  }
  return GT_OK;
}

// The passes of one adjoint nonlinear evaluation after pass A: work fields
// 0-5 hold pass A of the linearisation point; x is the cotangent (ze: nu is
// Z emk x) scaled by s.  post == nullptr: out = dN/da^T (s nu); else out =
// the step cotangent formed by k_zfwd_adj from post = lam and that g.
int adjoint_nlt_tail(gt_ctx* c, const KP& kp, const cx<double>* x, double s, bool ze, cx<double>* out,
                     const cx<double>* post) {
  gt_ctx::Adjoint& A = c->adj;
  hipStream_t st = c->stream;
  HIPCHK(launch_zinv_adj(kp, x, s, ze, A.work + 6 * c->fs, c->tw64[2], st));
  HIPCHK(launch_yfft(0, +1, 9, kp, A.work, c->tw64[1], st));
  HIPCHK(launch_xadj(kp, A.work, c->tw64[0], st));
  HIPCHK(launch_yfft(0, -1, 6, kp, A.work, c->tw64[1], st));
  HIPCHK(launch_zfwd_adj(kp, A.work, out, c->tw64[2], post, st));
  return GT_OK;
}

// lam_out = (d b_{n+1} / d b_n)^T lam_in, b_n = bn (the forcing does not
// depend on b).  step2: b+ = Z emk (emk b + dt N(bh; b)) (src/rkstep.f90:86-104),
// stage 1: bh = Z emk (b + dt2 N(b; b)) (src/rkstep.f90:30-48).
int step_adjoint(gt_ctx* c, const KP& kp, const cx<double>* bn, const cx<double>* lam_in, cx<double>* lam_out) {
  gt_ctx::Adjoint& A = c->adj;
  hipStream_t st = c->stream;
  const long long f3 = 3 * c->fs;
  if (c->dt2 == 0.0) {
    // reference mode, fused: pass A of bh = Z emk bn, nu = Z emk lam on load,
    // the whole step cotangent in the last z pass
    HIPCHK(launch_zinv_emk(kp, bn, A.work, c->tw64[2], st));
    return adjoint_nlt_tail(c, kp, lam_in, c->p.dt, true, lam_out, lam_in);
  }
  cx<double>* nu = A.vec + 2 * f3;
  cx<double>* g = A.vec + 3 * f3;
  cx<double>* bh = A.vec + 4 * f3;
  cx<double>* h = A.vec + 5 * f3;
  cx<double>* g1 = A.vec + 6 * f3;
  HIPCHK(launch_adj_ze(kp, lam_in, nu, st));
  if (int rc = nonlinear_stage(c, kp, bn, EPI_STEP1, bh, st, bn)) return rc;
  HIPCHK(launch_zinv(0, 6, kp, bh, A.work, c->tw64[2], st));
  if (int rc = adjoint_nlt_tail(c, kp, nu, c->p.dt, false, g, nullptr)) return rc;
  HIPCHK(launch_adj_post(kp, nu, g, lam_out, st));
  HIPCHK(launch_adj_ze(kp, g, h, st));
  HIPCHK(launch_zinv(0, 6, kp, bn, A.work, c->tw64[2], st));
  if (int rc = adjoint_nlt_tail(c, kp, h, c->dt2, false, g1, nullptr)) return rc;
  HIPCHK(launch_adj_stage1(kp, h, g1, lam_out, st));
  return GT_OK;
}

int check_adjoint_ctx(gt_ctx* c, const char* what) {
  if (int rc = check_ctx(c)) return rc;
  if (c->nmem != 1) return fail(GT_EINVAL, std::string(what) + ": single-member contexts only");
  if (c->prec != GT_PREC_FP64) return fail(GT_EINVAL, std::string(what) + ": FP64 contexts only");
  if (!adjoint_nx_supported(c->nx))
    return fail(GT_ESIZE, std::string(what) + ": NX = " + std::to_string(c->nx) + " not compiled (NX <= 512)");
  return GT_OK;
}

}  // namespace

extern "C" {

int gt_forcing_shells(gt_ctx* c, int* nshells, double* kk, int cap) {
  if (int rc = check_ctx(c)) return rc;
  if (!nshells) return fail(GT_EINVAL, "gt_forcing_shells: null count");
  *nshells = (int)c->shell_k2.size();
  if (kk)
    for (int s = 0; s < std::min(cap, *nshells); s++) kk[s] = std::sqrt(c->shell_k2[s]);
  return GT_OK;
}

int gt_set_shell_amps(gt_ctx* c, const double* amps, int nshells) {
  if (int rc = check_ctx(c)) return rc;
  if (!amps) {
    c->mode_amp.clear();
  } else {
    if (nshells != (int)c->shell_k2.size())
      return fail(GT_EINVAL, "gt_set_shell_amps: expected " + std::to_string(c->shell_k2.size()) + " shells");
    c->mode_amp.resize(c->forced.size());
    for (size_t im = 0; im < c->forced.size(); im++) c->mode_amp[im] = amps[c->mode_shell[im]];
  }
  c->adj.ready = false;
  return GT_OK;
}

int gt_adjoint_forward(gt_ctx* c, int nsteps, int32_t* seed, int every) {
  if (int rc = check_adjoint_ctx(c, "gt_adjoint_forward")) return rc;
  if (!seed) return fail(GT_EINVAL, "gt_adjoint_forward: null seed");
  if (nsteps < 1 || every < 1) return fail(GT_EINVAL, "gt_adjoint_forward: need nsteps >= 1 and every >= 1");
  gt_ctx::Adjoint& A = c->adj;
  A.ready = false;
  const int K = std::min(every, nsteps);
  const int nseg = (nsteps + K - 1) / K;
  const size_t per = entries_per_step(c);
  // the segment buffer at the backward's size (two halves of K states)
  if (int rc = adj_reserve(c, (size_t)nseg + 1, 2 * (size_t)K, std::max<size_t>(per * K, 1))) return rc;
  A.last_kept = false;
  const long long f3 = 3 * c->fs;
  const size_t st = sizeof(cx<double>) * f3;
  A.seeds.assign(nseg, 0);
  for (int s = 0; s < nseg; s++) {
    HIPCHK(hipMemcpyAsync(A.ckpt + (size_t)s * f3, c->b, st, hipMemcpyDeviceToDevice, c->stream));
    A.seeds[s] = *seed;
    const int len = std::min(K, nsteps - s * K);
    if (s + 1 < nseg) {
      if (int rc = step_m(c, len, seed)) return rc;
      continue;
    }
    // The last segment, the first the backward sweeps, runs out of place
    // through its slots of the segment buffer (state i of the segment in
    // slot i, as the backward's recompute leaves them), so the backward
    // starts at once instead of recomputing it first.  Pass E in pruned
    // mode writes only the kept modes: the slots are cleared first.
    if (c->ext_stream) HIPCHK(hipStreamSynchronize(c->ext_stream));
    cx<double>* half = A.seg + (size_t)(s & 1) * K * f3;
    if (len > 1) HIPCHK(hipMemsetAsync(half + f3, 0, (size_t)(len - 1) * st, c->stream));
    if (int rc = stage_forcing(c, len, seed)) return rc;
    for (int i = 0; i < len; i++) {
      const cx<double>* in = i == 0 ? A.ckpt + (size_t)s * f3 : half + (size_t)i * f3;
      cx<double>* out = i + 1 < len ? half + (size_t)(i + 1) * f3 : c->b;
      if (int rc = enqueue_step_to(c, c->stream, in, out, c->dforce + per * i)) return rc;
    }
  }
  HIPCHK(hipMemcpyAsync(A.ckpt + (size_t)nseg * f3, c->b, st, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  A.nsteps = nsteps;
  A.every = K;
  A.nseg = nseg;
  A.ready = true;
  A.last_kept = true;
  return GT_OK;
}

int gt_adjoint_backward(gt_ctx* c, const double* lx, const double* ly, const double* lz, double* gx, double* gy,
                        double* gz, double* grad_amps, int nshells) {
  if (int rc = check_adjoint_ctx(c, "gt_adjoint_backward")) return rc;
  gt_ctx::Adjoint& A = c->adj;
  if (!A.ready) return fail(GT_EINVAL, "gt_adjoint_backward: no forward run to differentiate (gt_adjoint_forward)");
  if (!lx || !ly || !lz || !gx || !gy || !gz) return fail(GT_EINVAL, "gt_adjoint_backward: null field");
  if (grad_amps && nshells != (int)c->shell_k2.size())
    return fail(GT_EINVAL, "gt_adjoint_backward: expected " + std::to_string(c->shell_k2.size()) + " shells");
  const KP kp = make_kp1(c, false, 0);
  const long long f3 = 3 * c->fs;
  const size_t sbytes = sizeof(cx<double>) * f3;
  const size_t per = entries_per_step(c);
  const int K = A.every, nseg = A.nseg, nsteps = A.nsteps;
  // two buffers of segment states (the recompute of one segment fills one
  // while the sweep of the next reads the other), and the unit increments
  // and forcing dots of every step
  // the half of the segment buffer holding the forward's last segment (-1:
  // none); a call that fails leaves it unset, so the next one recomputes
  const cx<double>* seg0 = A.seg;
  int kept = A.last_kept ? ((nseg - 1) & 1) : -1;
  A.last_kept = false;
  if (int rc = adj_reserve(c, (size_t)nseg + 1, 2 * (size_t)K, std::max<size_t>(per * nsteps, 1))) return rc;
  if (A.seg != seg0) kept = -1;
  if (per > 0)
    if (int rc = ensure_force_cap(c, per * K)) return rc;
  hipStream_t st = c->stream;
  // In the reference stage-1 mode the sweep of segment s (context stream)
  // and the recompute of segment s - 1 (a second stream) touch disjoint
  // buffers, so they run side by side.  The midpoint sweep evaluates the
  // stage-1 term through the context's work fields, which the recompute
  // uses too: there both run on the context stream.
  const bool overlap = c->dt2 == 0.0;
  hipStream_t rs = overlap ? A.rstream : st;
  const bool prof = c->prof;
  c->prof = false;  // its events go to the context stream
  struct ProfRestore {
    gt_ctx* c;
    bool on;
    ~ProfRestore() { c->prof = on; }
  } prof_restore{c, prof};
  HIPCHK(hipMemsetAsync(A.work, 0, sizeof(cx<double>) * 9 * c->fs, st));
  HIPCHK(hipMemsetAsync(A.vec, 0, 7 * sbytes, st));
  // the recomputed states are stored by pass E, which in pruned mode writes
  // only the kept modes: the rest of every slot must be zero
  for (int h = 0; h < 2; h++)
    if (h != kept) HIPCHK(hipMemsetAsync(A.seg + (size_t)h * K * f3, 0, (size_t)K * sbytes, st));
  // unit-amplitude increments of every step, from each segment's seed
  std::vector<FEnt> hunit(std::max<size_t>(per * nsteps, 1));
  for (int s = 0; s < nseg; s++) {
    int32_t sd = A.seeds[s];
    const int len = std::min(K, nsteps - s * K);
    for (int i = 0; i < len; i++)
      forcing_entries(c, &sd, per ? hunit.data() + per * ((size_t)s * K + i) : nullptr, 0, true);
  }
  if (per) HIPCHK(hipMemcpyAsync(A.dunit, hunit.data(), sizeof(FEnt) * per * nsteps, hipMemcpyHostToDevice, st));
  cx<double>* lam = A.vec;  // the cotangent of the current state
  cx<double>* lam_next = A.vec + f3;
  const double* src[3] = {lx, ly, lz};
  if (int rc = fields_to_device(c, src, lam, A.work, st)) return rc;
  if (kp.pruned) HIPCHK(launch_adj_project(kp, lam, st));
  if (overlap) {
    HIPCHK(hipEventRecord(A.ev[2], st));  // the cleared slots
    HIPCHK(hipStreamWaitEvent(rs, A.ev[2], 0));
  }
  // state i of segment s: the checkpoint, or slot i of the segment's buffer
  auto slot = [&](int s, int i) -> cx<double>* {
    return i == 0 ? A.ckpt + (size_t)s * f3 : A.seg + ((size_t)(s & 1) * K + i) * f3;
  };
  auto recompute = [&](int s) -> int {
    const int len = std::min(K, nsteps - s * K);
    int32_t sr = A.seeds[s];
    if (len > 1)
      if (int rc = stage_forcing(c, len - 1, &sr, rs)) return rc;
    for (int i = 0; i + 1 < len; i++)
      if (int rc = enqueue_step_to(c, rs, slot(s, i), slot(s, i + 1), c->dforce + per * i)) return rc;
    HIPCHK(hipEventRecord(A.ev[s & 1], rs));
    return GT_OK;
  };
  if (kept < 0) {
    if (int rc = recompute(nseg - 1)) return rc;
  } else if (overlap) {
    HIPCHK(hipEventRecord(A.ev[(nseg - 1) & 1], rs));
  }
  for (int s = nseg - 1; s >= 0; s--) {
    const int len = std::min(K, nsteps - s * K);
    if (overlap) HIPCHK(hipStreamWaitEvent(st, A.ev[s & 1], 0));
    if (s > 0) {
      // segment s - 1 refills the buffer that the sweep of segment s + 1 read
      if (overlap && s + 1 < nseg) HIPCHK(hipStreamWaitEvent(rs, A.ev[2 + ((s + 1) & 1)], 0));
      if (int rc = recompute(s - 1)) return rc;
    }
    for (int i = len - 1; i >= 0; i--) {
      const size_t n = (size_t)s * K + i;
      if (per) HIPCHK(launch_adj_fdot(lam, c->fs, A.dunit + per * n, (int)per, A.ddots + per * n, st));
      if (int rc = step_adjoint(c, kp, slot(s, i), lam, lam_next)) return rc;
      std::swap(lam, lam_next);
    }
    if (overlap) HIPCHK(hipEventRecord(A.ev[2 + (s & 1)], st));
  }
  std::vector<double> hdots(std::max<size_t>(per * nsteps, 1));
  if (per) HIPCHK(hipMemcpyAsync(hdots.data(), A.ddots, sizeof(double) * per * nsteps, hipMemcpyDeviceToHost, st));
  double* dst[3] = {gx, gy, gz};
  if (int rc = fields_to_host(c, lam, dst, A.work, st)) return rc;
  // leave the context at the forward run's final state
  HIPCHK(hipMemcpyAsync(c->b, A.ckpt + (size_t)nseg * f3, sbytes, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  // the recompute of segments 0 .. nseg - 2 refilled the other half only
  // when nseg <= 2, so then the last segment's states are still in place
  A.last_kept = nseg <= 2;
  std::vector<double> gsh(c->shell_k2.size(), 0.0);
  for (int n = 0; n < nsteps; n++)
    for (size_t e = 0; e < per; e++) gsh[c->entry_shell[e]] += hdots[per * n + e];
  if (grad_amps)
    for (size_t q = 0; q < gsh.size(); q++) grad_amps[q] = gsh[q];
  return GT_OK;
}

}  // extern "C"
