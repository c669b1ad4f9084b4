/*
 * gturbo3d.h -- C ABI of the MI355X pseudo-spectral step library
 * (libgturbo3d.so, built from gturbo3d_amd/csrc/).
 *
 * It replaces Gturbo3D's device step: the time-loop body of src/main.f90:124-150
 * (bnl <- b; nlt; step1; nlt; step2; forcing) and the routines it calls,
 * nlt (src/nlt.f90:1), step1/step2 (src/rkstep.f90:1,57), forcing
 * (src/forcing_3d3c.f90:11, src/forcing_2d2c.f90:11), fft_dir/fft_inv
 * (src/fft_dir.f90:1, src/fft_inv.f90:1) and the trunc/curl/invcurl/cross
 * point operations.  The reference passes module globals and device arrays
 * by implicit interface; here everything is an argument or lives in an
 * opaque context, and every array crossing the ABI is a host array in the
 * reference layout:
 *
 *   spectral field: complex128 (NX/2+1, NY, NZ) column-major = C order
 *                   [NZ][NY][NX/2+1] of interleaved (re, im) doubles
 *                   (src/main.f90:14);
 *   physical field: real [NZ][NY][NX] doubles (the unpadded part of the
 *                   in-place array of src/fft_inv.f90:9).
 *
 * Every function returns 0 on success and a nonzero GT_E* code otherwise;
 * gt_last_error() returns a thread-local message for the last failure.  No
 * C++ exception or abort crosses the ABI.  Calls on one context must be
 * serialised by the caller; distinct contexts are independent.  The
 * library owns all device memory; the caller owns host buffers and nothing
 * is retained across calls.
 */
#ifndef GTURBO3D_H
#define GTURBO3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_OK 0
#define GT_EINVAL 1   /* bad argument */
#define GT_EHIP 2     /* HIP runtime error */
#define GT_ENOMEM 3   /* device or host allocation failed */
#define GT_ENFORC 4   /* forced-mode table overflow ('NFORC too small !', src/iniforcing_3d3c.f90:33) */
#define GT_ESIZE 5    /* grid size not supported by the compiled kernels */
#define GT_ENONFINITE 6 /* the state holds a NaN or an infinity (its energy sums are not finite) */

#define GT_FORCING_3D3C 3
#define GT_FORCING_2D2C 2
#define GT_DT2_REFERENCE 0 /* dt2 never assigned (src/modules.f90:32): bh = emk*b */
#define GT_DT2_MIDPOINT 1  /* dt2 = dt/2 */
#define GT_PREC_FP64 0
#define GT_PREC_FP32_BULK 1

/* Run parameters: params.dat (src/inieuler.f90:14-20) plus the grid
 * (src/modules.f90:4), the forcing variant (Makefile:20-33), the stage-1
 * coefficient mode and the precision. */
typedef struct gt_params {
  int nx, ny, nz;
  double xlx, xly, xlz;
  double nu, alpnu, mu, alpmu;
  double dt, omega;
  double rtrunc; /* unsquared, as in params.dat */
  int truncate;
  double famp, k1f, k2f;
  int variant;
  int dt2_mode;
  int precision;
} gt_params;

typedef struct gt_ctx gt_ctx;

/* Context with the tables of defk (src/defk.f90:1-94) and iniforcing
 * (src/iniforcing_3d3c.f90:4-195 / src/iniforcing_2d2c.f90:4-88), the state
 * b and its work arrays (src/main.f90:14, :71-85). */
int gt_create(const gt_params* p, gt_ctx** out);
int gt_destroy(gt_ctx* c);

/* Copy the state b (3 spectral fields) in / out (src/main.f90:85, :202). */
int gt_set_state(gt_ctx* c, const double* bx, const double* by, const double* bz);
int gt_get_state(gt_ctx* c, double* bx, double* by, double* bz);

/* nsteps iterations of the step body of src/main.f90:124-150.  *seed is the
 * forcing LCG state (src/main.f90:42-45, src/forcing_3d3c.f90:82-93) on
 * entry and is advanced on return.  Returns after enqueueing; results are
 * visible to gt_get_state / gt_sync. */
int gt_step(gt_ctx* c, int nsteps, int32_t* seed);

/* N(b; b) = nlt(b, u, b) (src/nlt.f90:1-95) of the current state. */
int gt_nonlinear(gt_ctx* c, double* nx, double* ny, double* nz);

/* Physical velocity u = fft_inv(curl b) (src/main.f90:190-201). */
int gt_velocity(gt_ctx* c, double* ux, double* uy, double* uz);

/* Set b from a physical velocity: fft_dir x3, incompr, invcurl, trunc
 * (src/main.f90:91-106). */
int gt_init_velocity(gt_ctx* c, const double* ux, const double* uy, const double* uz);

/* Diagnostics of the current state on the device (src/diagnostics.f90).
 * sp, sp2d, hel: NX/2 doubles each; bin ib-1 holds the shell ib = int(|k|)+1
 *   (the spectrum of src/diagnostics.f90:22-48: sum fac |k|^2 |b|^2 over the
 *   shell, and the kz = 0 plane's |k|^2 |bz|^2; hel is sum fac Re(u*.w) with
 *   u = curl b, w = curl u).  glob: 18 doubles, the nine sums of glob
 *   (src/diagnostics.f90:126-153) for u followed by the nine for w
 *   (:164-191).  Any output pointer may be NULL.
 * Returns GT_ENONFINITE, with every output written, when one of the 18 glob
 * sums is not finite: they cover every mode with k != 0, so the state holds
 * a NaN or an infinity.  (The reference has no such check; it is the
 * divergence flag of a run, at the cadence the caller asks for diagnostics.) */
int gt_spectra(gt_ctx* c, double* sp, double* sp2d, double* hel, double* glob);

/* gt_spectra without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory, such as tight serving loops or
 * code that must not call hipMalloc.  The per-block partial sums live in a
 * fixed device array of the library (16.9 MB, held on every GPU the library
 * runs on until the process ends) instead of buffers gt_spectra allocates on
 * its first call; the outputs are those of gt_spectra, bit for bit on a GPU
 * with at most 256 compute units.  Single-member contexts; NX/2 <= 512 bins
 * and at most 1024 partial-sum blocks (4 per compute unit), else GT_ESIZE.
 * The context's device must be current.  Calls that overlap on one GPU
 * share the array and race: serialise them. */
int gt_spectra_noalloc(gt_ctx* c, double* sp, double* sp2d, double* hel, double* glob);

/* Host only (no device work): GT_OK when the n values are all finite, else
 * GT_ENONFINITE with gt_last_error() naming the first one that is not.  It is
 * the test gt_spectra applies to its glob sums. */
int gt_check_finite(const double* v, int n);

/* Energy flux spectrum of fluxes (src/diagnostics.f90:62-103): spf[i-1] =
 * sum over shells q >= i of fac*|k|^2 Re(b.conj N), N = nlt(b,u,b), for
 * i = 1..NX/2. */
int gt_fluxes(gt_ctx* c, double* spf);

/* The forced-mode table (1-based indices as in the reference) and the mean
 * energy / enstrophy input inpe, inpz (src/iniforcing_3d3c.f90:37-39). */
int gt_forcing_table(gt_ctx* c, int* nf, int* ikxf, int* ikyf, int* ikzf, double* ff, double* inpe,
                     double* inpz, int cap);

/* Host only (no device work): the number of forced modes nf and the mean
 * energy / enstrophy input that gt_create would build for *p (used to pick
 * famp for a given injection rate, inpe = nf*famp^2). */
int gt_forcing_count(const gt_params* p, int* nf, double* inpe, double* inpz);

/* Wait for all work on the context's stream. */
int gt_sync(gt_ctx* c);

/* The context's HIP stream (hipStream_t), for callers that order their own
 * work after the step. */
int gt_stream(gt_ctx* c, void** stream);

/* 1 if the state is known to vanish outside the truncation ellipsoid and
 * the step moves only the alive modes (pruned), 0 for the dense path. */
int gt_pruned(gt_ctx* c, int* pruned);

/* Per-pass timing with HIP events on the context stream.  Pass ids:
 * 0 A (z inverse + curl, k^2), 1 B (y inverse), 2 C (x C2R, u x w, x R2C),
 * 3 D (y forward), 4 E1 (z forward + step1), 5 E2 (z forward + step2),
 * 6 F (forcing), 7 EA (E1 fused with stage 2's pass A), 8 EF (E2 fused
 * with the forcing and the next step's pass A; gt_set_fusion).
 * gt_prof_read waits for the stream and returns the summed
 * milliseconds and launch counts since enable; bytes[] is the algorithmic
 * HBM traffic of one launch of each pass (every element the pass must
 * read or write, once; for passes B-D run in z chunks, the mean over the
 * chunks). */
int gt_prof_enable(gt_ctx* c, int on);
int gt_prof_read(gt_ctx* c, double* ms, int64_t* launches, double* bytes, int npass);

/* Launch-shape autotuner (no counterpart in the reference): times nrep
 * evaluations of the nonlinear term of the current state for each pencil
 * width (8 or 16 pencils per block) of the z-inverse and z-forward passes
 * and keeps the fastest of each, then the width of the y passes together
 * with the z chunk below, by whole-evaluation time, and last, where the
 * fused z passes exist (gt_set_fusion), whether to use them, by the time of
 * a step's passes.  Results do not depend on the widths or the chunk, and
 * on fusion only to rounding.
 * gt_set_tiles forces a triple (0 = default); gt_get_tiles reports it and
 * the tuned time per evaluation in ms (0 before tuning). */
int gt_autotune(gt_ctx* c, int nrep);
int gt_set_tiles(gt_ctx* c, int tz, int tze, int ty);
int gt_get_tiles(gt_ctx* c, int* tz, int* tze, int* ty, double* ms);
/* The y-inverse and x passes (and the y-forward pass when with_d is not 0)
 * run in chunks of `planes` z planes (0: all at once), so that each chunk's
 * intermediates are read back while they sit in the device's memory-side
 * cache; with with_d = 0 the y-forward pass runs once after the chunks.
 * gt_autotune picks both, with the y width (from whole-evaluation times;
 * gt_get_tiles then reports that time).  Results do not depend on them.
 * with_d may be NULL in gt_get_zchunk.  No counterpart in the reference. */
int gt_set_zchunk(gt_ctx* c, int planes, int with_d);
int gt_get_zchunk(gt_ctx* c, int* planes, int* with_d);
/* The z chunks on one stream (1, the default) or alternating between the
 * context stream and a second one (2), so that two chunks overlap; it
 * applies when the chunks are fewer planes than NZ and include the
 * y-forward pass.  At 512^3 FP64, two streams with chunks of 16 or 20
 * planes measured 2% faster steps than one with 40, but a pass's launches
 * then share the device with the other stream's, so per-pass timing
 * (gt_prof_*, events on each launch's stream) no longer separates the
 * passes; gt_autotune keeps one stream.  Results do not depend on it.
 * No counterpart in the reference. */
int gt_set_zstreams(gt_ctx* c, int nstreams);
int gt_get_zstreams(gt_ctx* c, int* nstreams);

/* Cross-stage fusion (no counterpart in the reference): the z-forward pass
 * of stage 1 keeps bh = emk (b + dt2 N1) in registers and runs stage 2's
 * z-inverse pass (curl bh, k^2 bh) on it in the same kernel, so bh never
 * goes through device memory; within one gt_step call, the z-forward pass
 * of stage 2 likewise adds the step's forcing, stores b and runs the next
 * step's z-inverse pass on it, so b is not read back.  Results equal the
 * unfused step to rounding (contraction order).  Compiled with simple
 * viscosity (alpnu = 1, mu = 0) for NZ = 512 at the default z-forward width
 * (which gt_autotune keeps while fusion is on) and for NZ = 128 and 256 at
 * both.
 * On by default in FP64, off in FP32-bulk; gt_autotune keeps whichever runs
 * a step faster, and gt_set_fusion after it forces a choice.  *active (may
 * be NULL) reports whether the current context uses it. */
int gt_set_fusion(gt_ctx* c, int on);
int gt_get_fusion(gt_ctx* c, int* on, int* active);

/* Batches: nmembers independent members on one grid with the same
 * parameters except omega (omegas[m], or p->omega when omegas is NULL) and
 * the forcing seed.  Every pass covers all members in one launch.  No
 * counterpart in the reference (one run per executable there). */
int gt_batch_create(const gt_params* p, int nmembers, const double* omegas, gt_ctx** out);
int gt_batch_size(gt_ctx* c, int* nmembers);
int gt_batch_set_state(gt_ctx* c, int member, const double* bx, const double* by, const double* bz);
int gt_batch_get_state(gt_ctx* c, int member, double* bx, double* by, double* bz);
/* seeds[nmembers]: each member's LCG state, advanced on return */
int gt_batch_step(gt_ctx* c, int nsteps, int32_t* seeds);
int gt_batch_spectra(gt_ctx* c, int member, double* sp, double* sp2d, double* hel, double* glob);

/* Ensembles: batches of different grids stepped together on one stream,
 * chunk steps of all of them captured in hipGraphs and replayed.  Seeds are
 * held by the ensemble (gt_ensemble_set_seeds / _get_seeds per batch).
 * The batches stay owned by the caller and must outlive the ensemble. */
typedef struct gt_ensemble gt_ensemble;
int gt_ensemble_create(gt_ctx** batches, int nbatches, int chunk, gt_ensemble** out);
int gt_ensemble_destroy(gt_ensemble* e);
int gt_ensemble_set_seeds(gt_ensemble* e, int batch, const int32_t* seeds);
int gt_ensemble_get_seeds(gt_ensemble* e, int batch, int32_t* seeds);
int gt_ensemble_step(gt_ensemble* e, int nsteps);
int gt_ensemble_sync(gt_ensemble* e);

/* Forcing shells: the forced modes (src/iniforcing_3d3c.f90:26-40,
 * src/iniforcing_2d2c.f90:12-68) grouped by |k|, ascending.  gt_forcing_shells
 * returns their count and, when kk is not NULL, the first cap |k| values.
 * gt_set_shell_amps replaces famp (src/forcing_3d3c.f90:34) by amps[s] for
 * the modes of shell s (amps NULL: famp again); with every amps[s] = famp the
 * step is bit-identical to the plain one.  No counterpart in the reference
 * (one famp for all modes there). */
int gt_forcing_shells(gt_ctx* c, int* nshells, double* kk, int cap);
int gt_set_shell_amps(gt_ctx* c, const double* amps, int nshells);

/* Adjoint of the step (no counterpart in the reference): reverse-mode
 * derivatives of a run of the step of src/main.f90:124-150, FP64 contexts
 * with one member and NX <= 512.
 * gt_adjoint_forward runs nsteps steps from the current state like gt_step
 * (advancing *seed) and keeps the state every `every` steps.
 * gt_adjoint_backward takes lam = dJ/db_N of the final state (three fields
 * in the gt_set_state layout, (re, im) = (dJ/dRe, dJ/dIm)) and returns
 * dJ/db_0 in g[xyz] and, when grad_amps is not NULL, dJ/damps[s] for the
 * nshells forcing shells (the derivative with respect to the per-shell
 * amplitudes of gt_set_shell_amps, at the amplitudes of the forward run).
 * It recomputes each segment of states from its checkpoint, may be called
 * again with another lam, and leaves the context at the final state.  In
 * pruned mode the derivatives are those of the map restricted to the modes
 * trunc keeps (src/trunc.f90:10-28): components outside them read zero. */
int gt_adjoint_forward(gt_ctx* c, int nsteps, int32_t* seed, int every);
int gt_adjoint_backward(gt_ctx* c, const double* lx, const double* ly, const double* lz, double* gx, double* gy,
                        double* gz, double* grad_amps, int nshells);

const char* gt_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GTURBO3D_H */
