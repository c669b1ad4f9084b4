#!/usr/bin/env python3
"""Throughput of the pseudo-spectral RK2 step on one MI355X.

Workload: 512^3 triply periodic box, FP64, nu = 5e-5, Omega = 2, 3D3C
forcing on the shell (2, 3] with injection rate 0.1, dt = 1e-3, seeded
initial spectrum k^4 exp(-2 (k/4)^2) with E = 0.5.  A step is one full
pass of the reference time-loop body (nlt, step1, nlt, step2, forcing) on
the state resident in HBM.  Prints one JSON line whose value is the FP64
step; the same line carries the FP32-bulk step of the same workload
("fp32_bulk": speed-up over FP64 and its own roofline) and the adjoint
step of the 128^3 gradient workload ("adjoint": seconds per step, its
fraction of HBM peak on the bytes its passes must move, a central-difference
check).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32-bulk]
                    [--legs fp64,fp32-bulk,adjoint]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt·steps/s at 512³ RK2 (FP64 vs FP32-bulk), % HBM peak; adjoint step s"
HBM_PEAK_GBS = 8000.0


def workload_params(n, precision):
    from gturbo3d_amd.params import FORCING_3D3C, PREC_FP32_BULK, PREC_FP64, Params
    p = Params(nx=n, ny=n, nz=n, nu=5e-5, alpnu=1.0, mu=0.0, alpmu=0.0, dt=1e-3, omega=2.0,
               rtrunc=2.0 / 3.0, truncate=1, famp=0.0, k1f=2.0, k2f=3.0, variant=FORCING_3D3C,
               precision=PREC_FP32_BULK if precision == "fp32-bulk" else PREC_FP64)
    return p


def cpu_baseline(p, b, seed, threads, reps=1, warmup=0):
    """Oracle steps of the same workload on the host cores (the oracle is the
    CPU restatement of the reference step, OpenMP over pencils): `warmup`
    untimed steps, then the median seconds of `reps` timed steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as ol
    ol.set_threads(threads)
    o = ol.Oracle(p)
    B = [np.ascontiguousarray(f).copy() for f in b]
    W = [np.ones(p.shape, np.complex128) for _ in range(6)]  # pre-faulted work fields
    s = ctypes.c_int32(seed)
    ptrs = [ctypes.c_void_p(x.ctypes.data) for x in B + W]
    for _ in range(warmup):
        ol.lib().or_step(o.h, *ptrs, ctypes.byref(s), 1)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ol.lib().or_step(o.h, *ptrs, ctypes.byref(s), 1)
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def cpu_fft_gflops(n, threads, reps=3):
    """Nominal rate of the oracle's 3D real-to-complex transform (fft_dir) at
    n^3 on `threads` threads: 2.5 N log2 N flop for N = n^3 real points over
    the median time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as ol
    from gturbo3d_amd.params import Params
    ol.set_threads(threads)
    o = ol.Oracle(Params(nx=n, ny=n, nz=n))
    rng = np.random.default_rng(1)
    z = np.zeros((n, n, n // 2 + 1), np.complex128)
    z.view(np.float64).reshape(n, n, n + 2)[:, :, :n] = rng.standard_normal((n, n, n))
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ol.lib().or_fft_dir(o.h, ctypes.c_void_p(z.ctypes.data))
        ts.append(time.perf_counter() - t0)
    t = sorted(ts)[len(ts) // 2]
    return 2.5 * n ** 3 * math.log2(n ** 3) / t / 1e9


def host_cpu():
    """Model name, logical CPU and socket counts of the host (/proc/cpuinfo),
    and the CPUs this process may run on."""
    model, count, sockets = "unknown", 0, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                if line.startswith("processor"):
                    count += 1
                if line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else count
    return {"model": model, "logical_cpus": count, "sockets": len(sockets) or None, "cpus_available": avail}


def ensemble_bench(args):
    """32 members at 128^3 plus 16 at 96^3, FP64, nu = 1e-3, dt = 4e-3, 3D3C forcing (2,3]
    rate 0.1, seeds 0..47, Omega ~ U[0.5, 10] from PCG64(48); chunk-step hipGraphs."""
    import numpy as np
    import torch
    from gturbo3d_amd import initial, lib
    from gturbo3d_amd.params import Params
    rng = np.random.default_rng(48)
    omegas = rng.uniform(0.5, 10.0, 48)
    groups = [(128, 32, 0), (96, 16, 32)]
    batches, tiles = [], {}
    for n, nm, off in groups:
        p = Params(nx=n, ny=n, nz=n, nu=1e-3, dt=4e-3, omega=1.0, famp=0.0, k1f=2.0, k2f=3.0)
        p.famp = math.sqrt(0.1 / lib.forcing_count(p)[0])
        bt = lib.Batch(p, nm, omegas[off:off + nm])
        for m in range(nm):
            bt.set_member(m, initial.initial_state(p, seed=off + m))
        tiles[f"{n}^3"] = dict(zip(("tz", "tze", "ty", "ms_per_nonlinear"), bt.autotune(3)))
        batches.append(bt)
    ens = lib.Ensemble(batches, chunk=10)
    ens.set_seeds(0, np.arange(0, 32, dtype=np.int32))
    ens.set_seeds(1, np.arange(32, 48, dtype=np.int32))
    ens.step(args.warmup * 10)
    ens.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(args.steps)
    ens.sync()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    pts = 32 * 128 ** 3 + 16 * 96 ** 3
    finite = all(np.isfinite(f).all() for bt in batches for f in bt.get_member(0))
    print(json.dumps({
        "metric": "ensemble grid-pt·steps/s (32 x 128^3 + 16 x 96^3, FP64, one hipGraph per 10 steps)",
        "value": pts * args.steps / t, "unit": "grid-pt·steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup * 10, "ms_per_step": 1e3 * t / args.steps, "member_steps_per_s": 48 * args.steps / t,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched ensemble", "members": {"128^3": 32, "96^3": 16},
                   "omegas_pcg64_seed48": [round(float(o), 6) for o in omegas], "tiles": tiles},
        "state_finite": bool(finite)}))
    ens.close()
    for bt in batches:
        bt.close()


def adjoint_params():
    from gturbo3d_amd import lib
    from gturbo3d_amd.params import FORCING_3D3C, Params
    p = Params(nx=128, ny=128, nz=128, nu=1e-3, dt=4e-3, omega=2.0, famp=0.0, k1f=2.0, k2f=3.0,
               variant=FORCING_3D3C)
    p.famp = math.sqrt(0.1 / lib.forcing_count(p)[0])
    return p


def adjoint_bench(args):
    """Gradient of the kinetic energy in 8 <= |k| <= 12 after 200 steps at
    128^3 (FP64, nu = 1e-3, Omega = 2, 3D3C forcing (2,3] rate 0.1, seed 7,
    dt = 4e-3) with respect to the initial vector potential and the per-shell
    forcing amplitudes, checkpoints every 20 steps, through the
    torch.autograd.Function.  --steps gradients are timed; the value is the
    backward time (checkpoint recompute + adjoint) per step.  Central
    differences on 16 random directions (eps = 1e-6, random-phase fields of
    the initial spectrum) check the gradient."""
    import numpy as np
    import torch
    from gturbo3d_amd import adjoint, initial, lib
    p = adjoint_params()
    nsteps, every, seed = 200, 20, 7
    b0 = initial.initial_state(p, seed=0)
    g = lib.Context(p)
    g.set_state(b0)
    tiles = g.autotune(3)
    wt = torch.from_numpy(adjoint.band_weights(p, 8.0, 12.0))
    nsh = len(g.forcing_shells())
    t_fwd, t_bwd = [], []
    for it in range(args.warmup + args.steps):
        x = adjoint.to_real(b0).requires_grad_(True)
        amps = torch.full((nsh,), p.famp, dtype=torch.float64, requires_grad=True)
        t0 = time.perf_counter()
        bN = adjoint.run(g, x, amps, nsteps, seed, every)
        loss = adjoint.band_energy(bN, wt)
        t1 = time.perf_counter()
        loss.backward()
        t2 = time.perf_counter()
        if it >= args.warmup:
            t_fwd.append(t1 - t0)
            t_bwd.append(t2 - t1)
    grad = adjoint.to_complex(x.grad)
    ga = amps.grad.numpy().copy()
    w = wt.numpy()[..., 0]

    def J(b):
        g.set_state(b)
        g.step(nsteps, seed)
        return sum(float(np.sum(w * np.abs(f) ** 2)) for f in g.get_state())

    # random directions: random-phase fields of the initial spectrum
    nb = math.sqrt(sum(float(np.sum(np.abs(f) ** 2)) for f in b0))
    worst, eps = 0.0, 1e-6
    t0 = time.perf_counter()
    for k in range(args.fd_directions):
        d = initial.initial_state(p, seed=100 + k)
        sc = nb / math.sqrt(sum(float(np.sum(np.abs(f) ** 2)) for f in d))
        d = [sc * f for f in d]
        fd = (J([a + eps * c for a, c in zip(b0, d)]) - J([a - eps * c for a, c in zip(b0, d)])) / (2 * eps)
        ad = sum(float(np.sum(np.real(np.conj(a) * c))) for a, c in zip(grad, d))
        worst = max(worst, abs(ad - fd) / abs(fd))
    t_fd = time.perf_counter() - t0
    bwd = float(np.median(t_bwd))
    print(json.dumps({
        "metric": "adjoint step s (128^3 FP64, 200 steps, checkpoints every 20)",
        "value": bwd / nsteps, "unit": "s per adjoint step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "forward_s_per_run": float(np.median(t_fwd)), "backward_s_per_run": bwd,
        "gradient_s": float(np.median(t_fwd)) + bwd,
        "fd_check": {"directions": args.fd_directions, "eps": eps, "worst_rel_err": worst, "seconds": t_fd,
                     "directions_from": "initial_state seeds 100..115 scaled to |b0|"},
        "grad_amps": [float(v) for v in ga], "shells_k": [float(v) for v in g.forcing_shells()],
        "config": {"workload": "adjoint 128^3", "nsteps": nsteps, "checkpoint_every": every, "seed": seed,
                   "objective": "kinetic energy in 8 <= |k| <= 12 at the final time",
                   "famp": p.famp, "tiles": dict(zip(("tz", "tze", "ty", "ms_per_nonlinear"), tiles))}}))
    g.close()


def adjoint_leg(args, ngrad=2, fd_directions=4):
    """The adjoint step of the 128^3 gradient workload (adjoint_bench's
    configuration) for the main bench line: one warm-up gradient, then the
    median backward time of ngrad gradients per step, a central-difference
    check on fd_directions random directions, and the step's fraction of HBM
    peak on the bytes its passes must move.

    Those bytes come from the forward passes of the same context, counted
    the way gt_prof_read counts them (every element a pass must read or
    write, once; pruned extents).  With A, B, C, E one launch of the z-inverse,
    y-inverse (6 fields), x and z-forward passes of one nonlinear evaluation:
    the adjoint's z-inverse of Z emk b is A, the z-inverse of the cotangent
    moves (A + E) / 3 (3 fields in at the kept modes, 3 out at the alive
    columns), the y passes 9 and 6 fields (1.5 B + B), the x pass 9 in and
    6 out (5/3 C), the z-forward 6 fields in and the cotangent in and out
    (2 (A + E) / 3): 2A + E + 2.5 B + 5/3 C per step in reference mode.  The
    recompute of each segment adds (every - 1) / every forward steps per
    step, and keeping the segment's states one write and one read of three
    fields at the kept modes."""
    import numpy as np
    import torch
    from gturbo3d_amd import adjoint, initial, lib
    p = adjoint_params()
    nsteps, every, seed = 200, 20, 7
    b0 = initial.initial_state(p, seed=0)
    g = lib.Context(p)
    g.set_state(b0)
    tiles = g.autotune(3)
    # forward pass bytes of this context (all passes whole: no z chunks)
    g.set_zchunk(0)
    g.prof_enable(True)
    g.step(2, seed)
    ms, launches, by = g.prof_read()
    g.prof_enable(False)
    g.set_state(b0)
    per = {q: by[q] * launches[q] / 2.0 for q in range(lib.NPASS)}  # per forward step
    A, B, C, E = by[0], by[1], by[2], by[4]
    X = (2.0 * E - A) / 9.0  # 16 bytes x kept modes of one field
    fwd_step = sum(per.values())
    adj_bytes = 2 * A + E + 2.5 * B + (5.0 / 3.0) * C + (every - 1) / every * fwd_step + 2 * 3 * X
    wt = torch.from_numpy(adjoint.band_weights(p, 8.0, 12.0))
    nsh = len(g.forcing_shells())
    t_fwd, t_bwd = [], []
    for it in range(1 + ngrad):
        x = adjoint.to_real(b0).requires_grad_(True)
        amps = torch.full((nsh,), p.famp, dtype=torch.float64, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = adjoint.band_energy(adjoint.run(g, x, amps, nsteps, seed, every), wt)
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it >= 1:
            t_fwd.append(t1 - t0)
            t_bwd.append(t2 - t1)
    grad = adjoint.to_complex(x.grad)
    w = wt.numpy()[..., 0]

    def J(b):
        g.set_state(b)
        g.step(nsteps, seed)
        return sum(float(np.sum(w * np.abs(f) ** 2)) for f in g.get_state())

    nb = math.sqrt(sum(float(np.sum(np.abs(f) ** 2)) for f in b0))
    worst, eps = 0.0, 1e-6
    for k in range(fd_directions):
        d = initial.initial_state(p, seed=100 + k)
        sc = nb / math.sqrt(sum(float(np.sum(np.abs(f) ** 2)) for f in d))
        d = [sc * f for f in d]
        fd = (J([a + eps * c for a, c in zip(b0, d)]) - J([a - eps * c for a, c in zip(b0, d)])) / (2 * eps)
        ad = sum(float(np.sum(np.real(np.conj(a) * c))) for a, c in zip(grad, d))
        worst = max(worst, abs(ad - fd) / abs(fd))
    g.close()
    bwd = float(np.median(t_bwd))
    s_step = bwd / nsteps
    return {
        "metric": "adjoint step s (128^3 FP64, 200 steps, checkpoints every 20)",
        "value": s_step, "unit": "s per adjoint step", "higher_is_better": False,
        "backward_s_per_run": bwd, "forward_s_per_run": float(np.median(t_fwd)), "gradients_timed": ngrad,
        "alg_bytes_per_step": adj_bytes, "alg_hbm_frac": adj_bytes / s_step / (HBM_PEAK_GBS * 1e9),
        "fd_check": {"directions": fd_directions, "eps": eps, "worst_rel_err": worst},
        "config": {"workload": "gradient of the kinetic energy in 8 <= |k| <= 12 after 200 steps at 128^3, "
                               "nu=1e-3, Omega=2, 3D3C (2,3] rate 0.1 seed 7, dt=4e-3, reference stage 1",
                   "tiles": dict(zip(("tz", "tze", "ty", "ms_per_nonlinear"), tiles))},
    }


def precision_bench(args):
    """FP32-bulk against FP64 on the GPU: the 512^3 / 1024^3 workload (nu = 5e-5,
    Omega = 2, 3D3C forcing (2,3] rate 0.1, dt = 1e-3) run --steps steps in each
    precision from the same state and seed; per-field relative L2 of b and the
    relative errors of the energy and helicity spectra and of glob.  The runs
    use the midpoint stage 1: with dt2 = 0 (the reference's) the nonlinear term
    advances by forward Euler and this workload diverges within 500 steps."""
    import numpy as np
    from gturbo3d_amd import initial, lib
    from gturbo3d_amd.params import DT2_MIDPOINT
    n = args.size
    res = {}
    for prec in ("fp64", "fp32-bulk"):
        p = workload_params(n, prec)
        p.dt2_mode = DT2_MIDPOINT
        p.famp = math.sqrt(0.1 / lib.forcing_count(p)[0])
        b0 = initial.initial_state(p, seed=0)
        g = lib.Context(p)
        g.set_state(b0)
        del b0
        g.autotune(2)
        t0 = time.perf_counter()
        seed = g.step(args.steps, 0)
        g.sync()
        t = time.perf_counter() - t0
        sp, _, hel, glob = g.spectra()
        res[prec] = dict(b=g.get_state(), sp=sp, hel=hel, glob=glob, seed=seed, ms=1e3 * t / args.steps)
        g.close()
    a, r = res["fp32-bulk"], res["fp64"]
    fields = [float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(a["b"], r["b"])]
    m = r["sp"] > 1e-12 * r["sp"].max()
    e_sp = float(np.max(np.abs(a["sp"][m] - r["sp"][m]) / r["sp"][m]))
    mh = np.abs(r["hel"]) > 1e-6 * np.abs(r["hel"]).max()
    e_hel = float(np.max(np.abs(a["hel"][mh] - r["hel"][mh]) / np.abs(r["hel"][mh])))
    e_E = abs(a["glob"][:3].sum() - r["glob"][:3].sum()) / r["glob"][:3].sum()
    e_D = abs(a["glob"][3:6].sum() - r["glob"][3:6].sum()) / r["glob"][3:6].sum()
    print(json.dumps({
        "metric": f"FP32-bulk vs FP64 after {args.steps} steps at {n}^3", "unit": "relative error",
        "field_rel_l2": {"bx": fields[0], "by": fields[1], "bz": fields[2]},
        "energy_spectrum_max_rel": e_sp, "helicity_spectrum_max_rel": e_hel,
        "energy_rel": e_E, "dissipation_rel": e_D, "seeds_equal": a["seed"] == r["seed"],
        "ms_per_step": {"fp64": r["ms"], "fp32-bulk": a["ms"]}, "speedup": r["ms"] / a["ms"],
        "config": {"workload": f"{n}^3, nu=5e-5, Omega=2, 3D3C (2,3] rate 0.1, dt=1e-3, midpoint stage 1",
                   "steps": args.steps}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--precision", choices=["fp64", "fp32-bulk"], default="fp64")
    ap.add_argument("--cpu-baseline", choices=["auto", "skip"], default="auto")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default pencil widths")
    ap.add_argument("--tiles", default="", help="tz,tze,ty[,zchunk[,with_d[,zstreams]]]: force the pencil widths (and the "
                         "z chunk of passes B-D and its stream count) instead of autotuning")
    ap.add_argument("--fusion", choices=["auto", "on", "off"], default="auto",
                    help="the z-forward passes fused with the z-inverse passes that follow them where compiled "
                         "(auto: as the autotuner picks, or the library's default with --tiles)")
    ap.add_argument("--fd-directions", type=int, default=16, help="adjoint mode: central-difference checks")
    ap.add_argument("--legs", default=None,
                    help="step mode: comma-separated legs, the first giving the line's value (default "
                         "fp64,fp32-bulk,adjoint; with --precision fp32-bulk only that precision)")
    ap.add_argument("--mode", choices=["step", "copy", "ensemble", "adjoint", "precision"], default="step",
                    help="copy: a plain 1 GiB device copy, the reference kernel for scaling rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures the single-GPU step only: this build has no multi-GPU path "
              "(the slab decomposition with its all-to-all is a design, not code). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from gturbo3d_amd import initial, lib

    if args.mode == "copy":
        nbytes = 1 << 30
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                          "launches": 5}))
        return

    if args.mode == "ensemble":
        ensemble_bench(args)
        return

    if args.mode == "adjoint":
        adjoint_bench(args)
        return

    if args.mode == "precision":
        precision_bench(args)
        return

    if args.legs:
        legs = [x.strip() for x in args.legs.split(",") if x.strip()]
    else:
        legs = ["fp64", "fp32-bulk", "adjoint"] if args.precision == "fp64" else [args.precision]
    line = step_leg(args, legs[0], with_cpu=args.cpu_baseline == "auto")
    for leg in legs[1:]:
        if leg in ("fp64", "fp32-bulk"):
            sub = step_leg(args, leg, with_cpu=False)
            key = leg.replace("-", "_")
            line[key] = {k: sub[k] for k in ("value", "unit", "ms_per_step", "dtype", "alg_bytes_per_step",
                                              "alg_hbm_frac", "pass_ms_per_step", "roofline", "state_finite")}
            line[key]["config"] = {k: sub["config"][k] for k in ("tiles", "pruned", "fused_stage_boundary")}
            line[key]["speedup_vs_" + legs[0].replace("-", "_")] = line["ms_per_step"] / sub["ms_per_step"]
        elif leg == "adjoint":
            line["adjoint"] = adjoint_leg(args)
    print(json.dumps(line))


def step_leg(args, precision, with_cpu):
    """One precision of the step workload: autotune, warm-up, the timed step
    loop, a profiled repeat (per-pass times, the roofline of the dominant
    pass) and, with with_cpu, the host baseline."""
    import numpy as np
    import torch
    from gturbo3d_amd import initial, lib
    args.precision = precision
    n = args.size
    p0 = workload_params(n, args.precision)
    # famp = sqrt(0.1 / nf): mean energy input inpe = nf * famp^2 = 0.1
    nf = lib.forcing_count(p0)[0]
    p = workload_params(n, args.precision)
    p.famp = math.sqrt(0.1 / nf)
    seed0 = 0
    b = initial.initial_state(p, seed=seed0)

    torch.cuda.init()
    ctx = lib.Context(p)
    ctx.set_state(b)
    tiles = None
    if args.tiles:
        tv = [int(v) for v in args.tiles.split(",")]
        ctx.set_tiles(*tv[:3])
        if len(tv) > 3:
            ctx.set_zchunk(tv[3], len(tv) < 5 or bool(tv[4]))
        if len(tv) > 5:
            ctx.set_zstreams(tv[5])
        tiles = dict(zip(("tz", "tze", "ty", "ms_per_nonlinear"), ctx.tiles()))
    elif not args.no_autotune:
        # pencil widths of the z-inverse, z-forward and y passes and the z
        # chunk of the y/x/y passes (bit-neutral launch shapes)
        tiles = dict(zip(("tz", "tze", "ty", "ms_per_nonlinear"), ctx.autotune(3)))
    if args.fusion != "auto":
        ctx.set_fusion(args.fusion == "on")  # after the autotuner, which picks it otherwise
    if tiles is not None:
        tiles["zchunk"], tiles["zchunk_d"] = ctx.zchunk()
        tiles["zstreams"] = ctx.zstreams()
    seed = seed0
    seed = ctx.step(args.warmup, seed)
    ctx.sync()

    # the timed region: the step loop alone (no per-launch events, which cost
    # GPU time at every kernel boundary once passes B-D run in z chunks)
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    seed = ctx.step(args.steps, seed)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the same number of steps again with an event pair around every kernel
    # launch on the context stream: per-pass times and the roofline figures
    ctx.prof_enable(True)
    tp0 = time.perf_counter()
    seed = ctx.step(args.steps, seed)
    ctx.sync()
    tp1 = time.perf_counter()
    ms, launches, bytes_per = ctx.prof_read()
    ctx.prof_enable(False)
    pruned = ctx.pruned()
    fused = ctx.fusion()[1]
    out = ctx.get_state()
    ctx.close()
    finite = all(np.isfinite(f).all() for f in out)

    t = t1 - t0
    gps = n ** 3 * args.steps / t
    ms_step = 1e3 * t / args.steps
    M = (n // 2 + 1) * n * n
    S = 16.0 * M
    dense_bytes = (90.0 if args.precision == "fp64" else 54.0) * S
    alg_bytes_step = float(sum(bytes_per[q] * launches[q] for q in range(lib.NPASS))) / max(args.steps, 1)

    dom = int(np.argmax(ms))
    avg_ms = ms[dom] / max(launches[dom], 1)
    achieved = bytes_per[dom] / (avg_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": lib.PASS_NAMES[dom],
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_ms": round(avg_ms, 4),
        "alg_bytes_per_launch": bytes_per[dom],
    }
    # HBM bytes per launch from the committed rocprofv3 counter runs (FETCH_SIZE
    # and WRITE_SIZE, calibrated on a device copy); passes B-D launch once per
    # z chunk, so their figure scales with the chunk size of this run
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{args.precision}_{n}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        name = lib.PASS_NAMES[dom]
        tb = tr.get("per_launch", {}).get(name)
        if tb is not None and name in ("B_yinv", "C_x_cross", "D_yfwd"):
            t_run, t_prof = tiles or {}, tr.get("tiles", {"zchunk": tr.get("zchunk"), "zchunk_d": True})
            def planes(t):
                if not t.get("zchunk") or (name == "D_yfwd" and not t.get("zchunk_d", True)):
                    return n
                return t["zchunk"]
            tb *= planes(t_run) / planes(t_prof)
        roofline["traffic"] = tb

    cpu = None
    if with_cpu:
        threads = args.cpu_threads or min(len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1), 16)
        reps = 3 if n <= 512 else 1
        tc = cpu_baseline(p, b, seed0, threads, reps=reps, warmup=1)
        cpu = {"value": n ** 3 / tc, "unit": "grid-pt·steps/s", "cores": threads, "kind": "port",
               "sample": f"the same {n}^3 FP64 workload, oracle (C + OpenMP) on {threads} threads: 1 warm-up "
                         f"step, then the median of {reps} timed steps, {tc:.2f} s per step",
               "fft": "the oracle's own mixed-radix Stockham (long-double twiddles); FFTW is not installed",
               "host": host_cpu()}
        # one core, on a bounded sample: one step of the same workload at half the size
        h = max(n // 2, 16)
        ph = workload_params(h, "fp64")
        ph.famp = math.sqrt(0.1 / lib.forcing_count(ph)[0])
        bh = initial.initial_state(ph, seed=seed0)
        t1 = cpu_baseline(ph, bh, seed0, 1)
        cpu["single_thread"] = {"value": h ** 3 / t1, "unit": "grid-pt·steps/s", "cores": 1,
                                "sample": f"1 step of the {h}^3 workload on 1 thread, {t1:.1f} s"}
        nf = min(n, 256)
        cpu["fft_gflops_1_thread"] = round(cpu_fft_gflops(nf, 1), 3)
        cpu["fft_gflops_sample"] = f"fft_dir at {nf}^3, 2.5 N log2 N nominal flop, median of 3"

    return {
        "metric": METRIC,
        "value": gps,
        "unit": "grid-pt·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-bulk (f64 state, IF, RK, reductions)",
        "data": "synthetic (seeded initial spectrum k^4 exp(-2(k/4)^2), E=0.5; 3D3C forcing seed 0)",
        "config": {
            "workload": f"{n}^3 RK2 integrating-factor step, nu=5e-5, Omega=2, 3D3C forcing (2,3] rate 0.1, "
                        f"dt=1e-3, {args.precision}",
            "grid": [n, n, n],
            "pruned": pruned,
            "fused_stage_boundary": fused,
            "tiles": tiles,
        },
        "hbm_frac_dense_model": dense_bytes * args.steps / t / (HBM_PEAK_GBS * 1e9),
        "alg_bytes_per_step": alg_bytes_step,
        "alg_hbm_frac": alg_bytes_step * args.steps / t / (HBM_PEAK_GBS * 1e9),
        "profiled_ms_per_step": 1e3 * (tp1 - tp0) / args.steps,
        "pass_ms_per_step": {lib.PASS_NAMES[q]: ms[q] / args.steps for q in range(lib.NPASS)},
        "pass_alg_gbs": {lib.PASS_NAMES[q]: (bytes_per[q] / (ms[q] / launches[q] * 1e-3) / 1e9 if launches[q] else 0.0)
                         for q in range(lib.NPASS)},
        "pass_alg_bytes_per_launch": {lib.PASS_NAMES[q]: float(bytes_per[q]) for q in range(lib.NPASS)},
        "pass_launches_per_step": {lib.PASS_NAMES[q]: float(launches[q]) / max(args.steps, 1) for q in range(lib.NPASS)},
        "state_finite": bool(finite),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


if __name__ == "__main__":    main()
