"""The hand-written adjoint of the step on the GPU.

* against the numpy adjoint (tests/adjoint_ref.py, itself pinned by central
  differences of the oracle) in every mode;
* against central differences of the GPU forward: 16 directions in the
  state, eps = 1e-6, plus every forcing-shell amplitude, relative error
  below 1e-6;
* against central differences of the FP64 oracle's forward;
* the torch.autograd.Function, the per-shell amplitudes, error paths.
"""
import math
import sys

import numpy as np
import pytest
import torch

from gturbo3d_amd import adjoint, initial, lib
from gturbo3d_amd.params import FORCING_3D3C, PREC_FP32_BULK, Params
import oracle_lib as ol
from adjoint_ref import AdjointRef
from test_adjoint_ref import CASES, J, band_w, case_params, directions, dot

pytestmark = pytest.mark.gpu

EPS = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    num = math.sqrt(sum(float(np.sum(np.abs(x - y) ** 2)) for x, y in zip(a, b)))
    den = math.sqrt(sum(float(np.sum(np.abs(y) ** 2)) for y in b))
    return num / den


def gpu_gradient(g, b0, seed, nsteps, every, w):
    g.set_state(b0)
    g.adjoint_forward(nsteps, seed, every)
    bN = g.get_state()
    return g.adjoint_backward([2.0 * w * f for f in bN])


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_matches_numpy_adjoint(require_gpu, name):
    p = case_params(CASES[name])
    say(f"adjoint vs numpy adjoint: {name} {p.nx}x{p.ny}x{p.nz}")
    b0 = initial.initial_state(p, seed=5, energy=0.5, k0=2.0)
    w = band_w(p, 2.0, 5.0)
    ref = AdjointRef(p)
    bN, _ = ref.o.step(b0, 17, 3)
    rg, ra = ref.gradient(b0, 17, 3, [2.0 * w * f for f in bN])
    with lib.Context(p) as g:
        gg, ga = gpu_gradient(g, b0, 17, 3, 2, w)
    e = rel_l2(gg, rg)
    say(f"  gradient rel L2 {e:.2e}; amplitude gradient {ga} vs {ra}")
    assert e < 1e-11
    assert np.allclose(ga, ra, rtol=1e-11, atol=1e-13 * np.abs(ra).max())


def test_adjoint_central_differences_16_directions(require_gpu):
    p = Params(nx=32, ny=32, nz=32, nu=2e-3, dt=5e-3, omega=2.0, famp=0.1, k1f=2.0, k2f=3.0,
               variant=FORCING_3D3C)
    say("adjoint vs central differences, 32^3, 10 steps, checkpoints every 3")
    b0 = initial.initial_state(p, seed=9, energy=0.5, k0=3.0)
    w = band_w(p, 4.0, 8.0)
    nsteps, seed = 10, 123
    with lib.Context(p) as g:
        assert g.pruned()
        gg, ga = gpu_gradient(g, b0, seed, nsteps, 3, w)

        def Jrun(b, amps=None):
            g.set_shell_amps(amps)
            g.set_state(b)
            g.step(nsteps, seed)
            return J(g.get_state(), w)

        worst = 0.0
        for d in directions(p, b0, 16, 2):
            fd = (Jrun([x + EPS * y for x, y in zip(b0, d)]) - Jrun([x - EPS * y for x, y in zip(b0, d)])) / (2 * EPS)
            err = abs(dot(gg, d) - fd) / abs(fd)
            worst = max(worst, err)
        say(f"  state: worst relative error over 16 directions {worst:.2e}")
        assert worst < 1e-6
        nsh = len(g.forcing_shells())
        assert nsh == len(ga) and nsh >= 2
        worst = 0.0
        for s in range(nsh):
            h = 1e-4 * p.famp  # amplitude steps: rounding error ~ 1e-16 J / h
            ap = np.full(nsh, p.famp)
            am = ap.copy()
            ap[s] += h
            am[s] -= h
            fd = (Jrun(b0, ap) - Jrun(b0, am)) / (2 * h)
            worst = max(worst, abs(ga[s] - fd) / abs(fd))
        say(f"  shell amplitudes ({nsh} shells): worst relative error {worst:.2e}")
        assert worst < 1e-6


def test_adjoint_against_oracle_differences(require_gpu):
    p = Params(nx=24, ny=24, nz=24, nu=3e-3, dt=5e-3, omega=4.0, famp=0.1, k1f=1.0, k2f=2.5, dt2_mode=1)
    say("adjoint vs oracle central differences, 24^3 midpoint mode, 4 steps")
    b0 = initial.initial_state(p, seed=3, energy=0.5, k0=2.0)
    w = band_w(p, 3.0, 6.0)
    o = ol.Oracle(p)
    with lib.Context(p) as g:
        gg, ga = gpu_gradient(g, b0, 77, 4, 4, w)
    for d in directions(p, b0, 4, 3):
        fd = (J(o.step([x + EPS * y for x, y in zip(b0, d)], 77, 4)[0], w)
              - J(o.step([x - EPS * y for x, y in zip(b0, d)], 77, 4)[0], w)) / (2 * EPS)
        assert abs(dot(gg, d) - fd) < 1e-6 * abs(fd)
    h = 1e-4 * p.famp  # amplitude steps: rounding error ~ 1e-16 J / h
    fd = (J(ol.Oracle(Params(**{**p.__dict__, "famp": p.famp + h})).step(b0, 77, 4)[0], w)
          - J(ol.Oracle(Params(**{**p.__dict__, "famp": p.famp - h})).step(b0, 77, 4)[0], w)) / (2 * h)
    assert abs(ga.sum() - fd) < 1e-6 * abs(fd)


def test_shell_amplitudes_equal_to_famp_are_bit_identical(require_gpu):
    p = Params(nx=32, ny=32, nz=32, nu=1e-3, dt=5e-3, omega=1.0, famp=0.07, k1f=2.0, k2f=3.0)
    say("shell amplitudes = famp: bit-identical steps")
    b0 = initial.initial_state(p, seed=1)
    with lib.Context(p) as g:
        g.set_state(b0)
        s1 = g.step(5, 42)
        r1 = g.get_state()
        kk = g.forcing_shells()
        assert np.all(np.diff(kk) > 0) and kk[0] > 2.0 and kk[-1] <= 3.0 + 1e-12
        g.set_shell_amps(np.full(len(kk), p.famp))
        g.set_state(b0)
        s2 = g.step(5, 42)
        r2 = g.get_state()
    assert s1 == s2
    for x, y in zip(r1, r2):
        assert np.array_equal(x, y)


def test_autograd_function_and_reuse(require_gpu):
    p = Params(nx=32, ny=32, nz=32, nu=2e-3, dt=5e-3, omega=2.0, famp=0.1, k1f=2.0, k2f=3.0)
    say("torch.autograd.Function over the adjoint, 32^3, 6 steps")
    b0 = initial.initial_state(p, seed=2)
    wt = torch.from_numpy(adjoint.band_weights(p, 4.0, 8.0))
    with lib.Context(p) as g:
        nsh = len(g.forcing_shells())
        x = adjoint.to_real(b0).requires_grad_(True)
        amps = torch.full((nsh,), p.famp, dtype=torch.float64, requires_grad=True)
        bN = adjoint.run(g, x, amps, 6, 11, every=4)
        loss = adjoint.band_energy(bN, wt)
        loss.backward()
        final = g.get_state()
        assert rel_l2(final, adjoint.to_complex(bN)) == 0.0
        # the same gradient straight from the library, twice (checkpoints reused)
        lam = [2.0 * wt.numpy()[..., 0] * f for f in final]
        g1, a1 = g.adjoint_backward(lam)
        g2, a2 = g.adjoint_backward(lam)
        for u, v in zip(g1, g2):
            assert np.array_equal(u, v)
        assert np.array_equal(a1, a2)
        assert rel_l2(adjoint.to_complex(x.grad), g1) < 1e-14
        assert np.allclose(amps.grad.numpy(), a1, rtol=1e-14, atol=0)
        # the context is left at the final state
        assert rel_l2(g.get_state(), final) == 0.0
        # a new state invalidates the checkpoints
        g.set_state(b0)
        with pytest.raises(lib.GturboError):
            g.adjoint_backward(lam)


def test_autograd_second_run_on_one_context_is_refused(require_gpu):
    """Two live runs on one context: the first node's checkpoints were replaced,
    so its backward raises instead of differentiating the second run."""
    p = Params(nx=32, ny=32, nz=32, nu=2e-3, dt=5e-3, omega=2.0, famp=0.1, k1f=2.0, k2f=3.0)
    say("two autograd runs on one context")
    b0 = initial.initial_state(p, seed=2)
    wt = torch.from_numpy(adjoint.band_weights(p, 4.0, 8.0))
    with lib.Context(p) as g:
        nsh = len(g.forcing_shells())
        amps = torch.full((nsh,), p.famp, dtype=torch.float64, requires_grad=True)
        x1 = adjoint.to_real(b0).requires_grad_(True)
        x2 = adjoint.to_real(b0).requires_grad_(True)
        l1 = adjoint.band_energy(adjoint.run(g, x1, amps, 3, 11, every=2), wt)
        l2 = adjoint.band_energy(adjoint.run(g, x2, amps, 4, 12, every=2), wt)
        with pytest.raises(RuntimeError, match="one live run per context"):
            l1.backward()
        l2.backward()  # the last run is the one the context holds
        assert x2.grad is not None and torch.isfinite(x2.grad).all()


def test_adjoint_rejects_unsupported_contexts(require_gpu):
    p = Params(nx=32, ny=32, nz=32, precision=PREC_FP32_BULK)
    with lib.Context(p) as g:
        with pytest.raises(lib.GturboError, match="FP64"):
            g.adjoint_forward(2, 0, 1)
    bt = lib.Batch(Params(nx=32, ny=32, nz=32), 2)
    with pytest.raises(lib.GturboError, match="single-member"):
        bt.adjoint_forward(2, 0, 1)
    bt.close()
    with lib.Context(Params(nx=32, ny=32, nz=32)) as g:
        with pytest.raises(lib.GturboError, match="no forward run"):
            g.adjoint_backward([np.zeros(g.shape, np.complex128)] * 3)
        with pytest.raises(lib.GturboError, match="nsteps"):
            g.adjoint_forward(0, 0, 1)


def shaped_directions(p, b0, n, seed0):
    """Random directions drawn from the initial-condition ensemble (random
    phases, the same spectrum), scaled to |b0|."""
    nb = math.sqrt(dot(b0, b0))
    out = []
    for s in range(n):
        d = initial.initial_state(p, seed=seed0 + s)
        sc = nb / math.sqrt(dot(d, d))
        out.append([sc * x for x in d])
    return out


def test_adjoint_128_200_steps_16_directions(require_gpu):
    """The 128^3 differentiation case: 200 steps, checkpoints every 20,
    J = kinetic energy in 8 <= |k| <= 12, 16 random directions, eps = 1e-6.

    Directions are random-phase fields of the initial spectrum.  Along white-
    noise directions (most of their norm at |k| ~ 40) the central difference
    has a third-order error of about 1e8 eps^2 here (7.7e-3 at eps = 1e-5,
    7.7e-5 at 1e-6), so two of them are checked with the Richardson
    combination (4 D(eps) - D(2 eps)) / 3, which cancels the eps^2 term."""
    import bench

    p = bench.adjoint_params()
    say(f"adjoint 128^3, 200 steps, famp {p.famp:.4f}")
    b0 = initial.initial_state(p, seed=0)
    w = band_w(p, 8.0, 12.0)
    with lib.Context(p) as g:
        gg, ga = gpu_gradient(g, b0, 7, 200, 20, w)

        def Jrun(b):
            g.set_state(b)
            g.step(200, 7)
            return J(g.get_state(), w)

        def central(d, eps):
            return (Jrun([x + eps * y for x, y in zip(b0, d)]) - Jrun([x - eps * y for x, y in zip(b0, d)])) / (2 * eps)

        worst = 0.0
        for d in shaped_directions(p, b0, 16, 100):
            worst = max(worst, abs(dot(gg, d) - central(d, EPS)) / abs(dot(gg, d)))
        say(f"  worst relative error over 16 directions {worst:.2e}")
        assert worst < 1e-6
        for d in directions(p, b0, 2, 16):
            rich = (4.0 * central(d, EPS) - central(d, 2 * EPS)) / 3.0
            err = abs(dot(gg, d) - rich) / abs(dot(gg, d))
            say(f"  white direction, Richardson: relative error {err:.2e}")
            assert err < 1e-6


def test_final_cotangent_is_projected_in_pruned_mode(require_gpu):
    """Entries of lam_N on modes trunc removes do not reach the gradient, and
    the gradient vanishes there (the state lives on the kept modes)."""
    p = Params(nx=32, ny=32, nz=32, nu=2e-3, dt=5e-3, omega=2.0, famp=0.1, k1f=2.0, k2f=3.0)
    say("pruned mode: lam_N projected on the kept modes")
    b0 = initial.initial_state(p, seed=4)
    w = band_w(p, 4.0, 8.0)
    masked = initial.trunc_mask(p)
    rng = np.random.default_rng(5)
    with lib.Context(p) as g:
        g.set_state(b0)
        g.adjoint_forward(5, 9, 2)
        bN = g.get_state()
        lam = [2.0 * w * f for f in bN]
        g1, a1 = g.adjoint_backward(lam)
        noisy = [x + np.where(masked, rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape), 0.0)
                 for x in lam]
        g2, a2 = g.adjoint_backward(noisy)
    for u, v in zip(g1, g2):
        assert np.array_equal(u, v)
        assert not np.any(u[masked])
    assert np.array_equal(a1, a2)


@pytest.mark.parametrize("dt2_mode", [0, 1])
def test_segment_layouts_give_one_gradient(require_gpu, dt2_mode):
    """The same 5-step gradient with checkpoints every 1, 2, 3 and 5 steps
    (five segments, a ragged last segment of one step, of two steps, and a
    single segment whose states all come from the forward run): every layout
    gives the gradient of every other, and a second backward on each run (the
    last segment recomputed where the recompute of earlier segments refilled
    its buffer) equals the first bit for bit."""
    p = Params(nx=32, ny=32, nz=32, nu=2e-3, dt=5e-3, omega=2.0, famp=0.1, k1f=2.0, k2f=3.0, dt2_mode=dt2_mode)
    say(f"segment layouts, 5 steps, dt2 mode {dt2_mode}")
    b0 = initial.initial_state(p, seed=6)
    w = band_w(p, 4.0, 8.0)
    out = {}
    with lib.Context(p) as g:
        for every in (1, 2, 3, 5):
            g.set_state(b0)
            g.adjoint_forward(5, 21, every)
            lam = [2.0 * w * f for f in g.get_state()]
            g1, a1 = g.adjoint_backward(lam)
            g2, a2 = g.adjoint_backward(lam)
            assert all(np.array_equal(u, v) for u, v in zip(g1, g2)), every
            assert np.array_equal(a1, a2), every
            out[every] = (g1, a1)
    g5, a5 = out[5]
    for every in (1, 2, 3):
        gg, ga = out[every]
        e = rel_l2(gg, g5)
        say(f"  every {every} against one segment: {e:.2e}")
        assert e < 1e-13
        assert np.allclose(ga, a5, rtol=1e-13, atol=1e-15 * np.abs(a5).max())
