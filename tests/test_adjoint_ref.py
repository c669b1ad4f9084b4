"""The numpy adjoint of the step (tests/adjoint_ref.py) against central
differences of the oracle's forward step, on the CPU.

This pins the hand derivation (transposed transforms with their half-
spectrum weights, the cross-product and curl adjoints, Coriolis, invcurl,
trunc, the integrating factors, both stage-1 modes, both forcing variants,
dense and truncated runs, a non-cubic grid) before it is compared with the
GPU adjoint.  J = band energy of the final state, dJ/db_N = 2 w b_N.
"""
import math

import numpy as np
import pytest

from gturbo3d_amd import initial
from gturbo3d_amd.params import FORCING_2D2C, FORCING_3D3C, Params
import oracle_lib as ol
from adjoint_ref import AdjointRef

EPS = 1e-6


def band_w(p, kmin, kmax):
    fkx, fky, fkz = initial.wavenumbers(p)
    k2 = fkz[:, None, None] ** 2 + fky[None, :, None] ** 2 + fkx[None, None, :] ** 2
    kk = np.sqrt(k2)
    fac = np.ones(fkx.shape)
    fac[0] = fac[-1] = 0.5
    return np.where((kk >= kmin) & (kk <= kmax), k2 * fac[None, None, :], 0.0)


def J(b, w):
    return sum(float(np.sum(w * np.abs(f) ** 2)) for f in b)


def dot(a, b):
    return sum(float(np.sum(np.real(np.conj(x) * y))) for x, y in zip(a, b))


def directions(p, b0, n, seed):
    rng = np.random.default_rng(seed)
    keep = np.ones((p.nz, p.ny, p.nx // 2 + 1), bool)
    if p.truncate == 1:
        keep &= ~initial.trunc_mask(p)
        keep[0, 0, 0] = False
    nb = math.sqrt(dot(b0, b0))
    out = []
    for _ in range(n):
        d = [np.where(keep, rng.standard_normal(keep.shape) + 1j * rng.standard_normal(keep.shape), 0.0)
             for _ in range(3)]
        s = nb / math.sqrt(dot(d, d))
        out.append([s * x for x in d])
    return out


CASES = {
    "trunc_ref_3d3c": dict(n=(16, 16, 16), truncate=1, dt2_mode=0, variant=FORCING_3D3C),
    "trunc_mid_3d3c": dict(n=(16, 16, 16), truncate=1, dt2_mode=1, variant=FORCING_3D3C),
    "dense_ref_2d2c": dict(n=(16, 16, 16), truncate=0, dt2_mode=0, variant=FORCING_2D2C),
    "dense_mid_noncubic": dict(n=(24, 16, 32), truncate=0, dt2_mode=1, variant=FORCING_3D3C),
}


def case_params(c):
    nx, ny, nz = c["n"]
    return Params(nx=nx, ny=ny, nz=nz, nu=5e-3, dt=1e-2, omega=3.0, famp=0.2, k1f=1.0, k2f=2.5,
                  truncate=c["truncate"], dt2_mode=c["dt2_mode"], variant=c["variant"])


@pytest.mark.parametrize("name", list(CASES))
def test_numpy_adjoint_matches_oracle_differences(name):
    p = case_params(CASES[name])
    o = ol.Oracle(p)
    ref = AdjointRef(p, o)
    b0 = initial.initial_state(p, seed=5, energy=0.5, k0=2.0)
    nsteps, seed = 3, 17
    w = band_w(p, 2.0, 5.0)
    bN, _ = o.step(b0, seed, nsteps)
    g0, ga = ref.gradient(b0, seed, nsteps, [2.0 * w * f for f in bN])
    worst = 0.0
    for d in directions(p, b0, 4, 1):
        jp = J(o.step([x + EPS * y for x, y in zip(b0, d)], seed, nsteps)[0], w)
        jm = J(o.step([x - EPS * y for x, y in zip(b0, d)], seed, nsteps)[0], w)
        fd = (jp - jm) / (2 * EPS)
        ad = dot(g0, d)
        worst = max(worst, abs(ad - fd) / abs(fd))
    assert worst < 1e-6, worst
    # amplitudes: all shells at famp, so dJ/dfamp = sum_s dJ/damp_s
    h = 1e-4 * p.famp  # amplitude steps: rounding error ~ 1e-16 J / h
    jp = J(ol.Oracle(Params(**{**p.__dict__, "famp": p.famp + h})).step(b0, seed, nsteps)[0], w)
    jm = J(ol.Oracle(Params(**{**p.__dict__, "famp": p.famp - h})).step(b0, seed, nsteps)[0], w)
    fd = (jp - jm) / (2 * h)
    assert abs(ga.sum() - fd) < 1e-6 * abs(fd), (ga.sum(), fd)


def test_transposed_transforms():
    """<F^-1 X, g> = <X, wx rfftn g> and <F g, m> = <g, F^T m> with the oracle's transforms."""
    p = Params(nx=12, ny=8, nz=10)
    o = ol.Oracle(p)
    ref = AdjointRef(p, o)
    rng = np.random.default_rng(0)
    shp = (p.nz, p.ny, p.nx // 2 + 1)
    X = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    g = rng.standard_normal((p.nz, p.ny, p.nx))
    lhs = float(np.sum(o.fft_inv(X) * g))
    rhs = float(np.sum(np.real(np.conj(X) * (ref.wx * ref.fu(g)))))
    assert lhs == pytest.approx(rhs, rel=1e-13)
    m = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    lhs = float(np.sum(np.real(np.conj(m) * o.fft_dir(g))))
    rhs = float(np.sum(g * ref.finv(m / ref.wx) / ref.ntot))
    assert lhs == pytest.approx(rhs, rel=1e-13)


@pytest.mark.parametrize("name", list(CASES))
def test_numpy_adjoint_matches_torch_autograd(name):
    """The hand derivation against torch's autograd through a PyTorch
    restatement of the step (tests/torch_step.py); in truncated runs the
    comparison is on the kept modes (the projection of torch's gradient)."""
    import torch
    from torch_step import TorchStep

    p = case_params(CASES[name])
    o = ol.Oracle(p)
    ref = AdjointRef(p, o)
    ts = TorchStep(p, o)
    b0 = initial.initial_state(p, seed=5, energy=0.5, k0=2.0)
    w = band_w(p, 2.0, 5.0)
    sh = ref.shells()
    x = torch.from_numpy(np.stack(b0).view(np.float64).reshape((3,) + b0[0].shape + (2,)).copy())
    x.requires_grad_(True)
    amps = torch.full((len(sh),), p.famp, dtype=torch.float64, requires_grad=True)
    bN = ts.run(x, amps, 17, 3, sh)
    obN, _ = o.step(b0, 17, 3)
    tb = [bN.detach().numpy()[f].view(np.complex128)[..., 0] for f in range(3)]
    assert math.sqrt(dot([a - b for a, b in zip(tb, obN)], [a - b for a, b in zip(tb, obN)])) \
        < 1e-12 * math.sqrt(dot(obN, obN))
    (torch.from_numpy(w)[..., None] * bN * bN).sum().backward()
    g0, ga = ref.gradient(b0, 17, 3, [2.0 * w * f for f in obN])
    gt = [x.grad.numpy()[f].view(np.complex128)[..., 0] for f in range(3)]
    if p.truncate == 1:
        keep = ~initial.trunc_mask(p)
        keep[0, 0, 0] = False
        gt = [np.where(keep, f, 0.0) for f in gt]
    diff = [a - b for a, b in zip(g0, gt)]
    assert math.sqrt(dot(diff, diff)) < 1e-11 * math.sqrt(dot(gt, gt))
    assert np.allclose(ga, amps.grad.numpy(), rtol=1e-10, atol=0)
