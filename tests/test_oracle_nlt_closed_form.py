"""Closed-form pins of the oracle's nonlinear term and of the step around it.

nlt (src/nlt.f90:16-89) forms u = F^-1[curl a] (src/produ.f90:34-66) and
w = F^-1[k^2 a] (src/nlt.f90:26-40), the product u x w (src/produ.f90:9-28),
then N = trunc(invcurl(F[u x w] + i*omega*kz*b)) (src/nlt.f90:67-89,
src/produ.f90:69-104).  The invariants in test_oracle_ops.py (Beltrami zero,
energy conservation) hold for any vorticity factor and either orientation of
the cross product, so these tests fix them with fields whose u, w, u x w and
N are written out by hand in physical space:

  a = (sin z, 0, cos 2y): u = curl a = (-2 sin 2y, cos z, 0),
      w = curl u = (sin z, 0, 4 cos 2y) (= k^2 a mode by mode),
      u x w = (4 cos z cos 2y, 8 sin 2y cos 2y, -sin z cos z).
  The y and z components are gradients, which invcurl removes; the x
  component gives, with invcurl(X) = i k x X / |k|^2 on the four modes
  (0, +-2, +-1),
      N = (0, -0.8 sin z cos 2y, 1.6 cos z sin 2y).
  A vorticity k^p a instead of k^2 a scales N by 2^(p-2), -k^2 a or w x u
  flips its sign, and a wrong transform scale changes its size.

  Coriolis with b = (cos z, sin z, 0) (curl b = b): i*omega*kz*b is
  F[omega * d/dz b] = omega F[(-sin z, cos z, 0)], whose curl is minus
  itself, so invcurl adds omega * (sin z, -cos z, 0) to N.

The expectations use numpy's FFT (rfftn scaled by 1/(NX NY NZ), the fft_dir
normalisation of src/fft_dir.f90:16-26) and the analytic fields only; no oracle
routine enters them.
"""
import math

import numpy as np
import pytest

from gturbo3d_amd import initial
from gturbo3d_amd.params import DT2_MIDPOINT, DT2_REFERENCE, Params
import oracle_lib as ol


def _grid(p):
    z, y, x = np.meshgrid(np.arange(p.nz) * p.xlz / p.nz, np.arange(p.ny) * p.xly / p.ny,
                          np.arange(p.nx) * p.xlx / p.nx, indexing="ij")
    return x, y, z


def _spec(p, f):
    """fft_dir of a physical field by numpy: forward DFT / (NX NY NZ)."""
    return np.fft.rfftn(f) / (p.nx * p.ny * p.nz)


def _two_shear(p):
    x, y, z = _grid(p)
    zero = 0.0 * x
    a = [np.sin(z), zero, np.cos(2 * y)]
    N = [zero, -0.8 * np.sin(z) * np.cos(2 * y), 1.6 * np.cos(z) * np.sin(2 * y)]
    return a, N


def _rel(got, expect):
    num = math.sqrt(sum(float(np.sum(np.abs(g - e) ** 2)) for g, e in zip(got, expect)))
    den = math.sqrt(sum(float(np.sum(np.abs(e) ** 2)) for e in expect))
    return num / den


def err_two_shear(make, n=16):
    """Relative errors of N and of v = F[u x w] against the two-shear closed form."""
    p = Params(nx=n, ny=n, nz=n, omega=0.0)
    o = make(p)
    a, N_phys = _two_shear(p)
    a_hat = [_spec(p, f) for f in a]
    N, v = o.nlt(a_hat, a_hat)
    x, y, z = _grid(p)
    uxw = [4 * np.cos(z) * np.cos(2 * y), 8 * np.sin(2 * y) * np.cos(2 * y), -np.sin(z) * np.cos(z)]
    return _rel(N, [_spec(p, f) for f in N_phys]), _rel(v, [_spec(p, f) for f in uxw])


def err_coriolis(make, omega=1.5):
    """Relative errors of nlt(a; b) and nlt(b; a) against their closed forms."""
    p = Params(nx=16, ny=16, nz=16, omega=omega)
    o = make(p)
    x, y, z = _grid(p)
    a, N_phys = _two_shear(p)
    b = [np.cos(z), np.sin(z), 0.0 * z]
    N, _ = o.nlt([_spec(p, f) for f in a], [_spec(p, f) for f in b])
    cor = [omega * np.sin(z), -omega * np.cos(z), 0.0 * z]
    e1 = _rel(N, [_spec(p, f + c) for f, c in zip(N_phys, cor)])
    # with the roles exchanged the Coriolis term follows the second argument:
    # b is Beltrami (u x w = 0); i omega kz a = omega F[(cos z, 0, 0)], whose
    # invcurl is omega F[curl (cos z, 0, 0)] / 1 = omega (0, -sin z, 0)
    N2, _ = o.nlt([_spec(p, f) for f in b], [_spec(p, f) for f in a])
    e2 = _rel(N2, [_spec(p, f) for f in (0.0 * z, -omega * np.sin(z), 0.0 * z)])
    return e1, e2


def err_helicity(make, seed=3):
    """|sum fac k^2 Re(conj(u) . N)| relative to sum fac k^2 |u| |N|, omega = 0."""
    p = Params(nx=32, ny=32, nz=32, omega=0.0)
    o = make(p)
    b = initial.initial_state(p, seed=seed)
    # a helical part, so that the transfer terms do not cancel by symmetry
    b = [bb + 0.5 * cc for bb, cc in zip(b, o.invcurl(b))]
    b = [o.trunc(f) for f in b]
    N, _ = o.nlt(b, b)
    k2 = initial.k_squared(p)
    fac = np.ones(p.nxh)
    fac[0] = fac[-1] = 0.5
    u = o.curl(b)
    dH = sum(np.sum(fac * k2 * np.real(np.conj(uu) * nn)) for uu, nn in zip(u, N))
    scale = sum(np.sum(fac * k2 * np.abs(uu) * np.abs(nn)) for uu, nn in zip(u, N))
    return abs(dH) / scale


BELTRAMI = dict(A=1.0, B=0.7, C=0.4, nu=0.5, dt=0.01, omega=3.0)


def err_beltrami_step(make, dt2_mode):
    """Relative error of one step of a rotating ABC flow against its closed form."""
    A, B, C = BELTRAMI["A"], BELTRAMI["B"], BELTRAMI["C"]
    nu, dt, omega = BELTRAMI["nu"], BELTRAMI["dt"], BELTRAMI["omega"]
    p = Params(nx=16, ny=16, nz=16, nu=nu, dt=dt, omega=omega, famp=0.0, dt2_mode=dt2_mode)
    o = make(p)
    x, y, z = _grid(p)
    b = [A * np.sin(z) + C * np.cos(y), B * np.sin(x) + A * np.cos(z), C * np.sin(y) + B * np.cos(x)]
    dbdz = [A * np.cos(z), -A * np.sin(z), 0.0 * z]
    e = math.exp(-0.5 * dt * nu)
    out, _ = o.step([_spec(p, f) for f in b], seed=1, nsteps=1)
    return _rel(out, [_spec(p, e * e * f + e * dt * omega * g) for f, g in zip(b, dbdz)])


def all_errors(make):
    """Every closed-form error of this module for one oracle build (name -> error)."""
    e = {}
    e["two_shear_N"], e["two_shear_v"] = err_two_shear(make)
    e["coriolis_ab"], e["coriolis_ba"] = err_coriolis(make)
    e["helicity"] = err_helicity(make)
    e["beltrami_step_ref"] = err_beltrami_step(make, DT2_REFERENCE)
    e["beltrami_step_mid"] = err_beltrami_step(make, DT2_MIDPOINT)
    return e


# the bar every closed form meets on the oracle (FP64 rounding only)
TOL = 1e-13


@pytest.mark.parametrize("n", [16, 24])
def test_nlt_two_shear_modes_closed_form(n):
    eN, ev = err_two_shear(ol.Oracle, n)
    assert eN < TOL
    assert ev < TOL


def test_nlt_closed_form_rejects_plausible_mistakes():
    """The closed form separates the reference's nlt from its likely misreadings,
    rebuilt here from the oracle's separately pinned pieces."""
    p = Params(nx=16, ny=16, nz=16, omega=0.0)
    o = ol.Oracle(p)
    a, N_phys = _two_shear(p)
    expect = [_spec(p, f) for f in N_phys]
    x, y, z = _grid(p)
    u = [-2 * np.sin(2 * y), np.cos(z), 0.0 * z]

    def n_of(w, flip=False):
        c = o.cross(w, u) if flip else o.cross(u, w)
        return [o.trunc(f) for f in o.invcurl([o.fft_dir(q) for q in c])]

    w_k2 = [np.sin(z), 0.0 * z, 4 * np.cos(2 * y)]
    w_k1 = [np.sin(z), 0.0 * z, 2 * np.cos(2 * y)]       # |k| a instead of k^2 a
    assert _rel(n_of(w_k2), expect) < TOL
    assert _rel(n_of(w_k1), expect) > 0.4
    assert _rel(n_of(w_k2, flip=True), expect) > 1.9    # w x u
    assert _rel([-f for f in n_of(w_k2)], expect) > 1.9  # -k^2 a


def test_nlt_coriolis_uses_second_argument_closed_form():
    e1, e2 = err_coriolis(ol.Oracle)
    assert e1 < TOL
    assert e2 < TOL


@pytest.mark.parametrize("seed", [3, 11])
def test_nlt_conserves_helicity_without_rotation(seed):
    """Sum fac k^2 Re(conj(u) . N) = 0 for a truncated b at omega = 0.

    With u = curl b and N solenoidal, dH/dt = 2 sum fac Re(conj(N) . curl curl u)
    = 2 sum fac k^2 Re(conj(u) . N); helicity is an invariant of u x w, and
    of its Galerkin truncation when b is truncated."""
    assert err_helicity(ol.Oracle, seed) < TOL


@pytest.mark.parametrize("dt2_mode", [DT2_REFERENCE, DT2_MIDPOINT])
def test_step_rotating_beltrami_closed_form(dt2_mode):
    """One step of an ABC flow (|k| = 1, curl b = b, so u x w = 0 at every stage).

    With e = exp(-dt nu / 2) the integrating factor on |k| = 1
    (src/defk.f90:57-68) and N(.; b) = invcurl(i omega kz b) = omega d/dz b for
    this b, both stage-1 modes give b+ = e^2 b + e dt omega d/dz b
    (src/rkstep.f90:30-40,86-96 with the stage-2 Coriolis on the old b,
    src/main.f90:144).  A Coriolis term on bh = e b would carry e^2 instead
    of e; emk applied once, or outside the dt term, changes the decay."""
    assert err_beltrami_step(ol.Oracle, dt2_mode) < TOL
