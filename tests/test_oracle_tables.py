"""Pins the oracle's host tables: the LCG, defk, iniforcing and trunc.

Each expectation comes from a closed form or an independent brute-force
enumeration, never from re-running the oracle's own formula.
"""
import math

import numpy as np
import pytest

from gturbo3d_amd.params import FORCING_2D2C, FORCING_3D3C, Params
from gturbo3d_amd import initial
import oracle_lib as ol


def lcg_python(seed, n):
    """(69069 s + 1) mod 2^31 with Python integers (src/forcing_3d3c.f90:88-90)."""
    out = []
    s = seed
    for _ in range(n):
        s = (69069 * s + 1) % (2 ** 31)
        out.append(s)
    return out


def test_lcg_closed_form():
    vals, final = ol.rann_sequence(8299396, 1000)
    states = lcg_python(8299396, 1000)
    assert final == states[-1]
    expect = np.array([np.float32(s / (2 ** 31 - 1)) for s in states], dtype=np.float32)
    assert np.array_equal(vals, expect)


def test_lcg_known_values():
    # reference seed file src/seed.0:1 = 8299396
    assert lcg_python(8299396, 3) == [2000331957, 419960306, 176741579]
    vals, _ = ol.rann_sequence(8299396, 3)
    assert np.array_equal(vals, np.array([0.9314771, 0.19555926, 0.08230171], dtype=np.float32))


def test_lcg_wraps_for_negative_products():
    # states near 2^31 - 1 overflow 32-bit signed arithmetic in the reference
    vals, final = ol.rann_sequence(2 ** 31 - 2, 5)
    assert final == lcg_python(2 ** 31 - 2, 5)[-1]


def _params(n, variant=FORCING_3D3C, k1=2.0, k2=3.0, **kw):
    return Params(nx=n, ny=n, nz=n, variant=variant, k1f=k1, k2f=k2, famp=0.05, **kw)


def brute_forced_set(p):
    """All half-space modes with k1f^2 < k^2 <= k2f^2 inside the trunc test, as a set."""
    fkx, fky, fkz = initial.wavenumbers(p)
    fxs, fys, fzs = fkx ** 2, fky ** 2, fkz ** 2
    out = set()
    for k in range(p.nz):
        for j in range(p.ny):
            for i in range(p.nxh):
                if p.variant == FORCING_2D2C and k != 0:
                    continue
                # half space: kx > 0, or kx in {0, N/2} with (ky, kz) in the upper half
                if i in (0, p.nxh - 1):
                    ky = j if j <= p.ny // 2 else j - p.ny
                    kz = k if k <= p.nz // 2 else k - p.nz
                    if p.variant == FORCING_3D3C:
                        upper = (ky > 0 and ky < p.ny // 2) or (ky == 0 and 0 < kz < p.nz // 2) or \
                                (ky == p.ny // 2 and 0 < kz < p.nz // 2)
                    else:
                        upper = 0 < ky < p.ny // 2
                    if not upper:
                        continue
                k2 = fxs[i] + fys[j] + fzs[k]
                k2n = fxs[i] / fxs[-1] + fys[j] / fys[p.ny // 2] + fzs[k] / fzs[p.nz // 2]
                if k2n <= p.rtrunc ** 2 and p.k1f ** 2 < k2 <= p.k2f ** 2:
                    out.add((i + 1, j + 1, k + 1))
    return out


@pytest.mark.parametrize("n,variant,k1,k2,nf", [
    (16, FORCING_3D3C, 2.0, 3.0, 45),
    (32, FORCING_3D3C, 2.0, 3.0, 45),
    (64, FORCING_3D3C, 2.0, 3.0, 45),
    (32, FORCING_2D2C, 3.0, 4.0, 10),
    (64, FORCING_2D2C, 3.0, 4.0, 10),
])
def test_iniforcing_matches_brute_force(n, variant, k1, k2, nf):
    p = _params(n, variant, k1, k2)
    o = ol.Oracle(p)
    ikx, iky, ikz, ff, inpe, inpz = o.forcing_table()
    assert len(ikx) == nf
    got = set(zip(ikx.tolist(), iky.tolist(), ikz.tolist()))
    assert got == brute_forced_set(p)
    # ff = 1/|k| and inpe = nf * famp^2 (ff*|k| = 1)
    fkx, fky, fkz = initial.wavenumbers(p)
    kk = np.sqrt(fkx[ikx - 1] ** 2 + fky[iky - 1] ** 2 + fkz[ikz - 1] ** 2)
    assert np.allclose(ff * kk, 1.0, rtol=1e-15)
    assert inpe == pytest.approx(nf * p.famp ** 2, rel=1e-12)


def test_iniforcing_order_3d3c():
    """Reference loop order: kx = 0 plane (ky = 0 row, ky > 0 rows, Nyquist row), interior kx
    ascending with ky then kz ascending, then the kx = Nyquist plane (iniforcing_3d3c.f90:22-184)."""
    p = _params(32)
    ikx, iky, ikz, *_ = ol.Oracle(p).forcing_table()
    key = []
    for i, j, k in zip(ikx, iky, ikz):
        plane = 0 if i == 1 else (2 if i == p.nxh else 1)
        if plane != 1:
            row = 0 if j == 1 else (2 if j == p.ny // 2 + 1 else 1)
            key.append((plane, 0, row, j, k))
        else:
            key.append((plane, i, 0, j, k))
    assert key == sorted(key)


def test_default_reference_grid_has_96_forced_modes():
    # src/params.dat:1-6 on the compile-time grid 1024 x 1024 x 16 (src/modules.f90:4)
    p = Params(nx=1024, ny=1024, nz=16, xlx=6.2831853071795864769, xly=6.2831853071795864769,
               xlz=0.09817477042468103, nu=1e-17, alpnu=4.0, mu=0.6, alpmu=-1.0, dt=1e-4, omega=0.0,
               rtrunc=0.6666666666666666667, truncate=1, famp=0.1, k1f=15.0, k2f=17.0)
    assert ol.Oracle(p).nf == 96


def test_wavenumber_tables_exact():
    p = Params(nx=16, ny=24, nz=32, xlx=2.0, xly=3.0, xlz=6.283185307179586)
    t = ol.Oracle(p).tables()
    fkx, fky, fkz = initial.wavenumbers(p)
    assert np.array_equal(t["fkx"], fkx)
    assert np.array_equal(t["fky"], fky)
    assert np.array_equal(t["fkz"], fkz)
    # fky(j) = -fky(jc(j)) and the Nyquist row carries +NY/2 (src/defk.f90:32-37)
    assert t["fky"][12] == 12 * (2 * math.pi / 3.0) or t["fky"][12] == pytest.approx(12 * 2 * math.pi / 3.0)
    assert np.array_equal(t["fky"][1:12], -t["fky"][:12:-1])


@pytest.mark.parametrize("alpnu,mu,alpmu", [(1.0, 0.0, 0.0), (4.0, 0.6, -1.0)])
def test_emk_closed_form(alpnu, mu, alpmu):
    p = Params(nx=16, ny=16, nz=16, nu=1e-3 if alpnu == 1.0 else 1e-9, alpnu=alpnu, mu=mu, alpmu=alpmu,
               dt=5e-3)
    emk = ol.Oracle(p).tables()["emk"]
    k2 = initial.k_squared(p)
    with np.errstate(divide="ignore"):
        rate = p.nu * k2 ** alpnu + p.mu * np.where(k2 > 0, k2, 1.0) ** alpmu
    expect = np.exp(-0.5 * p.dt * np.where(k2 > 0, rate, 0.0))
    assert emk[0, 0, 0] == 1.0
    assert np.allclose(emk, expect, rtol=2e-15, atol=0)


@pytest.mark.parametrize("n", [16, 24, 32, 96])
def test_trunc_mask_brute_force(n):
    p = Params(nx=n, ny=n, nz=n)
    m = ol.Oracle(p).mask()
    fkx, fky, fkz = initial.wavenumbers(p)
    expect = np.zeros(p.shape, bool)
    for k in range(p.nz):
        for j in range(p.ny):
            for i in range(p.nxh):
                k3 = fkz[k] ** 2 / fkz[p.nz // 2] ** 2
                k2 = k3 + fky[j] ** 2 / fky[p.ny // 2] ** 2
                k1 = k2 + fkx[i] ** 2 / fkx[-1] ** 2
                expect[k, j, i] = k1 > p.rtrunc * p.rtrunc
    assert np.array_equal(m, expect)
    kept = 1.0 - m.mean()
    # the ellipsoid (2/3)^3 * pi/6 of the box, within lattice effects
    assert abs(kept - math.pi / 6 * (2 / 3) ** 3 * 8 / 8 * 1.0) < 0.06
