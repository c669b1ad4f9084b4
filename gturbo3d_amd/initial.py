"""Seeded initial conditions and the wavenumber tables they need.

The reference starts either from a restart (src/inieuler.f90:76-77) or from
``init_noise`` with an uninitialised seed (src/inieuler.f90:11,84), which
is not reproducible.  This module builds the reproducible initial state the
benchmark configurations use: a solenoidal, truncated random velocity with
energy spectrum E(k) ~ k^4 exp(-2 (k/k0)^2), normalised so the energy of
``glob`` (src/diagnostics.f90:131-153, sum of fac*|u|^2 with fac = 1/2 on
the kx = 0 and kx = NX/2 planes) equals ``energy``, returned as the vector
potential b = invcurl(u) that the step evolves.  The random numbers come
from numpy's PCG64 seeded with the configuration seed.
"""
from __future__ import annotations

import numpy as np

from .params import Params

PI_REF = 3.14159265358979323846  # the literal of src/defk.f90:24-26


def wavenumbers(p: Params):
    """fkx (NX/2+1), fky (NY), fkz (NZ) exactly as defk computes them (src/defk.f90:24-43)."""
    dkx = 2.0 * PI_REF / p.xlx
    dky = 2.0 * PI_REF / p.xly
    dkz = 2.0 * PI_REF / p.xlz
    fkx = np.arange(p.nxh, dtype=np.float64) * dkx
    jy = np.arange(p.ny)
    fky = np.where(jy <= p.ny // 2, jy, jy - p.ny).astype(np.float64) * dky
    kz = np.arange(p.nz)
    fkz = np.where(kz <= p.nz // 2, kz, kz - p.nz).astype(np.float64) * dkz
    return fkx, fky, fkz


def k_squared(p: Params):
    fkx, fky, fkz = wavenumbers(p)
    return (fkx[None, None, :] ** 2 + fky[None, :, None] ** 2) + fkz[:, None, None] ** 2


def trunc_mask(p: Params):
    """True where trunc zeroes a mode (src/trunc.f90:14-17, rtrunc squared at src/defk.f90:91)."""
    fkx, fky, fkz = wavenumbers(p)
    fxs, fys, fzs = fkx * fkx, fky * fky, fkz * fkz
    k3 = (fzs / fzs[p.nz // 2])[:, None, None]
    k2 = k3 + (fys / fys[p.ny // 2])[None, :, None]
    k1 = k2 + (fxs / fxs[p.nxh - 1])[None, None, :]
    return k1 > p.rtrunc * p.rtrunc


def glob_energy(p: Params, u_hat) -> float:
    """sum fac*|u|^2 over k != 0 (src/diagnostics.f90:131-153)."""
    fac = np.ones(p.nxh)
    fac[0] = 0.5
    fac[-1] = 0.5
    e = 0.0
    for f in u_hat:
        e += float(np.sum(fac[None, None, :] * np.abs(f) ** 2))
    return e


def curl(p: Params, a):
    fkx, fky, fkz = wavenumbers(p)
    kx, ky, kz = fkx[None, None, :], fky[None, :, None], fkz[:, None, None]
    ax, ay, az = a
    return [1j * (ky * az - kz * ay), 1j * (kz * ax - kx * az), 1j * (kx * ay - ky * ax)]


def invcurl(p: Params, a):
    k2 = k_squared(p)
    k2s = np.where(k2 == 0.0, 1.0, k2)
    c = curl(p, a)
    out = [np.where(k2 == 0.0, x, y / k2s) for x, y in zip(a, c)]
    return out


def _box(p: Params, kcut: int):
    mx = min(kcut, p.nx // 2)
    my = min(kcut, p.ny // 2 - 1)
    mz = min(kcut, p.nz // 2 - 1)
    iy = np.array([j % p.ny for j in range(-my, my + 1)])
    iz = np.array([k % p.nz for k in range(-mz, mz + 1)])
    ix = np.arange(mx + 1)
    return ix, iy, iz


def _box_velocity(p: Params, seed: int, energy: float, k0: float, kcut: int):
    """The velocity on the box |kx|, |ky|, |kz| <= kcut (indices ix, iy, iz)."""
    rng = np.random.default_rng(seed)
    fkx, fky, fkz = wavenumbers(p)
    ix, iy, iz = _box(p, kcut)
    shape_box = (len(iz), len(iy), len(ix))
    ub = [rng.standard_normal(shape_box) + 1j * rng.standard_normal(shape_box) for _ in range(3)]
    kx = fkx[ix][None, None, :]
    ky = fky[iy][None, :, None]
    kz = fkz[iz][:, None, None]
    k2 = (kx * kx + ky * ky) + kz * kz
    k = np.sqrt(k2)
    amp = k * np.exp(-((k / k0) ** 2))
    ub = [f * amp for f in ub]
    k2s = np.where(k2 == 0.0, 1.0, k2)
    div = kx * ub[0] + ky * ub[1] + kz * ub[2]
    ub = [f - kk * div / k2s for f, kk in zip(ub, (kx, ky, kz))]
    # the kx = 0 and kx = NX/2 planes of a real field are Hermitian:
    # u(kx, -ky, -kz) = conj u(kx, ky, kz); the box is symmetric in ky and kz
    ny_b, nz_b = len(iy), len(iz)
    sy = np.arange(ny_b) - ny_b // 2
    sz = np.arange(nz_b) - nz_b // 2
    upper = (sy[None, :] > 0) | ((sy[None, :] == 0) & (sz[:, None] > 0))
    for f in ub:
        pl = f[:, :, 0]
        mirror = np.conj(pl[::-1, ::-1])
        pl[~upper] = mirror[~upper]
        f[:, :, 0] = pl
        # curl cannot keep the kx = NX/2 plane Hermitian (its +NX/2 wavenumber
        # has no negative partner in the half spectrum), so it stays empty
        if ix[-1] == p.nx // 2:
            f[:, :, -1] = 0.0
    if p.truncate == 1:
        fxs, fys, fzs = fkx * fkx, fky * fky, fkz * fkz
        k3 = (fzs[iz] / fzs[p.nz // 2])[:, None, None]
        kk2 = k3 + (fys[iy] / fys[p.ny // 2])[None, :, None]
        k1 = kk2 + (fxs[ix] / fxs[p.nxh - 1])[None, None, :]
        mask = k1 > p.rtrunc * p.rtrunc
        for f in ub:
            f[mask] = 0.0
    for f in ub:
        f[nz_b // 2, ny_b // 2, 0] = 0.0
    fac = np.ones(len(ix))
    fac[0] = 0.5
    if ix[-1] == p.nx // 2:
        fac[-1] = 0.5
    e = sum(float(np.sum(fac[None, None, :] * np.abs(f) ** 2)) for f in ub)
    if e > 0:
        ub = [f * np.sqrt(energy / e) for f in ub]
    return ub, (ix, iy, iz), (kx, ky, kz, k2)


def _scatter(p: Params, fields, idx):
    ix, iy, iz = idx
    out = []
    for f in fields:
        g = np.zeros(p.shape, np.complex128)
        g[np.ix_(iz, iy, ix)] = f
        out.append(g)
    return out


def initial_velocity_hat(p: Params, seed: int, energy: float = 0.5, k0: float = 4.0, kcut: int = 32):
    """Solenoidal, truncated, Hermitian-consistent random velocity (spectral).

    Random complex vectors are drawn (PCG64, ``seed``) for the wavenumbers
    |kx|, |ky|, |kz| <= kcut only (the spectrum k^4 exp(-2 (k/k0)^2) is below
    1e-25 of its peak there for k0 = 4), so the realisation does not depend
    on the grid once the grid holds that box, and no 3D FFT is needed.
    """
    ub, idx, _ = _box_velocity(p, seed, energy, k0, kcut)
    return _scatter(p, ub, idx)


def initial_state(p: Params, seed: int, energy: float = 0.5, k0: float = 4.0, kcut: int = 32):
    """Vector potential b = invcurl(u) of initial_velocity_hat (3 complex fields)."""
    ub, idx, (kx, ky, kz, k2) = _box_velocity(p, seed, energy, k0, kcut)
    k2s = np.where(k2 == 0.0, 1.0, k2)
    bx = np.where(k2 == 0.0, ub[0], 1j * (ky * ub[2] - kz * ub[1]) / k2s)
    by = np.where(k2 == 0.0, ub[1], 1j * (kz * ub[0] - kx * ub[2]) / k2s)
    bz = np.where(k2 == 0.0, ub[2], 1j * (kx * ub[1] - ky * ub[0]) / k2s)
    return _scatter(p, [bx, by, bz], idx)
