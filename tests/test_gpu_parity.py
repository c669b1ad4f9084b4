"""Parity of the HIP step (through the C ABI) with the CPU oracle.

FP64 tolerance: relative L2 <= 1e-12 over the three fields after 10 steps
(the bar of the hot path); single evaluations are held to 1e-13.
"""
import math
import sys

import numpy as np
import pytest

from gturbo3d_amd import initial, lib
from gturbo3d_amd.params import (DT2_MIDPOINT, DT2_REFERENCE, FORCING_2D2C, FORCING_3D3C, PREC_FP32_BULK,
                                 PREC_FP64, Params)
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    num = math.sqrt(sum(float(np.sum(np.abs(x - y) ** 2)) for x, y in zip(a, b)))
    den = math.sqrt(sum(float(np.sum(np.abs(y) ** 2)) for y in b))
    return num / den


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cfg(n, **kw):
    base = dict(nx=n, ny=n, nz=n, nu=1e-3, dt=5e-3, omega=1.0, famp=0.05, k1f=2.0, k2f=3.0)
    if isinstance(n, tuple):
        base.update(nx=n[0], ny=n[1], nz=n[2])
    base.update(kw)
    return Params(**base)


@pytest.mark.parametrize("n", [16, 24, 32, 48, 64, 96, 128, 256, (32, 16, 64), (64, 32, 16), (128, 24, 32), (256, 32, 16)])
def test_nonlinear_matches_oracle(require_gpu, n):
    p = cfg(n, omega=2.0)
    say(f"nonlinear {p.nx}x{p.ny}x{p.nz}")
    b = initial.initial_state(p, seed=1)
    with lib.Context(p) as g:
        g.set_state(b)
        assert g.pruned()
        N = g.nonlinear()
        assert rel_l2(g.get_state(), b) == 0.0
    No, _ = ol.Oracle(p).nlt(b, b)
    assert rel_l2(N, No) < 1e-13


@pytest.mark.parametrize("variant,dt2", [(FORCING_3D3C, DT2_REFERENCE), (FORCING_3D3C, DT2_MIDPOINT),
                                         (FORCING_2D2C, DT2_REFERENCE), (FORCING_2D2C, DT2_MIDPOINT)])
@pytest.mark.parametrize("n", [32, 64])
def test_ten_steps_match_oracle(require_gpu, n, variant, dt2):
    kw = dict(variant=variant, dt2_mode=dt2)
    if variant == FORCING_2D2C:
        kw.update(k1f=3.0, k2f=4.0, famp=0.1, omega=5.0, nu=2e-4, dt=2e-3)
    p = cfg(n, **kw)
    say(f"10 steps {n}^3 variant={variant} dt2={dt2}")
    b = initial.initial_state(p, seed=0)
    with lib.Context(p) as g:
        g.set_state(b)
        s1 = g.step(1, 8299396)
        b1 = g.get_state()
        s10 = g.step(9, s1)
        b10 = g.get_state()
    o = ol.Oracle(p)
    ob1, os1 = o.step(b, 8299396, 1)
    ob10, os10 = o.step(ob1, os1, 9)
    assert s1 == os1 and s10 == os10
    assert rel_l2(b1, ob1) < 1e-13
    assert rel_l2(b10, ob10) < 1e-12


def test_plain_64_workload_100_steps(require_gpu):
    """The plain 64^3 workload at its full length (nu = 1e-3, Omega = 1,
    3D3C forcing (2,3] at injection rate 0.1 from seed 0, dt = 5e-3, the
    reference's stage 1, 100 steps) against the oracle: fields 1e-12 after
    10 and after 100 steps (measured 2.0e-15 and 1.9e-14), energy spectrum
    and glob 1e-12, the seed equal."""
    p = cfg(64, famp=0.0)
    p.famp = math.sqrt(0.1 / lib.forcing_count(p)[0])
    say("plain 64^3 workload, 100 steps")
    b = initial.initial_state(p, seed=0)
    with lib.Context(p) as g:
        g.set_state(b)
        s10 = g.step(10, 0)
        b10 = g.get_state()
        s100 = g.step(90, s10)
        b100 = g.get_state()
        sp, _, _, glob = g.spectra()
    o = ol.Oracle(p)
    ob10, os10 = o.step(b, 0, 10)
    ob100, os100 = o.step(ob10, os10, 90)
    assert s10 == os10 and s100 == os100
    e10, e100 = rel_l2(b10, ob10), rel_l2(b100, ob100)
    osp, _ = o.spectrum(ob100)
    m = osp > 1e-12 * osp.max()
    e_sp = float(np.max(np.abs(sp[m] - osp[m]) / osp[m]))
    og = o.glob(ob100)
    e_g = float(np.max(np.abs(glob - og) / np.maximum(np.abs(og), 1e-12 * np.abs(og).max())))
    say(f"  fields after 10 steps {e10:.2e}, after 100 {e100:.2e}; energy spectrum {e_sp:.2e}; glob {e_g:.2e}")
    assert e10 < 1e-12
    assert e100 < 1e-12
    assert e_sp < 1e-12 and e_g < 1e-12


@pytest.mark.parametrize("n", [128, 256])
def test_one_step_large_matches_oracle(require_gpu, n):
    p = cfg(n, nu=2e-4, dt=2e-3, omega=5.0)
    say(f"1 step {n}^3")
    b = initial.initial_state(p, seed=3)
    with lib.Context(p) as g:
        g.set_state(b)
        s = g.step(1, 7)
        b1 = g.get_state()
    ob1, os1 = ol.Oracle(p).step(b, 7, 1)
    assert s == os1
    assert rel_l2(b1, ob1) < 1e-13


def test_nonlinear_512_matches_oracle(require_gpu):
    """The full 512^3 shape against the oracle (the x pass runs three
    wavefronts per row pair at this size)."""
    p = cfg(512, omega=2.0, nu=5e-5, dt=1e-3)
    say("nonlinear 512^3")
    b = initial.initial_state(p, seed=5)
    with lib.Context(p) as g:
        g.set_state(b)
        assert g.pruned()
        N = g.nonlinear()
        # other pencil widths and chunked passes B-D give the same bits
        g.set_tiles(16, 16, 16)
        g.set_zchunk(32)
        N4 = g.nonlinear()
    assert all(np.array_equal(x, y) for x, y in zip(N, N4))
    No, _ = ol.Oracle(p).nlt(b, b)
    err = rel_l2(N, No)
    say(f"  relative L2 {err:.2e}")
    assert err < 1e-13


def test_dense_path_without_truncation(require_gpu):
    p = cfg(32, truncate=0, omega=1.0)
    say("dense path truncate=0")
    rng = np.random.default_rng(4)
    b = initial.initial_state(cfg(32), seed=4)
    # a little energy in modes the 2/3 rule would remove (the transform of a real field)
    noise = [np.fft.rfftn(1e-4 * rng.standard_normal((32, 32, 32))) / 32 ** 3 for _ in range(3)]
    b = [x + y for x, y in zip(b, noise)]
    with lib.Context(p) as g:
        g.set_state(b)
        assert not g.pruned()
        N = g.nonlinear()
        s = g.step(3, 11)
        b3 = g.get_state()
    o = ol.Oracle(p)
    No, _ = o.nlt(b, b)
    ob3, os3 = o.step(b, 11, 3)
    assert rel_l2(N, No) < 1e-13
    assert s == os3
    assert rel_l2(b3, ob3) < 1e-12


def test_untruncated_state_with_truncation_uses_dense_path(require_gpu):
    p = cfg(32, omega=1.0)
    say("untruncated state, truncate=1")
    b = initial.initial_state(p, seed=6)
    b[0] = b[0].copy()
    b[0][0, 0, 15] = 1e-3  # kx = 15 is outside the ellipsoid at 32^3
    with lib.Context(p) as g:
        g.set_state(b)
        assert not g.pruned()
        s = g.step(2, 5)
        b2 = g.get_state()
    ob2, os2 = ol.Oracle(p).step(b, 5, 2)
    assert s == os2
    assert rel_l2(b2, ob2) < 1e-12


def test_reference_default_physics(require_gpu):
    # src/params.dat:1-6 physics (hyperviscosity k^8, friction k^-2, thin layer) on 64 x 64 x 16
    p = Params(nx=64, ny=64, nz=16, xlz=0.09817477042468103, nu=1e-17, alpnu=4.0, mu=0.6, alpmu=-1.0,
               dt=1e-4, omega=0.0, famp=0.1, k1f=15.0, k2f=17.0)
    say("reference default physics 64x64x16")
    b = initial.initial_state(p, seed=2, k0=8.0)
    with lib.Context(p) as g:
        nf = len(g.forcing_table()[0])
        g.set_state(b)
        s = g.step(5, 8299396)
        b5 = g.get_state()
    o = ol.Oracle(p)
    assert nf == o.nf
    ob5, os5 = o.step(b, 8299396, 5)
    assert s == os5
    assert rel_l2(b5, ob5) < 1e-12


def test_forcing_table_matches_oracle(require_gpu):
    p = cfg(64)
    with lib.Context(p) as g:
        got = g.forcing_table()
    exp = ol.Oracle(p).forcing_table()
    for a, e in zip(got[:4], exp[:4]):
        assert np.array_equal(a, e)
    assert got[4] == exp[4] and got[5] == exp[5]


@pytest.mark.parametrize("n", [16, 32, 96])
def test_velocity_matches_oracle(require_gpu, n):
    p = cfg(n)
    say(f"velocity {n}^3")
    b = initial.initial_state(p, seed=7)
    with lib.Context(p) as g:
        g.set_state(b)
        u = g.velocity()
    uo = ol.Oracle(p).velocity(b)
    assert rel_l2(u, uo) < 1e-13


@pytest.mark.parametrize("n", [16, 32, 96])
def test_init_velocity_matches_oracle(require_gpu, n):
    p = cfg(n)
    say(f"init from velocity {n}^3")
    rng = np.random.default_rng(8)
    u = [rng.standard_normal((n, n, n)) for _ in range(3)]
    with lib.Context(p) as g:
        g.init_velocity(u)
        assert g.pruned()
        b = g.get_state()
    bo = ol.Oracle(p).init_from_velocity(u)
    assert rel_l2(b, bo) < 1e-13


def lowmode_state(p, seed, kmax):
    """Solenoidal random b on the interior kx >= 1 modes with |k| <= kmax (no
    Hermitian constraint there), built without a 3D FFT."""
    rng = np.random.default_rng(seed)
    fkx, fky, fkz = initial.wavenumbers(p)
    b = [np.zeros(p.shape, np.complex128) for _ in range(3)]
    ks = [k for k in range(p.nz) if abs(fkz[k]) <= kmax]
    js = [j for j in range(p.ny) if abs(fky[j]) <= kmax]
    for k in ks:
        for j in js:
            for i in range(1, kmax + 1):
                kv = np.array([fkx[i], fky[j], fkz[k]])
                k2 = kv @ kv
                if k2 > kmax * kmax:
                    continue
                v = rng.standard_normal(3) + 1j * rng.standard_normal(3)
                v -= kv * (kv @ v) / k2
                v *= np.exp(-k2 / 16.0)
                for c in range(3):
                    b[c][k, j, i] = v[c]
    return b


def test_energy_conservation_at_512(require_gpu):
    """Size-independent property at the full 512^3 shape: with nu = 0, no forcing,
    N conserves the energy sum fac k^2 Re(conj(b) N) = 0 (Galerkin truncation)."""
    p = cfg(512, omega=2.0, nu=5e-5, dt=1e-3)
    say("energy conservation 512^3")
    b = lowmode_state(p, seed=0, kmax=12)
    with lib.Context(p) as g:
        g.set_state(b)
        N = g.nonlinear()
    k2 = initial.k_squared(p)
    fac = np.ones(p.nxh)
    fac[0] = fac[-1] = 0.5
    dE = sum(float(np.sum(fac * k2 * np.real(np.conj(x) * y))) for x, y in zip(b, N))
    sc = sum(float(np.sum(fac * k2 * np.abs(x) * np.abs(y))) for x, y in zip(b, N))
    assert abs(dE) < 1e-12 * sc
    assert all(np.all(f[initial.trunc_mask(p)] == 0) for f in N)


def test_state_round_trip_and_seed_only_forcing(require_gpu):
    p = cfg(16, famp=0.0)
    b = initial.initial_state(p, seed=1)
    with lib.Context(p) as g:
        g.set_state(b)
        assert rel_l2(g.get_state(), b) == 0.0
        s = g.step(0, 42)
        assert s == 42
    # famp = 0 still draws two numbers per forced mode
    with lib.Context(p) as g:
        g.set_state(b)
        s = g.step(1, 42)
    assert s == ol.rann_sequence(42, 2 * ol.Oracle(p).nf)[1]



@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32_BULK])
def test_quadratic_support_and_energy_at_1024(require_gpu, prec):
    """The largest compiled size, through size-independent properties of N(b)
    for a state b on the modes |k| <= 10: N is quadratic, so it lives on
    |k_i| <= 20 up to transform roundoff (a wrong twiddle, index or stride in
    any pass spreads O(1) energy outside), and it conserves the energy
    sum fac k^2 Re(conj(b) N) = 0 (Galerkin truncation)."""
    kmax = 10
    p = cfg(1024, omega=2.0, nu=5e-5, dt=1e-3, precision=prec)
    say(f"quadratic support and energy conservation 1024^3, precision {prec}")
    b = lowmode_state(p, seed=1, kmax=kmax)
    with lib.Context(p) as g:
        g.set_state(b)
        assert g.pruned()
        N = g.nonlinear()
    q = 2 * kmax
    zin = np.r_[0:q + 1, p.nz - q:p.nz]
    yin = np.r_[0:q + 1, p.ny - q:p.ny]
    e_in = e_out = 0.0
    for f in N:
        e_all = 0.0
        # |kz| > q: contiguous slabs, no copies
        e_out += float(np.vdot(f[q + 1:p.nz - q].ravel(), f[q + 1:p.nz - q].ravel()).real)
        for k in zin:
            e_out += float(np.vdot(f[k, q + 1:p.ny - q].ravel(), f[k, q + 1:p.ny - q].ravel()).real)
            blk = f[k][yin]
            e_out += float(np.sum(np.abs(blk[:, q + 1:]) ** 2))
            e_all += float(np.sum(np.abs(blk[:, :q + 1]) ** 2))
        e_in += e_all
    ratio = e_out / e_in
    say(f"  energy of N outside |k_i| <= {q}: {ratio:.2e} of the inside")
    assert e_in > 0 and ratio < (1e-24 if prec == PREC_FP64 else 1e-10)
    # energy: b vanishes outside its box, so the sum runs over that box
    fkx, fky, fkz = initial.wavenumbers(p)
    kb = np.r_[0:kmax + 1, p.nz - kmax:p.nz]
    jb = np.r_[0:kmax + 1, p.ny - kmax:p.ny]
    k2 = (fkx[None, None, :kmax + 1] ** 2 + fky[jb][None, :, None] ** 2) + fkz[kb][:, None, None] ** 2
    fac = np.ones(kmax + 1)
    fac[0] = 0.5
    dE = sc = 0.0
    for x, y in zip(b, N):
        xs, ys = x[np.ix_(kb, jb, np.arange(kmax + 1))], y[np.ix_(kb, jb, np.arange(kmax + 1))]
        dE += float(np.sum(fac * k2 * np.real(np.conj(xs) * ys)))
        sc += float(np.sum(fac * k2 * np.abs(xs) * np.abs(ys)))
    say(f"  energy transfer {abs(dE) / sc:.2e} of its scale")
    assert sc > 0 and abs(dE) < (1e-12 if prec == PREC_FP64 else 1e-6) * sc
    # the kx planes past the truncation ellipsoid are zeroed
    kxcut = int(math.floor(p.rtrunc * (p.nxh - 1)))
    assert all(not np.any(f[zin][:, :, kxcut + 1:]) for f in N)
