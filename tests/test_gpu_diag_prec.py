"""Device diagnostics against the oracle, and the FP32-bulk precision variant
against the FP64 oracle (error stated per field and per diagnostic)."""
import math
import sys

import numpy as np
import pytest

from gturbo3d_amd import initial, lib
from gturbo3d_amd.params import FORCING_2D2C, PREC_FP32_BULK, Params
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
    base = dict(nx=n, ny=n, nz=n, nu=1e-3, dt=5e-3, omega=2.0, famp=0.05, k1f=2.0, k2f=3.0)
    base.update(kw)
    return Params(**base)


@pytest.mark.parametrize("n", [32, 64])
def test_spectra_and_glob_match_oracle(require_gpu, n):
    p = cfg(n)
    say(f"spectra {n}^3")
    b = initial.initial_state(p, seed=2)
    with lib.Context(p) as g:
        g.set_state(b)
        g.step(3, 11)
        b3 = g.get_state()
        sp, sp2d, hel, glob = g.spectra()
    o = ol.Oracle(p)
    osp, osp2d = o.spectrum(b3)
    oglob = o.glob(b3)
    assert np.allclose(sp, osp, rtol=1e-12, atol=1e-14 * osp.max())
    assert np.allclose(sp2d, osp2d, rtol=1e-12, atol=1e-14 * max(osp2d.max(), 1e-300))
    assert np.allclose(glob, oglob, rtol=1e-12, atol=1e-14 * np.abs(oglob).max())
    # helicity: sum over shells = <u.w>/2 in physical space (w = curl u)
    u = o.velocity(b3)
    w = o.velocity(o.curl(b3))
    h_phys = 0.5 * float(np.mean(u[0] * w[0] + u[1] * w[1] + u[2] * w[2]))
    assert hel.sum() == pytest.approx(h_phys, rel=1e-10, abs=1e-14)
    # energy in the spectrum equals glob's energy (solenoidal b)
    assert sp.sum() == pytest.approx(glob[0] + glob[1] + glob[2], rel=1e-12)


def test_fluxes_match_oracle(require_gpu):
    p = cfg(32, omega=0.0)
    say("fluxes 32^3")
    b = initial.initial_state(p, seed=3)
    with lib.Context(p) as g:
        g.set_state(b)
        spf = g.fluxes()
        b_after = g.get_state()
    assert rel_l2(b_after, b) == 0.0  # fluxes leaves the state alone
    o = ol.Oracle(p)
    N, _ = o.nlt(b, b)
    ospf = o.fluxes(b, N)
    assert np.allclose(spf, ospf, rtol=0, atol=1e-11 * np.abs(ospf).max())
    # the nonlinear term conserves energy: the flux through k = 0 vanishes
    assert abs(spf[0]) < 1e-10 * np.abs(spf).max()


@pytest.mark.parametrize("n,variant", [(32, 3), (64, 3), (64, FORCING_2D2C)])
def test_fp32_bulk_against_fp64_oracle(require_gpu, n, variant):
    kw = dict(precision=PREC_FP32_BULK, variant=variant)
    if variant == FORCING_2D2C:
        kw.update(k1f=3.0, k2f=4.0, famp=0.1, omega=5.0, nu=2e-4, dt=2e-3)
    p = cfg(n, **kw)
    say(f"fp32-bulk 10 steps {n}^3 variant {variant}")
    b = initial.initial_state(p, seed=0)
    with lib.Context(p) as g:
        g.set_state(b)
        N = g.nonlinear()
        s = g.step(10, 8299396)
        b10 = g.get_state()
        sp, _, hel, glob = g.spectra()
    o = ol.Oracle(p)
    No, _ = o.nlt(b, b)
    ob10, os10 = o.step(b, 8299396, 10)
    assert s == os10
    e_nl = rel_l2(N, No)
    e_fields = [rel_l2([x], [y]) for x, y in zip(b10, ob10)]
    osp, _ = o.spectrum(ob10)
    m = osp > 1e-12 * osp.max()
    e_sp = float(np.max(np.abs(sp[m] - osp[m]) / osp[m]))
    oglob = o.glob(ob10)
    e_E = abs(sum(glob[:3]) - sum(oglob[:3])) / sum(oglob[:3])
    say(f"  nlt rel L2 {e_nl:.2e}; fields after 10 steps {[f'{e:.2e}' for e in e_fields]}; "
        f"spectrum max rel {e_sp:.2e}; energy rel {e_E:.2e}")
    assert e_nl < 1e-5
    assert max(e_fields) < 1e-5
    assert e_sp < 1e-5
    assert e_E < 1e-6


def test_fp32_bulk_512_workload(require_gpu):
    """The FP32-bulk step at the 512^3 shape of its workload (nu = 5e-5,
    Omega = 2, 3D3C forcing (2,3] at rate 0.1, dt = 1e-3): one nonlinear
    evaluation against the FP64 oracle, then 10 steps against the FP64 GPU
    run from the same state and seed, per field and per spectrum."""
    import math as _m
    base = dict(nu=5e-5, omega=2.0, dt=1e-3, k1f=2.0, k2f=3.0, famp=0.0)
    p64 = cfg(512, **base)
    p64.famp = _m.sqrt(0.1 / lib.forcing_count(p64)[0])
    p32 = cfg(512, **{**base, "famp": p64.famp, "precision": PREC_FP32_BULK})
    say("fp32-bulk 512^3: nonlinear term and 10 steps against FP64")
    b = initial.initial_state(p64, seed=0)
    with lib.Context(p32) as g:
        g.set_state(b)
        assert g.pruned()
        N32 = g.nonlinear()
        s32 = g.step(10, 0)
        b32 = g.get_state()
        sp32, _, hel32, _ = g.spectra()
    with lib.Context(p64) as g:
        g.set_state(b)
        s64 = g.step(10, 0)
        b64 = g.get_state()
        sp64, _, hel64, _ = g.spectra()
    assert s32 == s64
    No, _ = ol.Oracle(p64).nlt(b, b)
    e_nl = rel_l2(N32, No)
    del No
    e_fields = [rel_l2([x], [y]) for x, y in zip(b32, b64)]
    m = sp64 > 1e-12 * sp64.max()
    e_sp = float(np.max(np.abs(sp32[m] - sp64[m]) / sp64[m]))
    mh = np.abs(hel64) > 1e-6 * np.abs(hel64).max()
    e_hel = float(np.max(np.abs(hel32[mh] - hel64[mh]) / np.abs(hel64[mh])))
    say(f"  nlt vs oracle {e_nl:.2e}; fields after 10 steps {[f'{e:.2e}' for e in e_fields]}; "
        f"energy spectrum {e_sp:.2e}; helicity spectrum {e_hel:.2e}")
    assert e_nl < 1e-5
    assert max(e_fields) < 1e-5
    assert e_sp < 1e-5
    assert e_hel < 1e-4


def test_fp32_bulk_512_500_steps(require_gpu):
    """The precision workload at its full length: 500 steps at 512^3 (nu =
    5e-5, Omega = 2, 3D3C forcing (2,3] at rate 0.1, dt = 1e-3) in FP32-bulk
    and in FP64 from the same state and seed, stage 1 in the midpoint mode
    (with dt2 = 0 this workload diverges within 500 steps in both
    precisions).  Error per field and per diagnostic: fields 1e-5, energy
    spectrum 1e-5 (shells above 1e-12 of the peak), helicity spectrum 1e-4,
    energy 1e-6, dissipation 1e-5."""
    import math as _m
    from gturbo3d_amd.params import DT2_MIDPOINT
    base = dict(nu=5e-5, omega=2.0, dt=1e-3, k1f=2.0, k2f=3.0, famp=0.0, dt2_mode=DT2_MIDPOINT)
    p64 = cfg(512, **base)
    p64.famp = _m.sqrt(0.1 / lib.forcing_count(p64)[0])
    p32 = cfg(512, **{**base, "famp": p64.famp, "precision": PREC_FP32_BULK})
    say("fp32-bulk 512^3: 500 midpoint steps against FP64")
    b = initial.initial_state(p64, seed=0)
    out = {}
    for name, p in (("fp32", p32), ("fp64", p64)):
        with lib.Context(p) as g:
            g.set_state(b)
            s = g.step(500, 0)
            st = g.get_state()
            sp, _, hel, glob = g.spectra()
        out[name] = (s, st, sp, hel, glob)
    (s32, b32, sp32, hel32, gl32), (s64, b64, sp64, hel64, gl64) = out["fp32"], out["fp64"]
    assert s32 == s64
    assert all(np.isfinite(f).all() for f in b64)
    e_fields = [rel_l2([x], [y]) for x, y in zip(b32, b64)]
    m = sp64 > 1e-12 * sp64.max()
    e_sp = float(np.max(np.abs(sp32[m] - sp64[m]) / sp64[m]))
    mh = np.abs(hel64) > 1e-6 * np.abs(hel64).max()
    e_hel = float(np.max(np.abs(hel32[mh] - hel64[mh]) / np.abs(hel64[mh])))
    # energy and dissipation: the sums of the first three and the next three glob terms
    e_E = abs(gl32[:3].sum() - gl64[:3].sum()) / gl64[:3].sum()
    e_D = abs(gl32[3:6].sum() - gl64[3:6].sum()) / gl64[3:6].sum()
    say(f"  fields {[f'{e:.2e}' for e in e_fields]}; energy spectrum {e_sp:.2e}; helicity spectrum {e_hel:.2e}; "
        f"energy {e_E:.2e}; dissipation {e_D:.2e}")
    assert max(e_fields) < 1e-5
    assert e_sp < 1e-5
    assert e_hel < 1e-4
    assert e_E < 1e-6
    assert e_D < 1e-5


def test_fp32_bulk_1024_100_steps(require_gpu):
    """The precision workload at 1024^3 (nu = 5e-5, Omega = 2, 3D3C forcing
    (2,3] at rate 0.1, dt = 1e-3, midpoint stage 1): 100 steps in FP32-bulk
    and in FP64 from the same state and seed.  With dt = 1e-3 the resolved
    high-k modes of this grid grow by about 1.17 per step (kmax u_max dt ~ 1.1),
    so FP32 rounding noise there reaches O(1) before step 500; 100 steps is
    the run both precisions keep bounded.  Bars: fields 1e-6, energy and
    helicity spectra 1e-4 (shells above 1e-12 and 1e-6 of the peak), energy
    spectrum 1e-5 on the shells above 1e-6 of the peak, energy 1e-6,
    dissipation 1e-5.  Fields are compared plane by plane to bound the host
    memory (three 8.6 GB fields per state)."""
    import math as _m
    from gturbo3d_amd.params import DT2_MIDPOINT
    base = dict(nu=5e-5, omega=2.0, dt=1e-3, k1f=2.0, k2f=3.0, famp=0.0, dt2_mode=DT2_MIDPOINT)
    p64 = cfg(1024, **base)
    p64.famp = _m.sqrt(0.1 / lib.forcing_count(p64)[0])
    p32 = cfg(1024, **{**base, "famp": p64.famp, "precision": PREC_FP32_BULK})
    say("fp32-bulk 1024^3: 100 midpoint steps against FP64")
    b = initial.initial_state(p64, seed=0)
    out = {}
    for name, p in (("fp64", p64), ("fp32", p32)):
        with lib.Context(p) as g:
            g.set_state(b)
            if name == "fp32":
                del b
            s = g.step(100, 0)
            sp, _, hel, glob = g.spectra()
            st = g.get_state()
        out[name] = (s, st, sp, hel, glob)
        say(f"  {name}: 100 steps done, E = {sp.sum():.6f}")
    (s32, b32, sp32, hel32, gl32), (s64, b64, sp64, hel64, gl64) = out["fp32"], out["fp64"]
    del out
    assert s32 == s64
    e_fields = []
    for x, y in zip(b32, b64):
        num = den = 0.0
        for k in range(0, y.shape[0], 64):
            yk = y[k:k + 64]
            assert np.isfinite(yk).all()
            num += float(np.sum(np.abs(x[k:k + 64] - yk) ** 2))
            den += float(np.sum(np.abs(yk) ** 2))
        e_fields.append(math.sqrt(num / den))
    del b32, b64
    m = sp64 > 1e-12 * sp64.max()
    e_sp = float(np.max(np.abs(sp32[m] - sp64[m]) / sp64[m]))
    m6 = sp64 > 1e-6 * sp64.max()
    e_sp6 = float(np.max(np.abs(sp32[m6] - sp64[m6]) / sp64[m6]))
    mh = np.abs(hel64) > 1e-6 * np.abs(hel64).max()
    e_hel = float(np.max(np.abs(hel32[mh] - hel64[mh]) / np.abs(hel64[mh])))
    e_E = abs(gl32[:3].sum() - gl64[:3].sum()) / gl64[:3].sum()
    e_D = abs(gl32[3:6].sum() - gl64[3:6].sum()) / gl64[3:6].sum()
    say(f"  fields {[f'{e:.2e}' for e in e_fields]}; energy spectrum {e_sp:.2e} (shells above 1e-6 of the peak "
        f"{e_sp6:.2e}); helicity spectrum {e_hel:.2e}; energy {e_E:.2e}; dissipation {e_D:.2e}")
    assert max(e_fields) < 1e-6
    assert e_sp < 1e-4
    assert e_sp6 < 1e-5
    assert e_hel < 1e-4
    assert e_E < 1e-6
    assert e_D < 1e-5
