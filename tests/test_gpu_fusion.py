"""Cross-stage fusion: stage 1's z-forward pass keeps bh in registers and runs
stage 2's z-inverse pass on it, and within one gt_step call stage 2's
z-forward pass adds the step's forcing and runs the next step's z-inverse
pass on the new b (compiled for NZ = 512).  The fused steps are held against
the unfused steps of the same library and against the oracle.

Tolerances: FP64 fused against unfused 1e-14 relative L2 (the two differ
only in the contraction order of the curl and k^2 products); FP64 against
the oracle 1e-13 after one step (the bar of a single evaluation); FP32-bulk
fused against unfused 1e-6 (an FP32 rounding of an intermediate may differ).
"""
import math
import sys

import numpy as np
import pytest

from gturbo3d_amd import initial, lib
from gturbo3d_amd.params import DT2_MIDPOINT, DT2_REFERENCE, FORCING_2D2C, PREC_FP32_BULK, Params
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    num = math.sqrt(sum(float(np.sum(np.abs(x - y) ** 2)) for x, y in zip(a, b)))
    den = math.sqrt(sum(float(np.sum(np.abs(y) ** 2)) for y in b))
    return num / den


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def params512(**kw):
    base = dict(nx=512, ny=512, nz=512, nu=5e-5, dt=1e-3, omega=2.0, k1f=2.0, k2f=3.0)
    base.update(kw)
    p = Params(**base)
    if "famp" not in kw:
        p.famp = math.sqrt(0.1 / lib.forcing_count(p)[0])
    return p


def run(p, b, seed, nsteps, fused, tiles=None):
    with lib.Context(p) as g:
        g.set_fusion(fused)
        if tiles:
            g.set_tiles(*tiles)
        g.set_state(b)
        on, active = g.fusion()
        assert on == fused and active == fused
        s = g.step(nsteps, seed)
        return g.get_state(), s, g.pruned()


@pytest.mark.parametrize("dt2", [DT2_REFERENCE, DT2_MIDPOINT])
def test_fused_step_512_fp64(require_gpu, dt2):
    p = params512(dt2_mode=dt2)
    say(f"fused step 512^3 FP64, dt2 mode {dt2}")
    b = initial.initial_state(p, seed=5)
    bf, sf, pruned = run(p, b, 8299396, 2, True)
    assert pruned
    bu, su, _ = run(p, b, 8299396, 2, False)
    assert sf == su
    err = rel_l2(bf, bu)
    say(f"  fused vs unfused, 2 steps: {err:.2e}")
    assert err < 1e-14
    if dt2 == DT2_REFERENCE:
        # two steps in one call: the second step's z-inverse pass ran in the
        # first step's last kernel, with the forcing
        ob2, os2 = ol.Oracle(p).step(b, 8299396, 2)
        assert sf == os2
        err = rel_l2(bf, ob2)
        say(f"  fused vs oracle, 2 steps: {err:.2e}")
        assert err < 1e-13


def test_fused_steps_one_call_equal_separate_calls(require_gpu):
    """Three steps in one call (two stage-2 passes fused with the next step's
    z-inverse pass and the forcing) against three one-step calls (none)."""
    p = params512()
    say("fused 3 steps in one call vs 3 calls")
    b = initial.initial_state(p, seed=9)
    b3, s3, _ = run(p, b, 4242, 3, True)
    with lib.Context(p) as g:
        g.set_state(b)
        s = 4242
        for _ in range(3):
            s = g.step(1, s)
        c3 = g.get_state()
    assert s == s3
    err = rel_l2(b3, c3)
    say(f"  one call vs three: {err:.2e}")
    assert err < 1e-14


def test_fused_step_512_2d2c(require_gpu):
    """2D2C forcing (bz only, conjugate entries on the kx = 0 plane)."""
    p = params512(variant=FORCING_2D2C, k1f=3.0, k2f=4.0, famp=0.1, omega=5.0)
    say("fused steps 512^3 2D2C")
    b = initial.initial_state(p, seed=10)
    bf, sf, _ = run(p, b, 77, 3, True)
    bu, su, _ = run(p, b, 77, 3, False)
    assert sf == su
    err = rel_l2(bf, bu)
    say(f"  fused vs unfused, 3 steps: {err:.2e}")
    assert err < 1e-14


def test_fused_step_512_fp32_bulk(require_gpu):
    p = params512(precision=PREC_FP32_BULK)
    say("fused step 512^3 FP32-bulk")
    with lib.Context(p) as g:
        # off by default in FP32-bulk, where the fused kernels measured slower
        assert g.fusion() == (False, False)
    b = initial.initial_state(params512(), seed=6)
    bf, sf, _ = run(p, b, 17, 2, True)
    bu, su, _ = run(p, b, 17, 2, False)
    assert sf == su
    err = rel_l2(bf, bu)
    say(f"  fused vs unfused, 2 steps: {err:.2e}")
    assert err < 1e-6


def test_fused_step_512_dense(require_gpu):
    """An untruncated state runs the dense path: full extents, no pruned slots."""
    p = params512(truncate=0)
    say("fused step 512^3 dense")
    rng = np.random.default_rng(8)
    b = initial.initial_state(params512(), seed=8)
    b[0] = b[0] + 1e-7 * (rng.standard_normal(b[0].shape) + 1j * rng.standard_normal(b[0].shape))
    b[0][0, 0, 0] = 0.0
    bf, sf, pruned = run(p, b, 3, 1, True)
    assert not pruned
    bu, su, _ = run(p, b, 3, 1, False)
    assert sf == su
    err = rel_l2(bf, bu)
    say(f"  fused vs unfused, 1 step: {err:.2e}")
    assert err < 1e-14


def test_fusion_inactive_at_other_widths_and_sizes(require_gpu):
    """The fused kernel exists for NZ = 512 at the default z-forward width;
    elsewhere the setting is kept and the two passes run (the same bits as
    with fusion off)."""
    p = params512()
    with lib.Context(p) as g:
        g.set_tiles(0, 16, 0)
        assert g.fusion() == (True, False)
    p = Params(nx=64, ny=64, nz=64, nu=1e-3, dt=5e-3, omega=1.0, famp=0.05, k1f=2.0, k2f=3.0)
    b = initial.initial_state(p, seed=2)
    out = []
    for fused in (True, False):
        with lib.Context(p) as g:
            g.set_fusion(fused)
            assert g.fusion() == (fused, False)
            g.set_state(b)
            g.step(2, 5)
            out.append(g.get_state())
    assert all(np.array_equal(x, y) for x, y in zip(*out))


def test_fused_batch_members_equal_solo_runs(require_gpu):
    """Two members at 512^3 in one batch, two steps (the member index in the
    fused kernels' work, state and forcing offsets) against solo fused runs,
    bit for bit."""
    p = params512()
    say("fused batch 2 x 512^3")
    b0 = initial.initial_state(p, seed=11)
    b1 = initial.initial_state(p, seed=12)
    with lib.Batch(p, 2, omegas=[2.0, 3.5]) as g:
        assert g.fusion() == (True, True)
        g.set_member(0, b0)
        g.set_member(1, b1)
        seeds = g.step_batch(2, [21, 22])
        m0, m1 = g.get_member(0), g.get_member(1)
    s0, ss0, _ = run(p, b0, 21, 2, True)
    q = params512(omega=3.5)
    q.famp = p.famp
    s1, ss1, _ = run(q, b1, 22, 2, True)
    assert list(seeds) == [ss0, ss1]
    assert all(np.array_equal(x, y) for x, y in zip(m0, s0))
    assert all(np.array_equal(x, y) for x, y in zip(m1, s1))


def test_fused_ensemble_graphs_512(require_gpu):
    """A captured ensemble of one 512^3 batch (two members), two steps per
    graph: the graphs hold the fused kernels, their forcing entries come from
    the ensemble's own buffers.  Five steps (two replays and an eager
    remainder) against solo fused runs of five steps in one call."""
    p = params512()
    say("fused ensemble 2 x 512^3, chunk 2")
    oms, seeds = [2.0, 3.5], [31, 32]
    ics = [initial.initial_state(p, seed=40 + m) for m in range(2)]
    A = lib.Batch(p, 2, oms)
    assert A.fusion() == (True, True)
    for m in range(2):
        A.set_member(m, ics[m])
    ens = lib.Ensemble([A], chunk=2)
    ens.set_seeds(0, seeds)
    ens.step(5)
    ens.sync()
    got_seeds = ens.get_seeds(0)
    for m in range(2):
        q = params512(omega=oms[m])
        q.famp = p.famp
        ref, rs, _ = run(q, ics[m], seeds[m], 5, True)
        assert got_seeds[m] == rs
        err = rel_l2(A.get_member(m), ref)
        say(f"  member {m}: {err:.2e}")
        assert err < 1e-14
    ens.close()
    A.close()


@pytest.mark.parametrize("n,tze", [(128, 8), (128, 16), (256, 8), (256, 16)])
def test_fused_steps_128_256(require_gpu, n, tze):
    """The fused kernels of 128- and 256-point z transforms, at both z-forward
    widths: three fused steps against three unfused ones and against the
    oracle."""
    p = Params(nx=n, ny=n, nz=n, nu=1e-3, dt=4e-3, omega=2.0, k1f=2.0, k2f=3.0)
    p.famp = math.sqrt(0.1 / lib.forcing_count(p)[0])
    say(f"fused steps {n}^3, {tze} pencils")
    b = initial.initial_state(p, seed=13)
    bf, sf, pruned = run(p, b, 7, 3, True, tiles=(0, tze, 0))
    assert pruned
    bu, su, _ = run(p, b, 7, 3, False, tiles=(0, tze, 0))
    assert sf == su
    err = rel_l2(bf, bu)
    say(f"  fused vs unfused, 3 steps: {err:.2e}")
    assert err < 1e-14
    ob, os_ = ol.Oracle(p).step(b, 7, 3)
    assert sf == os_
    err = rel_l2(bf, ob)
    say(f"  fused vs oracle, 3 steps: {err:.2e}")
    assert err < 1e-12


def test_autotune_keeps_a_fusion_choice(require_gpu):
    """gt_autotune times the step with and without fusion where the fused
    kernels exist and keeps one; the result does not depend on it."""
    p = Params(nx=128, ny=128, nz=128, nu=1e-3, dt=4e-3, omega=2.0, famp=0.05, k1f=2.0, k2f=3.0)
    b = initial.initial_state(p, seed=14)
    out = []
    with lib.Context(p) as g:
        g.set_state(b)
        g.autotune(2)
        on, active = g.fusion()
        assert active == on
        say(f"  autotune at 128^3: fusion {'on' if on else 'off'}, tiles {g.tiles()}")
        g.set_state(b)
        s1 = g.step(2, 9)
        out.append(g.get_state())
        g.set_fusion(not on)
        g.set_state(b)
        s2 = g.step(2, 9)
        out.append(g.get_state())
    assert s1 == s2
    assert rel_l2(out[0], out[1]) < 1e-14
