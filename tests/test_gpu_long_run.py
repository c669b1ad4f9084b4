"""Long-run parity: the 256^3 rotating run (nu = 2e-4, Omega = 5, 2D2C forcing
on (3, 4], dt = 2e-3, 2000 steps, midpoint stage 1) against the CPU oracle's
spectra recorded every 100 steps (tests/golden/rotating256_midpoint_spectra.npz, written
by tests/golden/make_rotating256_spectra.py).

Bar: time-averaged energy and kz = 0 spectra within 1% on every shell that
holds energy, energy and dissipation (from glob) within 1% at every record,
and the forcing seed after 2000 steps equal.

The same run in the reference's stage-1 mode (dt2 = 0) diverges between
steps 300 and 400 in both codes; over its first 300 steps (records every
100, tests/golden/rotating256_reference_spectra.npz) the spectra and glob of every
record must match the oracle's to 1e-6.

The batched ensemble (32 x 128^3 + 16 x 96^3, captured graphs) runs its
1000 steps with the midpoint stage 1 (in the reference's mode its member 0
diverges between steps 300 and 400, in the oracle too); two members are
held against the oracle's spectra and glob every 100 steps
(tests/golden/ensemble1000_spectra.npz) to 1e-6.
"""
import os
import sys

import numpy as np
import pytest

from gturbo3d_amd import initial, lib

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_rotating256_spectra import rotating256_params  # noqa: E402

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(HERE, "golden", "rotating256_midpoint_spectra.npz")
FIXTURE_REF = os.path.join(HERE, "golden", "rotating256_reference_spectra.npz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="long-run fixture not generated")
def test_long_run_spectra_match_oracle(require_gpu):
    fx = np.load(FIXTURE)
    steps = fx["steps"]
    p = rotating256_params(int(fx["n"]))
    say(f"long run {p.nx}^3, {int(steps[-1])} steps, records every {int(steps[1] - steps[0])}")
    b = initial.initial_state(p, seed=0)
    sps, sp2ds, globs = [], [], []
    with lib.Context(p) as g:
        g.set_state(b)
        seed, done = 0, 0
        for s in steps:
            seed = g.step(int(s) - done, seed)
            done = int(s)
            sp, sp2d, _, glob = g.spectra()
            sps.append(sp)
            sp2ds.append(sp2d)
            globs.append(glob)
    assert seed == int(fx["seed"])
    sps, sp2ds, globs = np.array(sps), np.array(sp2ds), np.array(globs)
    for name, got, ref in (("energy", sps, fx["sp"]), ("kz=0", sp2ds, fx["sp2d"])):
        a, r = got.mean(axis=0), ref.mean(axis=0)
        m = r > 1e-12 * r.max()
        err = np.abs(a[m] - r[m]) / r[m]
        say(f"  time-averaged {name} spectrum: max relative error {err.max():.2e} over {m.sum()} shells")
        assert err.max() < 1e-2
    for name, sl in (("energy", slice(0, 3)), ("dissipation", slice(3, 6))):
        a, r = globs[:, sl].sum(axis=1), fx["glob"][:, sl].sum(axis=1)
        err = np.abs(a - r) / np.abs(r)
        say(f"  {name} at every record: max relative error {err.max():.2e}")
        assert err.max() < 1e-2


@pytest.mark.skipif(not os.path.exists(FIXTURE_REF), reason="reference-mode fixture not generated")
def test_reference_mode_run_matches_oracle(require_gpu):
    fx = np.load(FIXTURE_REF)
    steps = fx["steps"]
    p = rotating256_params(int(fx["n"]), dt2_mode=int(fx["dt2_mode"]))
    assert p.dt2_mode == 0
    say(f"reference-mode run {p.nx}^3, {int(steps[-1])} steps, records every {int(steps[1] - steps[0])}")
    b = initial.initial_state(p, seed=0)
    with lib.Context(p) as g:
        g.set_state(b)
        seed, done = 0, 0
        for i, s in enumerate(steps):
            seed = g.step(int(s) - done, seed)
            done = int(s)
            sp, sp2d, _, glob = g.spectra()
            for name, got, ref in (("energy", sp, fx["sp"][i]), ("kz=0", sp2d, fx["sp2d"][i])):
                m = ref > 1e-12 * ref.max()
                err = np.abs(got[m] - ref[m]) / ref[m]
                say(f"  step {int(s)}: {name} spectrum max relative error {err.max():.2e} over {m.sum()} shells, "
                    f"E = {got.sum():.6f} (oracle {ref.sum():.6f})")
                assert err.max() < 1e-6
            gr = fx["glob"][i]
            gerr = np.abs(glob - gr) / np.maximum(np.abs(gr), 1e-12 * np.abs(gr).max())
            say(f"  step {int(s)}: glob max relative error {gerr.max():.2e}")
            assert gerr.max() < 1e-6
    assert seed == int(fx["seed"])


FIXTURE_ENS = os.path.join(HERE, "golden", "ensemble1000_spectra.npz")


@pytest.mark.skipif(not os.path.exists(FIXTURE_ENS), reason="ensemble fixture not generated")
def test_ensemble_1000_steps_match_oracle_and_solo_runs(require_gpu):
    """The batched-ensemble workload at its full length: 32 members at 128^3
    and 16 at 96^3 in one ensemble of captured graphs (10 steps per graph),
    1000 steps with the midpoint stage 1.  At every record (every 100
    steps) the device spectra and glob of members 0 and 45 match the
    oracle's (tests/golden/ensemble1000_spectra.npz, written by
    tests/golden/make_ensemble_spectra.py) to 1e-6; after 1000 steps their
    seeds equal the oracle's, every member is finite, and members 0, 31, 32
    and 45 equal their solo runs bit for bit."""
    from make_ensemble_spectra import GROUPS, ensemble_omegas, ensemble_params
    fx = np.load(FIXTURE_ENS)
    steps = [int(s) for s in fx["steps"]]
    members = [int(m) for m in fx["members"]]
    omegas = ensemble_omegas()
    say(f"ensemble 32 x 128^3 + 16 x 96^3, {steps[-1]} steps, records every {steps[1] - steps[0]}")
    batches, ics, where = [], {}, {}
    for q, (n, nm, off) in enumerate(GROUPS):
        bt = lib.Batch(ensemble_params(n), nm, omegas[off:off + nm])
        for m in range(nm):
            ics[off + m] = initial.initial_state(ensemble_params(n), seed=off + m)
            bt.set_member(m, ics[off + m])
            where[off + m] = (q, m)
        batches.append(bt)
    ens = lib.Ensemble(batches, chunk=10)
    for q, (n, nm, off) in enumerate(GROUPS):
        ens.set_seeds(q, np.arange(off, off + nm, dtype=np.int32))
    # device diagnostics of a member's state, through a context of its grid
    checkers = {n: lib.Context(ensemble_params(n)) for n, _, _ in GROUPS}
    done, worst = 0, {m: 0.0 for m in members}
    for i, s in enumerate(steps):
        ens.step(s - done)
        ens.sync()
        done = s
        for m in members:
            q, mm = where[m]
            n = GROUPS[q][0]
            g = checkers[n]
            g.set_state(batches[q].get_member(mm))
            sp, sp2d, _, glob = g.spectra()
            for name, got, ref in (("energy", sp, fx[f"sp_{m}"][i]), ("kz=0", sp2d, fx[f"sp2d_{m}"][i])):
                mk = ref > 1e-12 * ref.max()
                err = np.abs(got[mk] - ref[mk]) / ref[mk]
                assert err.max() < 1e-6, (m, s, name, err.max())
                worst[m] = max(worst[m], float(err.max()))
            gr = fx[f"glob_{m}"][i]
            gerr = np.abs(glob - gr) / np.maximum(np.abs(gr), 1e-12 * np.abs(gr).max())
            assert gerr.max() < 1e-6, (m, s, gerr.max())
            worst[m] = max(worst[m], float(gerr.max()))
            if s % 500 == 0:
                say(f"  step {s}, member {m} ({n}^3): E = {sp.sum():.6f} (oracle {fx[f'sp_{m}'][i].sum():.6f}), "
                    f"largest relative error of the spectra and glob so far {worst[m]:.2e}")
    for g in checkers.values():
        g.close()
    seeds = [ens.get_seeds(q) for q in range(len(GROUPS))]
    for m in members:
        q, mm = where[m]
        assert seeds[q][mm] == int(fx[f"seed_{m}"])
    for q, (n, nm, off) in enumerate(GROUPS):
        for m in range(nm):
            assert all(np.isfinite(f).all() for f in batches[q].get_member(m)), (n, m)
    say("  every member finite after 1000 steps")
    for m in (0, 31, 32, 45):
        q, mm = where[m]
        n = GROUPS[q][0]
        with lib.Context(ensemble_params(n, omegas[m])) as g:
            g.set_state(ics[m])
            rs = g.step(steps[-1], m)
            ref = g.get_state()
        assert seeds[q][mm] == rs
        got = batches[q].get_member(mm)
        assert all(np.array_equal(x, y) for x, y in zip(got, ref)), m
    say("  members 0, 31, 32, 45 equal their solo runs bit for bit")
    ens.close()
    for bt in batches:
        bt.close()
