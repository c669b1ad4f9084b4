"""Batched members and hipGraph-captured ensembles against solo runs."""
import sys

import numpy as np
import pytest

from gturbo3d_amd import initial, lib
from gturbo3d_amd.params import Params
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def params(n, omega=1.0, **kw):
    base = dict(nx=n, ny=n, nz=n, nu=1e-3, dt=4e-3, omega=omega, famp=0.05, k1f=2.0, k2f=3.0)
    base.update(kw)
    return Params(**base)


def solo(p, b, seed, nsteps):
    with lib.Context(p) as g:
        g.set_state(b)
        s = g.step(nsteps, seed)
        return g.get_state(), s


def test_batch_members_equal_solo_runs(require_gpu):
    n, omegas, seeds = 32, [0.5, 2.0, 7.0], [0, 1, 2]
    say("batch 3 x 32^3")
    p = params(n)
    ics = [initial.initial_state(p, seed=10 + m) for m in range(3)]
    bt = lib.Batch(p, 3, omegas)
    for m in range(3):
        bt.set_member(m, ics[m])
    out_seeds = bt.step_batch(4, seeds)
    for m in range(3):
        pm = params(n, omega=omegas[m])
        ref, rs = solo(pm, ics[m], seeds[m], 4)
        got = bt.get_member(m)
        assert out_seeds[m] == rs
        for x, y in zip(got, ref):
            assert np.array_equal(x, y)  # same kernels, same data: bit-identical
    # and member 1 against the oracle
    ob, os_ = ol.Oracle(params(n, omega=omegas[1])).step(ics[1], seeds[1], 4)
    got = bt.get_member(1)
    num = sum(float(np.sum(np.abs(x - y) ** 2)) for x, y in zip(got, ob))
    den = sum(float(np.sum(np.abs(y) ** 2)) for y in ob)
    assert (num / den) ** 0.5 < 1e-13
    sp, _, _, glob = bt.member_spectra(2)
    with lib.Context(params(n, omega=omegas[2])) as g:
        g.set_state(bt.get_member(2))
        sp2, _, _, glob2 = g.spectra()
    # diagnostics accumulate with LDS atomics: equal up to summation order
    assert np.allclose(sp, sp2, rtol=1e-13, atol=0) and np.allclose(glob, glob2, rtol=1e-13, atol=0)
    bt.close()


def test_ensemble_graph_replays_match_eager(require_gpu):
    say("ensemble: 3 x 32^3 + 2 x 24^3, chunk 3")
    pa, pb = params(32), params(24)
    om_a, om_b = [1.0, 3.0, 9.0], [0.7, 5.0]
    sa, sb = [5, 6, 7], [8, 9]
    ics_a = [initial.initial_state(pa, seed=20 + m) for m in range(3)]
    ics_b = [initial.initial_state(pb, seed=30 + m) for m in range(2)]
    A = lib.Batch(pa, 3, om_a)
    B = lib.Batch(pb, 2, om_b)
    # passes B-D of batch A in z chunks (a ragged last one), pass D of B whole:
    # the captured graphs hold every chunk launch
    A.set_zchunk(12)
    A.set_zstreams(2)  # the chunks of A alternate between two streams in the graphs
    B.set_zchunk(8, False)
    for m in range(3):
        A.set_member(m, ics_a[m])
    for m in range(2):
        B.set_member(m, ics_b[m])
    ens = lib.Ensemble([A, B], chunk=3)
    ens.set_seeds(0, sa)
    ens.set_seeds(1, sb)
    done = 0
    for nsteps in (7, 6, 3):  # two replays + a remainder, two replays, one replay
        ens.step(nsteps)
        done += nsteps
        ens.sync()
        for batch, ics, oms, seeds, pq in ((A, ics_a, om_a, sa, pa), (B, ics_b, om_b, sb, pb)):
            got_seeds = ens.get_seeds(0 if batch is A else 1)
            for m in range(batch.n):
                pm = Params(**{**pq.__dict__, "omega": oms[m]})
                ref, rs = solo(pm, ics[m], seeds[m], done)
                assert got_seeds[m] == rs
                got = batch.get_member(m)
                for x, y in zip(got, ref):
                    assert np.array_equal(x, y)
        say(f"  after {done} steps: all members equal the eager solo runs")
    ens.close()
    A.close()
    B.close()


@pytest.mark.parametrize("n", [32, 96, 128])
def test_tile_widths_do_not_change_results(require_gpu, n):
    say(f"tile widths {n}^3")
    p = params(n, omega=2.0)
    b = initial.initial_state(p, seed=4)
    results = {}
    with lib.Context(p) as g:
        g.set_state(b)
        tz0, tze0, ty0, _ = g.tiles()
        for tz in (8, 16):
            for tze in (8, 16):
                for ty in (8, 16):
                    g.set_tiles(tz, tze, ty)
                    g.set_state(b)
                    g.step(2, 3)
                    results[(tz, tze, ty)] = g.get_state()
        # z chunks of passes B-D: equal chunks, a ragged last chunk, one plane
        g.set_tiles(tz0, tze0, ty0)
        # and the chunks alternating between two streams
        for zc, wd, ns in ((n // 4, True, 1), (3 * n // 8 + 1, True, 1), (1, True, 1), (n // 4, False, 1),
                           (n // 4, True, 2), (3 * n // 8 + 1, True, 2), (1, True, 2)):
            g.set_zchunk(zc, wd)
            g.set_zstreams(ns)
            assert g.zchunk() == (zc, wd) and g.zstreams() == ns
            g.set_state(b)
            g.step(2, 3)
            results[("zchunk", zc, wd, ns)] = g.get_state()
        g.set_zchunk(0)
        g.set_zstreams(1)
        tz, tze, ty, ms = g.autotune(2)
        assert tz in (8, 16) and tze in (8, 16) and ty in (8, 16) and ms > 0
        # the autotuner's candidates: NZ/16 .. NZ/2 and 3, 5, 6 NZ/64 planes, one stream
        cands = {0} | {n // d for d in (16, 8, 4, 2) if n % d == 0 and n // d >= 8}
        cands |= {m * n // 64 for m in (3, 5, 6) if (m * n) % 64 == 0 and m * n // 64 >= 8}
        assert g.zchunk()[0] in cands and g.zstreams() == 1
    ref = results[(tz0, tze0, ty0)]
    for key, st in results.items():
        for x, y in zip(st, ref):
            assert np.array_equal(x, y), key


def test_ensemble_follows_batch_changes_after_capture(require_gpu):
    """What the captured graphs bake in can change after gt_ensemble_create: a
    member set to an untruncated state switches its batch to the dense path, a
    batch stepped on its own past the chunk reallocates its forcing buffers,
    and a new z chunk changes the launches.  The ensemble recaptures its graphs
    and every member still equals its eager solo run, bit for bit."""
    say("ensemble after batch changes: 2 x 32^3, chunk 2")
    p = params(32)
    oms, seeds = [1.0, 4.0], [3, 4]
    ics = [initial.initial_state(p, seed=40 + m) for m in range(2)]
    A = lib.Batch(p, 2, oms)
    for m in range(2):
        A.set_member(m, ics[m])
    assert A.pruned()
    ens = lib.Ensemble([A], chunk=2)
    ens.set_seeds(0, seeds)
    # 1. member 1 gets energy outside the truncation ellipsoid: dense path
    rng = np.random.default_rng(7)
    noisy = [f.copy() for f in ics[1]]
    for f in noisy:  # interior kx planes hold free complex values of a real field
        f[:, :, 1:-1] += 1e-3 * (rng.standard_normal(f[:, :, 1:-1].shape) +
                                 1j * rng.standard_normal(f[:, :, 1:-1].shape))
    A.set_member(1, noisy)
    assert not A.pruned()
    start = [ics[0], noisy]
    ens.step(4)
    ens.sync()
    for m in range(2):
        with lib.Context(Params(**{**p.__dict__, "omega": oms[m]})) as g:
            g.set_state(start[m])
            assert g.pruned() == (m == 0)
            rs = g.step(4, seeds[m])
            ref = g.get_state()
        assert ens.get_seeds(0)[m] == rs
        got = A.get_member(m)
        if m == 1:  # dense in both runs: the same kernels
            for x, y in zip(got, ref):
                assert np.array_equal(x, y)
        else:  # dense in the batch, pruned solo: the same values up to rounding
            num = sum(float(np.sum(np.abs(x - y) ** 2)) for x, y in zip(got, ref))
            den = sum(float(np.sum(np.abs(y) ** 2)) for y in ref)
            assert (num / den) ** 0.5 < 1e-14
    say("  dense member after capture: equal")
    # 2. the batch stepped on its own past the chunk, then a new z chunk
    #    (member 0 back on the ellipsoid: the batch returns to the pruned path)
    A.set_member(0, ics[0])
    assert not A.pruned()
    A.set_member(1, ics[1])
    assert A.pruned()
    s_after = A.step_batch(5, seeds)
    A.set_zchunk(8)
    ens.set_seeds(0, s_after)
    ens.step(3)
    ens.sync()
    for m in range(2):
        ref, rs = solo(Params(**{**p.__dict__, "omega": oms[m]}), ics[m], seeds[m], 8)
        assert ens.get_seeds(0)[m] == rs
        for x, y in zip(A.get_member(m), ref):
            assert np.array_equal(x, y)
    say("  batch steps past the chunk and a new z chunk: equal")
    ens.close()
    A.close()


def test_ensemble_128_and_96_workload_shapes(require_gpu):
    """The batched-ensemble workload at its shapes: 32 members at 128^3 and 16
    at 96^3 (FP64, nu = 1e-3, dt = 4e-3, 3D3C (2,3] at rate 0.1, seeds 0..47,
    Omega ~ U[0.5, 10] from PCG64(48)) in one ensemble of captured graphs.
    After one replay and a remainder every member equals its eager solo run
    bit for bit, and one member of each grid matches the oracle."""
    import math as _m
    rng = np.random.default_rng(48)
    omegas = rng.uniform(0.5, 10.0, 48)
    groups = [(128, 32, 0), (96, 16, 32)]
    batches, ps, ics = [], [], []
    say("ensemble 32 x 128^3 + 16 x 96^3, chunk 5, 6 steps")
    for n, nm, off in groups:
        p = params(n, famp=0.0)
        p.famp = _m.sqrt(0.1 / lib.forcing_count(p)[0])
        bt = lib.Batch(p, nm, omegas[off:off + nm])
        ic = [initial.initial_state(p, seed=off + m) for m in range(nm)]
        for m in range(nm):
            bt.set_member(m, ic[m])
        batches.append(bt)
        ps.append(p)
        ics.append(ic)
    ens = lib.Ensemble(batches, chunk=5)
    ens.set_seeds(0, np.arange(0, 32, dtype=np.int32))
    ens.set_seeds(1, np.arange(32, 48, dtype=np.int32))
    ens.step(6)
    ens.sync()
    for q, (n, nm, off) in enumerate(groups):
        got_seeds = ens.get_seeds(q)
        for m in range(nm):
            pm = Params(**{**ps[q].__dict__, "omega": float(omegas[off + m])})
            ref, rs = solo(pm, ics[q][m], off + m, 6)
            assert got_seeds[m] == rs
            got = batches[q].get_member(m)
            for x, y in zip(got, ref):
                assert np.array_equal(x, y), (n, m)
        say(f"  {nm} members at {n}^3 equal their solo runs")
        m = nm - 1
        pm = Params(**{**ps[q].__dict__, "omega": float(omegas[off + m])})
        ob, os_ = ol.Oracle(pm).step(ics[q][m], off + m, 6)
        assert got_seeds[m] == os_
        got = batches[q].get_member(m)
        num = sum(float(np.sum(np.abs(x - y) ** 2)) for x, y in zip(got, ob))
        den = sum(float(np.sum(np.abs(y) ** 2)) for y in ob)
        say(f"  member {m} at {n}^3 against the oracle: {(num / den) ** 0.5:.2e}")
        assert (num / den) ** 0.5 < 1e-12
    ens.close()
    for bt in batches:
        bt.close()
