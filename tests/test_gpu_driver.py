"""The Fortran driver end to end: a run directory in the reference's layout
(params.dat, curframe.dat, files/seed.000, fields/u*.000), the driver
binary on the GPU, and every file it writes compared with the oracle:

  spectra.000 / fluxes.000  of the state at time t, before the step
                            (src/main.f90:133-138)
  globalE.000 / globalZ.000 11 columns after the step (src/main.f90:152-161,
                            src/diagnostics.f90:155,193)
  fields/u?.000.001         the writef2 snapshot every iout steps
                            (src/main.f90:163-184)
  fields/u?.001, seed.001, curframe.dat   the restart output (:188-215)
"""
import math
import os
import subprocess

import numpy as np
import pytest

from gturbo3d_amd import initial
from gturbo3d_amd.params import Params, write_params_dat
from gturbo3d_amd.records import read_record, write_record
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gturbo3d_amd", "driver", "gturb_mi355x")
SEED0 = 8299396


def test_driver_binary_is_built():
    assert os.path.exists(EXE), "build() compiles the Fortran driver"


def test_records_round_trip_with_subrecords(tmp_path):
    a = np.arange(16 * 16 * 16, dtype=np.float32).reshape(16, 16, 16)
    for maxsub in (5000, 4 * a.size, 2147483639):
        path = str(tmp_path / f"r{maxsub}")
        write_record(path, a, maxsub)
        assert np.array_equal(read_record(path, a.shape), a)
        nsub = -(-4 * a.size // maxsub)
        assert os.path.getsize(path) == 4 * a.size + 8 * nsub
    # one plain record: length, data, length
    raw = open(str(tmp_path / "r2147483639"), "rb").read()
    assert np.frombuffer(raw[:4], np.int32)[0] == np.frombuffer(raw[-4:], np.int32)[0] == 4 * a.size
    # subrecords: leading markers negative but the last, trailing negative but the first
    raw = open(str(tmp_path / "r5000"), "rb").read()
    assert np.frombuffer(raw[:4], np.int32)[0] == -5000
    assert np.frombuffer(raw[5004:5008], np.int32)[0] == 5000
    assert np.frombuffer(raw[-4:], np.int32)[0] == -(4 * a.size - 3 * 5000)


def rel(got, expect):
    num = math.sqrt(sum(float(np.sum(np.abs(g - e) ** 2)) for g, e in zip(got, expect)))
    den = math.sqrt(sum(float(np.sum(np.abs(e) ** 2)) for e in expect))
    return num / den


def make_run(run, p, u32, maxsub=2147483639):
    (run / "files").mkdir()
    (run / "fields").mkdir()
    write_params_dat(p, str(run / "params.dat"))
    (run / "curframe.dat").write_text("0 0.0\n")
    (run / "files" / "seed.000").write_text(f"{SEED0}\n")
    # Fortran (NX, NY, NZ) column-major == C [NZ][NY][NX]
    for name, f in zip("xyz", u32):
        write_record(str(run / "fields" / f"u{name}.000"), f, maxsub)


def run_driver(run, n, env=None):
    print(f"driver {n}^3 in {run}", flush=True)
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([EXE, str(n), str(n), str(n)], cwd=str(run), capture_output=True, text=True, timeout=120,
                       env=e)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    return r


def close(got, ref, rtol):
    ref = np.asarray(ref)
    return np.allclose(got, ref, rtol=rtol, atol=rtol * np.abs(ref).max())


@pytest.mark.gpu
def test_driver_run_matches_oracle(require_gpu, tmp_path):
    n, steps, imix, iout = 32, 6, 3, 4
    p = Params(nx=n, ny=n, nz=n, nu=1e-3, dt=5e-3, omega=1.0, famp=0.05, k1f=2.0, k2f=3.0,
               npas=steps, iout=iout, imix=imix)
    o = ol.Oracle(p)
    u = o.velocity(initial.initial_state(p, seed=5))
    u32 = [np.asarray(f, np.float32) for f in u]
    run = tmp_path / "plain"
    run.mkdir()
    make_run(run, p, u32)
    r = run_driver(run, n)
    assert r.returncode == 0
    # the oracle: the same restart path, then the same steps, state kept after each
    b = o.init_from_velocity([f.astype(np.float64) for f in u32])
    states, seed = [b], SEED0
    for _ in range(steps):
        b, seed = o.step(b, seed, 1)
        states.append(b)
    nb = n // 2
    _, _, _, _, inpe, inpz = o.forcing_table()
    sp = np.loadtxt(str(run / "files" / "spectra.000"))
    fl = np.loadtxt(str(run / "files" / "fluxes.000"))
    gE = np.loadtxt(str(run / "files" / "globalE.000"))
    gZ = np.loadtxt(str(run / "files" / "globalZ.000"))
    assert sp.shape == (2 * nb, 3) and fl.shape == (2 * nb, 2)
    assert gE.shape == (2, 11) and gZ.shape == (2, 11)
    for row, ipas in enumerate((3, 6)):
        before, after = states[ipas - 1], states[ipas]
        # spectra and fluxes of the state at t, before step ipas
        spo, sp2o = o.spectrum(before)
        blk = sp[row * nb:(row + 1) * nb]
        assert np.array_equal(blk[:, 0], np.arange(nb))
        assert close(blk[:, 1], spo, 1e-10) and close(blk[:, 2], sp2o, 1e-10)
        N, _ = o.nlt(before, before)
        spf = o.fluxes(before, N)
        blk = fl[row * nb:(row + 1) * nb]
        assert np.array_equal(blk[:, 0], np.arange(1, nb + 1))
        assert close(blk[:, 1], spf, 1e-9)
        # glob of the state after the step, energy then enstrophy
        g = o.glob(after)
        assert gE[row, 0] == pytest.approx(ipas * p.dt, rel=1e-14)
        assert close(gE[row, 1:10], g[:9], 1e-10) and gE[row, 10] == pytest.approx(inpe, rel=1e-14)
        assert gZ[row, 0] == gE[row, 0]
        assert close(gZ[row, 1:10], g[9:], 1e-10) and gZ[row, 10] == pytest.approx(inpz, rel=1e-14)
    # the writef2 snapshot of step 4, and the restart output
    for name, idx in (("000.001", 4), ("001", steps)):
        got = [read_record(str(run / "fields" / f"u{c}.{name}"), (n, n, n)) for c in "xyz"]
        assert rel([g.astype(np.float64) for g in got], o.velocity(states[idx])) < 1e-6  # float32 records
    assert not (run / "fields" / "ux.000.002").exists()  # no snapshot at the last step
    assert int((run / "files" / "seed.001").read_text().split()[0]) == seed
    ifr, t = (run / "curframe.dat").read_text().split()
    assert int(ifr) == 1 and float(t) == pytest.approx(steps * p.dt)

    # the same run with records split into subrecords of 5000 bytes (the
    # layout of fields larger than 2 GiB) gives the same fields, bit for bit
    run2 = tmp_path / "subrecords"
    run2.mkdir()
    make_run(run2, p, u32, maxsub=5000)
    r = run_driver(run2, n, {"GT_MAX_SUBRECORD": "5000"})
    assert r.returncode == 0
    raw = open(str(run2 / "fields" / "ux.001"), "rb").read()
    assert np.frombuffer(raw[:4], np.int32)[0] == -5000
    for c in "xyz":
        for name in ("000.001", "001"):
            a = read_record(str(run / "fields" / f"u{c}.{name}"), (n, n, n))
            b2 = read_record(str(run2 / "fields" / f"u{c}.{name}"), (n, n, n))
            assert np.array_equal(a, b2)


@pytest.mark.gpu
def test_driver_stops_on_a_non_finite_state(require_gpu, tmp_path):
    n = 16
    p = Params(nx=n, ny=n, nz=n, nu=1e-3, dt=5e-3, omega=1.0, famp=0.05, k1f=2.0, k2f=3.0,
               npas=6, iout=10 ** 6, imix=3)
    o = ol.Oracle(p)
    u32 = [np.asarray(f, np.float32) for f in o.velocity(initial.initial_state(p, seed=5))]
    u32[1][3, 4, 5] = np.nan
    make_run(tmp_path, p, u32)
    r = run_driver(tmp_path, n)
    assert r.returncode == 3
    assert "non-finite state at step 3" in r.stderr
    # nothing of the diverged step was written
    assert (tmp_path / "files" / "spectra.000").read_text() == ""
