"""The Fortran driver end to end: a run directory in the reference's layout
(params.dat, curframe.dat, files/seed.000, fields/u*.000), the driver
binary on the GPU, and its restart output compared with the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

from gturbo3d_amd import initial
from gturbo3d_amd.params import Params, write_params_dat
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gturbo3d_amd", "driver", "gturb_mi355x")


def write_record(path, a32):
    """One Fortran unformatted sequential record (4-byte markers) of float32."""
    raw = np.ascontiguousarray(a32, dtype=np.float32).tobytes()
    n = np.array([len(raw)], np.int32).tobytes()
    with open(path, "wb") as f:
        f.write(n + raw + n)


def read_record(path, shape):
    data = open(path, "rb").read()
    n = int(np.frombuffer(data[:4], np.int32)[0])
    assert n == int(np.prod(shape)) * 4 and len(data) == n + 8
    return np.frombuffer(data[4:4 + n], np.float32).reshape(shape)


def test_driver_binary_is_built():
    assert os.path.exists(EXE), "build() compiles the Fortran driver"


@pytest.mark.gpu
def test_driver_run_matches_oracle(require_gpu, tmp_path):
    n, steps = 32, 6
    p = Params(nx=n, ny=n, nz=n, nu=1e-3, dt=5e-3, omega=1.0, famp=0.05, k1f=2.0, k2f=3.0,
               npas=steps, iout=10 ** 6, imix=3)
    run = tmp_path
    (run / "files").mkdir()
    (run / "fields").mkdir()
    write_params_dat(p, str(run / "params.dat"))
    (run / "curframe.dat").write_text("0 0.0\n")
    (run / "files" / "seed.000").write_text("8299396\n")
    o = ol.Oracle(p)
    u = o.velocity(initial.initial_state(p, seed=5))
    u32 = [np.asarray(f, np.float32) for f in u]
    # Fortran (NX, NY, NZ) column-major == C [NZ][NY][NX]
    for name, f in zip("xyz", u32):
        write_record(str(run / "fields" / f"u{name}.000"), f)
    print(f"driver {n}^3 {steps} steps in {run}", flush=True)
    r = subprocess.run([EXE, str(n), str(n), str(n)], cwd=str(run), capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    # oracle: the same restart path, then the same steps
    b0 = o.init_from_velocity([f.astype(np.float64) for f in u32])
    b, seed = o.step(b0, 8299396, steps)
    uo = o.velocity(b)
    got = [read_record(str(run / "fields" / f"u{c}.001"), (n, n, n)) for c in "xyz"]
    num = math.sqrt(sum(float(np.sum((g.astype(np.float64) - e) ** 2)) for g, e in zip(got, uo)))
    den = math.sqrt(sum(float(np.sum(e ** 2)) for e in uo))
    assert num / den < 1e-6  # float32 output records
    assert int((run / "files" / "seed.001").read_text().split()[0]) == seed
    ifr, t = (run / "curframe.dat").read_text().split()
    assert int(ifr) == 1 and float(t) == pytest.approx(steps * p.dt)
    rows = np.loadtxt(str(run / "files" / "globalE.000"))
    assert rows.shape == (2, 4)
    sp = np.loadtxt(str(run / "files" / "spectra.000"))
    assert sp.shape == (2 * (n // 2), 3)
    # the last spectrum is the oracle's spectrum of the final state
    spo, sp2o = o.spectrum(b)
    assert np.allclose(sp[n // 2:, 1], spo, rtol=1e-9, atol=1e-18)
