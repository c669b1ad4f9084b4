"""CPU checks of the boundary: the C-ABI library loads and exports every
function include/gturbo3d.h declares; host-side argument checking."""
import ctypes
import os
import re

import pytest

import numpy as np

from gturbo3d_amd import initial, lib
from gturbo3d_amd.params import Params, read_params_dat, write_params_dat
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gturbo3d.h")).read()
    return sorted(set(re.findall(r"\b(gt_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = lib.load()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_last_error_is_callable_without_gpu():
    L = lib.load()
    assert isinstance(L.gt_last_error(), bytes)


def test_create_rejects_unsupported_grid_without_touching_the_gpu():
    p = Params(nx=20, ny=20, nz=20)
    with pytest.raises(lib.GturboError, match="not compiled"):
        lib.Context(p)


def test_create_rejects_bad_variant():
    p = Params(nx=16, ny=16, nz=16)
    cp = lib.to_c(p)
    cp.variant = 7
    h = ctypes.c_void_p()
    assert lib.load().gt_create(ctypes.byref(cp), ctypes.byref(h)) == 1
    assert b"variant" in lib.load().gt_last_error()


def test_params_dat_round_trip(tmp_path):
    p = Params(nx=64, ny=64, nz=64, nu=2e-4, dt=2e-3, omega=5.0, famp=0.1, k1f=3.0, k2f=4.0,
               npas=2000, iout=10 ** 9, imix=100)
    path = str(tmp_path / "params.dat")
    write_params_dat(p, path)
    q = read_params_dat(path, 64, 64, 64)
    for f in ("xlx", "xly", "xlz", "nu", "alpnu", "mu", "alpmu", "dt", "omega", "npas", "iout", "imix",
              "rtrunc", "truncate", "famp", "k1f", "k2f"):
        assert getattr(q, f) == getattr(p, f), f


def test_reads_the_reference_params_dat_layout(tmp_path):
    # the six lines of src/params.dat:1-6 followed by its legend
    text = ("6.2831853071795864769d0\t6.2831853071795864769d0\t0.09817477042468103d0\n"
            "1d-17\t4.0d0\t6.0d-1\t-1.0d0\n1.0d-4\t0.0d0\n10000\t10000\t1000\n"
            "0.6666666666666666667d0\t1\n1.0d-1\t15.0d0\t17.0d0\n\nxlx   yly\nnu    alpnu\n")
    path = tmp_path / "params.dat"
    path.write_text(text)
    q = read_params_dat(str(path), 1024, 1024, 16)
    assert q.nu == 1e-17 and q.alpnu == 4.0 and q.mu == 0.6 and q.alpmu == -1.0
    assert q.npas == 10000 and q.imix == 1000 and q.truncate == 1
    assert q.famp == 0.1 and q.k1f == 15.0 and q.k2f == 17.0


def test_non_finite_state_check_on_the_host():
    """The host check behind gt_spectra's GT_ENONFINITE status, fed the glob
    sums (src/diagnostics.f90:126-191) of fields with a NaN or an infinity."""
    p = Params(nx=16, ny=16, nz=16)
    o = ol.Oracle(p)
    b = initial.initial_state(p, seed=1)
    lib.check_finite(o.glob(b))
    for bad in (np.nan, np.inf, -np.inf):
        c = [f.copy() for f in b]
        c[1][2, 3, 4] = bad
        with pytest.raises(lib.GturboError) as e:
            lib.check_finite(o.glob(c))
        assert e.value.code == lib.GT_ENONFINITE
        assert b"non-finite value" in lib.load().gt_last_error()
