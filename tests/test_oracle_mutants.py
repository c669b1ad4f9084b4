"""The closed-form pins reject plausible misreadings of the reference.

oracle/gt_oracle.c carries compile-time mutation switches (OR_MUTANT), each
planting one mistake in nlt or in the step: a vorticity |k| a or -k^2 a
instead of k^2 a (src/nlt.f90:26-40), w x u instead of u x w
(src/produ.f90:14-16), the Coriolis term with the wrong sign or on the wrong
argument (src/nlt.f90:72-77), the stage-2 Coriolis on bh instead of the old b
(src/main.f90:144), and emk outside the dt term of step2
(src/rkstep.f90:86-96).  Each mutant is compiled into a temporary directory
with the Makefile's flags; the closed forms of test_oracle_nlt_closed_form.py
must pass on the real oracle and fail on every mutant.
"""
import os
import subprocess

import pytest

import oracle_lib as ol
from test_oracle_nlt_closed_form import TOL, all_errors

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "gt_oracle.c")
CFLAGS = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c11"]

MUTANTS = {
    1: "vorticity |k| a",
    2: "w x u",
    3: "vorticity -k^2 a",
    4: "Coriolis sign",
    5: "Coriolis on the first argument",
    6: "stage-2 Coriolis on bh",
    7: "step2 without emk on the dt term",
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    procs = {}
    for k in MUTANTS:
        out = str(d / f"liboracle_m{k}.so")
        procs[k] = (out, subprocess.Popen(["gcc", *CFLAGS, f"-DOR_MUTANT={k}", "-shared", "-o", out, SRC, "-lm"]))
    libs = {}
    for k, (out, pr) in procs.items():
        assert pr.wait() == 0, f"mutant {k} failed to build"
        libs[k] = out
    return libs


def test_real_oracle_meets_every_closed_form():
    errs = all_errors(ol.Oracle)
    bad = {k: v for k, v in errs.items() if not v < TOL}
    assert not bad, bad


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_mutant_fails_a_closed_form(mutant_libs, k):
    errs = all_errors(lambda p: ol.Oracle(p, mutant_libs[k]))
    caught = {name: v for name, v in errs.items() if v > 1e3 * TOL}
    print(f"mutant {k} ({MUTANTS[k]}): caught by {sorted(caught)}")
    assert caught, f"mutant {k} ({MUTANTS[k]}) passes every closed form: {errs}"
