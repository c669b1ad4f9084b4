"""gt_spectra_noalloc (the diagnostics sweep with its partial sums in the
library's fixed device array) against gt_spectra, through the C ABI: two
calls in a row on two contexts, the second on a smaller grid, each on a
context whose gt_spectra has not allocated anything yet."""
import sys

import numpy as np
import pytest

from gturbo3d_amd import initial, lib
from gturbo3d_amd.params import Params

pytestmark = pytest.mark.gpu


def test_spectra_noalloc_matches_spectra(require_gpu):
    for n in (64, 32):
        print(f"spectra_noalloc {n}^3", flush=True)
        sys.stdout.flush()
        p = Params(nx=n, ny=n, nz=n, nu=1e-3, dt=5e-3, omega=1.5, famp=0.05, k1f=2.0, k2f=3.0)
        with lib.Context(p) as g:
            g.set_state(initial.initial_state(p, seed=n))
            g.step(3, 7)
            got = g.spectra_noalloc()  # before gt_spectra's first (allocating) call
            ref = g.spectra()
            again = g.spectra_noalloc()
        # the per-block sums accumulate with LDS atomics: equal up to their order
        for x, y in zip(got, ref):
            assert np.allclose(x, y, rtol=1e-13, atol=0)
        for x, y in zip(again, ref):
            assert np.allclose(x, y, rtol=1e-13, atol=0)
        assert ref[0].sum() > 0 and ref[3][0] > 0
