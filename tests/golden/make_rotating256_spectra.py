"""Golden fixture for the long-run spectra check: the 256^3 rotating run
(nu = 2e-4, Omega = 5, 2D2C forcing on (3, 4], dt = 2e-3, 2000 steps, seed 0)
stepped by the CPU oracle, with the energy spectrum, the kz = 0 spectrum and
glob recorded every 100 steps (imix = 100).

The run uses the midpoint stage 1 (dt2 = dt/2).  With dt2 = 0, the value the
reference leaves it at (src/modules.f90:32), stage 1 is bh = emk*b and the
nonlinear and Coriolis terms advance by forward Euler; this configuration
then diverges between steps 300 and 400 (E = 0.5694 at step 300, 6.69e82 at
step 400, identically in the oracle and on the GPU).

    python tests/golden/make_rotating256_spectra.py [--steps 2000] [--threads 6]

Writes tests/golden/rotating256_midpoint_spectra.npz.  Takes about 1.8 h on 6 cores.

With --mode reference the same run uses the reference's stage 1 (dt2 = 0)
for 300 steps, records every 100, and writes
tests/golden/rotating256_reference_spectra.npz (about 16 minutes on 8 cores):

    python tests/golden/make_rotating256_spectra.py --mode reference
"""
import argparse
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from gturbo3d_amd import initial  # noqa: E402
from gturbo3d_amd.params import FORCING_2D2C, Params  # noqa: E402
import oracle_lib as ol  # noqa: E402


def rotating256_params(n=256, dt2_mode=1):
    # famp = sqrt(0.1 / nf) with nf = 10 forced modes on (3, 4]; dt2_mode 1:
    # midpoint stage 1, 0: the reference's (dt2 = 0)
    return Params(nx=n, ny=n, nz=n, nu=2e-4, alpnu=1.0, mu=0.0, alpmu=0.0, dt=2e-3, omega=5.0,
                  rtrunc=2.0 / 3.0, truncate=1, famp=math.sqrt(0.1 / 10), k1f=3.0, k2f=4.0,
                  variant=FORCING_2D2C, dt2_mode=dt2_mode, npas=2000, imix=100)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["midpoint", "reference"], default="midpoint")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--imix", type=int, default=100)
    ap.add_argument("--threads", type=int, default=6)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    ref = a.mode == "reference"
    if a.steps is None:
        a.steps = 300 if ref else 2000
    if a.out is None:
        a.out = os.path.join(HERE, "rotating256_reference_spectra.npz" if ref else "rotating256_midpoint_spectra.npz")
    ol.set_threads(a.threads)
    p = rotating256_params(a.n, dt2_mode=0 if ref else 1)
    o = ol.Oracle(p)
    assert o.nf == 10, o.nf
    b = initial.initial_state(p, seed=0)
    seed = 0
    sps, sp2ds, globs, steps = [], [], [], []
    t0 = time.time()
    for s in range(a.imix, a.steps + 1, a.imix):
        b, seed = o.step(b, seed, a.imix)
        sp, sp2d = o.spectrum(b)
        sps.append(sp)
        sp2ds.append(sp2d)
        globs.append(o.glob(b))
        steps.append(s)
        print(f"step {s}: E = {sp.sum():.6f}  ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(a.out, steps=np.array(steps), sp=np.array(sps), sp2d=np.array(sp2ds),
                        glob=np.array(globs), seed=seed, n=a.n, dt2_mode=p.dt2_mode)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
