"""Golden fixture for the long-run check of the batched ensemble: 32 members
at 128^3 and 16 at 96^3 (FP64, nu = 1e-3, dt = 4e-3, 3D3C forcing on (2, 3]
at injection rate 0.1, member m with forcing seed m and the seeded initial
state of seed m, Omega_m the m-th draw of U[0.5, 10] from numpy's PCG64(48)),
1000 steps with the midpoint stage 1 (dt2 = dt/2).  In the reference's
stage-1 mode (dt2 = 0: forward Euler in the nonlinear and Coriolis terms)
member 0 diverges between steps 300 and 400 (E = 0.6064 at step 300, NaN at
step 400 in the oracle), as the 256^3 rotating run does.

Two members are stepped by the CPU oracle, with the energy spectrum, the
kz = 0 spectrum and glob recorded every 100 steps: member 0 (128^3,
Omega = 4.18) and member 45 (96^3, Omega = 9.89, the fastest rotation of
the draw).

    python tests/golden/make_ensemble_spectra.py [--threads 7]

Writes tests/golden/ensemble1000_spectra.npz (about 12 minutes on 7 cores).
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

from gturbo3d_amd import initial, lib  # noqa: E402
from gturbo3d_amd.params import DT2_MIDPOINT, Params  # noqa: E402

GROUPS = [(128, 32, 0), (96, 16, 32)]  # (grid, members, first member index)
MEMBERS = [0, 45]


def ensemble_omegas():
    return np.random.default_rng(48).uniform(0.5, 10.0, 48)


def ensemble_params(n, omega=1.0, dt2_mode=DT2_MIDPOINT):
    p = Params(nx=n, ny=n, nz=n, nu=1e-3, dt=4e-3, omega=float(omega), famp=0.0, k1f=2.0, k2f=3.0,
               dt2_mode=dt2_mode)
    p.famp = math.sqrt(0.1 / lib.forcing_count(p)[0])
    return p


def member_grid(m):
    for n, nm, off in GROUPS:
        if off <= m < off + nm:
            return n
    raise ValueError(m)


def main():
    import oracle_lib as ol
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--imix", type=int, default=100)
    ap.add_argument("--threads", type=int, default=7)
    ap.add_argument("--out", default=os.path.join(HERE, "ensemble1000_spectra.npz"))
    a = ap.parse_args()
    ol.set_threads(a.threads)
    omegas = ensemble_omegas()
    out = {"steps": np.arange(a.imix, a.steps + 1, a.imix), "members": np.array(MEMBERS)}
    t0 = time.time()
    for m in MEMBERS:
        n = member_grid(m)
        p = ensemble_params(n, omegas[m])
        o = ol.Oracle(p)
        b = initial.initial_state(p, seed=m)
        seed = m
        sps, sp2ds, globs = [], [], []
        for s in out["steps"]:
            b, seed = o.step(b, seed, a.imix)
            sp, sp2d = o.spectrum(b)
            sps.append(sp)
            sp2ds.append(sp2d)
            globs.append(o.glob(b))
            print(f"member {m} ({n}^3): step {s}: E = {sp.sum():.6f}  ({time.time() - t0:.0f} s)", flush=True)
        out[f"sp_{m}"] = np.array(sps)
        out[f"sp2d_{m}"] = np.array(sp2ds)
        out[f"glob_{m}"] = np.array(globs)
        out[f"seed_{m}"] = seed
    np.savez_compressed(a.out, **out)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
