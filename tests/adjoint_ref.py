"""Numpy restatement of the adjoint of the step (test infrastructure only).

The library's adjoint kernels are hand-derived; this module states the same
derivation with numpy FFTs so that it can be pinned on the CPU against
central differences of the oracle's forward step (or_step), and the GPU
adjoint compared with it.  Forward pieces (curl, invcurl, step1, the steps
themselves, the forcing) come from the oracle.

Conventions: fields are three complex (NZ, NY, NX/2+1) arrays; the inner
product is sum Re(conj(x) y) over every stored coefficient; F^-1 is the
unnormalised C2R of fft_inv (src/fft_inv.f90:12), F the normalised R2C of
fft_dir (src/fft_dir.f90:12-25).  With wx = 1 on the kx = 0 and kx = NX/2
planes and 2 elsewhere:
    (F^-1)^T g = wx * rfftn(g),        F^T m = F^-1(m / wx) / (NX NY NZ).
"""
from __future__ import annotations

import numpy as np

import oracle_lib as ol


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


class AdjointRef:
    def __init__(self, p, o: ol.Oracle | None = None):
        self.p = p
        self.o = o or ol.Oracle(p)
        t = self.o.tables()
        self.kx = t["fkx"][None, None, :]
        self.ky = t["fky"][None, :, None]
        self.kz = t["fkz"][:, None, None]
        self.k2 = (self.kx ** 2 + self.ky ** 2) + self.kz ** 2
        self.emk = t["emk"]
        self.masked = self.o.mask()
        wx = np.full(p.nx // 2 + 1, 2.0)
        wx[0] = wx[-1] = 1.0
        self.wx = wx[None, None, :]
        self.ntot = float(p.nx * p.ny * p.nz)
        self.s = (p.nz, p.ny, p.nx)
        self.dt2 = 0.5 * p.dt if p.dt2_mode == 1 else 0.0

    # ---- linear pieces ----
    def T(self, f):
        if self.p.truncate != 1:
            return f
        g = f.copy()
        g[self.masked] = 0.0
        g[0, 0, 0] = 0.0
        return g

    def Z(self, f):
        g = f.copy()
        g[0, 0, 0] = 0.0
        return g

    def finv(self, X):
        return np.fft.irfftn(X, s=self.s, axes=(0, 1, 2), norm="forward")

    def fu(self, g):
        return np.fft.rfftn(g, axes=(0, 1, 2))

    def coriolis_T(self, mu):
        return [-1j * self.p.omega * self.kz * m for m in mu]

    def dN_da_T(self, a, lam):
        """(dN/da^T lam, mu = I T lam) at a."""
        u = [self.finv(x) for x in self.o.curl(a)]
        w = [self.finv(self.k2 * x) for x in a]
        mu = self.o.invcurl([self.T(x) for x in lam])
        pp = [self.finv(m / self.wx) / self.ntot for m in mu]
        A = [self.fu(c) for c in _cross(w, pp)]
        B = [self.fu(c) for c in _cross(pp, u)]
        cA = self.o.curl(A)
        return [self.wx * (x + self.k2 * y) for x, y in zip(cA, B)], mu

    def step_adjoint(self, b, lam):
        """(d b_{n+1}/d b_n)^T lam at b_n = b (projected on the kept modes when truncate = 1)."""
        e = self.emk
        nu = [self.Z(e * x) for x in lam]
        if self.dt2 == 0.0:
            bh = [self.Z(e * x) for x in b]
        else:
            c1, _ = self.o.nlt(b, b)
            bh = self.o.step1(b, c1)
        g, mu2 = self.dN_da_T(bh, [self.p.dt * x for x in nu])
        co = self.coriolis_T(mu2)
        out = [e * v + c + self.Z(e * gg) for v, c, gg in zip(nu, co, g)]
        if self.dt2 != 0.0:
            h = [self.Z(e * gg) for gg in g]
            g1, mu1 = self.dN_da_T(b, [self.dt2 * x for x in h])
            co1 = self.coriolis_T(mu1)
            out = [o + x + c for o, x, c in zip(out, g1, co1)]
        return [self.T(x) for x in out]

    # ---- forcing shells ----
    def shells(self):
        """Distinct k^2 of the forced modes (ascending, merged at 1e-9 relative)."""
        ikx, iky, ikz, _, _, _ = self.o.forcing_table()
        k2 = [self.k2[k - 1, j - 1, i - 1] for i, j, k in zip(ikx, iky, ikz)]
        out = []
        for v in sorted(k2):
            if not out or v > out[-1] * (1.0 + 1e-9):
                out.append(v)
        return np.array(out)

    def unit_increments(self, seed):
        """The forcing increments of one step at famp = 1, and the advanced seed."""
        from gturbo3d_amd.params import Params

        p1 = Params(**{**self.p.__dict__, "famp": 1.0})
        o1 = ol.Oracle(p1)
        z = [np.zeros(self.s[:2] + (self.p.nx // 2 + 1,), np.complex128) for _ in range(3)]
        inc, s = o1.forcing(z, seed)
        return inc, s

    def gradient(self, b0, seed, nsteps, lamN):
        """dJ/db_0 and dJ/d(shell amplitudes) (all shells at famp) for lam_N = dJ/db_N."""
        states, seeds = [b0], [seed]
        b, s = b0, seed
        for _ in range(nsteps):
            b, s = self.o.step(b, s, 1)
            states.append(b)
            seeds.append(s)
        sh = self.shells()
        ga = np.zeros(len(sh))
        lam = [np.array(x, dtype=np.complex128) for x in lamN]
        for n in range(nsteps - 1, -1, -1):
            inc, _ = self.unit_increments(seeds[n])
            for f in range(3):
                nz = np.nonzero(inc[f])
                for idx in zip(*nz):
                    k2 = self.k2[idx]
                    q = int(np.argmin(np.abs(sh - k2)))
                    ga[q] += float(np.real(np.conj(lam[f][idx]) * inc[f][idx]))
            lam = self.step_adjoint(states[n], lam)
        return lam, ga
