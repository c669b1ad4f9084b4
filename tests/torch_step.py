"""A small PyTorch FP64 restatement of the step, differentiated by torch's
own autograd (test infrastructure only).

It is independent of the hand derivation of the adjoint: the transposes of
the transforms, the cross product and the pointwise operators all come from
torch.  Tables (wavenumbers, emk, the trunc mask) and the unit forcing
increments come from the oracle.  Used for small grids on the CPU.
"""
from __future__ import annotations

import numpy as np
import torch

import oracle_lib as ol


class TorchStep:
    def __init__(self, p, o: ol.Oracle | None = None):
        self.p = p
        self.o = o or ol.Oracle(p)
        t = self.o.tables()
        f = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))
        self.kx = f(t["fkx"])[None, None, :]
        self.ky = f(t["fky"])[None, :, None]
        self.kz = f(t["fkz"])[:, None, None]
        self.k2 = (self.kx ** 2 + self.ky ** 2) + self.kz ** 2
        self.emk = f(t["emk"])
        keep = np.ones(self.o.shape, np.float64)
        if p.truncate == 1:
            keep[self.o.mask()] = 0.0
            keep[0, 0, 0] = 0.0
        self.keep = f(keep)
        z = np.ones(self.o.shape)
        z[0, 0, 0] = 0.0
        self.z = f(z)
        k2s = self.k2.clone()
        k2s[0, 0, 0] = 1.0
        self.k2s = k2s
        self.s = (p.nz, p.ny, p.nx)
        self.ntot = float(p.nx * p.ny * p.nz)
        self.dt2 = 0.5 * p.dt if p.dt2_mode == 1 else 0.0

    def curl(self, a):
        ax, ay, az = a
        j = 1j
        return [j * (self.ky * az - self.kz * ay), j * (self.kz * ax - self.kx * az), j * (self.kx * ay - self.ky * ax)]

    def invcurl(self, a):
        c = self.curl(a)
        out = [x / self.k2s for x in c]
        # k = 0 left as it is
        return [torch.where(self.z == 0.0, y, x) for x, y in zip(out, a)]

    def nlt(self, a, b):
        u = [torch.fft.irfftn(x, s=self.s, dim=(0, 1, 2), norm="forward") for x in self.curl(a)]
        w = [torch.fft.irfftn(self.k2 * x, s=self.s, dim=(0, 1, 2), norm="forward") for x in a]
        c = [u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]]
        v = [torch.fft.rfftn(x, dim=(0, 1, 2)) / self.ntot for x in c]
        v = [x + 1j * self.p.omega * self.kz * y for x, y in zip(v, b)]
        return [self.keep * x for x in self.invcurl(v)] if self.p.truncate == 1 else self.invcurl(v)

    def step(self, b, inc):
        e = self.emk
        n1 = self.nlt(b, b)
        bh = [self.z * e * (x + self.dt2 * y) for x, y in zip(b, n1)]
        n2 = self.nlt(bh, b)
        out = [self.z * e * (e * x + self.p.dt * y) for x, y in zip(b, n2)]
        return [x + y for x, y in zip(out, inc)]

    def run(self, b0_real, amps, seed, nsteps, shells):
        """b0_real: (3, NZ, NY, NXH, 2) float64; amps: per-shell amplitudes."""
        from adjoint_ref import AdjointRef

        ref = AdjointRef(self.p, self.o)
        b = [torch.view_as_complex(b0_real[f]) for f in range(3)]
        s = seed
        for _ in range(nsteps):
            unit, s = ref.unit_increments(s)
            inc = []
            for f in range(3):
                acc = torch.zeros(self.o.shape, dtype=torch.complex128)
                for q, k2 in enumerate(shells):
                    sel = np.isclose(ref.k2, k2, rtol=1e-9, atol=0.0)
                    part = torch.from_numpy(np.where(sel, unit[f], 0.0))
                    acc = acc + amps[q] * part
                inc.append(acc)
            b = self.step(b, inc)
        return torch.stack([torch.view_as_real(x) for x in b])
