"""Run parameters of the pseudo-spectral step, mirroring Gturbo3D's ``params.dat``.

The six-line ``params.dat`` layout is the one read by ``inieuler``
(src/inieuler.f90:14-20)::

    xlx xly xlz
    nu alpnu mu alpmu
    dt omega
    npas iout imix
    rtrunc truncate
    famp k1f k2f

The grid size (compile-time ``NX, NY, NZ`` in src/modules.f90:4), the
forcing variant (chosen at link time in Makefile:20-33) and the stage-1
coefficient mode are run-time fields here.
"""
from __future__ import annotations

import dataclasses
import math
import re

# Stage-1 coefficient modes.  The reference never assigns dt2
# (src/modules.f90:32, read at src/rkstep.f90:34-36), so its half step is
# bh = emk*b; MIDPOINT sets dt2 = dt/2.
DT2_REFERENCE = 0
DT2_MIDPOINT = 1

FORCING_3D3C = 3
FORCING_2D2C = 2

PREC_FP64 = 0
PREC_FP32_BULK = 1

TWO_PI = 6.283185307179586


@dataclasses.dataclass
class Params:
    nx: int = 64
    ny: int = 64
    nz: int = 64
    xlx: float = TWO_PI
    xly: float = TWO_PI
    xlz: float = TWO_PI
    nu: float = 1e-3
    alpnu: float = 1.0
    mu: float = 0.0
    alpmu: float = 0.0
    dt: float = 5e-3
    omega: float = 1.0
    npas: int = 100
    iout: int = 10**9
    imix: int = 10**9
    rtrunc: float = 2.0 / 3.0
    truncate: int = 1
    famp: float = 0.0
    k1f: float = 2.0
    k2f: float = 3.0
    variant: int = FORCING_3D3C
    dt2_mode: int = DT2_REFERENCE
    precision: int = PREC_FP64

    @property
    def nxh(self) -> int:
        return self.nx // 2 + 1

    @property
    def shape(self):
        """Shape of one spectral field, C order [NZ][NY][NX/2+1]."""
        return (self.nz, self.ny, self.nxh)

    def validate(self) -> None:
        for n in (self.nx, self.ny, self.nz):
            if n < 2 or n % 2:
                raise ValueError(f"grid sizes must be even and >= 2, got {n}")
        if self.variant not in (FORCING_2D2C, FORCING_3D3C):
            raise ValueError(f"unknown forcing variant {self.variant}")
        if self.dt2_mode not in (DT2_REFERENCE, DT2_MIDPOINT):
            raise ValueError(f"unknown dt2 mode {self.dt2_mode}")


_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][-+]?\d+)?")


def _numbers(line: str):
    return [float(t.replace("d", "e").replace("D", "e")) for t in _NUM.findall(line)]


def read_params_dat(path: str, nx: int, ny: int, nz: int, variant: int = FORCING_3D3C,
                    dt2_mode: int = DT2_REFERENCE, precision: int = PREC_FP64) -> Params:
    """Read a ``params.dat`` (src/inieuler.f90:14-20) for a given grid."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines()]
    rows = []
    for ln in lines:
        nums = _numbers(ln)
        if nums:
            rows.append(nums)
        if len(rows) == 6:
            break
    if len(rows) < 6 or [len(r) for r in rows[:6]] < [3, 4, 2, 3, 2, 3]:
        raise ValueError(f"{path}: expected 6 parameter lines (inieuler.f90:14-20)")
    (xlx, xly, xlz), (nu, alpnu, mu, alpmu), (dt, omega), (npas, iout, imix), \
        (rtrunc, truncate), (famp, k1f, k2f) = [r[:n] for r, n in zip(rows, (3, 4, 2, 3, 2, 3))]
    p = Params(nx=nx, ny=ny, nz=nz, xlx=xlx, xly=xly, xlz=xlz, nu=nu, alpnu=alpnu, mu=mu,
               alpmu=alpmu, dt=dt, omega=omega, npas=int(npas), iout=int(iout), imix=int(imix),
               rtrunc=rtrunc, truncate=int(truncate), famp=famp, k1f=k1f, k2f=k2f,
               variant=variant, dt2_mode=dt2_mode, precision=precision)
    p.validate()
    return p


def write_params_dat(p: Params, path: str) -> None:
    """Write ``p`` in the six-line layout of src/params.dat:1-6."""
    def d(x):
        return repr(float(x)).replace("e", "d") if "e" in repr(float(x)) else repr(float(x)) + "d0"
    with open(path, "w") as f:
        f.write(f"{d(p.xlx)}\t{d(p.xly)}\t{d(p.xlz)}\n")
        f.write(f"{d(p.nu)}\t{d(p.alpnu)}\t{d(p.mu)}\t{d(p.alpmu)}\n")
        f.write(f"{d(p.dt)}\t{d(p.omega)}\n")
        f.write(f"{p.npas}\t{p.iout}\t{p.imix}\n")
        f.write(f"{d(p.rtrunc)}\t{p.truncate}\n")
        f.write(f"{d(p.famp)}\t{d(p.k1f)}\t{d(p.k2f)}\n")


def famp_for_injection(rate: float, nf: int) -> float:
    """Forcing amplitude giving a mean energy input ``inpe = nf*famp**2 = rate``
    (src/iniforcing_3d3c.f90:39: inpe sums k2*(famp*ff)**2 with ff = 1/|k|)."""
    return math.sqrt(rate / nf)
