"""gturbo3d_amd -- MI355X-native pseudo-spectral step of Gturbo3D.

The hot path (src/main.f90:124-150 of the reference: nlt, step1, nlt,
step2, forcing) runs as hand-written HIP kernels for gfx950 in
libgturbo3d.so behind the C ABI of include/gturbo3d.h.  This package holds
the kernels (csrc/), the build (build.py), the ctypes binding (lib.py), the
params.dat mirror (params.py), the seeded initial condition (initial.py)
and the Fortran driver (driver/).
"""
from .params import (DT2_MIDPOINT, DT2_REFERENCE, FORCING_2D2C, FORCING_3D3C, PREC_FP32_BULK,
                     PREC_FP64, Params, famp_for_injection, read_params_dat, write_params_dat)

__all__ = [
    "Params", "read_params_dat", "write_params_dat", "famp_for_injection",
    "DT2_REFERENCE", "DT2_MIDPOINT", "FORCING_3D3C", "FORCING_2D2C", "PREC_FP64", "PREC_FP32_BULK",
]
