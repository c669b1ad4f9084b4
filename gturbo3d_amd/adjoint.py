"""Differentiable runs of the step for PyTorch.

``StepRun`` is a ``torch.autograd.Function`` whose forward runs nsteps steps
of the library (checkpointing the state every ``every`` steps) and whose
backward is the library's hand-written adjoint (gt_adjoint_forward /
gt_adjoint_backward, through ctypes).  Inputs and outputs are float64
tensors holding the real view of the three spectral fields,
shape (3, NZ, NY, NX/2+1, 2), so torch's gradient of a real loss is
(dJ/dRe, dJ/dIm) for every stored coefficient, and a float64 tensor of the
per-shell forcing amplitudes.  The step itself has no counterpart in this
form in the reference (src/main.f90:124-150 runs it forward only).
"""
from __future__ import annotations

import numpy as np
import torch

from . import lib
from .params import Params


def to_real(b) -> torch.Tensor:
    """Three complex (NZ, NY, NXH) arrays -> float64 tensor (3, NZ, NY, NXH, 2)."""
    a = np.stack([np.ascontiguousarray(f, dtype=np.complex128) for f in b])
    return torch.from_numpy(a.view(np.float64).reshape(a.shape + (2,)).copy())


def to_complex(t: torch.Tensor):
    """float64 tensor (3, NZ, NY, NXH, 2) -> three complex128 arrays."""
    a = np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())
    c = a.view(np.complex128)[..., 0]
    return [np.ascontiguousarray(c[f]) for f in range(3)]


def _host_empty(shape, dtype=torch.float64) -> torch.Tensor:
    """An uninitialised CPU tensor from torch's caching host allocator when a
    GPU is present: pinned, so the library's copies to and from it go by DMA,
    and reused from call to call, so that filling it does not page-fault
    through fresh memory (51 MB per state at 128^3)."""
    return torch.empty(shape, dtype=dtype, pin_memory=torch.cuda.is_available())


def _fields_of(t: torch.Tensor):
    """The three complex fields viewing the memory of a CPU float64 tensor
    (3, NZ, NY, NXH, 2): the library reads or writes them in place."""
    return [t[f].numpy().view(np.complex128)[..., 0] for f in range(3)]


class StepRun(torch.autograd.Function):
    """b_N = run(b_0, amps): nsteps steps on context ``g`` from forcing seed ``seed``.

    The checkpoints live in the native context, which holds one run at a
    time: a second run on the same context before this node's backward
    replaces them, and this node's backward then raises instead of
    differentiating the other run.  Use one context per live run."""

    @staticmethod
    def forward(ctx, b0, amps, g, nsteps, seed, every):
        g.set_state(to_complex(b0))
        g.set_shell_amps(amps.detach().to("cpu", torch.float64).numpy())
        ctx.seed_out = g.adjoint_forward(int(nsteps), int(seed), int(every))
        ctx.g = g
        ctx.run = g.adjoint_run
        out = _host_empty((3,) + tuple(g.shape) + (2,))
        g.get_state(out=_fields_of(out))
        return out.to(b0.device)

    @staticmethod
    def backward(ctx, grad):
        if ctx.g.adjoint_run != ctx.run:
            raise RuntimeError(f"StepRun.backward: the context's checkpoints belong to run {ctx.g.adjoint_run}, "
                               f"not to this node's run {ctx.run} (one live run per context)")
        # the library writes dJ/db_0 straight into the result tensor's memory
        # (its real view is the three complex fields back to back)
        res = _host_empty(grad.shape)
        _, ga = ctx.g.adjoint_backward(to_complex(grad), out=_fields_of(res))
        dev = grad.device
        return res.to(dev), torch.from_numpy(ga.copy()).to(dev), None, None, None, None


def run(g: lib.Context, b0: torch.Tensor, amps: torch.Tensor, nsteps: int, seed: int, every: int = 20):
    """Differentiable b_N of nsteps steps (see StepRun)."""
    return StepRun.apply(b0, amps, g, nsteps, seed, every)


def band_weights(p: Params, kmin: float, kmax: float) -> np.ndarray:
    """Weights w with sum(w * |b|^2) = energy in kmin <= |k| <= kmax of a
    solenoidal b: fac * k^2, fac = 1/2 on the kx = 0 and kx = NX/2 planes
    (the half-spectrum count of src/diagnostics.f90:22-35).  Shape
    (NZ, NY, NXH, 1), for the real view."""
    from .initial import wavenumbers

    kx, ky, kz = wavenumbers(p)
    k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    kk = np.sqrt(k2)
    fac = np.ones(kx.shape)
    fac[0] = 0.5
    fac[-1] = 0.5
    w = np.where((kk >= kmin) & (kk <= kmax), k2 * fac[None, None, :], 0.0)
    return w[..., None]


class _BandEnergy(torch.autograd.Function):
    """J = sum w b^2 and its gradient 2 w b, each one pass over b into a
    buffer of the caching host allocator (_host_empty) for CPU tensors."""

    @staticmethod
    def forward(ctx, b, w):
        ctx.save_for_backward(b, w)
        s = _host_empty(b.shape, b.dtype) if b.device.type == "cpu" else torch.empty_like(b)
        torch.mul(b, b, out=s)
        s.mul_(w)
        return s.sum()

    @staticmethod
    def backward(ctx, gout):
        b, w = ctx.saved_tensors
        g = _host_empty(b.shape, b.dtype) if b.device.type == "cpu" else torch.empty_like(b)
        torch.mul(b, w, out=g)
        g.mul_(2.0 * gout)
        return g, None


def band_energy(b: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """J = sum_k w |b|^2 over the real view b (3, NZ, NY, NXH, 2)."""
    return _BandEnergy.apply(b, w.to(b.device, b.dtype))
