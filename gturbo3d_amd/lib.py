"""ctypes binding of libgturbo3d.so (include/gturbo3d.h).

The library must have been built in-tree (``python -m gturbo3d_amd.build``
or ``__graft_entry__.build()``).  There is no fallback: without the HIP
library, or without a GPU, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import Params

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GT_LIB_PATH") or os.path.join(PKG, "libgturbo3d.so")

NPASS = 9
PASS_NAMES = ["A_zinv_curl_k2", "B_yinv", "C_x_cross", "D_yfwd", "E1_zfwd_step1", "E2_zfwd_step2", "F_forcing",
              "EA_zfwd_step1_zinv", "EF_zfwd_step2_force_zinv"]

# Every exported symbol of include/gturbo3d.h.
EXPORTS = [
    "gt_create", "gt_destroy", "gt_set_state", "gt_get_state", "gt_step", "gt_nonlinear",
    "gt_velocity", "gt_init_velocity", "gt_forcing_table", "gt_forcing_count", "gt_sync", "gt_stream", "gt_pruned",
    "gt_prof_enable", "gt_prof_read", "gt_last_error", "gt_spectra", "gt_fluxes",
    "gt_batch_create", "gt_batch_size", "gt_batch_set_state", "gt_batch_get_state", "gt_batch_step",
    "gt_batch_spectra", "gt_ensemble_create", "gt_ensemble_destroy", "gt_ensemble_set_seeds",
    "gt_ensemble_get_seeds", "gt_ensemble_step", "gt_ensemble_sync", "gt_autotune", "gt_set_tiles",
    "gt_get_tiles", "gt_set_zchunk", "gt_get_zchunk", "gt_set_zstreams", "gt_get_zstreams", "gt_set_fusion", "gt_get_fusion", "gt_forcing_shells", "gt_set_shell_amps", "gt_adjoint_forward", "gt_adjoint_backward",
    "gt_check_finite", "gt_spectra_noalloc",
]

GT_ENONFINITE = 6


class GtParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xlx", ctypes.c_double), ("xly", ctypes.c_double), ("xlz", ctypes.c_double),
        ("nu", ctypes.c_double), ("alpnu", ctypes.c_double), ("mu", ctypes.c_double),
        ("alpmu", ctypes.c_double), ("dt", ctypes.c_double), ("omega", ctypes.c_double),
        ("rtrunc", ctypes.c_double), ("truncate", ctypes.c_int),
        ("famp", ctypes.c_double), ("k1f", ctypes.c_double), ("k2f", ctypes.c_double),
        ("variant", ctypes.c_int), ("dt2_mode", ctypes.c_int), ("precision", ctypes.c_int),
    ]


class GturboError(RuntimeError):
    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


def check_finite(v) -> None:
    """Raise GturboError (code GT_ENONFINITE) unless every value is finite: the
    host check gt_spectra applies to its glob sums (no device work)."""
    a = np.ascontiguousarray(v, dtype=np.float64).ravel()
    _check(load().gt_check_finite(_ptr(a), int(a.size)), "gt_check_finite")


_lib = None


def load():
    """Load libgturbo3d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GturboError(f"{LIB_PATH} not found: build it with `python -m gturbo3d_amd.build` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip = ctypes.c_void_p, ctypes.c_int
    L.gt_create.argtypes = [ctypes.POINTER(GtParams), ctypes.POINTER(vp)]
    L.gt_destroy.argtypes = [vp]
    L.gt_set_state.argtypes = [vp, vp, vp, vp]
    L.gt_get_state.argtypes = [vp, vp, vp, vp]
    L.gt_step.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_int32)]
    L.gt_nonlinear.argtypes = [vp, vp, vp, vp]
    L.gt_velocity.argtypes = [vp, vp, vp, vp]
    L.gt_init_velocity.argtypes = [vp, vp, vp, vp]
    L.gt_forcing_table.argtypes = [vp, ctypes.POINTER(ip), vp, vp, vp, vp, vp, vp, ip]
    L.gt_forcing_count.argtypes = [ctypes.POINTER(GtParams), ctypes.POINTER(ip), vp, vp]
    L.gt_sync.argtypes = [vp]
    L.gt_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.gt_pruned.argtypes = [vp, ctypes.POINTER(ip)]
    L.gt_prof_enable.argtypes = [vp, ip]
    L.gt_prof_read.argtypes = [vp, vp, vp, vp, ip]
    L.gt_spectra.argtypes = [vp, vp, vp, vp, vp]
    L.gt_fluxes.argtypes = [vp, vp]
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.gt_autotune.argtypes = [vp, ip]
    L.gt_set_tiles.argtypes = [vp, ip, ip, ip]
    L.gt_set_zchunk.argtypes = [vp, ip, ip]
    L.gt_set_zstreams.argtypes = [vp, ip]
    L.gt_get_zstreams.argtypes = [vp, ctypes.POINTER(ip)]
    L.gt_set_fusion.argtypes = [vp, ip]
    L.gt_get_fusion.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip)]
    L.gt_get_zchunk.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip)]
    L.gt_get_tiles.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.POINTER(ip),
                               ctypes.POINTER(ctypes.c_double)]
    L.gt_batch_create.argtypes = [ctypes.POINTER(GtParams), ip, vp, ctypes.POINTER(vp)]
    L.gt_batch_size.argtypes = [vp, ctypes.POINTER(ip)]
    L.gt_batch_set_state.argtypes = [vp, ip, vp, vp, vp]
    L.gt_batch_get_state.argtypes = [vp, ip, vp, vp, vp]
    L.gt_batch_step.argtypes = [vp, ip, i32p]
    L.gt_batch_spectra.argtypes = [vp, ip, vp, vp, vp, vp]
    L.gt_ensemble_create.argtypes = [ctypes.POINTER(vp), ip, ip, ctypes.POINTER(vp)]
    L.gt_ensemble_destroy.argtypes = [vp]
    L.gt_ensemble_set_seeds.argtypes = [vp, ip, i32p]
    L.gt_ensemble_get_seeds.argtypes = [vp, ip, i32p]
    L.gt_ensemble_step.argtypes = [vp, ip]
    L.gt_ensemble_sync.argtypes = [vp]
    L.gt_forcing_shells.argtypes = [vp, ctypes.POINTER(ip), vp, ip]
    L.gt_set_shell_amps.argtypes = [vp, vp, ip]
    L.gt_adjoint_forward.argtypes = [vp, ip, i32p, ip]
    L.gt_adjoint_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ip]
    L.gt_check_finite.argtypes = [vp, ip]
    L.gt_spectra_noalloc.argtypes = [vp, vp, vp, vp, vp]
    L.gt_last_error.restype = ctypes.c_char_p
    L.gt_last_error.argtypes = []
    _lib = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().gt_last_error().decode(errors="replace")
        raise GturboError(f"{what} failed (code {rc}): {msg}", rc)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def to_c(p: Params) -> GtParams:
    p.validate()
    return GtParams(p.nx, p.ny, p.nz, p.xlx, p.xly, p.xlz, p.nu, p.alpnu, p.mu, p.alpmu, p.dt,
                    p.omega, p.rtrunc, p.truncate, p.famp, p.k1f, p.k2f, p.variant, p.dt2_mode,
                    p.precision)


def forcing_count(p: Params):
    """(nf, inpe, inpz) of the forced-mode table for p, computed on the host."""
    nf = ctypes.c_int(0)
    inpe, inpz = ctypes.c_double(), ctypes.c_double()
    cp = to_c(p)
    _check(load().gt_forcing_count(ctypes.byref(cp), ctypes.byref(nf), ctypes.byref(inpe), ctypes.byref(inpz)),
           "gt_forcing_count")
    return nf.value, inpe.value, inpz.value


class Context:
    """One device context: the state b and everything the step needs.

    Fields are numpy complex128 arrays of shape (NZ, NY, NX/2+1), i.e. the
    reference's column-major complex*16 (NX2P1, NY, NZ) arrays.
    """

    def __init__(self, p: Params):
        self.p = p
        h = ctypes.c_void_p()
        cp = to_c(p)
        _check(load().gt_create(ctypes.byref(cp), ctypes.byref(h)), "gt_create")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            load().gt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return self.p.shape

    def _fields(self, fs, what):
        out = []
        for f in fs:
            a = np.ascontiguousarray(f, dtype=np.complex128)
            if a.shape != self.shape:
                raise ValueError(f"{what}: expected shape {self.shape}, got {a.shape}")
            out.append(a)
        return out

    def set_state(self, b) -> None:
        b = self._fields(b, "set_state")
        _check(load().gt_set_state(self.h, *[_ptr(x) for x in b]), "gt_set_state")

    def get_state(self, out=None):
        """The three fields; out: three contiguous complex128 arrays of the
        state's shape to write them into (else new arrays)."""
        b = self._out_fields(out, "get_state")
        _check(load().gt_get_state(self.h, *[_ptr(x) for x in b]), "gt_get_state")
        return b

    def _out_fields(self, out, what):
        if out is None:
            return [np.empty(self.shape, np.complex128) for _ in range(3)]
        b = list(out)
        if len(b) != 3 or any(x.shape != self.shape or x.dtype != np.complex128 or not x.flags.c_contiguous
                              for x in b):
            raise ValueError(f"{what}: out must be three contiguous complex128 arrays of shape {self.shape}")
        return b

    def step(self, nsteps: int, seed: int) -> int:
        """Run nsteps steps; returns the advanced forcing seed."""
        s = ctypes.c_int32(seed)
        _check(load().gt_step(self.h, int(nsteps), ctypes.byref(s)), "gt_step")
        return s.value

    def sync(self) -> None:
        _check(load().gt_sync(self.h), "gt_sync")

    def nonlinear(self):
        n = [np.empty(self.shape, np.complex128) for _ in range(3)]
        _check(load().gt_nonlinear(self.h, *[_ptr(x) for x in n]), "gt_nonlinear")
        return n

    def velocity(self):
        p = self.p
        u = [np.empty((p.nz, p.ny, p.nx), np.float64) for _ in range(3)]
        _check(load().gt_velocity(self.h, *[_ptr(x) for x in u]), "gt_velocity")
        return u

    def init_velocity(self, u) -> None:
        p = self.p
        us = [np.ascontiguousarray(f, dtype=np.float64) for f in u]
        for f in us:
            if f.shape != (p.nz, p.ny, p.nx):
                raise ValueError(f"init_velocity: expected {(p.nz, p.ny, p.nx)}, got {f.shape}")
        _check(load().gt_init_velocity(self.h, *[_ptr(x) for x in us]), "gt_init_velocity")

    def forcing_table(self):
        nf = ctypes.c_int(0)
        _check(load().gt_forcing_table(self.h, ctypes.byref(nf), None, None, None, None, None, None, 0),
               "gt_forcing_table")
        n = nf.value
        ikx, iky, ikz = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        ff = np.zeros(max(n, 1))
        inpe, inpz = ctypes.c_double(), ctypes.c_double()
        _check(load().gt_forcing_table(self.h, ctypes.byref(nf), _ptr(ikx), _ptr(iky), _ptr(ikz), _ptr(ff),
                                       ctypes.byref(inpe), ctypes.byref(inpz), n), "gt_forcing_table")
        return ikx[:n], iky[:n], ikz[:n], ff[:n], inpe.value, inpz.value

    def spectra(self):
        """(sp, sp2d, hel, glob) of the current state (see gt_spectra)."""
        nb = self.p.nx // 2
        sp, sp2d, hel, glob = np.zeros(nb), np.zeros(nb), np.zeros(nb), np.zeros(18)
        _check(load().gt_spectra(self.h, _ptr(sp), _ptr(sp2d), _ptr(hel), _ptr(glob)), "gt_spectra")
        return sp, sp2d, hel, glob

    def spectra_noalloc(self):
        """spectra() with the partial sums in the library's fixed device array
        (gt_spectra_noalloc): no device allocation."""
        nb = self.p.nx // 2
        sp, sp2d, hel, glob = np.zeros(nb), np.zeros(nb), np.zeros(nb), np.zeros(18)
        _check(load().gt_spectra_noalloc(self.h, _ptr(sp), _ptr(sp2d), _ptr(hel), _ptr(glob)), "gt_spectra_noalloc")
        return sp, sp2d, hel, glob

    def fluxes(self):
        spf = np.zeros(self.p.nx // 2)
        _check(load().gt_fluxes(self.h, _ptr(spf)), "gt_fluxes")
        return spf

    def autotune(self, nrep: int = 3):
        """Pick the fastest pencil widths and z chunk; returns (tz, tze, ty, ms per nonlinear evaluation)."""
        _check(load().gt_autotune(self.h, int(nrep)), "gt_autotune")
        return self.tiles()

    def set_tiles(self, tz: int, tze: int, ty: int) -> None:
        """Pencil widths of the z-inverse, z-forward and y passes (0: default)."""
        _check(load().gt_set_tiles(self.h, int(tz), int(tze), int(ty)), "gt_set_tiles")

    def set_zchunk(self, planes: int, with_d: bool = True) -> None:
        """z planes per chunk of the y-inverse and x passes (0: all at once); with_d:
        the y-forward pass runs in the chunks too (else once after them)."""
        _check(load().gt_set_zchunk(self.h, int(planes), 1 if with_d else 0), "gt_set_zchunk")

    def set_zstreams(self, n: int) -> None:
        """The z chunks of passes B-D on one stream or alternating between two."""
        _check(load().gt_set_zstreams(self.h, int(n)), "gt_set_zstreams")

    def zstreams(self) -> int:
        v = ctypes.c_int()
        _check(load().gt_get_zstreams(self.h, ctypes.byref(v)), "gt_get_zstreams")
        return v.value

    def set_fusion(self, on: bool) -> None:
        """Fuse the z-forward passes with the z-inverse passes that follow them
        (bh kept in registers; b not read back within one step call) where the
        library compiles it.  On by default in FP64, off in FP32-bulk; autotune()
        keeps whichever runs a step faster."""
        _check(load().gt_set_fusion(self.h, 1 if on else 0), "gt_set_fusion")

    def fusion(self):
        """(on, active): the setting, and whether the current context uses it."""
        v, a = ctypes.c_int(), ctypes.c_int()
        _check(load().gt_get_fusion(self.h, ctypes.byref(v), ctypes.byref(a)), "gt_get_fusion")
        return bool(v.value), bool(a.value)

    def zchunk(self):
        """(planes, with_d)"""
        v, d = ctypes.c_int(), ctypes.c_int()
        _check(load().gt_get_zchunk(self.h, ctypes.byref(v), ctypes.byref(d)), "gt_get_zchunk")
        return v.value, bool(d.value)

    def tiles(self):
        tz, tze, ty, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(load().gt_get_tiles(self.h, ctypes.byref(tz), ctypes.byref(tze), ctypes.byref(ty), ctypes.byref(ms)),
               "gt_get_tiles")
        return tz.value, tze.value, ty.value, ms.value

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(load().gt_stream(self.h, ctypes.byref(s)), "gt_stream")
        return s.value or 0

    def pruned(self) -> bool:
        v = ctypes.c_int()
        _check(load().gt_pruned(self.h, ctypes.byref(v)), "gt_pruned")
        return bool(v.value)

    def forcing_shells(self) -> np.ndarray:
        """|k| of each forcing shell (forced modes grouped by |k|), ascending."""
        n = ctypes.c_int(0)
        _check(load().gt_forcing_shells(self.h, ctypes.byref(n), None, 0), "gt_forcing_shells")
        kk = np.zeros(max(n.value, 1))
        _check(load().gt_forcing_shells(self.h, ctypes.byref(n), _ptr(kk), n.value), "gt_forcing_shells")
        return kk[: n.value]

    def set_shell_amps(self, amps=None) -> None:
        """Per-shell forcing amplitudes (None: famp for every shell)."""
        if amps is None:
            _check(load().gt_set_shell_amps(self.h, None, 0), "gt_set_shell_amps")
            return
        a = np.ascontiguousarray(amps, dtype=np.float64)
        _check(load().gt_set_shell_amps(self.h, _ptr(a), int(a.size)), "gt_set_shell_amps")

    def adjoint_forward(self, nsteps: int, seed: int, every: int) -> int:
        """nsteps steps kept for the adjoint (a checkpoint every `every` steps);
        returns the advanced seed.  The context keeps the checkpoints of its
        last forward run only: ``adjoint_run`` counts the runs, so that a
        holder of an older run can tell its checkpoints were replaced."""
        s = ctypes.c_int32(seed)
        self.adjoint_run = getattr(self, "adjoint_run", 0) + 1
        _check(load().gt_adjoint_forward(self.h, int(nsteps), ctypes.byref(s), int(every)), "gt_adjoint_forward")
        return s.value

    def adjoint_backward(self, lam, out=None):
        """dJ/db_0 (three complex fields) and dJ/d(shell amplitudes) from
        lam = dJ/db_N of the last adjoint_forward run.  out: three contiguous
        complex128 arrays of the state's shape to write dJ/db_0 into (else
        new arrays)."""
        lam = self._fields(lam, "adjoint_backward")
        g = self._out_fields(out, "adjoint_backward")
        nsh = len(self.forcing_shells())
        ga = np.zeros(max(nsh, 1))
        _check(load().gt_adjoint_backward(self.h, *[_ptr(x) for x in lam], *[_ptr(x) for x in g], _ptr(ga), nsh),
               "gt_adjoint_backward")
        return g, ga[:nsh]

    def prof_enable(self, on: bool = True) -> None:
        _check(load().gt_prof_enable(self.h, 1 if on else 0), "gt_prof_enable")

    def prof_read(self):
        ms = np.zeros(NPASS)
        n = np.zeros(NPASS, np.int64)
        by = np.zeros(NPASS)
        _check(load().gt_prof_read(self.h, _ptr(ms), _ptr(n), _ptr(by), NPASS), "gt_prof_read")
        return ms, n, by


class Batch(Context):
    """nmembers independent members on one grid (per-member omega and seed)."""

    def __init__(self, p: Params, nmembers: int, omegas=None):
        self.p = p
        self.n = int(nmembers)
        h = ctypes.c_void_p()
        cp = to_c(p)
        om = None if omegas is None else np.ascontiguousarray(omegas, dtype=np.float64)
        if om is not None and om.shape != (self.n,):
            raise ValueError("omegas must have one entry per member")
        _check(load().gt_batch_create(ctypes.byref(cp), self.n, None if om is None else _ptr(om), ctypes.byref(h)),
               "gt_batch_create")
        self.h = h

    def set_member(self, m: int, b) -> None:
        b = self._fields(b, "set_member")
        _check(load().gt_batch_set_state(self.h, int(m), *[_ptr(x) for x in b]), "gt_batch_set_state")

    def get_member(self, m: int):
        b = [np.empty(self.shape, np.complex128) for _ in range(3)]
        _check(load().gt_batch_get_state(self.h, int(m), *[_ptr(x) for x in b]), "gt_batch_get_state")
        return b

    def step_batch(self, nsteps: int, seeds):
        s = np.ascontiguousarray(seeds, dtype=np.int32).copy()
        if s.shape != (self.n,):
            raise ValueError("one seed per member")
        _check(load().gt_batch_step(self.h, int(nsteps), s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "gt_batch_step")
        return s

    def member_spectra(self, m: int):
        nb = self.p.nx // 2
        sp, sp2d, hel, glob = np.zeros(nb), np.zeros(nb), np.zeros(nb), np.zeros(18)
        _check(load().gt_batch_spectra(self.h, int(m), _ptr(sp), _ptr(sp2d), _ptr(hel), _ptr(glob)),
               "gt_batch_spectra")
        return sp, sp2d, hel, glob


class Ensemble:
    """Batches of several grids stepped together, chunk steps per hipGraph replay."""

    def __init__(self, batches, chunk: int):
        self.batches = list(batches)
        arr = (ctypes.c_void_p * len(self.batches))(*[b.h.value for b in self.batches])
        h = ctypes.c_void_p()
        _check(load().gt_ensemble_create(arr, len(self.batches), int(chunk), ctypes.byref(h)), "gt_ensemble_create")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            load().gt_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_seeds(self, q: int, seeds) -> None:
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        _check(load().gt_ensemble_set_seeds(self.h, int(q), s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "gt_ensemble_set_seeds")

    def get_seeds(self, q: int):
        s = np.zeros(self.batches[q].n, np.int32)
        _check(load().gt_ensemble_get_seeds(self.h, int(q), s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "gt_ensemble_get_seeds")
        return s

    def step(self, nsteps: int) -> None:
        _check(load().gt_ensemble_step(self.h, int(nsteps)), "gt_ensemble_step")

    def sync(self) -> None:
        _check(load().gt_ensemble_sync(self.h), "gt_ensemble_sync")
