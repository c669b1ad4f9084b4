"""In-tree build of libgturbo3d.so (HIP, gfx950) and of the Fortran driver.

``python -m gturbo3d_amd.build`` compiles every source under csrc/ with
hipcc for gfx950, in parallel, and links gturbo3d_amd/libgturbo3d.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgturbo3d.so")
ROOT = os.path.dirname(PKG)

SOURCES = ["kernels_z.hip", "kernels_y.hip", "kernels_x.hip", "kernels_diag.hip", "kernels_adj.hip", "gt_api.cpp"]

ARCH = os.environ.get("GT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: the compiler emits no packed FP32 (v_pk_*_f32) instructions,
# some forms of which have given wrong results on this hardware
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-x", "hip",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def _stale(src: str, out: str) -> bool:
    """True when the object is older than its source or any header of csrc/ or include/."""
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    deps = [os.path.join(CSRC, src)]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    if "gturbo3d.h" in open(deps[0]).read():
        deps += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, obj_dir: str = OBJ, defines=(), incremental: bool = False) -> str:
    out = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(src, out):
        return out
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-I", os.path.join(ROOT, "include"), "-c",
           os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(ln for ln in (r.stderr or "").splitlines() if "packed-fp32-ops" not in ln)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{errs}")
    return out


def build_lib(verbose: bool = True, out: str = LIB, defines=(), incremental: bool = False) -> str:
    """Compile and link the library; ``defines`` are extra -D options (tile-shape overrides).
    ``incremental`` recompiles only sources older than their object files."""
    obj_dir = OBJ if not defines else os.path.join(OBJ, "_".join(defines).replace("=", ""))
    os.makedirs(obj_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir, defines, incremental), SOURCES))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build_driver(verbose: bool = True) -> str | None:
    """Compile the Fortran driver against libgturbo3d.so (amdflang)."""
    fc = shutil.which("amdflang") or "/opt/rocm/lib/llvm/bin/amdflang"
    src = os.path.join(PKG, "driver", "gturb_driver.f90")
    if not (os.path.exists(fc) and os.path.exists(src)):
        raise RuntimeError(f"cannot build the Fortran driver: {fc} or {src} missing")
    exe = os.path.join(PKG, "driver", "gturb_mi355x")
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    cmd = [fc, "-O2", "-module-dir", mod, "-o", exe, src, f"-L{PKG}", "-lgturbo3d", f"-Wl,-rpath,{PKG}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stderr}")
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    build_lib(incremental="--incremental" in sys.argv)
    if "--no-driver" not in sys.argv:
        build_driver()
