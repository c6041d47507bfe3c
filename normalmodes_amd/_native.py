"""Build and load the native libraries of the package.

``libnmasm.so``  CPU (g++/OpenMP) finite-element producer of A and B.
``libnmgpu.so``  the gfx950 HIP library behind the C-ABI declared in
                 include/nmgpu.h (filtered Lanczos and its operators).

Both are built in-tree next to their sources; loading never falls back to any
other implementation: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
INCLUDE = ROOT / "include"
ASM_SO = PKG / "libnmasm.so"
GPU_SO = PKG / "libnmgpu.so"

HIP_ARCH = os.environ.get("NMGPU_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# hipcc flags: gfx950 only, no packed-FP32 instructions
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={HIP_ARCH}", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-fopenmp", f"-I{INCLUDE}"]

GPU_SOURCES = ["nmgpu_kernels.hip", "nmgpu_mfma.hip", "nmgpu_assemble.hip", "nmgpu_api.hip", "nmgpu_host.cpp"]
GPU_LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_asm(verbose=False):
    cmd = ["g++", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
           str(CSRC / "nm_assemble.cpp"), "-o", str(ASM_SO) + ".tmp"]
    _run(cmd)
    os.replace(str(ASM_SO) + ".tmp", ASM_SO)
    if verbose:
        print("built", ASM_SO)


def build_gpu(verbose=False):
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    srcs = [str(CSRC / s) for s in GPU_SOURCES if (CSRC / s).exists()]
    cmd = [HIPCC] + HIP_FLAGS + srcs + GPU_LIBS + ["-o", str(GPU_SO) + ".tmp"]
    _run(cmd)
    os.replace(str(GPU_SO) + ".tmp", GPU_SO)
    if verbose:
        print("built", GPU_SO)


FORTRAN_DIR = PKG / "fortran"
FDRIVE = FORTRAN_DIR / "nmgpu_drive"


def build_fortran(verbose=False):
    """Fortran host (ISO_C_BINDING module + driver) with amdflang, linked
    against libnmgpu.so.  Skipped (returns False) when no Fortran compiler."""
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    if not Path(fc).exists():
        return False
    bdir = FORTRAN_DIR / "_build"
    bdir.mkdir(exist_ok=True)
    _run([fc, "-O2", "-J", str(bdir), "-c", str(FORTRAN_DIR / "nmgpu_f90.f90"), "-o", str(bdir / "nmgpu_f90.o")])
    _run([fc, "-O2", "-J", str(bdir), "-I", str(bdir), str(FORTRAN_DIR / "nmgpu_drive.f90"),
          str(bdir / "nmgpu_f90.o"), f"-L{PKG}", "-lnmgpu", "-Wl,-rpath,$ORIGIN/..", "-o", str(FDRIVE)])
    if verbose:
        print("built", FDRIVE)
    return True


def build_all(verbose=False):
    build_asm(verbose)
    build_gpu(verbose)
    build_fortran(verbose)


_asm = None
_gpu = None


def load_asm():
    global _asm
    if _asm is None:
        if not ASM_SO.exists():
            build_asm()
        _asm = ctypes.CDLL(str(ASM_SO))
    return _asm


def load_gpu():
    """Load libnmgpu.so; raises if it is not built (no other code path exists)."""
    global _gpu
    if _gpu is None:
        if not GPU_SO.exists():
            raise RuntimeError(f"{GPU_SO} is not built: run __graft_entry__.build() "
                               "or python -m normalmodes_amd.build")
        _gpu = ctypes.CDLL(str(GPU_SO), mode=ctypes.RTLD_GLOBAL)
    return _gpu
