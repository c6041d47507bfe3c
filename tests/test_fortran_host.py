"""The Fortran host (ISO_C_BINDING module over the C-ABI) builds with
amdflang, and on the device reproduces the Python-driven solve."""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_fortran_module_compiles(tmp_path):
    from normalmodes_amd import _native
    fc = Path("/opt/rocm/bin/amdflang")
    if not fc.exists():
        pytest.skip("no Fortran compiler")
    r = subprocess.run([str(fc), "-J", str(tmp_path), "-c", str(_native.FORTRAN_DIR / "nmgpu_f90.f90"),
                        "-o", str(tmp_path / "m.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _write_problem(P, path):
    A, B = P.A, P.B
    with open(path, "wb") as f:
        f.write(struct.pack("<iii", P.n, A.nnz, B.nnz))
        for M in (A, B):
            f.write(M.rowptr.astype("<i4").tobytes())
            f.write(M.col.astype("<i4").tobytes())
            f.write(M.val.astype("<f8").tobytes())


@pytest.mark.gpu
def test_fortran_driver_on_device(small_p2, tmp_path, dense_eigs):
    from normalmodes_amd import _native
    if not _native.FDRIVE.exists():
        pytest.skip("Fortran driver not built")
    f = tmp_path / "tiny.bin"
    _write_problem(small_p2, f)
    r = subprocess.run([str(_native.FDRIVE), str(f), "1.0", "12.0", "60"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l.split() for l in r.stdout.splitlines() if l.startswith("eig")]
    lam = np.array([float(l[2]) for l in lines])
    res2 = np.array([float(l[4]) for l in lines])
    w = dense_eigs(small_p2)
    ref = w[(w >= 1.0) & (w <= 12.0)]
    assert lam.size == ref.size
    assert np.max(np.abs(lam - ref) / ref) < 1e-12
    assert np.all(res2 <= 1e-10)
