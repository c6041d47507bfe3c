"""The MVFunc-shaped callbacks (nmg_amv / nmg_bmv / nmg_bsol) driven by a C
program that uses nothing else after set-up: its own Lanczos loop calls them
through function pointers, as a pEVSL caller would (src/mod_pevsl.f90:69-80).
The products are checked against scipy on the same scaled pencil, and the
extreme Ritz values of its Lanczos tridiagonal against dense eigenvalues."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

ROOT = Path(__file__).resolve().parent.parent


def _build_caller(tmp):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = tmp / "mvfunc_caller"
    pkg = ROOT / "normalmodes_amd"
    subprocess.run([cc, "-O2", "-I", str(ROOT / "include"), str(ROOT / "tests" / "mvfunc_caller.c"),
                    "-L", str(pkg), "-lnmgpu", f"-Wl,-rpath,{pkg}", "-lm", "-o", str(exe)], check=True)
    return exe


def test_caller_builds_against_header(tmp_path):
    """The callbacks are declared with pEVSL's MVFunc signature and exported."""
    if not (ROOT / "normalmodes_amd" / "libnmgpu.so").exists():
        pytest.skip("libnmgpu.so not built")
    _build_caller(tmp_path)


@pytest.mark.gpu
def test_mvfunc_callbacks_drive_lanczos(small_p2, dense_eigs, tmp_path):
    P = small_p2
    n = P.n
    for name, M in (("A", P.A), ("B", P.B)):
        M.rowptr.astype(np.int64).tofile(tmp_path / f"{name}_rowptr.i64")
        M.col.astype(np.int32).tofile(tmp_path / f"{name}_col.i32")
        M.val.astype(np.float64).tofile(tmp_path / f"{name}_val.f64")
    x = np.random.default_rng(3).standard_normal(n)
    x.tofile(tmp_path / "x.f64")
    exe = _build_caller(tmp_path)
    m = 80
    print(f"mvfunc caller: n={n}, {m} steps", flush=True)
    r = subprocess.run([str(exe), str(tmp_path), str(n), str(m)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    A, B = P.A.to_scipy(), P.B.to_scipy()
    D = sp.diags(1.0 / np.sqrt(B.diagonal()))
    At, Bt = (D @ A @ D).tocsr(), (D @ B @ D).tocsr()
    ya = np.fromfile(tmp_path / "y_amv.f64")
    yb = np.fromfile(tmp_path / "y_bmv.f64")
    ys = np.fromfile(tmp_path / "y_bsol.f64")
    ra, rb = At @ x, Bt @ x
    assert np.linalg.norm(ya - ra) <= 1e-13 * np.linalg.norm(ra)
    assert np.linalg.norm(yb - rb) <= 1e-14 * np.linalg.norm(rb)
    rs = sl.solve(Bt.toarray(), x, assume_a="pos")
    assert np.linalg.norm(ys - rs) <= 1e-11 * np.linalg.norm(rs)     # degree-58 Chebyshev
    al = np.fromfile(tmp_path / "alpha.f64")
    be = np.fromfile(tmp_path / "beta.f64")
    assert al.size == m and be.size == m - 1
    theta = sl.eigh_tridiagonal(al, be, eigvals_only=True)
    w = dense_eigs(P)
    # Ritz values interlace the pencil's spectrum; the largest converges first
    assert theta[0] >= w[0] - 1e-10 * w[-1] and theta[-1] <= w[-1] * (1 + 1e-12)
    assert abs(theta[-1] - w[-1]) <= 1e-8 * w[-1]
