"""The device fluid-solid functions on the CPU.

normalmodes_amd/csrc/nmgpu_fs_elem.h holds the per-element and per-node
functions the fluid-solid kernels run (element data, the fluid element
matrix on doubles and on dual numbers, the matrixstruct_general gather, the
adjoint's quadratic forms).  They are __host__ __device__; here they are
compiled for the host only (tests/fs_host_check.hip, hipcc
--offload-host-only) and run in host loops with the entry points' argument
lists, so that the arithmetic of the device producer and adjoint is pinned
without a GPU:
  - the producer against the CPU producer nm_assemble_fs (CGFSE3D_ISO,
    src/mod_cg_create_matrix.f90:103-977) on interface, free-surface and
    pure-fluid bodies: the same columns and values;
  - the adjoint by the same finite-difference and scaling-identity checks as
    the GPU tests of nmg_backward_fs."""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = Path("/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def hostlib(tmp_path_factory):
    if not HIPCC.exists():
        pytest.skip("hipcc is not installed")
    out = tmp_path_factory.mktemp("fs_host") / "libfs_host.so"
    subprocess.run([str(HIPCC), "-O2", "-std=c++17", "--offload-host-only", "-fPIC", "-shared",
                    f"-I{ROOT / 'normalmodes_amd' / 'csrc'}", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "fs_host_check.hip"), "-o", str(out)], check=True, timeout=600)
    return ctypes.CDLL(str(out))


class _Producer:
    """The CPU producer's library with nm_assemble_fs taken from the host build."""

    def __init__(self, real, host):
        self.real, self.host = real, host

    def __getattr__(self, name):
        return self.host.host_assemble_fs if name == "nm_assemble_fs" else getattr(self.real, name)


@pytest.mark.parametrize("case", ["core-p2-job1", "3layer-p2-job2", "3layer-p1-job2", "fluidball-job2",
                                  "fluidball-job2-pure"])
def test_host_producer_matches_cpu_producer(hostlib, monkeypatch, case):
    from normalmodes_amd import assemble as asm
    from test_device_assembly import _fs_body
    me, rho, vp, vs, g0, pure = _fs_body(case)
    job = 2 if g0 is not None else 1
    ref = asm.assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=g0, job=job, purefluid=pure)
    real = asm.load_asm()
    monkeypatch.setattr(asm, "load_asm", lambda: _Producer(real, hostlib))
    got = asm.assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=g0, job=job, purefluid=pure)
    assert got.stats == ref.stats
    for k in ("Ad", "B", "E", "ET", "Ap"):
        X, Y = getattr(got, k), getattr(ref, k)
        assert np.array_equal(X.rowptr, Y.rowptr) and np.array_equal(X.col, Y.col), k
        scale = np.abs(Y.val).max() if not (pure and k == "Ad") else 1.0
        assert np.abs(X.val - Y.val).max() <= 1e-15 * scale, k


class _Adjoint:
    def __init__(self, host):
        self.nmg_backward_fs = host.host_backward_fs


@pytest.mark.parametrize("case", ["core-p1-job2", "3layer-p1-job2"])
def test_host_adjoint_matches_finite_differences(hostlib, monkeypatch, case):
    import normalmodes_amd.adjoint as adj
    import test_fluid_adjoint as T
    monkeypatch.setattr(adj, "load_gpu", lambda: _Adjoint(hostlib))
    T.test_fluid_adjoint_matches_finite_differences(case)


def test_host_adjoint_scaling_identities(hostlib, monkeypatch):
    import normalmodes_amd.adjoint as adj
    import test_fluid_adjoint as T
    monkeypatch.setattr(adj, "load_gpu", lambda: _Adjoint(hostlib))
    T.test_fluid_adjoint_scaling_identities_p2()
