import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libnmgpu.so")
    config.addinivalue_line("markers", "slow: longer CPU test")
    config.addinivalue_line("markers", "gpu_long: GPU parity test of several minutes (run on its own with "
                                       "-m gpu_long -s, so that its progress lines show)")


def pytest_collection_modifyitems(config, items):
    """gpu_long tests run only when the marker expression names them."""
    expr = config.getoption("-m") or ""
    if "gpu_long" in expr:
        return
    skip = pytest.mark.skip(reason="several-minute GPU test: run with -m gpu_long -s")
    for item in items:
        if item.get_closest_marker("gpu_long"):
            item.add_marker(skip)


@pytest.fixture(scope="session")
def small_p2():
    """P2 homogeneous ball, ~300 tets (1.7k DoF): the dense oracle runs in seconds."""
    from normalmodes_amd.mesh import ball_mesh
    from normalmodes_amd.models import homogeneous
    from normalmodes_amd.assemble import assemble
    return assemble(ball_mesh(300, porder=2, seed=1019), homogeneous(), name="ball300_p2")


@pytest.fixture(scope="session")
def small_p1():
    from normalmodes_amd.mesh import ball_mesh
    from normalmodes_amd.models import homogeneous
    from normalmodes_amd.assemble import assemble
    return assemble(ball_mesh(3000, porder=1, seed=1019), homogeneous(), name="ball3k_p1")


@pytest.fixture(scope="session")
def layered_p2():
    from normalmodes_amd.mesh import ball_mesh
    from normalmodes_amd.models import three_layer_solid
    from normalmodes_amd.assemble import assemble
    return assemble(ball_mesh(600, porder=2, seed=1019), three_layer_solid(), name="layered600_p2")


@pytest.fixture(scope="session")
def dense_eigs():
    cache = {}

    def get(problem):
        key = id(problem)
        if key not in cache:
            import scipy.linalg as sl
            cache[key] = sl.eigh(problem.A.to_scipy().toarray(), problem.B.to_scipy().toarray(),
                                 eigvals_only=True)
        return cache[key]
    return get
