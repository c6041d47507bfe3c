"""Writes tests/golden/c1_eigs.json: the 45 lowest eigenvalues of the C1
pencil (homogeneous solid unit ball rho=1, vp=1.73, vs=1; seeded mesh of
~12k P2 tetrahedra, mesh seed 1019) by scipy shift-invert Lanczos (ARPACK)
on the unscaled pencil (A, B), and the C1 interval: a = lambda_7 / 2 (above
the six rigid-body zero modes), b = midpoint of the first gap of relative
width > 1e-3 at or after the 36th eigenvalue.

Run: python tests/golden/make_c1_golden.py
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from normalmodes_amd.workloads import build_problem  # noqa: E402


def main():
    import scipy.sparse.linalg as sla
    t0 = time.time()
    P = build_problem("c1")
    A = P.A.to_scipy()
    B = P.B.to_scipy()
    sigma = -1.0
    w, _ = sla.eigsh(A, k=45, M=B, sigma=sigma, which="LM", tol=1e-14, ncv=120)
    w = np.sort(w)
    a = 0.5 * w[6]
    b = None
    for i in range(35, len(w) - 1):
        if (w[i + 1] - w[i]) / w[i] > 1e-3:
            b = 0.5 * (w[i] + w[i + 1])
            break
    out = dict(workload="c1", n_dof=P.n, nnz_A=P.A.nnz, nnz_B=P.B.nnz, n_tets=P.mesh.n_elem,
               sigma=sigma, eigenvalues=w.tolist(), interval=[a, b],
               n_in_interval=int(np.sum((w >= a) & (w <= b))), seconds=time.time() - t0,
               method="scipy.sparse.linalg.eigsh shift-invert (sigma=-1), tol 1e-14")
    (Path(__file__).parent / "c1_eigs.json").write_text(json.dumps(out, indent=1))
    print(json.dumps({k: v for k, v in out.items() if k != "eigenvalues"}))


if __name__ == "__main__":
    main()
