"""Writes tests/golden/tiny.json: checksums of the assembled tiny problem
(homogeneous unit ball, ~300 P2 tets, mesh seed 1019) and of its P1
counterpart (~3000 tets), and all their generalized eigenvalues from dense
scipy.linalg.eigh(A, B).

Run: python tests/golden/make_tiny_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np
import scipy.linalg as sl

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from normalmodes_amd.assemble import assemble  # noqa: E402
from normalmodes_amd.mesh import ball_mesh  # noqa: E402
from normalmodes_amd.models import homogeneous  # noqa: E402


def summary(P):
    A, B = P.A, P.B
    w = sl.eigh(A.to_scipy().toarray(), B.to_scipy().toarray(), eigvals_only=True)
    return dict(n_dof=P.n, n_elem=P.mesh.n_elem, nnz_A=A.nnz, nnz_B=B.nnz,
                sum_A=float(A.val.sum()), sum_abs_A=float(np.abs(A.val).sum()),
                sum_B=float(B.val.sum()), fro_A=float(np.linalg.norm(A.val)),
                fro_B=float(np.linalg.norm(B.val)), col_hash=int(np.sum(A.col.astype(np.int64) *
                                                                        np.arange(A.nnz) % 1000003)),
                eigenvalues=w.tolist())


def main():
    out = {
        "p2": summary(assemble(ball_mesh(300, porder=2, seed=1019), homogeneous())),
        "p1": summary(assemble(ball_mesh(3000, porder=1, seed=1019), homogeneous())),
        "method": "dense scipy.linalg.eigh(A, B)",
    }
    (Path(__file__).parent / "tiny.json").write_text(json.dumps(out))
    print({k: {kk: vv for kk, vv in v.items() if kk != "eigenvalues"} if isinstance(v, dict) else v
           for k, v in out.items()})


if __name__ == "__main__":
    main()
