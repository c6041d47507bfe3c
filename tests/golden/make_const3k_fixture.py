"""Copy the reference's CONST3k demo input (P1 files: the mesh, its
neighbours, rho/vp/vs and reference gravity) into tests/golden/const3k/ as a
data fixture.  Run once in a checkout that has the reference at
/root/reference; the GPU box and the tests only read the copies."""
import shutil
from pathlib import Path

SRC = Path("/root/reference/demos/models/input/CONST3k")
DST = Path(__file__).resolve().parent / "const3k"
BASE = "CONST_1L_3k.1"
FILES = ["_mesh.header", "_ele.dat", "_neigh.dat", "_node.dat", "_rho_pod_1_true.dat", "_vp_pod_1_true.dat",
         "_vs_pod_1_true.dat", "_pod_1_potential_acceleration_true.dat"]

if __name__ == "__main__":
    DST.mkdir(exist_ok=True)
    for f in FILES:
        shutil.copyfile(SRC / f"{BASE}{f}", DST / f"{BASE}{f}")
        print(DST / f"{BASE}{f}")
