"""Host-side parts of bench.py: the one-GPU guard, and the counter summaries
under profiles/ from which the bench line's roofline traffic is read."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


@pytest.mark.parametrize("tag", ["block4_chebstep:c1", "block8_chebstep:c1", "mass3_chebstep:c1"])
def test_traffic_of_each_dominant_kernel_is_committed(tag):
    """Block 8 on two streams (4 columns a part), block 8 on one stream and
    the single vector: each has a FETCH_SIZE/WRITE_SIZE summary whose tag the
    bench finds."""
    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))
    import bench
    traffic, src = bench.traffic_from_profiles(tag)
    assert traffic is not None and traffic > 1e6, tag
    assert (ROOT / "profiles" / src).is_file()
    assert bench.traffic_from_profiles(tag + "x") == (None, None)
