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
    bench = _bench()
    traffic, src = bench.traffic_from_profiles(tag)
    assert traffic is not None and traffic > 1e6, tag
    assert (ROOT / "profiles" / src).is_file()
    assert bench.traffic_from_profiles(tag + "x") == (None, None)


def _bench():
    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))
    import bench
    return bench


def test_profile_commands_use_valid_layout_specs():
    """Every fixed-layout command recorded under profiles/ parses with the
    bench's current layout syntax, so the recorded commands can be rerun."""
    import re
    bench = _bench()
    specs = []
    for f in sorted((ROOT / "profiles").glob("*.md")):
        specs += re.findall(r"--layouts\s+(\S+)", f.read_text())
    assert specs
    for spec in specs:
        parsed = bench.parse_layouts(spec)
        assert parsed and all(name in "AB" and K in (1, 2, 4, 8, 16) for name, K, _, _ in parsed), spec


@pytest.mark.parametrize("spec", ["A:32", "C1:lanes:16", "A1:tiled:231", "A3:rows:1", "B4:wide", "B8:twide:200"])
def test_bad_layout_specs_are_refused(spec):
    with pytest.raises(ValueError):
        _bench().parse_layouts(spec)


@pytest.mark.parametrize("workload,block,split,warmup,pieces", [("c2s", 4, 2, 0, 1), ("c1", 8, 2, 1, 0)])
def test_workload_defaults(workload, block, split, warmup, pieces):
    """The headline (c2s) runs block 4 as two chains, plain FP64, without a
    warm-up solve and with the fluid ball's pieces; c1 keeps block 8."""
    bench = _bench()
    args = bench.build_parser().parse_args(["--workload", workload])
    assert bench.resolve_defaults(args) is False
    assert (args.block, args.split, args.warmup, args.fluid_pieces, args.refine) == (block, split, warmup, pieces, 0)
    assert bench.build_parser().parse_args([]).workload == "c2s"
    a2 = bench.build_parser().parse_args(["--workload", "c2"])
    assert bench.resolve_defaults(a2) is True and a2.warmup == 0


@pytest.mark.parametrize("workload,capped", [("c2s", True), ("c2", True), ("c1", False)])
def test_long_solves_time_one_solve(workload, capped):
    """A 2M-DoF whole solve takes minutes: asked for 20 timed and 5 warm-up
    solves, the bench times one with no warm-up and records what was asked;
    c1 keeps the steps as asked, and so does --long-solves 1."""
    bench = _bench()
    args = bench.build_parser().parse_args(["--workload", workload, "--steps", "20", "--warmup", "5"])
    bench.resolve_defaults(args)
    assert args.requested == {"steps": 20, "warmup": 5}
    assert args.capped is capped
    assert (args.steps, args.warmup) == ((1, 0) if capped else (20, 5))
    keep = bench.build_parser().parse_args(["--workload", workload, "--steps", "3", "--warmup", "1",
                                            "--long-solves", "1"])
    bench.resolve_defaults(keep)
    assert (keep.steps, keep.warmup, keep.capped) == (3, 1, False)
    counters = bench.build_parser().parse_args(["--workload", workload, "--steps", "20", "--kernel-only"])
    bench.resolve_defaults(counters)
    assert counters.steps == 20
