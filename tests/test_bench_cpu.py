"""The benchmark's command line without a GPU: one GPU only, and the
contract's fixed fields."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_bytes_per_site_model():
    """Algorithmic bytes per site update: read and write every population
    once (and the ternary dipole field), 2*(19*C + 3*ternary)*sizeof(real)."""
    sys.path.insert(0, ROOT)
    import bench

    assert bench.bytes_per_site(2, False, 4) == 304
    assert bench.bytes_per_site(2, False, 8) == 608
    assert bench.bytes_per_site(3, True, 4) == 480
    assert bench.bytes_per_site(3, True, 8) == 960


def test_batch_plan_lines_parse():
    """The batch tuner's report lines become the plan the traffic probe re-applies."""
    sys.path.insert(0, ROOT)
    import bench

    p = bench.batch_plan("chosen: persistent nwg=51104 xg=16 groups=18 (192 MiB) 2.2412 ms")
    assert p == dict(kind="persistent", nwg=51104, xg=16, group_mib=192)
    p = bench.batch_plan("chosen: collide-moments cols=62 threads=512 lz=8 xg=16 2.5400 ms")
    assert p == dict(kind="collide-moments", cols=62, threads=512, lz=8, xg=16)
    assert bench.batch_plan("chosen: sequential nwg=0 xg=0 groups=0 (0 MiB) 5.1 ms")["kind"] == "sequential"
    assert bench.ENS_BYTES == bench.bytes_per_site(3, True, 4)
