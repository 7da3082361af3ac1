"""bench.py's command line (CPU only): more than one GPU is refused with a
message and exit status 2 before any GPU work starts."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_host_cores_is_the_affinity_capped_by_the_share():
    sys.path.insert(0, str(ROOT))
    import os

    import bench

    cores, info = bench.host_cores()
    assert info["affinity"] == len(os.sched_getaffinity(0))
    assert 1 <= cores <= info["affinity"]
    for cap in (info["cgroup_quota"], info["omp_num_threads"]):
        if cap is not None:
            assert cores <= cap
