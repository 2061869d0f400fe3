"""bench.py's contract parts that need no GPU: more than one GPU is refused
before any device work, and the roofline's traffic figure is read from the
committed rocprofv3 PMC summary."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert r.stdout == ""  # no JSON line
    assert "one GPU" in r.stderr


def test_traffic_figures_come_from_the_committed_profile():
    sym = bench.pmc_traffic("pp32", "halo256", "traffic_bytes_symmetric")
    plain = bench.pmc_traffic("pp32", "halo256", "traffic_bytes")
    assert sym is not None and plain is not None
    # each-pair-once variant: the fixed-point j-side atomics dominate
    assert 1e10 < plain < sym < 1e12
    assert bench.pmc_traffic("pp32", "no_such_workload") is None
