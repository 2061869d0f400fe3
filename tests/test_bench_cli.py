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


def test_compulsory_stage_bytes_at_256_cubed():
    """The HBM roofline's byte counts for 256^3 particles on a 512^3 mesh:
    the three forward axis passes (z rows, y lines, and the x pass that also
    holds the Green's function and the inverse x transform) move
    (R + C) + 2C + 2C, R = 512^3 floats, C = 512^2*257 complex floats."""
    N, M = 256 ** 3, 512
    kb = bench.morton_bits(M)
    assert kb == 27
    b = bench.stage_bytes(N, M, kb, 1 << kb)
    R, C = 4 * M ** 3, 8 * M * M * (M // 2 + 1)
    assert b["fft_fwd"] + b["green_x"] == (R + C) + 2 * C + 2 * C
    assert abs(b["fft_fwd"] + b["green_x"] - 3.23e9) < 0.01e9
    assert b["fft_inv"] == b["fft_fwd"]
    assert b["sort"] == 4 * 20 * N
    assert b["deposit"] == 20 * N + R
    r = bench.hbm_roofline({"deposit": 1.0, "interp": 0.0}, b, ["deposit", "interp"])
    assert set(r) == {"deposit", "total"}
    assert abs(r["deposit"]["achieved"] - b["deposit"] / 1e-3 / 1e9) < 1e-6
    assert abs(r["deposit"]["frac"] - r["deposit"]["achieved"] / 8000.0) < 1e-12
