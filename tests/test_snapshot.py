"""Snapshot files (io.c savepositions / read_ic.c read_ic) and the output
schedule (begrun.c read_outputlist, run.c find_next_outputtime).  Host-only
entry points: no device needed.

`tests/golden/list_32.txt` is the output list shipped with the reference
(output_32/list_32.txt), kept as a data fixture."""
import math
import os
import struct

import numpy as np
import pytest

from cuco_amd import engine as E

HERE = os.path.dirname(os.path.abspath(__file__))
LIST = os.path.join(HERE, "golden", "list_32.txt")


def random_records(n, seed=3):
    rng = np.random.default_rng(seed)
    P = np.zeros(n, E.PARTICLE_DTYPE)
    for f in ("Pos", "Vel", "GravAccel", "GravPM"):
        P[f] = rng.standard_normal((n, 3)).astype(np.float32)
    P["Mass"] = rng.random(n).astype(np.float32)
    P["Ti_endstep"] = rng.integers(0, 1 << 28, n)
    P["Ti_begstep"] = rng.integers(0, 1 << 28, n)
    return P


def header(n):
    return E.Header(n, 1.5, 0.25, 3.0, 10000.0, 0.3, 0.7, 0.7)


def test_round_trip_is_byte_exact(tmp_path):
    P = random_records(1000)
    path = str(tmp_path / "snap_0.cuco")
    E.write_snapshot(path, header(1000), P)
    raw = open(path, "rb").read()
    assert len(raw) == 32 + 60 * 1000
    # HEADER: int NumPart then seven floats, native order (allvars.h:47-57)
    fields = struct.unpack("<i7f", raw[:32])
    assert fields[0] == 1000
    assert fields[1:] == struct.unpack("<7f", struct.pack("<7f", 1.5, 0.25, 3.0,
                                                          10000.0, 0.3, 0.7, 0.7))
    assert raw[32:] == P.tobytes()
    h, Q = E.read_snapshot(path)
    assert h.NumPart == 1000 and h.BoxSize == 10000.0
    assert Q.tobytes() == P.tobytes()
    h2, none = E.read_snapshot(path, header_only=True)
    assert none is None and h2.NumPart == 1000


def test_empty_snapshot(tmp_path):
    path = str(tmp_path / "empty.cuco")
    E.write_snapshot(path, header(0), np.zeros(0, E.PARTICLE_DTYPE))
    h, Q = E.read_snapshot(path)
    assert h.NumPart == 0 and Q.shape == (0,)


def test_truncated_or_padded_file_is_rejected(tmp_path):
    P = random_records(10)
    path = str(tmp_path / "s.cuco")
    E.write_snapshot(path, header(10), P)
    raw = open(path, "rb").read()
    for bad in (raw[:-1], raw + b"\0", raw[:20]):
        open(path, "wb").write(bad)
        with pytest.raises(E.CucoError):
            E.read_snapshot(path)
    with pytest.raises(E.CucoError):
        E.read_snapshot(str(tmp_path / "missing.cuco"))


def test_small_buffer_is_rejected(tmp_path):
    P = random_records(10)
    path = str(tmp_path / "s.cuco")
    E.write_snapshot(path, header(10), P)
    h = E.Header()
    buf = np.zeros(5, E.PARTICLE_DTYPE)
    import ctypes as C
    rc = E.lib().cuco_read_snapshot(path.encode(), C.byref(h),
                                    buf.ctypes.data_as(C.c_void_p), 5)
    assert rc == 1


def test_output_list_of_the_reference_run():
    t = E.read_output_list(LIST)
    expect = np.array([0.166667, 0.25, 0.333333, 0.4, 0.5, 0.6, 0.7, 0.8, 0.85,
                       0.9, 0.95, 1.0], dtype=np.float32).astype(np.float64)
    assert np.array_equal(t, expect)          # read as float, as "%g" into float
    assert E.read_output_list(LIST, max_times=3).shape == (3,)


def test_next_output_time_brackets_each_listed_time():
    # begrun.c:44-45 and init.c:22: a in [0.0909, 1], TIMEBASE = 2^28 steps in ln a
    t = E.read_output_list(LIST)
    a0, a1 = float(np.float32(0.0909)), 1.0
    dti = (math.log(a1) - math.log(a0)) / E.TIMEBASE
    ti, seen = 0, []
    while True:
        nxt = E.find_next_output_time(t, a0, a1, dti, ti)
        if nxt == 2 * E.TIMEBASE:
            break
        assert nxt > ti
        # the integer time just below the listed scale factor
        a_lo = a0 * math.exp(nxt * dti)
        a_hi = a0 * math.exp((nxt + 1) * dti)
        assert any(a_lo <= x * (1 + 1e-12) and x < a_hi * (1 + 1e-12) for x in t)
        seen.append(nxt)
        ti = nxt
    assert len(seen) == len(t)                  # every listed time, in order
    assert seen[-1] == E.TIMEBASE or seen[-1] == E.TIMEBASE - 1
    # a shuffled list gives the same schedule (run.c:72-73)
    rng = np.random.default_rng(1)
    sh = t[rng.permutation(len(t))]
    assert E.find_next_output_time(sh, a0, a1, dti, seen[2]) == seen[3]
    # times outside [TimeBegin, TimeMax] are ignored
    assert E.find_next_output_time(np.array([0.05, 2.0]), a0, a1, dti, 0) == 2 * E.TIMEBASE


def test_init_records_rescales_peculiar_velocities():
    """init.c:27-36: an IC file's peculiar velocities become a^(3/2) times
    themselves (float *= double, rounded once to float), and the force and
    timeline fields start at zero.  Checked against the closed form in
    float64 and against the rounding of a single product."""
    P = random_records(1000, seed=11)
    Q = P.copy()
    a = 0.02
    E.init_records(Q, a)
    fac = math.sqrt(a) * a
    expect = (P["Vel"].astype(np.float64) * fac).astype(np.float32)
    assert np.array_equal(Q["Vel"], expect)
    # within one float rounding of the exact scaled value
    assert np.allclose(Q["Vel"], P["Vel"].astype(np.float64) * a ** 1.5,
                       rtol=2 ** -23, atol=0)
    for f in ("GravAccel", "GravPM", "Ti_endstep", "Ti_begstep"):
        assert not Q[f].any(), f
    assert np.array_equal(Q["Pos"], P["Pos"]) and np.array_equal(Q["Mass"], P["Mass"])
    E.init_records(Q[:0].copy(), a)          # empty array: nothing to do
    with pytest.raises(E.CucoError):
        E.init_records(Q, 0.0)
