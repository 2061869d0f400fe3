"""The C host program `cuco_amd/cuco_driver` (the reference's main/begrun/
run/accel call structure over the C ABI) against the same operations made
through the Python binding: identical snapshot records."""
import math
import os
import subprocess

import numpy as np
import pytest

from cuco_amd import engine as E
from cuco_amd import ics

from helpers import halo_box, make_engine

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(E.LIB_PATH), "cuco_driver")


def ic_records(box):
    pos, vel, mass = ics.generate(box)
    P = np.zeros(box.N, E.PARTICLE_DTYPE)
    P["Pos"] = pos.astype(np.float32)
    P["Pos"][P["Pos"] >= np.float32(box.L)] = 0   # float rounding up to L
    P["Vel"] = vel.astype(np.float32)
    P["Mass"] = mass.astype(np.float32)
    return P


def run_driver(args, timeout=120):
    print("driver", " ".join(args), flush=True)
    res = subprocess.run([DRIVER, *args], capture_output=True, text=True,
                         timeout=timeout)
    assert res.returncode == 0, res.stderr + res.stdout
    return res.stdout


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_forces_mode_matches_the_binding(tmp_path, prec):
    box = halo_box(16, 32, 50.0, n_halos=5, seed=31)
    P = ic_records(box)
    a = 0.5
    ic = str(tmp_path / "ic.cuco")
    E.write_snapshot(ic, E.Header(box.N, float(P["Mass"][0]), a, 1 / a - 1, box.L,
                                  0.3, 0.7, 0.7), P)
    out = str(tmp_path / "forces.cuco")
    run_driver(["forces", ic, out, str(box.M), prec, repr(box.softening)])
    h, Q = E.read_snapshot(out)
    assert h.NumPart == box.N and np.array_equal(Q["Pos"], P["Pos"])
    precision = E.FP64 if prec == "fp64" else E.FP32
    eng = E.Engine(pm_grid=box.M, box_size=float(np.float32(box.L)), G=ics.G_INTERNAL,
                   asmth=1.25 * float(np.float32(box.L)) / box.M,
                   rcut=4.5 * (1.25 * float(np.float32(box.L)) / box.M),
                   soft_h=2.8 * box.softening, max_particles=box.N)
    R = P.copy()
    eng.pmforce_particles(R, precision)
    eng.gravity_particles(R, precision)
    assert np.array_equal(Q["GravPM"], R["GravPM"])
    assert np.array_equal(Q["GravAccel"], R["GravAccel"])
    assert np.abs(Q["GravAccel"]).max() > 0 and np.abs(Q["GravPM"]).max() > 0


def test_run_mode_writes_the_scheduled_snapshots(tmp_path):
    box = halo_box(12, 24, 40.0, n_halos=4, seed=9)
    P = ic_records(box)
    ic = str(tmp_path / "ic.cuco")
    E.write_snapshot(ic, E.Header(box.N, float(P["Mass"][0]), 0.2, 4.0, box.L,
                                  0.3, 0.7, 0.7), P)
    lst = str(tmp_path / "list.txt")
    with open(lst, "w") as f:
        f.write("0.203\n0.201\n")                     # shuffled (run.c:72-73)
    soft, soft_max = box.softening, box.softening * 0.25
    max_dt = math.log(1.5) / 8
    out = run_driver(["run", ic, str(tmp_path / "snap"), str(box.M), "fp64", lst,
                      "0.2", "0.3", repr(soft), repr(soft_max), "0.025",
                      repr(max_dt), "40"])
    assert "output" in out
    # the same run through the binding
    L = float(np.float32(box.L))
    eng = E.Engine(pm_grid=box.M, box_size=L, G=ics.G_INTERNAL,
                   asmth=1.25 * L / box.M, rcut=4.5 * (1.25 * L / box.M),
                   soft_h=2.8 * soft, max_particles=box.N)
    # The driver applies init()'s velocity rescale (init.c:27-36) to the
    # IC's peculiar velocities; the binding side does the same host step.
    R0 = P.copy()
    E.init_records(R0, 0.2)
    assert not np.array_equal(R0["Vel"], P["Vel"])
    eng.upload_records(R0)
    tl = dict(time_begin=0.2, time_max=0.3, omega0=0.3, omega_lambda=0.7,
              hubble=ics.HUBBLE_INTERNAL, err_tol_int_accuracy=0.025,
              max_size_timestep=max_dt, min_size_timestep=0.0,
              softening_halo=soft, softening_halo_max_phys=soft_max)
    eng.timeline_init(**tl)
    times = E.read_output_list(lst)
    interval = math.log(1.5) / E.TIMEBASE
    ti_next = E.find_next_output_time(times, 0.2, 0.3, interval, 0)
    k, steps = 0, 0
    while steps < 40:
        st = eng.timeline_run(E.FP64, max_steps=40 - steps,
                              ti_stop=ti_next if ti_next < 2 * E.TIMEBASE else -1)
        steps = st["steps"]
        if st["at_stop"]:
            h, Q = E.read_snapshot(str(tmp_path / f"snap_{k}.cuco"))
            R = eng.download_records()
            assert Q.tobytes() == R.tobytes(), k
            assert h.Time == np.float32(st["time"])
            k += 1
            ti_next = E.find_next_output_time(times, 0.2, 0.3, interval, ti_next + 1)
            continue
        if st["finished"]:
            break
    assert k == 2
    h, Q = E.read_snapshot(str(tmp_path / f"snap_{k}.cuco"))
    assert Q.tobytes() == eng.download_records().tobytes()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_resume_from_a_checkpoint_writes_the_same_snapshots(tmp_path, prec):
    box = halo_box(12, 24, 40.0, n_halos=4, seed=9)
    P = ic_records(box)
    ic = str(tmp_path / "ic.cuco")
    E.write_snapshot(ic, E.Header(box.N, float(P["Mass"][0]), 0.2, 4.0, box.L,
                                  0.3, 0.7, 0.7), P)
    lst = str(tmp_path / "list.txt")
    with open(lst, "w") as f:
        f.write("0.201\n0.203\n0.206\n")
    args = [str(box.M), prec, lst]
    run_driver(["run", ic, str(tmp_path / "full"), *args, "0.2", "0.3",
                repr(box.softening), repr(box.softening * 0.25), "0.025",
                repr(math.log(1.5) / 8), "60"])
    ckpt = str(tmp_path / "full_0.ckpt")
    info = E.checkpoint_info(ckpt)
    assert info["timeline"] == 1 and info["n"] == box.N and info["state"]["at_stop"] == 1
    run_driver(["resume", ic, str(tmp_path / "part"), *args, ckpt, "1", "60"])
    n = 0
    for k in range(1, 5):
        full, part = tmp_path / f"full_{k}.cuco", tmp_path / f"part_{k}.cuco"
        assert full.exists() == part.exists(), k
        if full.exists():
            assert full.read_bytes() == part.read_bytes(), k
            n += 1
    assert n >= 2                                      # a listed output and the end
    assert not (tmp_path / "part_0.cuco").exists()
