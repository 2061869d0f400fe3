"""Checkpoints: a handle restored from a checkpoint continues exactly as the
handle that wrote it (integer timeline in FP64 and FP32, KDK steps of an
ensemble from the stored accelerations), and a checkpoint is refused by a
handle with other parameters without changing it.  The reference cannot
restart (savepositions keeps neither the internal velocities nor the
timeline, io.c:9-40, init.c:27-36), so the comparison is with an
uninterrupted run."""
import math

import numpy as np
import pytest

from cuco_amd import engine as E
from cuco_amd import ics

from helpers import halo_box, make_engine

pytestmark = pytest.mark.gpu


def timeline_params(box):
    return dict(time_begin=0.2, time_max=0.3, omega0=0.3, omega_lambda=0.7,
                hubble=ics.HUBBLE_INTERNAL, err_tol_int_accuracy=0.025,
                max_size_timestep=math.log(1.5) / 8, min_size_timestep=0.0,
                softening_halo=box.softening, softening_halo_max_phys=box.softening * 0.25)


def snapshot(eng):
    pos, vel = eng.download_particles()
    b, e = eng.timesteps()
    app, apm = eng.accelerations()
    return pos, vel, b, e, app, apm


@pytest.mark.parametrize("precision", [E.FP64, E.FP32], ids=["fp64", "fp32"])
def test_timeline_continues_bit_for_bit(tmp_path, precision):
    box = halo_box(12, 24, 40.0, n_halos=4, seed=9)
    pos, vel, mass = ics.generate(box)
    path = str(tmp_path / "run.ckpt")
    a = make_engine(box)
    a.upload(pos, vel, mass)
    a.timeline_init(**timeline_params(box))
    a.timeline_run(precision, max_steps=5)
    a.checkpoint_write(path)
    info = E.checkpoint_info(path)
    assert info["n"] == box.N and info["timeline"] == 1 and info["nboxes"] == 1
    assert info["pp_precision"] == precision and info["pm_precision"] == precision
    st_a = a.timeline_run(precision, max_steps=5)
    ref = snapshot(a)

    b = make_engine(box)
    b.checkpoint_read(path)
    st_b = b.timeline_run(precision, max_steps=5)
    assert st_a == st_b
    assert st_a["steps"] == 10
    for x, y in zip(ref, snapshot(b)):
        assert np.array_equal(x, y)


def test_state_without_timeline_round_trips(tmp_path):
    box = halo_box(10, 20, 30.0, n_halos=3, seed=5)
    pos, vel, mass = ics.generate(box)
    path = str(tmp_path / "forces.ckpt")
    a = make_engine(box)
    a.upload(pos, vel, mass)
    a.gravity_linklist(E.FP64)
    a.pmforce_periodic(E.FP32)
    a.checkpoint_write(path)
    info = E.checkpoint_info(path)
    assert info["timeline"] == 0 and info["pp_precision"] == E.FP64
    assert info["pm_precision"] == E.FP32
    b = make_engine(box, capacity=box.N + 100)
    b.checkpoint_read(path)
    for x, y in zip(a.download_particles(), b.download_particles()):
        assert np.array_equal(x, y)
    for x, y in zip(a.accelerations(), b.accelerations()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.download_records(), b.download_records())
    with pytest.raises(E.CucoError) as err:
        b.timeline_run(E.FP64, max_steps=1)
    assert err.value.status == 4


def test_ensemble_kdk_continues_from_stored_accelerations(tmp_path):
    boxes = [ics.small_box("ck0", 16, 32, 50.0, 41, "halos", n_halos=5),
             ics.small_box("ck1", 12, 24, 37.5, 42, "zeldovich")]
    data = [ics.generate(bx) for bx in boxes]
    lat = ics.WORKLOADS["lattice32"]
    d, k = E.init_drift_table(lat.a_start, lat.a_end, ics.OMEGA_M, ics.OMEGA_L,
                              ics.HUBBLE_INTERNAL)
    f = E.kdk_factors(d, k, lat.a_start, lat.a_end, lat.a_start, lat.a_end, 10)[:4]

    def ensemble():
        return E.Ensemble([bx.engine_kwargs() for bx in boxes], [bx.N for bx in boxes])

    path = str(tmp_path / "ens.ckpt")
    a = ensemble()
    a.upload(*[np.concatenate([dd[i] for dd in data]) for i in range(3)])
    a.compute_accelerations(E.FP32, True)
    a.kdk_steps(f[:2], E.FP32, use_graph=True)
    a.checkpoint_write(path)
    a.kdk_steps(f[2:], E.FP32, use_graph=True)
    b = ensemble()
    b.checkpoint_read(path)
    assert E.checkpoint_info(path)["nboxes"] == 2
    b.kdk_steps(f[2:], E.FP32, use_graph=True)
    for x, y in zip(a.download_particles(), b.download_particles()):
        assert np.array_equal(x, y)
    for x, y in zip(a.accelerations(), b.accelerations()):
        assert np.array_equal(x, y)


def test_other_handles_refuse_and_stay_unchanged(tmp_path):
    box = halo_box(10, 20, 30.0, n_halos=3, seed=6)
    pos, vel, mass = ics.generate(box)
    path = str(tmp_path / "a.ckpt")
    a = make_engine(box)
    a.upload(pos, vel, mass)
    a.gravity_linklist(E.FP64)
    a.checkpoint_write(path)

    other = halo_box(10, 24, 30.0, n_halos=3, seed=7)     # another mesh
    o = make_engine(other)
    o.upload(*ics.generate(other))
    before = o.download_particles()
    with pytest.raises(E.CucoError) as err:
        o.checkpoint_read(path)
    assert err.value.status == 1
    for x, y in zip(before, o.download_particles()):
        assert np.array_equal(x, y)
    with pytest.raises(E.CucoError) as err:
        make_engine(box, capacity=box.N - 1).checkpoint_read(path)
    assert err.value.status == 1

    short = tmp_path / "short.ckpt"
    short.write_bytes(open(path, "rb").read()[:-8])
    with pytest.raises(E.CucoError):
        make_engine(box).checkpoint_read(str(short))


def test_out_of_range_timeline_is_refused_before_any_change(tmp_path):
    box = halo_box(10, 20, 30.0, n_halos=3, seed=8)
    pos, vel, mass = ics.generate(box)
    a = make_engine(box)
    a.upload(pos, vel, mass)
    a.timeline_init(**timeline_params(box))
    a.timeline_run(E.FP64, max_steps=2)
    path = tmp_path / "t.ckpt"
    a.checkpoint_write(str(path))
    data = bytearray(path.read_bytes())
    off = E.CheckpointHeader.params.offset + E.Timeline.time_max.offset
    data[off:off + 8] = np.float64(0.0).tobytes()       # time_max <= time_begin
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(bytes(data))
    assert E.checkpoint_info(str(bad))["params"]["time_max"] == 0.0

    b = make_engine(box)
    b.upload(pos[::-1].copy(), vel, mass)
    before = b.download_particles()
    with pytest.raises(E.CucoError) as err:
        b.checkpoint_read(str(bad))
    assert err.value.status == 1
    for x, y in zip(before, b.download_particles()):
        assert np.array_equal(x, y)
