"""Individual power-of-two timesteps on the device (run.c / timestep.c)
against the oracle's restatement: the integer timeline (every particle's
Ti_begstep/Ti_endstep, Ti_Current, the PM step) must match exactly, the
trajectory to FP64 round-off; the active-set short-range force must equal a
full evaluation for the kicked particles; the state round-trips through a
snapshot file."""
import math

import numpy as np
import pytest

from cuco_amd import engine as E
from cuco_amd import ics

from helpers import halo_box, make_engine, oracle_params

pytestmark = pytest.mark.gpu


def setup(orc, seed=9, wide=0):
    box = halo_box(12, 24, 40.0, n_halos=4, seed=seed)
    pos, vel, mass = ics.generate(box)
    tl = dict(time_begin=0.2, time_max=0.3, omega0=0.3, omega_lambda=0.7,
              hubble=ics.HUBBLE_INTERNAL, err_tol_int_accuracy=0.025,
              max_size_timestep=math.log(1.5) / 8, min_size_timestep=0.0,
              softening_halo=box.softening,
              softening_halo_max_phys=box.softening * 0.25)
    # the engine builds its tables on the device; the oracle takes the same
    d, k = E.device_drift_table(tl["time_begin"], tl["time_max"], tl["omega0"],
                                tl["omega_lambda"], tl["hubble"])
    ref = orc.Timeline(oracle_params(box), tl, d, k, pos, vel, mass)
    eng = make_engine(box)
    if wide:
        eng.set_tuning(**dict(eng.tuning(), pp_wide_tiles=1))
    eng.upload(pos, vel, mass)
    eng.timeline_init(**tl)
    return box, tl, ref, eng


def compare(box, ref, eng, st, rs):
    for key in ("ti_current", "pm_ti_begstep", "pm_ti_endstep", "steps",
                "finished", "at_stop"):
        assert st[key] == rs[key], key
    assert st["time"] == pytest.approx(rs["time"], rel=1e-15)
    b, e = eng.timesteps()
    assert np.array_equal(b, ref.ti_beg) and np.array_equal(e, ref.ti_end)
    pos, vel = eng.download_particles()
    d = np.abs(pos - ref.pos)
    d = np.minimum(d, box.L - d)
    assert d.max() <= 1e-9 * box.L
    assert np.abs(vel - ref.vel).max() <= 1e-9 * np.abs(ref.vel).max()


def test_timeline_matches_oracle(orc):
    box, tl, ref, eng = setup(orc)
    levels = set()
    ti_out = -1
    for chunk in range(4):
        if chunk == 1:
            # An output time strictly between this sync point and the next:
            # the run must drift there, stop, and continue from it.
            nxt = min(int(ref.ti_end.min()), int(ref.state[2]))
            ti_out = st["ti_current"] + max(1, (nxt - st["ti_current"]) // 2)
        stop = ti_out if chunk == 1 else -1
        st = eng.timeline_run(E.FP64, max_steps=6, ti_stop=stop)
        rs = ref.run(6, stop)
        assert rs["flags"] == 0
        compare(box, ref, eng, st, rs)
        b, e = eng.timesteps()
        levels |= set(np.unique(e - b).tolist())
        if chunk == 1:
            assert st["at_stop"] == 1 and st["ti_current"] == ti_out
    assert len(levels) >= 2                   # a real hierarchy was exercised


def test_active_set_force_equals_full_evaluation(orc):
    for precision in (E.FP64, E.FP32):
        box, tl, ref, eng = setup(orc, seed=10)
        eng.timeline_run(precision, max_steps=3)
        st = eng.timeline_run(precision, max_steps=1)
        b, _ = eng.timesteps()
        active = b == st["ti_current"]        # kicked in this step: begstep moved here
        assert 0 < active.sum() == st["active"]
        app = eng.accelerations(pp=True, pm=False)
        pos, _ = eng.download_particles()
        full = make_engine(box)
        full.upload(pos, None, ics.generate(box)[2])
        full.set_softening(eng_soft(tl, st["time"]))
        full.gravity_linklist(precision)
        fa = full.accelerations(pp=True, pm=False)
        if precision == E.FP64:
            assert np.array_equal(app[active], fa[active])
        else:
            # FP32: the active targets run in tiles of their own, so their
            # candidates are chunked differently: equal to FP32 rounding.
            scale = np.linalg.norm(fa, axis=1).max()
            assert np.abs(app[active] - fa[active]).max() <= 1e-5 * scale


@pytest.mark.parametrize("wide", [0, 1])
def test_mostly_active_step_is_a_whole_evaluation(orc, wide):
    """A step whose active particles carry at least 4/5 of the estimated
    pair work (the first: all particles active) evaluates the FP32
    short-range force of every particle with the each-pair-once kernel:
    bit-identical to a full evaluation (tiles of 64 and of 128 targets)."""
    box, tl, ref, eng = setup(orc, seed=11, wide=wide)
    st = eng.timeline_run(E.FP32, max_steps=1)
    assert st["active"] == box.N
    pos, _ = eng.download_particles()
    full = make_engine(box)
    if wide:
        full.set_tuning(**dict(full.tuning(), pp_wide_tiles=1))
    full.upload(pos, None, ics.generate(box)[2])
    full.set_softening(eng_soft(tl, st["time"]))
    full.gravity_linklist(E.FP32)
    assert np.array_equal(eng.accelerations(pp=True, pm=False),
                          full.accelerations(pp=True, pm=False))


def eng_soft(tl, a):
    s = tl["softening_halo"]
    if s * a > tl["softening_halo_max_phys"]:
        s = tl["softening_halo_max_phys"] / a
    return 2.8 * s


def test_fp32_timeline_tracks_fp64(orc):
    box, tl, ref, e64 = setup(orc)
    _, _, _, e32 = setup(orc)
    s64 = e64.timeline_run(E.FP64, max_steps=8)
    s32 = e32.timeline_run(E.FP32, max_steps=8)
    assert s32["steps"] == 8 and s32["ti_current"] > 0
    p64, _ = e64.download_particles()
    p32, _ = e32.download_particles()
    d = np.abs(p64 - p32)
    d = np.minimum(d, box.L - d)
    assert np.isfinite(p32).all()
    assert np.median(d) <= 1e-6 * box.L
    assert s64["steps"] == 8


def active_work_share(box, pos, active):
    """Share of the active particles in the pair-work estimate: every
    particle weighs the occupancy of its aligned 16^3-cell block of the
    Morton key cube (cells sharing ix >> 4, iy >> 4, iz >> 4)."""
    c = np.minimum((pos * (box.M / box.L)).astype(np.int64), box.M - 1) >> 4
    k = (c[:, 0] * 64 + c[:, 1]) * 64 + c[:, 2]
    w = np.bincount(k)[k].astype(np.float64)
    return w[active].sum() / w.sum()


def test_force_path_follows_the_pair_work_rule(orc):
    """FP32 steps whose active set carries less than 4/5 of the estimated
    pair work leave the inactive particles' short-range accelerations
    untouched (active-set tiles); the others recompute them (whole
    evaluation)."""
    box, tl, ref, eng = setup(orc, seed=13)
    eng.timeline_run(E.FP32, max_steps=1)          # all active
    seen = set()
    for _ in range(10):
        before = eng.accelerations(pp=True, pm=False)
        st = eng.timeline_run(E.FP32, max_steps=1)
        b, _ = eng.timesteps()
        active = b == st["ti_current"]
        if active.all():
            continue
        pos, _ = eng.download_particles()
        whole = active_work_share(box, pos, active) >= 0.8
        after = eng.accelerations(pp=True, pm=False)
        untouched = np.array_equal(after[~active], before[~active])
        assert untouched != whole, (st["ti_current"], active.mean())
        seen.add(whole)
    assert False in seen


@pytest.mark.parametrize("wide", [0, 1])
def test_fp32_timeline_repeats_bit_for_bit(orc, wide):
    """Every FP32 force on the timeline is deterministic (fixed-point j sums,
    fixed per-target order, fixed-point mesh deposit) and the choice between
    active-set tiles and a whole evaluation depends on the particles only:
    two handles given the same particles agree bit for bit (whole
    evaluations with tiles of 64 or of 128 targets; the active-set tiles
    always hold 64)."""
    runs = []
    for _ in range(2):
        box, tl, ref, eng = setup(orc, seed=12, wide=wide)
        st = eng.timeline_run(E.FP32, max_steps=10)
        pos, vel = eng.download_particles()
        runs.append((st, pos, vel, *eng.timesteps()))
    (s0, *a0), (s1, *a1) = runs
    assert s0 == s1
    for x, y in zip(a0, a1):
        assert np.array_equal(x, y)


def test_records_and_snapshot_round_trip(orc, tmp_path):
    box, tl, ref, eng = setup(orc)
    st = eng.timeline_run(E.FP64, max_steps=4)
    P = eng.download_records()
    b, e = eng.timesteps()
    assert np.array_equal(P["Ti_begstep"], b) and np.array_equal(P["Ti_endstep"], e)
    pos, vel = eng.download_particles()
    assert np.array_equal(P["Pos"], pos.astype(np.float32))
    assert np.array_equal(P["Vel"], vel.astype(np.float32))
    a = st["time"]
    h = E.Header(P.shape[0], float(P["Mass"][0]), a, 1 / a - 1, box.L, 0.3, 0.7, 0.7)
    path = str(tmp_path / "snap_000.cuco")
    E.write_snapshot(path, h, P)
    h2, Q = E.read_snapshot(path)
    assert Q.tobytes() == P.tobytes() and h2.Time == np.float32(a)
    # A second handle restarts from the records.
    other = make_engine(box)
    other.upload_records(Q)
    p2, v2 = other.download_particles()
    assert np.array_equal(p2, np.mod(Q["Pos"].astype(np.float64), box.L))
    assert np.array_equal(v2, Q["Vel"].astype(np.float64))


def test_device_drift_tables(golden_dir):
    """init_drift_table built on the device (driftfac.c:14-42): the
    Einstein-de Sitter closed forms drift = 2/H0 (a0^-1/2 - a^-1/2) and
    kick = 2/H0 (a^1/2 - a0^1/2) to 1e-12 (the closed form's own
    difference of two nearly equal terms costs ~1e-13); for LambdaCDM the host tables
    (adaptive Simpson) to 1e-12 and the scipy.integrate.quad fixture through
    the reference's interpolation to 1e-9."""
    a0, a1, H0 = 0.02, 0.05, 0.1
    d, k = E.device_drift_table(a0, a1, 1.0, 0.0, H0)
    a = np.exp(np.log(a0) + (np.log(a1) - np.log(a0)) / 1000 * (np.arange(1000) + 1))
    np.testing.assert_allclose(d, 2 / H0 * (a0 ** -0.5 - a ** -0.5), rtol=1e-12)
    np.testing.assert_allclose(k, 2 / H0 * (a ** 0.5 - a0 ** 0.5), rtol=1e-12)
    g = np.load(f"{golden_dir}/drift_kick.npz")
    a0, a1 = float(g["a0"]), float(g["a1"])
    args = (a0, a1, float(g["omega0"]), float(g["omega_lambda"]), float(g["hubble"]))
    d, k = E.device_drift_table(*args)
    hd, hk = E.init_drift_table(*args)
    np.testing.assert_allclose(d, hd, rtol=1e-12)
    np.testing.assert_allclose(k, hk, rtol=1e-12)
    for av, dr, kr in zip(g["avals"], g["drift"], g["kick"]):
        assert E.table_factor(d, a0, a1, math.log(a0), math.log(av)) == pytest.approx(dr, rel=1e-9)
        assert E.table_factor(k, a0, a1, math.log(a0), math.log(av)) == pytest.approx(kr, rel=1e-9)
    with pytest.raises(E.CucoError):
        E.device_drift_table(0.05, 0.02, 0.3, 0.7, 0.1)
