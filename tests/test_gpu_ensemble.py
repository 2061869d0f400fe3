"""Ensembles of independent boxes: every box of a batched handle must give
the same keys, cell order, forces and KDK trajectory, bit for bit, as a
single-box handle holding only that box's particles (the batched kernels
count waves, tiles and mesh blocks from each box's own start)."""
import numpy as np
import pytest

from cuco_amd import engine as E
from cuco_amd import ics

from helpers import make_engine

pytestmark = pytest.mark.gpu

# (n_side, mesh, L in Mpc/h, seed, kind, halos): mixed meshes (radix 2, 3, 5),
# clustered and lattice boxes, particle counts not multiples of 64.
MIXED = [(16, 32, 50.0, 11, "halos", 6), (24, 48, 75.0, 12, "zeldovich", 0),
         (20, 40, 62.5, 13, "halos", 10), (12, 32, 37.5, 14, "zeldovich", 0)]
# Equal mean spacing (3.125 Mpc/h) and so one common particle mass.
UNIFORM = [(16, 32, 50.0, 21, "zeldovich", 0), (24, 48, 75.0, 22, "halos", 8),
           (20, 40, 62.5, 23, "zeldovich", 0)]


def make_boxes(specs, jitter_mass):
    boxes, data = [], []
    for k, (n, M, L, seed, kind, nh) in enumerate(specs):
        box = ics.small_box(f"ens{k}", n, M, L, seed, kind, n_halos=nh)
        pos, vel, mass = ics.generate(box)
        if jitter_mass:
            rng = np.random.default_rng(seed)
            mass = mass * (1.0 + 0.25 * rng.random(mass.shape[0]))
        boxes.append(box)
        data.append((pos, vel, mass))
    return boxes, data


def make_ensemble(boxes, data):
    ens = E.Ensemble([b.engine_kwargs() for b in boxes], [b.N for b in boxes])
    pos = np.concatenate([d[0] for d in data])
    vel = np.concatenate([d[1] for d in data])
    mass = np.concatenate([d[2] for d in data])
    ens.upload(pos, vel, mass)
    return ens


def singles(boxes, data):
    out = []
    for box, (pos, vel, mass) in zip(boxes, data):
        eng = make_engine(box)
        eng.upload(pos, vel, mass)
        out.append(eng)
    return out


@pytest.mark.parametrize("precision", [E.FP64, E.FP32])
@pytest.mark.parametrize("specs,jitter", [(MIXED, True), (UNIFORM, False)],
                         ids=["mixed-masses", "common-mass"])
def test_forces_match_single_boxes(precision, specs, jitter):
    boxes, data = make_boxes(specs, jitter)
    ens = make_ensemble(boxes, data)
    ens.compute_accelerations(precision, True)
    app, apm = ens.accelerations()
    for b, eng in enumerate(singles(boxes, data)):
        eng.compute_accelerations(precision, True)
        sp, sm = eng.accelerations()
        sl = ens.box_slice(b)
        assert np.array_equal(app[sl], sp), (b, "pp")
        assert np.array_equal(apm[sl], sm), (b, "pm")


def test_each_pair_once_matches_single_boxes():
    """The each-pair-once pair kernel (fixed-point j sums, one per-box scale)
    gives every box of an ensemble exactly what a single-box handle gives."""
    boxes, data = make_boxes(UNIFORM, False)
    ens = make_ensemble(boxes, data)
    ens.set_tuning(3, 8, 0, 1)
    ens.compute_accelerations(E.FP32, True)
    app = ens.accelerations(pp=True, pm=False)
    for b, eng in enumerate(singles(boxes, data)):
        eng.set_tuning(3, 8, 0, 1)
        eng.compute_accelerations(E.FP32, True)
        sl = ens.box_slice(b)
        assert np.array_equal(app[sl], eng.accelerations(pp=True, pm=False)), b


def test_sorted_order_and_meshes_match_single_boxes():
    boxes, data = make_boxes(MIXED, True)
    ens = make_ensemble(boxes, data)
    ens.compute_accelerations(E.FP64, True)
    keys, idx = ens.sorted_order()
    rho = ens.pm_grid(0)
    phi = ens.pm_grid(1)
    key_base, p0 = 0, 0
    for b, eng in enumerate(singles(boxes, data)):
        eng.compute_accelerations(E.FP64, True)
        sk, si = eng.sorted_order()
        sl = ens.box_slice(b)
        assert np.array_equal(keys[sl], sk + key_base), b
        assert np.array_equal(idx[sl], si + p0), b
        assert np.array_equal(rho[b], eng.pm_grid(0)), b
        assert np.array_equal(phi[b], eng.pm_grid(1)), b
        key_base += boxes[b].M ** 3
        p0 += boxes[b].N
    assert ens.size() == (len(boxes), [b.N for b in boxes])


def test_kdk_graph_matches_single_boxes():
    boxes, data = make_boxes(UNIFORM, False)
    lat = ics.WORKLOADS["lattice32"]
    d, k = E.init_drift_table(lat.a_start, lat.a_end, ics.OMEGA_M, ics.OMEGA_L,
                              ics.HUBBLE_INTERNAL)
    f = E.kdk_factors(d, k, lat.a_start, lat.a_end, lat.a_start, lat.a_end, 10)[:3]
    ens = make_ensemble(boxes, data)
    ens.compute_accelerations(E.FP32, True)
    ens.kdk_steps(f[:1], E.FP32, use_graph=True)
    ens.kdk_steps(f[1:], E.FP32, use_graph=True)
    pos, vel = ens.download_particles()
    app, apm = ens.accelerations()
    for b, eng in enumerate(singles(boxes, data)):
        eng.compute_accelerations(E.FP32, True)
        eng.kdk_steps(f, E.FP32, use_graph=True)
        sp, sv = eng.download_particles()
        sa, sm = eng.accelerations()
        sl = ens.box_slice(b)
        assert np.array_equal(pos[sl], sp) and np.array_equal(vel[sl], sv), b
        assert np.array_equal(app[sl], sa) and np.array_equal(apm[sl], sm), b
        assert (pos[sl] >= 0).all() and (pos[sl] < boxes[b].L).all()


def test_upload_must_fill_every_box():
    boxes, data = make_boxes(UNIFORM[:2], False)
    ens = E.Ensemble([b.engine_kwargs() for b in boxes], [b.N for b in boxes])
    with pytest.raises(E.CucoError):
        ens.upload(data[0][0])
    with pytest.raises(E.CucoError):
        E.Ensemble([boxes[0].engine_kwargs()], [0])


def test_set_params_changes_one_box():
    boxes, data = make_boxes(UNIFORM, False)
    ens = make_ensemble(boxes, data)
    b = boxes[1]
    ens.set_params(ics.G_INTERNAL * 2, b.asmth, b.rcut, b.soft_h, box=1)
    ens.compute_accelerations(E.FP64, True)
    app, apm = ens.accelerations()
    eng = singles(boxes, data)[1]
    eng.set_params(ics.G_INTERNAL * 2, b.asmth, b.rcut, b.soft_h)
    eng.compute_accelerations(E.FP64, True)
    sp, sm = eng.accelerations()
    sl = ens.box_slice(1)
    assert np.array_equal(app[sl], sp) and np.array_equal(apm[sl], sm)
    ref = singles(boxes, data)[0]
    ref.compute_accelerations(E.FP64, True)
    assert np.array_equal(app[ens.box_slice(0)], ref.accelerations()[0])
