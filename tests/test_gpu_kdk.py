"""Device-resident KDK (predict.c drift/wrap, timestep.c kicks) against the
oracle's integrator over the lattice workload's 10 steps, and hipGraph
replays against eager steps."""
import math

import numpy as np
import pytest

from cuco_amd import engine as E
from cuco_amd import ics

from helpers import make_engine, oracle_params, rel_err

pytestmark = pytest.mark.gpu


def factors(box):
    d, k = E.init_drift_table(box.a_start, box.a_end, ics.OMEGA_M, ics.OMEGA_L,
                              ics.HUBBLE_INTERNAL)
    return E.kdk_factors(d, k, box.a_start, box.a_end, box.a_start, box.a_end,
                         box.nsteps)


def test_ten_kdk_steps_match_oracle(orc):
    box = ics.WORKLOADS["lattice32"]
    pos, vel, mass = ics.generate(box)
    f = factors(box)
    p = oracle_params(box)
    opos, ovel, oapp, oapm = orc.kdk_run(p, pos, vel, mass, f)
    eng = make_engine(box)
    eng.upload(pos, vel, mass)
    eng.compute_accelerations(E.FP64, True)
    eng.kdk_steps(f, E.FP64, use_graph=False)
    gpos, gvel = eng.download_particles()
    app, apm = eng.accelerations()
    tot = oapp + oapm
    assert rel_err(app + apm, tot, tot).max() <= 1e-10
    assert np.abs(gpos - opos).max() <= 1e-9 * box.L
    assert np.abs(gvel - ovel).max() <= 1e-9 * np.abs(ovel).max()
    eng.compute_accelerations(E.FP64, True)
    keys, idx = eng.sorted_order()
    ok, oi = orc.calc_keys(p, opos)
    ok, oi = orc.sort_pairs(ok, oi)
    assert np.array_equal(keys, ok) and np.array_equal(idx, oi)


@pytest.mark.parametrize("precision", [E.FP64, E.FP32])
def test_graph_replays_match_eager_steps(precision):
    box = ics.small_box("graph", 24, 48, 60.0, 21)
    pos, vel, mass = ics.generate(box)
    f = factors(ics.WORKLOADS["lattice32"])[:4]
    eager = make_engine(box)
    eager.upload(pos, vel, mass)
    eager.compute_accelerations(precision, True)
    graph = make_engine(box)
    graph.upload(pos, vel, mass)
    graph.compute_accelerations(precision, True)
    for s in range(4):
        eager.kdk_steps(f[s:s + 1], precision, use_graph=False)
        ep, ev = eager.download_particles()
        ea = eager.accelerations()
        # one replay per call; eager work and synchronisation in between
        graph.kdk_steps(f[s:s + 1], precision, use_graph=True)
        gp, gv = graph.download_particles()
        ga = graph.accelerations()
        assert np.array_equal(ep, gp) and np.array_equal(ev, gv), s
        assert np.array_equal(ea[0], ga[0]) and np.array_equal(ea[1], ga[1]), s
    # a multi-step replay continues the same trajectory
    eager.kdk_steps(f, precision, use_graph=False)
    graph.kdk_steps(f, precision, use_graph=True)
    assert np.array_equal(eager.download_particles()[0], graph.download_particles()[0])


def test_drift_wraps_into_box():
    box = ics.small_box("wrap", 16, 32, 40.0, 3)
    pos, vel, mass = ics.generate(box)
    eng = make_engine(box)
    vel = np.full_like(pos, 1.0)
    eng.upload(pos, vel, mass)
    eng.move_particles(box.L * 2.5)
    gpos, _ = eng.download_particles()
    expect = pos + 2.5 * box.L
    for _ in range(4):
        expect = np.where(expect >= box.L, expect - box.L, expect)
    assert np.all(gpos >= 0) and np.all(gpos < box.L)
    np.testing.assert_allclose(gpos, expect, rtol=0, atol=1e-9 * box.L)
