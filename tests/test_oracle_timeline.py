"""Pins the oracle's integer-timeline integrator (orc_timeline_run, a
restatement of run.c / timestep.c) without a GPU:

* force-free particles drift by the closed-form Einstein-de Sitter integral
  and the system takes TIMEBASE / (largest power-of-two step) steps;
* with a tolerance so loose that every particle takes the maximal step the
  hierarchical integrator must reduce to the global KDK leapfrog
  (orc_kdk_run, an independently written loop) with the same factors;
* on a clustered box the steps form an aligned power-of-two hierarchy and
  every system step starts at the smallest Ti_endstep."""
import math

import numpy as np

from cuco_amd import ics

from helpers import halo_box

TIMEBASE = 1 << 28
H0 = ics.HUBBLE_INTERNAL


def timeline(**kw):
    tl = dict(time_begin=0.2, time_max=0.3, omega0=0.3, omega_lambda=0.7,
              hubble=H0, err_tol_int_accuracy=0.025,
              max_size_timestep=math.log(1.5) / 8, min_size_timestep=0.0,
              softening_halo=100.0, softening_halo_max_phys=1e30)
    tl.update(kw)
    return tl


def tables(orc, tl):
    return orc.drift_tables(tl["time_begin"], tl["time_max"], tl["omega0"],
                            tl["omega_lambda"], tl["hubble"])


def lattice(n, L):
    q = (np.indices((n, n, n)).reshape(3, -1).T + 0.5) * (L / n)
    return q


def test_force_free_drift_is_the_eds_integral(orc):
    # Perfect lattice: zero force by symmetry.  EdS: drift integral
    # int da / (H0 a^3 sqrt(a^-3)) = (2/H0)(a0^-1/2 - a1^-1/2).
    L, n, M = 10000.0, 8, 16
    tl = timeline(omega0=1.0, omega_lambda=0.0, softening_halo=L / n / 30)
    d, k = tables(orc, tl)
    pos = lattice(n, L)
    vel = np.tile([3.0, -2.0, 0.5], (n ** 3, 1))
    mass = np.full(n ** 3, 1.0)
    p = orc.params(M, L, ics.G_INTERNAL, 1.25 * L / M, 4.5 * 1.25 * L / M,
                   2.8 * tl["softening_halo"])
    run = orc.Timeline(p, tl, d, k, pos, vel, mass)
    st = run.run(10_000)
    assert st["finished"] == 1 and st["ti_current"] == TIMEBASE
    interval = math.log(1.5) / TIMEBASE
    step = 1
    while step * 2 <= tl["max_size_timestep"] / interval:
        step *= 2
    assert st["steps"] == TIMEBASE // step + 1       # + the closing half kick
    assert (run.ti_end == TIMEBASE).all()
    D = 2 / H0 * (0.2 ** -0.5 - 0.3 ** -0.5)
    expect = np.mod(pos + vel * D, L)
    assert np.abs(run.pos - expect).max() <= 1e-9 * L
    assert np.abs(run.vel - vel).max() <= 1e-9


def test_single_level_reduces_to_global_kdk(orc):
    box = ics.small_box("tl", 8, 16, 25.0, 5)
    pos, vel, mass = ics.generate(box)
    tl = timeline(err_tol_int_accuracy=1e30, softening_halo=box.softening)
    d, k = tables(orc, tl)
    p = orc.params(box.M, box.L, ics.G_INTERNAL, box.asmth, box.rcut, box.soft_h)
    run = orc.Timeline(p, tl, d, k, pos, vel, mass)
    st = run.run(10_000)
    assert st["finished"] == 1
    nsteps = st["steps"] - 1
    step = TIMEBASE // nsteps
    interval = math.log(1.5) / TIMEBASE
    lb, lm = math.log(0.2), math.log(0.3)

    def fac(table, t0, t1):
        return orc.table_factor(table, 0.2, 0.3, lb + t0 * interval, lb + t1 * interval)

    f = np.array([[fac(k, s * step, s * step + step // 2),
                   fac(d, s * step, (s + 1) * step),
                   fac(k, s * step + step // 2, (s + 1) * step)] for s in range(nsteps)])
    kpos, kvel, _, _ = orc.kdk_run(p, pos, vel, mass, f)
    assert np.abs(run.pos - kpos).max() <= 1e-9 * box.L
    assert np.abs(run.vel - kvel).max() <= 1e-9 * np.abs(kvel).max()


def test_clustered_box_has_an_aligned_power_of_two_hierarchy(orc):
    box = halo_box(12, 24, 40.0, n_halos=4, seed=9)
    pos, vel, mass = ics.generate(box)
    tl = timeline(softening_halo=box.softening,
                  softening_halo_max_phys=box.softening * 0.25)
    d, k = tables(orc, tl)
    p = orc.params(box.M, box.L, ics.G_INTERNAL, box.asmth, box.rcut, box.soft_h)
    run = orc.Timeline(p, tl, d, k, pos, vel, mass)
    levels = set()
    for _ in range(20):
        before = run.ti_end.copy()
        pm_end = int(run.state[2])
        st = run.run(1)
        assert st["ti_current"] == min(int(before.min()), pm_end)
        assert st["flags"] == 0
        step = run.ti_end - run.ti_beg
        assert (step > 0).all()
        assert ((step & (step - 1)) == 0).all()           # powers of two
        assert (run.ti_beg % step == 0).all()             # aligned
        assert (run.ti_end >= st["ti_current"]).all()
        levels |= set(np.unique(step).tolist())
    assert len(levels) >= 2
