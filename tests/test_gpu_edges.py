"""Edge cases of the force path (gravlist.cpp:148-193, pm_periodic.c:67-314):
fewer particles than one wave, one wave exactly, one and two waves plus one,
every particle crowded into one mesh cell, particles on the box faces.  FP64
short-range forces bit-identical to the oracle and PM within 1e-10; FP32,
both pair variants, within 2e-5 of a natural force unit (the larger of the
largest FP64 force and G m_max / cell^2), plus, per particle, an allowance
for each neighbour closer than 2h: the FP32 kernel holds coordinates
relative to a tile anchor up to rcut and a few cells away, rounded at that
magnitude (a few ulps of offset error per pair), and the softened law's
slope, at most 10.67 G m / h^3, turns that into a force error; two
particles on opposite box faces are periodic images 1e-11 apart, the
extreme case."""
import numpy as np
import pytest

from cuco_amd import engine as E
from cuco_amd import ics

from helpers import make_engine, oracle_params

pytestmark = pytest.mark.gpu

BOX = ics.small_box("edges", 16, 32, 40.0, 31)


def particles(case, rng):
    L, cell = BOX.L, BOX.L / BOX.M
    if case.startswith("n"):
        n = int(case[1:])
        pos = rng.uniform(0, L, (n, 3))
        if n >= 2:                       # one pair certainly inside rcut
            pos[1] = np.mod(pos[0] + 0.4 * BOX.rcut, L)
    elif case == "one_cell":
        corner = np.array([7, 20, 3]) * cell
        pos = np.concatenate([corner + rng.uniform(0, cell, (300, 3)),
                              rng.uniform(0, L, (200, 3))])
    else:                                # faces and corners
        pos = rng.uniform(0, L, (400, 3))
        top = np.nextafter(L, 0.0)
        pos[0] = [0.0, 0.0, 0.0]
        pos[1] = [top, top, top]
        pos[2] = [0.0, top, 0.5 * L]
        pos[3:40, 0] = 0.0
        pos[40:80, 1] = top
        pos[80:120, 2] = rng.choice([0.0, top], 40)
    mass = BOX.N * ics.generate(BOX)[2][0] / pos.shape[0] * rng.uniform(0.5, 1.5, pos.shape[0])
    return pos, mass


def periodic(d):
    return d - BOX.L * np.round(d / BOX.L)


CASES = ["n1", "n2", "n17", "n64", "n65", "n129", "one_cell", "faces"]


@pytest.mark.parametrize("case", CASES)
def test_edge_cases(orc, case):
    rng = np.random.default_rng(CASES.index(case) + 100)
    pos, mass = particles(case, rng)
    n = pos.shape[0]
    print(case, n, flush=True)
    p = oracle_params(BOX)
    ref_pp, npairs = orc.pp_force(p, pos, mass)
    ref_pm = orc.pm_force(p, pos, mass)
    eng = make_engine(BOX)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP64, True)
    app, apm = eng.accelerations()
    assert np.array_equal(app, ref_pp), "FP64 PP must be bit-identical"
    assert eng.count_pairs()[0] == npairs
    cell = BOX.L / BOX.M
    unit = max(np.linalg.norm(ref_pp + ref_pm, axis=1).max(),
               ics.G_INTERNAL * mass.max() / cell ** 2)
    assert np.abs(apm - ref_pm).max() <= 1e-10 * unit
    if n == 1:
        assert not app.any()
    h = BOX.soft_h
    delta = 4 * float(np.spacing(np.float32(BOX.rcut + 3 * cell)))
    per_pair = 10.67 * ics.G_INTERNAL * mass.max() * delta / h ** 3
    close = np.array([np.sum(np.linalg.norm(periodic(pos - x), axis=1) < 2 * h) - 1
                      for x in pos])
    tol = 2e-5 * unit + per_pair * close
    for sym in (0, 1):
        eng.set_tuning(**dict(eng.tuning(), pp_symmetric=sym))
        eng.compute_accelerations(E.FP32, True)
        a32, m32 = eng.accelerations()
        assert np.isfinite(a32).all() and np.isfinite(m32).all()
        err = np.abs(a32 - ref_pp).max(axis=1)
        assert (err <= tol).all(), (sym, "pp", err.max(), int(np.argmax(err - tol)))
        assert np.abs(m32 - ref_pm).max() <= 2e-5 * unit, (sym, "pm")
        if n == 1:
            assert not a32.any()


def test_isolated_particles_feel_no_short_range_force():
    """512 particles on a grid 8 cells apart (rcut is 5.6 cells), each at a
    random point of its cell: a particle's only pair inside rcut is itself,
    whose term must vanish exactly.  The kernel that visits every ordered
    pair stages each target among its own candidates, so this holds only if
    a target's coordinates and its candidate copy's are bit-identical (a
    build that contracted the two coordinate expressions differently gave
    every self pair a one-ulp separation and a spurious softened force).
    Every pair-kernel tuning; FP64 as well."""
    box = ics.small_box("isolated", 8, 64, 80.0, 41)
    rng = np.random.default_rng(41)
    cell = box.L / box.M
    g = np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (g * 8 + rng.uniform(0.0, 1.0, g.shape)) * cell
    mass = np.full(len(pos), ics.generate(box)[2][0])
    eng = make_engine(box, capacity=len(pos))
    eng.upload(pos, None, mass)
    assert eng.count_pairs()[0] == 0
    eng.compute_accelerations(E.FP64, True)
    assert not eng.accelerations()[0].any()
    for sym in (0, 1):
        for level in (0, 1, 2):
            for wide in ((0, 1) if sym else (0,)):
                eng.set_tuning(3, 8, 0, sym, pp_cell_level=level, pp_wide_tiles=wide)
                eng.compute_accelerations(E.FP32, True)
                a32 = eng.accelerations()[0]
                assert not a32.any(), (sym, level, wide, np.abs(a32).max())
    eng.close()
