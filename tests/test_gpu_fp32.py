"""Chosen FP32 path: cell-relative FP32 pairs with Kahan accumulation, FP32
mesh.  Per-particle relative acceleration error against FP64 (median and
99th percentile), separately for PM, PP and total, with |a_total| of the
FP64 reference as the denominator.  Bar: 99th percentile <= 1e-3."""
import numpy as np
import pytest

from cuco_amd import engine as E
from cuco_amd import ics

from helpers import halo_box, make_engine, oracle_params, percentiles, rel_err

pytestmark = pytest.mark.gpu

P99_BAR = 1e-3


def fp32_errors(box, pos, mass, orc=None, sym=None):
    eng = make_engine(box)
    eng.upload(pos, None, mass)
    if sym is not None:
        eng.set_tuning(**dict(eng.tuning(), pp_symmetric=sym))
    if orc is not None:
        p = oracle_params(box)
        ref_pp, _ = orc.pp_force(p, pos, mass)
        ref_pm = orc.pm_force(p, pos, mass)
    else:
        eng.compute_accelerations(E.FP64, True)
        ref_pp, ref_pm = eng.accelerations()
    eng.compute_accelerations(E.FP32, True)
    a_pp, a_pm = eng.accelerations()
    tot = ref_pp + ref_pm
    errs = {
        "pp": percentiles(rel_err(a_pp, ref_pp, tot)),
        "pm": percentiles(rel_err(a_pm, ref_pm, tot)),
        "total": percentiles(rel_err(a_pp + a_pm, tot, tot)),
    }
    return eng, errs


@pytest.mark.parametrize("sym", [0, 1])
def test_lattice_error_percentiles(orc, sym):
    box = ics.WORKLOADS["lattice32"]
    pos, vel, mass = ics.generate(box)
    eng, errs = fp32_errors(box, pos, mass, orc, sym)
    print(errs)
    for k in ("pp", "pm", "total"):
        assert errs[k]["p99"] <= P99_BAR, (k, errs)
        assert errs[k]["median"] <= 1e-4, (k, errs)


@pytest.mark.parametrize("sym", [0, 1])
def test_clustered_error_percentiles(orc, sym):
    box = halo_box()
    pos, vel, mass = ics.generate(box)
    eng, errs = fp32_errors(box, pos, mass, orc, sym)
    print(errs)
    assert errs["total"]["p99"] <= P99_BAR, errs


def test_halo128_error_percentiles():
    box = ics.WORKLOADS["halo128"]
    pos, vel, mass = ics.generate(box)
    eng, errs = fp32_errors(box, pos, mass)
    print(errs)
    assert errs["total"]["p99"] <= P99_BAR, errs
    assert errs["pp"]["p99"] <= P99_BAR, errs
    assert errs["pm"]["p99"] <= P99_BAR, errs


@pytest.mark.parametrize("sym", [0, 1])
def test_halo128_pair_force_error_both_variants(sym):
    """Short-range force alone, FP32 against FP64, for both pair variants
    (the each-pair-once one enumerates wide tiles by 2x2x2-cell blocks): a
    missing or doubled pair shows up far above FP32 rounding, which keeps
    the 99th percentile near 1e-6 of |a_total|."""
    box = ics.WORKLOADS["halo128"]
    pos, vel, mass = ics.generate(box)
    eng = make_engine(box)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP64, True)
    ref_pp, ref_pm = eng.accelerations()
    eng.set_tuning(3, 8, 0, sym)
    eng.compute_accelerations(E.FP32, True)
    a_pp = eng.accelerations()[0]
    err = rel_err(a_pp, ref_pp, ref_pp + ref_pm)
    print(sym, np.median(err), np.percentile(err, 99), err.max())
    assert np.percentile(err, 99) <= 2e-5
    assert np.median(err) <= 1e-6


@pytest.mark.parametrize("M", [40, 48, 96, 160, 192])
def test_error_percentiles_other_meshes(orc, M):
    """Meshes that are not powers of two (40, 48: LDS Stockham FFT; 96, 160,
    192: register FFT with radix-3 and radix-5 steps) or not multiples of
    the 16-cell deposit and interpolation bricks (the last brick wraps)."""
    box = ics.small_box(f"mesh{M}", 24, M, 60.0, 30 + M, "halos", n_halos=8)
    pos, vel, mass = ics.generate(box)
    eng, errs = fp32_errors(box, pos, mass, orc)
    print(M, errs)
    for k in ("pp", "pm", "total"):
        assert errs[k]["p99"] <= P99_BAR, (k, errs)


@pytest.mark.parametrize("M", [40, 48, 96])
def test_pair_variants_other_meshes(M):
    """Pair force alone for both variants on meshes whose Morton grid is
    larger than the mesh (the 2x2x2-cell blocks of the each-pair-once
    variant wrap on M/2 blocks).  Sparse boxes, so most tiles are wide and
    take blocks."""
    box = ics.small_box(f"mesh{M}", 24, M, 60.0, 50 + M, "halos", n_halos=8)
    pos, vel, mass = ics.generate(box)
    eng = make_engine(box)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP64, True)
    ref_pp, ref_pm = eng.accelerations()
    for sym in (0, 1):
        eng.set_tuning(3, 8, 0, sym)
        eng.compute_accelerations(E.FP32, True)
        err = rel_err(eng.accelerations()[0], ref_pp, ref_pp + ref_pm)
        print(M, sym, np.median(err), np.percentile(err, 99), err.max())
        assert np.percentile(err, 99) <= 2e-5, (M, sym)
        assert np.median(err) <= 1e-6, (M, sym)


def test_fp32_refuses_rcut_beyond_its_reach():
    """The FP32 pair kernel addresses cells up to 475 cells from a tile; the
    parameter check admits rcut up to M/2 - 1 cells, so a larger rcut on a
    fine mesh is refused with a status (nothing is computed)."""
    M, L = 960, 1000.0
    cell = L / M
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.0, L, (2000, 3))
    eng = E.Engine(max_particles=2000, pm_grid=M, box_size=L, G=ics.G_INTERNAL,
                   asmth=1.25 * cell, rcut=477.0 * cell, soft_h=0.1 * cell)
    eng.upload(pos, None, np.ones(2000))
    with pytest.raises(E.CucoError) as ei:
        eng.gravity_linklist(E.FP32)
    assert ei.value.status == 5
    assert "reach" in str(ei.value)


def test_fp32_is_deterministic_and_counts_cover_pairs():
    box = halo_box(seed=4)
    pos, vel, mass = ics.generate(box)
    eng = make_engine(box)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP32, True)
    a1 = eng.accelerations()
    eng.compute_accelerations(E.FP32, True)
    a2 = eng.accelerations()
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])
    assert eng.tuning()["pp_symmetric"] == 1        # the default variant
    pairs, cand_once = eng.count_pairs()
    assert 2 * cand_once >= pairs > 0               # unordered pairs, once each
    eng.set_tuning(**dict(eng.tuning(), pp_symmetric=0))
    pairs0, cand = eng.count_pairs()
    assert pairs0 == pairs and cand >= pairs        # every ordered pair


def test_every_tuning_gives_the_same_forces():
    box = halo_box(seed=12)
    pos, vel, mass = ics.generate(box)
    eng = make_engine(box)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP32, True)
    ref = eng.accelerations()[0]
    scale = np.linalg.norm(ref, axis=1).max()
    for shift in (2, 3, 4):
        for wg in (1, 4, 16):
            for sym in (0, 1):
                eng.set_tuning(shift, wg, 0, sym)
                eng.compute_accelerations(E.FP32, True)
                a = eng.accelerations()[0]
                assert np.abs(a - ref).max() <= 1e-5 * scale, (shift, wg, sym)
    best, ms = eng.autotune()
    assert best["pp_block_shift"] in (2, 3, 4) and ms > 0
    assert best["pp_cell_level"] in (0, 1, 2) and best["fft_split"] in (0, 1)
    assert best["pp_wide_tiles"] in (0, 1)
    assert eng.tuning() == best
    eng.compute_accelerations(E.FP32, True)
    assert np.abs(eng.accelerations()[0] - ref).max() <= 1e-5 * scale
    with pytest.raises(E.CucoError):
        eng.set_tuning(9, 4)
    with pytest.raises(E.CucoError):
        eng.set_tuning(3, 8, 0, 1, pp_cell_level=3)
    with pytest.raises(E.CucoError):
        eng.set_tuning(3, 8, 0, 1, fft_split=2)
    with pytest.raises(E.CucoError):
        eng.set_tuning(3, 8, 0, 1, pp_wide_tiles=2)


@pytest.mark.parametrize("M", [64, 128, 96])
def test_enumeration_level_and_fft_split_give_the_same_forces(M):
    """The autotuner's candidate-enumeration level of the each-pair-once
    kernel (mesh cells, 2x2x2-cell blocks for wide tiles, blocks for every
    tile), its tile width (64 or 128 targets) and the register-FFT line
    split (96: not a power of two, one register-FFT shape whatever the
    split) change only FP32 rounding: PP within 1e-5 and PM within 1e-5 of
    the force scale."""
    box = halo_box(32, M, 100.0, n_halos=20, seed=M)
    pos, vel, mass = ics.generate(box)
    eng = make_engine(box)
    eng.upload(pos, None, mass)
    eng.set_tuning(3, 8, 0, 1)
    eng.compute_accelerations(E.FP32, True)
    ref_pp, ref_pm = eng.accelerations()
    scale = np.linalg.norm(ref_pp + ref_pm, axis=1).max()
    for level in (0, 1, 2):
        for split in (0, 1):
            for wide in (0, 1):
                eng.set_tuning(3, 8, 0, 1, pp_cell_level=level, fft_split=split,
                               pp_wide_tiles=wide)
                t = eng.tuning()
                assert (t["pp_cell_level"], t["fft_split"], t["pp_wide_tiles"]) == \
                    (level, split, wide)
                eng.compute_accelerations(E.FP32, True)
                a, p = eng.accelerations()
                assert np.abs(a - ref_pp).max() <= 1e-5 * scale, (level, split, wide)
                assert np.abs(p - ref_pm).max() <= 1e-5 * scale, (level, split, wide)
                if split == 0:
                    assert np.array_equal(p, ref_pm)
    eng.close()


@pytest.mark.parametrize("case", ["equal", "unequal", "off_edge"])
def test_wide_tiles(case):
    """Each pair once with tiles of up to 128 targets, two per lane: the
    same error bar against the FP64 path as tiles of 64, identical results
    on every run, the same pairs counted, and the pair force within FP32
    rounding of the 64-target tiles.  Cases: uniform masses; per-particle
    masses; rcut = 4 asmth, off the table's bin edge (explicit cut and bin
    clamp).  Switching to the kernel that visits every ordered pair brings
    back tiles of 64."""
    box = halo_box(seed=17)
    pos, vel, mass = ics.generate(box)
    if case == "unequal":
        mass = mass * np.random.default_rng(17).uniform(0.5, 1.5, box.N)
    kw = box.engine_kwargs()
    if case == "off_edge":
        kw["rcut"] = 4.0 * box.asmth
    eng = E.Engine(max_particles=box.N, **kw)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP64, True)
    ref_pp, ref_pm = eng.accelerations()
    tot = ref_pp + ref_pm
    scale = np.linalg.norm(tot, axis=1).max()
    eng.set_tuning(3, 8, 0, 1)
    eng.compute_accelerations(E.FP32, True)
    narrow = eng.accelerations()[0]
    pairs64, cand64 = eng.count_pairs()
    eng.set_tuning(3, 8, 0, 1, pp_wide_tiles=1)
    runs = []
    for _ in range(2):
        eng.compute_accelerations(E.FP32, True)
        runs.append(eng.accelerations()[0])
    assert np.array_equal(runs[0], runs[1])
    err = rel_err(runs[0], ref_pp, tot)
    print(case, np.median(err), np.percentile(err, 99), err.max())
    assert np.percentile(err, 99) <= 2e-5 and np.median(err) <= 1e-6
    assert np.abs(runs[0] - narrow).max() <= 1e-5 * scale
    pairs128, cand128 = eng.count_pairs()
    assert pairs128 == pairs64 and 0 < cand128 and cand64 > 0
    eng.set_tuning(3, 8, 0, 0, pp_wide_tiles=1)
    eng.compute_accelerations(E.FP32, True)
    assert np.percentile(rel_err(eng.accelerations()[0], ref_pp, tot), 99) <= 2e-5
    eng.close()


@pytest.mark.parametrize("precision", [E.FP64, E.FP32])
def test_pm_overlap_gives_identical_forces_and_graphs(precision):
    """The PM chain on a second stream, concurrent with the pair force:
    same forces, and a captured KDK step (fork/join inside the graph)
    replays to the same trajectory as eager steps without overlap."""
    box = halo_box(seed=13)
    pos, vel, mass = ics.generate(box)
    lat = ics.WORKLOADS["lattice32"]
    d, k = E.init_drift_table(lat.a_start, lat.a_end, ics.OMEGA_M, ics.OMEGA_L,
                              ics.HUBBLE_INTERNAL)
    f = E.kdk_factors(d, k, lat.a_start, lat.a_end, lat.a_start, lat.a_end, 10)[:3]
    runs = []
    for overlap in (0, 1):
        eng = make_engine(box)
        eng.set_tuning(3, 4, overlap)
        eng.upload(pos, vel, mass)
        eng.compute_accelerations(precision, True)
        a = eng.accelerations()
        eng.kdk_steps(f[:1], precision, use_graph=bool(overlap))
        eng.kdk_steps(f[1:], precision, use_graph=bool(overlap))
        runs.append((a, eng.download_particles(), eng.accelerations()))
    (a0, s0, b0), (a1, s1, b1) = runs
    assert np.array_equal(a0[0], a1[0]) and np.array_equal(a0[1], a1[1])
    assert np.array_equal(s0[0], s1[0]) and np.array_equal(s0[1], s1[1])
    assert np.array_equal(b0[0], b1[0]) and np.array_equal(b0[1], b1[1])


def test_each_pair_once_is_deterministic_and_accurate():
    """The variant that evaluates each pair once for both particles (j side
    summed in 64-bit fixed point): identical results on every run, and the
    same error bar against the FP64 path as the default kernel."""
    box = halo_box(seed=14)
    pos, vel, mass = ics.generate(box)
    eng = make_engine(box)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP64, True)
    ref = eng.accelerations()[0]
    tot = ref + eng.accelerations()[1]
    eng.set_tuning(3, 8, 0, 1)
    runs = []
    for _ in range(3):
        eng.compute_accelerations(E.FP32, True)
        runs.append(eng.accelerations()[0])
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2])
    err = rel_err(runs[0], ref, tot)
    assert np.percentile(err, 99) <= 1e-4
    pairs, cand_once = eng.count_pairs()
    eng.set_tuning(3, 8, 0, 0)
    pairs0, cand_both = eng.count_pairs()
    assert pairs == pairs0 and 0 < cand_once < 0.7 * cand_both  # half the work


def test_each_pair_once_with_unequal_masses():
    """Each pair once with per-particle masses (the target side takes j's
    mass, the j side the target's; the fixed-point scale follows the
    largest mass): pair force within FP32 rounding of the FP64 path for both
    variants, identical on every run, and the each-pair-once kernel really
    runs (about half the candidates)."""
    box = halo_box(seed=15)
    pos, vel, mass = ics.generate(box)
    rng = np.random.default_rng(15)
    mass = mass * rng.uniform(0.5, 1.5, box.N)
    eng = make_engine(box)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP64, True)
    ref_pp, ref_pm = eng.accelerations()
    for sym in (0, 1):
        eng.set_tuning(3, 8, 0, sym)
        runs = []
        for _ in range(2):
            eng.compute_accelerations(E.FP32, True)
            runs.append(eng.accelerations()[0])
        assert np.array_equal(runs[0], runs[1]), sym
        err = rel_err(runs[0], ref_pp, ref_pp + ref_pm)
        print(sym, np.median(err), np.percentile(err, 99), err.max())
        assert np.percentile(err, 99) <= 2e-5, sym
        assert np.median(err) <= 1e-6, sym
    _, cand_once = eng.count_pairs()
    eng.set_tuning(3, 8, 0, 0)
    _, cand_both = eng.count_pairs()
    assert 0 < cand_once < 0.7 * cand_both


@pytest.mark.parametrize("unequal", [False, True])
def test_rcut_off_the_table_edge(unequal):
    """rcut = 4 asmth puts the cut inside table bin 666 (the reference's 4.5
    asmth sits on the edge of bin 750): both pair variants then mask pairs
    past rcut explicitly and clamp the bin.  Pair force within FP32 rounding
    of the FP64 path, equal and unequal masses."""
    box = halo_box(seed=16)
    pos, vel, mass = ics.generate(box)
    if unequal:
        mass = mass * np.random.default_rng(16).uniform(0.5, 1.5, box.N)
    kw = box.engine_kwargs()
    kw["rcut"] = 4.0 * box.asmth
    eng = E.Engine(max_particles=box.N, **kw)
    eng.upload(pos, None, mass)
    eng.compute_accelerations(E.FP64, True)
    ref_pp, ref_pm = eng.accelerations()
    for sym in (0, 1):
        eng.set_tuning(3, 8, 0, sym)
        eng.compute_accelerations(E.FP32, True)
        err = rel_err(eng.accelerations()[0], ref_pp, ref_pp + ref_pm)
        print(unequal, sym, np.median(err), np.percentile(err, 99), err.max())
        assert np.percentile(err, 99) <= 2e-5, sym
        assert np.median(err) <= 1e-6, sym
    eng.set_tuning(3, 8, 0, 1)
    _, cand_once = eng.count_pairs()
    eng.set_tuning(3, 8, 0, 0)
    _, cand_both = eng.count_pairs()
    assert 0 < cand_once < 0.7 * cand_both
