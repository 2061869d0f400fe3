"""Pins the FP64 CPU oracle before it is trusted as the parity reference.

Each test checks an oracle function against something the oracle's own code
does not restate: numpy/scipy golden vectors, closed forms from the physics
of the erfc force split, invariants (Newton's third law, mass conservation,
symmetry), or brute force on tiny inputs.
"""
import math

import numpy as np
import pytest
from scipy import special

from cuco_amd import ics


def make_params(orc, M=16, L=1000.0, G=1.0, n_side=16):
    asmth = 1.25 * L / M
    return orc.params(M, L, G, asmth, 4.5 * asmth, 2.8 * L / n_side / 30)


# ---------------------------------------------------------------- FFT
def test_fft1d_matches_numpy_golden(orc, golden_dir):
    g = np.load(f"{golden_dir}/fft_golden.npz")
    for n in (7, 8, 12, 30, 64):
        x = g[f"c1d_{n}_in"]
        ref = g[f"c1d_{n}_fwd"]
        np.testing.assert_allclose(orc.fft1d(x, -1), ref, rtol=0,
                                   atol=1e-12 * np.abs(ref).max())
        # inverse of the forward transform returns n*x
        np.testing.assert_allclose(orc.fft1d(ref, +1) / n, x, rtol=0, atol=1e-12)


def test_fft3d_r2c_matches_numpy_golden(orc, golden_dir):
    g = np.load(f"{golden_dir}/fft_golden.npz")
    for M in (8, 12, 16):
        a = g[f"r3d_{M}_in"]
        ref = g[f"r3d_{M}_rfftn"]
        c = orc.fft3d_r2c(a)
        np.testing.assert_allclose(c, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        back = orc.fft3d_c2r(c)
        np.testing.assert_allclose(back, M ** 3 * a, rtol=0,
                                   atol=1e-10 * M ** 3 * np.abs(a).max())


def test_fft3d_parseval(orc):
    rng = np.random.default_rng(3)
    M = 12
    a = rng.standard_normal((M, M, M))
    c = orc.fft3d_r2c(a)
    # Hermitian half: count interior kz planes twice.
    w = np.full(M // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    power = (np.abs(c) ** 2 * w).sum()
    assert power == pytest.approx(M ** 3 * (a ** 2).sum(), rel=1e-12)


# ---------------------------------------------------------------- table
def test_shortrange_table_matches_scipy(orc, golden_dir):
    g = np.load(f"{golden_dir}/erfc_table.npz")
    t = orc.shortrange_table()
    ref = g["table"].astype(np.float32)
    # float rounding of the same double values: equal or one float ulp apart
    assert np.all(np.abs(t.view(np.int32) - ref.view(np.int32)) <= 1)


def test_params_closed_forms(orc):
    M, L, G = 64, 1e5, 43007.1
    asmth = 1.25 * L / M
    p = orc.params(M, L, G, asmth, 4.5 * asmth, 300.0)
    assert p.asmthfac == pytest.approx(0.5 / asmth * 1000 / 3, rel=1e-15)
    # rcut sits on table bin 750 for RCUT/ASMTH = 4.5
    assert p.asmthfac * p.rcut == pytest.approx(750.0, rel=1e-14)
    assert p.h3_inv == pytest.approx(300.0 ** -3, rel=1e-14)
    assert p.boxhalf == L / 2 and p.to_grid_fac == M / L


# ---------------------------------------------------------------- keys / sort
def test_keys_match_bruteforce(orc):
    rng = np.random.default_rng(4)
    p = make_params(orc)
    pos = rng.uniform(0, 1000.0, (500, 3))
    pos[0] = [999.9999999999999, 0.0, 500.0]     # x*M/L may round to M
    keys, idx = orc.calc_keys(p, pos)
    c = np.minimum((pos * 16 / 1000.0).astype(int), 15)
    assert np.array_equal(keys, (c[:, 0] * 16 + c[:, 1]) * 16 + c[:, 2])
    assert np.array_equal(idx, np.arange(500))


@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 200])
def test_sort_is_stable_insertion_sort(orc, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 7, n).astype(np.uint32)   # many ties
    vals = np.arange(n, dtype=np.uint32)
    k2, v2 = orc.sort_pairs(keys, vals)
    # textbook stable insertion sort
    kk, vv = list(keys), list(vals)
    for i in range(1, n):
        j = i
        while j > 0 and kk[j - 1] > kk[j]:
            kk[j - 1], kk[j] = kk[j], kk[j - 1]
            vv[j - 1], vv[j] = vv[j], vv[j - 1]
            j -= 1
    assert np.array_equal(k2, kk) and np.array_equal(v2, vv)


def test_cell_table_invariants(orc):
    rng = np.random.default_rng(5)
    keys = np.sort(rng.integers(0, 64, 300).astype(np.uint32))
    s, e = orc.cell_table(keys, 64)
    empty = s == 0xFFFFFFFF
    assert np.array_equal(empty, e == 0xFFFFFFFF)
    assert (e[~empty] - s[~empty]).sum() == 300
    for c in np.nonzero(~empty)[0]:
        assert np.all(keys[s[c]:e[c]] == c)
    assert set(np.nonzero(~empty)[0]) == set(np.unique(keys))


# ---------------------------------------------------------------- PP
def test_pp_cell_list_equals_bruteforce(orc):
    rng = np.random.default_rng(6)
    p = make_params(orc)
    pos = rng.uniform(0, 1000.0, (700, 3))
    pos[:50] = 500.0 + rng.normal(0, 5.0, (50, 3))   # a clump inside h
    mass = rng.uniform(0.5, 1.5, 700)
    a, _ = orc.pp_force(p, pos, mass)
    b = orc.pp_force_bruteforce(p, pos, mass)
    scale = np.abs(b).max()
    assert np.abs(a - b).max() <= 1e-12 * scale


def test_pp_two_particle_erfc_closed_form(orc):
    """|a| = G m / r^2 * [erfc(u) + 2u/sqrt(pi) exp(-u^2)] at the bin centre."""
    M, L, G = 16, 1000.0, 2.5
    asmth = 1.25 * L / M
    p = orc.params(M, L, G, asmth, 4.5 * asmth, 1.0)
    for r in (3.0, 10.0, 55.0, 200.0, 350.0):
        pos = np.array([[100.0, 200.0, 300.0], [100.0 + r, 200.0, 300.0]])
        mass = np.array([1.0, 3.0])
        a, n_in = orc.pp_force(p, pos, mass)
        k = math.floor(0.5 / asmth * (1000 / 3.0) * r)
        u = 3.0 / 1000 * (k + 0.5)
        f = special.erfc(u) + 2 * u / math.sqrt(math.pi) * math.exp(-u * u)
        f = float(np.float32(f))
        assert a[0, 0] == pytest.approx(G * 3.0 / r ** 2 * f, rel=1e-12)
        assert a[1, 0] == pytest.approx(-G * 1.0 / r ** 2 * f, rel=1e-12)
        assert n_in == 2


def test_pp_beyond_rcut_is_zero(orc):
    p = make_params(orc)
    pos = np.array([[10.0, 10.0, 10.0], [10.0 + p.rcut * 1.0001, 10.0, 10.0]])
    a, n_in = orc.pp_force(p, pos, np.ones(2))
    assert n_in == 0 and np.all(a == 0)


def test_pp_spline_continuous_at_joins(orc):
    """The softened law joins itself at u=0.5 and Newton at r=h (1e-5 from
    the rounded spline constants)."""
    M, L = 16, 1000.0
    asmth = 1.25 * L / M
    h = 40.0
    p = orc.params(M, L, 1.0, asmth, 4.5 * asmth, h)

    def fmag(r):
        pos = np.array([[100.0, 100.0, 100.0], [100.0 + r, 100.0, 100.0]])
        a, _ = orc.pp_force(p, pos, np.ones(2))
        return a[0, 0] * r     # fac = |a|/r
    for r0 in (0.5 * h, h):
        lo, hi = fmag(r0 * (1 - 1e-9)), fmag(r0 * (1 + 1e-9))
        assert lo == pytest.approx(hi, rel=2e-5)


def test_pp_newton_third_law_and_periodic_images(orc):
    rng = np.random.default_rng(7)
    p = make_params(orc)
    pos = rng.uniform(0, 1000.0, (400, 3))
    pos[:40] = rng.uniform(0, 30.0, (40, 3))          # straddle the corner
    pos[40:80] = rng.uniform(970.0, 1000.0, (40, 3))
    mass = rng.uniform(0.5, 2.0, 400)
    a, _ = orc.pp_force(p, pos, mass)
    net = (a * mass[:, None]).sum(0)
    assert np.abs(net).max() <= 1e-12 * np.abs(a * mass[:, None]).sum()


def test_pp_perfect_lattice_has_no_force(orc):
    n = 16
    L = 1000.0
    q = (np.indices((n, n, n)).reshape(3, -1).T + 0.5) * (L / n)
    p = make_params(orc)
    a, _ = orc.pp_force(p, q, np.ones(n ** 3))
    assert np.abs(a).max() < 1e-12


# ---------------------------------------------------------------- PM
def test_deposit_conserves_mass_and_node_particle(orc):
    rng = np.random.default_rng(8)
    p = make_params(orc)
    pos = rng.uniform(0, 1000.0, (300, 3))
    mass = rng.uniform(0.1, 2.0, 300)
    rho = orc.pm_deposit(p, pos, mass)
    assert rho.sum() == pytest.approx(mass.sum(), rel=1e-13)
    # a particle exactly on node (3, 5, 7) deposits to that node only
    node = np.array([[3, 5, 7]]) * 1000.0 / 16
    r1 = orc.pm_deposit(p, node, np.array([2.0]))
    assert r1[3, 5, 7] == 2.0 and np.count_nonzero(r1) == 1


def test_deposit_cic_weights_closed_form(orc):
    p = make_params(orc)
    d = np.array([0.25, 0.5, 0.125])
    pos = ((np.array([15, 2, 9]) + d) * 1000.0 / 16)[None, :]
    rho = orc.pm_deposit(p, pos, np.array([1.0]))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = (d[0] if a else 1 - d[0]) * (d[1] if b else 1 - d[1]) * \
                    (d[2] if c else 1 - d[2])
                assert rho[(15 + a) % 16, 2 + b, 9 + c] == pytest.approx(w, rel=1e-14)


def test_potential_single_mode_closed_form(orc):
    """cos mode through r2c -> Green -> c2r: phi = M^3 G(k) cos, with
    G(k) = -exp(-k^2 (2 pi asmth/L)^2)/k^2 / sinc(pi k/M)^4."""
    M, L = 16, 1000.0
    asmth = 1.25 * L / M
    p = orc.params(M, L, 1.0, asmth, 4.5 * asmth, 1.0)
    x = np.arange(M)
    for axis, k0 in ((0, 1), (1, 3), (2, 2)):
        shape = [1, 1, 1]
        shape[axis] = M
        rho = np.broadcast_to(np.cos(2 * np.pi * k0 * x / M).reshape(shape),
                              (M, M, M)).copy()
        phi = orc.pm_potential(p, rho)
        g = -math.exp(-k0 ** 2 * (2 * math.pi * asmth / L) ** 2) / k0 ** 2 / \
            np.sinc(k0 / M) ** 4
        np.testing.assert_allclose(phi, M ** 3 * g * rho, rtol=0, atol=1e-9)


def test_gradient_of_sine_mode_closed_form(orc):
    """4-point stencil on sin(theta x): -(8/3 sin th - 1/3 sin 2th) cos(th x),
    interpolated at the nodes."""
    M, L, G = 16, 1000.0, 1.0
    asmth = 1.25 * L / M
    p = orc.params(M, L, G, asmth, 4.5 * asmth, 1.0)
    th = 2 * math.pi * 2 / M
    x = np.arange(M)
    phi = np.broadcast_to(np.sin(th * x)[None, :, None], (M, M, M)).copy()
    nodes = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1).reshape(-1, 3) * (L / M)
    acc = orc.pm_interpolate(p, phi, nodes)
    fac = G / (math.pi * L) * M / (2 * L)
    expect = fac * -(8 / 3 * math.sin(th) - 1 / 3 * math.sin(2 * th)) * \
        np.cos(th * nodes[:, 1] * M / L)
    np.testing.assert_allclose(acc[:, 1], expect, rtol=0, atol=1e-13)
    assert np.abs(acc[:, 0]).max() < 1e-14 and np.abs(acc[:, 2]).max() < 1e-14


def test_pm_no_self_force_single_particle(orc):
    p = make_params(orc)
    pos = np.array([[123.4, 567.8, 901.2]])
    a = orc.pm_force(p, pos, np.array([5.0]))
    assert np.abs(a).max() < 1e-12


def test_pm_plus_pp_recovers_newton(orc):
    """Gadget-2 split test: PM + PP for a pair reproduces G m / r^2 (plus the
    small periodic-image term) across the split scale."""
    M, L, G = 64, 1000.0, 1.0
    asmth = 1.25 * L / M
    p = orc.params(M, L, G, asmth, 4.5 * asmth, 0.5)
    for r in (5.0, 20.0, 40.0, 60.0):
        errs = []
        for off in ((0.11, 0.37, 0.73), (0.5, 0.5, 0.5), (0.9, 0.2, 0.6)):
            base = (np.array([20, 30, 40]) + np.array(off)) * L / M
            pos = np.array([base, base + [r, 0, 0]])
            mass = np.array([1.0, 1.0])
            a = orc.pp_force(p, pos, mass)[0] + orc.pm_force(p, pos, mass)
            errs.append(a[0, 0] / (G / r ** 2) - 1)
        assert abs(np.mean(errs)) < 0.02, (r, errs)


# ---------------------------------------------------------------- drift tables
def test_drift_tables_einstein_de_sitter(orc):
    a0, a1, H0 = 0.02, 0.05, 0.1
    d, k = orc.drift_tables(a0, a1, 1.0, 0.0, H0)
    a = np.exp(np.log(a0) + (np.log(a1) - np.log(a0)) / 1000 * (np.arange(1000) + 1))
    np.testing.assert_allclose(d, 2 / H0 * (a0 ** -0.5 - a ** -0.5), rtol=1e-12)
    np.testing.assert_allclose(k, 2 / H0 * (a ** 0.5 - a0 ** 0.5), rtol=1e-12)


def test_drift_tables_lcdm_match_scipy(orc, golden_dir):
    g = np.load(f"{golden_dir}/drift_kick.npz")
    a0, a1 = float(g["a0"]), float(g["a1"])
    d, k = orc.drift_tables(a0, a1, float(g["omega0"]), float(g["omega_lambda"]),
                            float(g["hubble"]))
    for av, dr, kr in zip(g["avals"], g["drift"], g["kick"]):
        la = math.log(av)
        assert orc.table_factor(d, a0, a1, math.log(a0), la) == pytest.approx(dr, rel=1e-9)
        assert orc.table_factor(k, a0, a1, math.log(a0), la) == pytest.approx(kr, rel=1e-9)


def test_table_factor_first_bin_branch(orc):
    """driftfac.c:79-82: below u=1 the factor is u*table[0]."""
    a0, a1 = 0.02, 0.05
    d, _ = orc.drift_tables(a0, a1, 1.0, 0.0, 0.1)
    la0, la1 = math.log(a0), math.log(a1)
    u = 0.5
    f = orc.table_factor(d, a0, a1, la0, la0 + u * (la1 - la0) / 1000)
    assert f == pytest.approx(u * d[0], rel=1e-15)


def test_lattice_workload_parameters():
    box = ics.WORKLOADS["lattice32"]
    assert box.N == 32768 and box.M == 64
    assert box.rcut / (box.L / box.M) == pytest.approx(5.625)
    assert box.soft_h == pytest.approx(2.8 * box.L / 32 / 30)
