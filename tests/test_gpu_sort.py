"""Device radix sort (RadixSort::sort, radixsort.h:170-187) against the
stable order, bit for bit."""
import numpy as np
import pytest

from cuco_amd import engine as E

pytestmark = pytest.mark.gpu


def stable(keys, bits):
    masked = keys & np.uint32((1 << bits) - 1) if bits < 32 else keys
    order = np.argsort(masked, kind="stable")
    return keys[order], order.astype(np.uint32)


@pytest.mark.parametrize("n,bits", [
    (1, 8), (2, 1), (31, 8), (32, 8), (33, 8), (1023, 10), (1024, 10),
    (1025, 10), (4095, 12), (4096, 12), (4097, 12), (8191, 18), (65537, 24),
    (262144, 27), (300001, 30), (100000, 32),
])
def test_sort_matches_stable_permutation(n, bits, orc):
    rng = np.random.default_rng(n * 31 + bits)
    keys = rng.integers(0, 1 << min(bits, 32), n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    ks, vs = E.radix_sort(keys, vals, bits)
    rk, rv = stable(keys, bits)
    assert np.array_equal(ks, rk) and np.array_equal(vs, rv)
    if bits <= 30 and n <= 5000:
        ok, ov = orc.sort_pairs(keys, vals)
        assert np.array_equal(vs, ov)


@pytest.mark.parametrize("case", ["equal", "max", "descending", "few"])
def test_sort_edge_cases(case):
    n, bits = 50000, 27
    if case == "equal":
        keys = np.full(n, 12345, np.uint32)
    elif case == "max":
        keys = np.full(n, (1 << bits) - 1, np.uint32)
        keys[::7] = 0
    elif case == "descending":
        keys = np.arange(n, 0, -1).astype(np.uint32)
    else:
        keys = np.random.default_rng(1).integers(0, 3, n).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    ks, vs = E.radix_sort(keys, vals, bits)
    rk, rv = stable(keys, bits)
    assert np.array_equal(ks, rk) and np.array_equal(vs, rv)


def test_sort_keys_only_and_ignores_high_bits():
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 1 << 32, 20000, dtype=np.uint64).astype(np.uint32)
    ks, vs = E.radix_sort(keys, None, 16)
    assert vs is None
    rk, _ = stable(keys, 16)
    assert np.array_equal(ks, rk)


@pytest.mark.parametrize("M", [512, 320])
def test_engine_sort_with_25_to_27_bit_keys(M, orc):
    """The engine's own sort on a 512^3 (27-bit keys) or 320^3 (25-bit)
    mesh with a dense clump of repeated keys: the order must be the
    oracle's (x-major keys) and a stable order of the Morton keys."""
    L, n = 100.0, 20000
    rng = np.random.default_rng(M)
    pos = rng.random((n, 3)) * L
    pos[:4000] = L * 0.5 + rng.random((4000, 3)) * (0.02 * L)  # a clump: repeated keys
    mass = np.ones(n)
    asmth = 1.25 * L / M
    kw = dict(pm_grid=M, box_size=L, G=1.0, asmth=asmth, rcut=4.5 * asmth, soft_h=0.1 * L / M)
    eng = E.Engine(max_particles=n, **kw)
    eng.upload(pos, None, mass)
    eng.gravity_linklist(E.FP64)
    keys, idx = eng.sorted_order()
    p = orc.params(M, L, 1.0, asmth, 4.5 * asmth, 0.1 * L / M)
    ok, oi = orc.calc_keys(p, pos)
    ok, oi = orc.sort_pairs(ok, oi)
    assert int(keys.max()) >= 1 << 24
    assert np.array_equal(keys, ok) and np.array_equal(idx, oi)
    eng.gravity_linklist(E.FP32)
    k32, i32 = eng.sorted_order()
    dk = np.diff(k32.astype(np.int64))
    assert (dk >= 0).all()
    assert (np.diff(i32.astype(np.int64))[dk == 0] > 0).all()
    assert np.array_equal(np.sort(i32), np.arange(n, dtype=np.uint32))


def test_cell_table_with_long_empty_ranges(orc):
    """Cell table of a 256^3-cell box holding 20000 particles in two small
    clumps: the empty key ranges between them run to millions of cells (the
    first and the last range included), so they go to the long-range list
    and are filled in pieces of 65,536 entries.  Every cell's start and end
    must be the oracle's.  A second particle set with the clumps elsewhere,
    on the same handle, shows that no entry keeps the first table's value."""
    M, L, n = 256, 100.0, 20000
    asmth = 1.25 * L / M
    kw = dict(pm_grid=M, box_size=L, G=1.0, asmth=asmth, rcut=4.5 * asmth, soft_h=0.1 * L / M)
    p = orc.params(M, L, 1.0, asmth, 4.5 * asmth, 0.1 * L / M)
    eng = E.Engine(max_particles=n, **kw)
    rng = np.random.default_rng(77)
    for centres in ((0.3, 0.7), (0.12, 0.55)):
        pos = np.empty((n, 3))
        for h, c in enumerate(centres):
            pos[h * n // 2:(h + 1) * n // 2] = L * c + rng.random((n // 2, 3)) * (0.02 * L)
        eng.upload(pos, None, np.ones(n))
        eng.gravity_linklist(E.FP64)
        ok, _ = orc.sort_pairs(*orc.calc_keys(p, pos))
        s, e = eng.cell_table()
        os_, oe = orc.cell_table(ok, M ** 3)
        gaps = np.diff(np.concatenate([[-1], np.unique(ok).astype(np.int64), [M ** 3]])) - 1
        assert gaps.max() > 65536 and gaps[0] > 2048 and gaps[-1] > 2048
        assert np.array_equal(s, os_) and np.array_equal(e, oe), centres
    eng.close()
