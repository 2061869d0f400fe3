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
