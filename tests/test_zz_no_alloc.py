"""cuco_radix_sort_noalloc (the allocation-free variant of cuco_radix_sort,
for callers that cannot allocate device memory) against cuco_radix_sort
through the C ABI: identical keys and values on two calls in a row, the
second with a smaller input (the static scratch keeps the first call's
contents), with an odd and an even number of 8-bit passes and keys only;
more keys than its maximum is an error, not a fault."""
import ctypes as C

import numpy as np
import pytest

from cuco_amd import engine as E

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 4194304


def sort_with(fn_name, keys, vals, key_bits):
    n = keys.shape[0]
    kb = E.DeviceBuffer(max(4, keys.nbytes))
    kb.upload(keys)
    vb = None
    if vals is not None:
        vb = E.DeviceBuffer(max(4, vals.nbytes))
        vb.upload(vals)
    fn = getattr(E.lib(), fn_name)
    rc = fn(kb.ptr, vb.ptr if vb else None, C.c_uint32(n), C.c_uint32(key_bits))
    if rc != 0:
        raise E.CucoError(rc, fn_name, E.lib().cuco_last_error().decode())
    ko = kb.download(np.empty(n, np.uint32))
    vo = vb.download(np.empty(n, np.uint32)) if vb is not None else None
    kb.free()
    if vb is not None:
        vb.free()
    return ko, vo


@pytest.mark.parametrize("key_bits", [27, 24])
def test_noalloc_sort_matches_allocating_sort(key_bits):
    rng = np.random.default_rng(key_bits)
    for n in (NOALLOC_MAX, 123457):        # second call smaller
        print(f"noalloc sort: n={n} key_bits={key_bits}", flush=True)
        keys = rng.integers(0, 1 << key_bits, n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 10] = keys[0]           # many ties
        vals = rng.permutation(n).astype(np.uint32)
        ka, va = sort_with("cuco_radix_sort", keys, vals, key_bits)
        kn, vn = sort_with("cuco_radix_sort_noalloc", keys, vals, key_bits)
        assert np.array_equal(kn, ka) and np.array_equal(vn, va)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(kn, keys[order]) and np.array_equal(vn, vals[order])
        ko, _ = sort_with("cuco_radix_sort_noalloc", keys, None, key_bits)
        assert np.array_equal(ko, ka)


def test_noalloc_sort_rejects_more_than_its_maximum():
    n = NOALLOC_MAX + 1
    keys = np.zeros(n, np.uint32)
    with pytest.raises(E.CucoError) as ei:
        sort_with("cuco_radix_sort_noalloc", keys, None, 8)
    assert ei.value.status == 1
    # still usable afterwards
    k, _ = sort_with("cuco_radix_sort_noalloc", np.array([3, 1, 2], np.uint32), None, 8)
    assert list(k) == [1, 2, 3]
