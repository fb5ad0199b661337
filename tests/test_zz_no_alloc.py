"""sph_sort_by_key_bits_noalloc (scratch in a file-scope device array, no
allocation) through the C ABI against sph_sort_by_key_bits, the entry point
with allocated scratch, and numpy's stable argsort: bit-identical on two
calls in a row, the second with a smaller input, and a size beyond the
array's limit rejected without work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")

MAX_PAIRS = 1 << 21  # SPH_SORT_NOALLOC_MAX_PAIRS


def _sort(entry, keys, vals, bits):
    kb, vb = sa.DeviceBuffer.from_array(keys), sa.DeviceBuffer.from_array(vals)
    sa.check(entry(kb.ptr, vb.ptr, len(keys), bits, None))
    return kb.to_array(np.uint32, len(keys)), vb.to_array(np.uint32, len(keys))


def test_noalloc_sort_matches_allocating_sort():
    L = sa.lib()
    rng = np.random.default_rng(2026)
    # 23-bit keys: three passes (an odd count ends in the scratch array and is
    # copied back); then 16-bit keys: two passes, a smaller input over the
    # contents the first call left in the array; then the largest size
    for n, bits in ((2_000_000, 23), (100_003, 16), (MAX_PAIRS, 32)):
        hi = 2 ** bits
        keys = rng.integers(0, min(hi, 4096 if bits == 16 else hi), n,
                            dtype=np.uint64).astype(np.uint32)
        vals = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        print(f"noalloc sort n={n} key_bits={bits}", flush=True)
        k0, v0 = _sort(L.sph_sort_by_key_bits_noalloc, keys, vals, bits)
        k1, v1 = _sort(L.sph_sort_by_key_bits, keys, vals, bits)
        assert np.array_equal(k0, k1) and np.array_equal(v0, v1)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k0, keys[order]) and np.array_equal(v0, vals[order])


def test_noalloc_sort_rejects_sizes_beyond_its_array():
    L = sa.lib()
    keys = sa.DeviceBuffer(4 * 8)
    rc = L.sph_sort_by_key_bits_noalloc(keys.ptr, keys.ptr, MAX_PAIRS + 1, 8, None)
    assert rc == -1  # SPH_ERR_ARG
    assert b"SPH_SORT_NOALLOC_MAX_PAIRS" in L.sph_last_error()
    assert L.sph_sort_by_key_bits_noalloc(keys.ptr, keys.ptr, 4, 0, None) == -1
