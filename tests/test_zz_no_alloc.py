"""fe_sort_pairs_noalloc (no reference counterpart; for callers that cannot allocate device
memory) against the existing entry point fe_sort, through the C ABI.

Its result must equal fe_sort's bit for bit on the same pairs (both are the stable radix sort
on all 32 bits), and the stable argsort of the keys. Two calls in a row share the one device
array, the second with a smaller input, so stale contents of the first call must not leak
into the second. The device buffer and the stream are made with the HIP runtime the library
itself links (libamdhip64), called through ctypes: plumbing only.
"""
import ctypes

import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu


def _fe_sort(pairs):
    """fe_sort on a context whose GRID_ENTRIES holds `pairs` (T = len(pairs), a power of two)."""
    n = pairs.shape[0]
    params = scene.make_params(0.01, 0.008, 0.04, 1000, 0.01, 0.5, 1e-3, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    sim = fe.Simulation(params, np.zeros((n, 3), np.float32), device=0, options=o)
    assert sim.table_size == n
    sim.upload("grid_entries", pairs)
    sim.sort()
    out = sim.download("grid_entries")
    sim.close()
    return out


class Hip:
    """The few HIP runtime calls the test needs, from the runtime the library is linked to."""

    def __init__(self):
        fe.load_library()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.stream = ctypes.c_void_p()
        self.ok(self.h.hipSetDevice(0), "hipSetDevice")
        self.ok(self.h.hipStreamCreate(ctypes.byref(self.stream)), "hipStreamCreate")

    @staticmethod
    def ok(rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def sort(self, lib, pairs):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
        dev = ctypes.c_void_p()
        self.ok(self.h.hipMalloc(ctypes.byref(dev), ctypes.c_size_t(pairs.nbytes)), "hipMalloc")
        try:
            self.ok(self.h.hipMemcpy(dev, pairs.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(pairs.nbytes), 1),
                    "hipMemcpy H2D")
            fe.check(lib.fe_sort_pairs_noalloc(0, dev, pairs.shape[0], self.stream), "fe_sort_pairs_noalloc")
            self.ok(self.h.hipStreamSynchronize(self.stream), "hipStreamSynchronize")
            out = np.empty_like(pairs)
            self.ok(self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, ctypes.c_size_t(pairs.nbytes), 2),
                    "hipMemcpy D2H")
        finally:
            self.h.hipFree(dev)
        return out

    def close(self):
        self.h.hipStreamDestroy(self.stream)


@pytest.fixture(scope="module")
def hip():
    h = Hip()
    yield h
    h.close()


def test_noalloc_sort_equals_fe_sort_twice(hip):
    lib = fe.load_library()
    rng = np.random.default_rng(2024)
    for log2n, keyspace in ((22, 2 ** 32), (12, 5000)):      # the second call is smaller
        n = 1 << log2n
        keys = rng.integers(0, keyspace, size=n, dtype=np.uint64).astype(np.uint32)
        pairs = np.stack([keys, np.arange(n, dtype=np.uint32)], axis=1)
        want = _fe_sort(pairs)
        print(f"noalloc sort: n={n} keyspace={keyspace}", flush=True)
        got = hip.sort(lib, pairs)
        assert np.array_equal(got, want)
        perm = np.argsort(keys, kind="stable")
        assert np.array_equal(got[:, 1], perm)
        assert np.array_equal(got[:, 0], keys[perm])


def test_noalloc_sort_ragged_and_limits(hip):
    lib = fe.load_library()
    rng = np.random.default_rng(7)
    for n in (100003, 1, 2, 4097, 2500003):   # the last: 512 x 16-key tiles, a partial one at the end
        keys = rng.integers(0, 1 << 20, size=n, dtype=np.uint64).astype(np.uint32)
        pairs = np.stack([keys, rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)], axis=1)
        got = hip.sort(lib, pairs)
        perm = np.argsort(keys, kind="stable")
        assert np.array_equal(got, pairs[perm]), n
    # beyond the maximum: refused before anything runs
    st = lib.fe_sort_pairs_noalloc(0, ctypes.c_void_p(1), (1 << 22) + 1, None)
    assert st == 1
    assert b"FE_SORT_NOALLOC_MAX_PAIRS" in lib.fe_last_error()
    assert lib.fe_sort_pairs_noalloc(0, None, 5, None) == 1
