"""GPU unit tests of the broadphase primitives (scans, stable LSD radix sort,
the square-root-free contact test) against numpy, through the library's
bgx_ test hooks."""
import ctypes as C

import numpy as np
import pytest

from blobs_amd import _lib as L

pytestmark = pytest.mark.gpu


def _hooks():
    lib = L.load()
    u32 = C.POINTER(C.c_uint32)
    lib.bgx_radix_sort_u32.argtypes = [u32, u32, C.c_uint32, C.c_int, u32, u32]
    lib.bgx_exclusive_scan_u32.argtypes = [u32, C.c_uint32, u32]
    lib.bgx_exclusive_scan_u32_segments.argtypes = [u32, C.c_uint32, u32, u32, C.c_uint32, u32]
    f32 = C.POINTER(C.c_float)
    lib.bgx_sqrt_lt.argtypes = [f32, f32, C.c_uint32, C.POINTER(C.c_uint8)]
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def gpu_sort(keys, vals, bits):
    lib = _hooks()
    ko = np.zeros_like(keys)
    vo = np.zeros_like(vals)
    assert lib.bgx_radix_sort_u32(_p(keys), _p(vals), keys.size, bits, _p(ko), _p(vo)) == 0
    return ko, vo


@pytest.mark.parametrize("n", [1, 2, 100, 4095, 4096, 4097, 65536 + 13, 1 << 20])
@pytest.mark.parametrize("bits", [8, 13, 22])
def test_radix_sort_random_is_stable_sort(n, bits):
    rng = np.random.default_rng(n * 31 + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    ko, vo = gpu_sort(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ko, keys[order])
    assert np.array_equal(vo, vals[order])


@pytest.mark.parametrize("kind", ["sorted", "nearly_sorted", "reversed", "all_equal", "few_distinct"])
def test_radix_sort_structured_inputs(kind):
    n = 1 << 20
    rng = np.random.default_rng(3)
    if kind == "sorted":
        keys = np.sort(rng.integers(0, 1 << 21, n)).astype(np.uint32)
    elif kind == "nearly_sorted":
        keys = np.sort(rng.integers(0, 1 << 21, n)).astype(np.uint32)
        idx = rng.integers(0, n - 8, n // 50)
        keys[idx], keys[idx + 7] = keys[idx + 7].copy(), keys[idx].copy()
    elif kind == "reversed":
        keys = np.sort(rng.integers(0, 1 << 21, n))[::-1].astype(np.uint32).copy()
    elif kind == "all_equal":
        keys = np.full(n, 12345, np.uint32)
    else:
        keys = rng.integers(0, 3, n).astype(np.uint32) * 700000
    vals = np.arange(n, dtype=np.uint32)
    ko, vo = gpu_sort(keys, vals, 22)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ko, keys[order])
    assert np.array_equal(vo, vals[order])


@pytest.mark.parametrize("n", [1, 17, 4096, 4097, 300000, 5_000_000])
def test_exclusive_scan(n):
    lib = _hooks()
    rng = np.random.default_rng(n)
    x = rng.integers(0, 100, n).astype(np.uint32)
    out = np.zeros_like(x)
    assert lib.bgx_exclusive_scan_u32(_p(x), n, _p(out)) == 0
    ref = np.concatenate([[0], np.cumsum(x, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("seed", [0, 1])
def test_segmented_scan_matches_numpy_on_used_cells(seed):
    """A batch-shaped cell table: segments of 8 cells per body (at least
    8,192) of which each uses a random prefix; unused cells hold zero counts
    and keep whatever the output held."""
    lib = _hooks()
    rng = np.random.default_rng(seed)
    sizes = rng.integers(64, 5000, 300)
    caps = np.maximum(8 * sizes, 8192)
    base = np.concatenate([[0], np.cumsum(caps)]).astype(np.uint32)
    used = np.array([rng.integers(1, c) for c in caps], np.uint32)
    used[::7] = caps[::7]  # some use every cell: their end sentinel is the next segment's first cell
    n = int(base[-1]) + 1
    x = np.zeros(n, np.uint32)
    for w in range(len(caps)):
        x[base[w]:base[w] + used[w]] = rng.integers(0, 4, used[w])
    out = np.full(n, 0xDEADBEEF, np.uint32)
    assert lib.bgx_exclusive_scan_u32_segments(_p(x), n, _p(base), _p(used), len(caps), _p(out)) == 0
    ref = np.concatenate([[0], np.cumsum(x, dtype=np.uint64)[:-1]]).astype(np.uint32)
    for w in range(len(caps)):
        lo, hi = int(base[w]), int(base[w]) + int(used[w]) + 1
        assert np.array_equal(out[lo:hi], ref[lo:hi]), w


def test_sqrt_free_contact_test_equals_sqrtf_on_the_device():
    from tests.sqrt_cases import cases, reference
    lib = _hooks()
    d2, r = cases()
    out = np.zeros(d2.size, np.uint8)
    f32 = C.POINTER(C.c_float)
    assert lib.bgx_sqrt_lt(d2.ctypes.data_as(f32), r.ctypes.data_as(f32), d2.size,
                           out.ctypes.data_as(C.POINTER(C.c_uint8))) == 0
    ref = reference(d2, r)
    dev_sqrt = (out & 1) != 0
    dev_mid = (out & 2) != 0
    assert np.array_equal(dev_sqrt, ref), np.flatnonzero(dev_sqrt != ref)[:5]
    assert np.array_equal(dev_mid, ref), np.flatnonzero(dev_mid != ref)[:5]
