"""bg_spatial_hash_query_noalloc (no device allocation; scratch in a
file-scope device array) against bg_spatial_hash_query through the C ABI:
identical offsets and hits, byte for byte, on two calls in a row -- the
second with fewer points and queries, so it runs on the array the first
call left behind -- and sizes beyond the stated maximum rejected before
anything runs."""
import ctypes as C

import numpy as np
import pytest

from blobs_amd import _lib as L
from blobs_amd.spatial import HIT, QUERY, SpatialHash
from tests.devbuf import DevBuf

pytestmark = pytest.mark.gpu


def _case(n, nq, seed):
    rng = np.random.default_rng(seed)
    side = np.sqrt(n) * 1.2
    h = SpatialHash(1.0)
    for p, r in zip(rng.uniform(-side, side, (n, 2)).astype(np.float32), rng.uniform(0, 0.3, n)):
        h.insert(p, r)
    q = np.zeros(nq, QUERY)
    q["position"] = rng.uniform(-side - 1, side + 1, (nq, 2))
    q["radius"] = rng.uniform(0, 0.8, nq)
    return h.memberships(), q


def _call(lib, fn, pts, qs, cap):
    dp, dq = DevBuf(lib, pts.nbytes, pts), DevBuf(lib, qs.nbytes, qs)
    off, hits = DevBuf(lib, 4 * (qs.shape[0] + 1)), DevBuf(lib, cap * HIT.itemsize)
    total = C.c_uint64()
    code = fn(1.0, dp.p, pts.shape[0], dq.p, qs.shape[0], off.p, hits.p, cap, C.byref(total))
    return code, off.get(np.uint32, qs.shape[0] + 1), hits.get(HIT, min(total.value, cap)), total.value


def test_no_alloc_query_equals_the_allocating_one_on_two_calls():
    lib = L.load()
    for n, nq, seed in ((60000, 30000, 1), (2500, 900, 2)):
        pts, qs = _case(n, nq, seed)
        ca, oa, ha, ta = _call(lib, lib.bg_spatial_hash_query, pts, qs, 1 << 18)
        L.check(ca, None)
        cb, ob, hb, tb = _call(lib, lib.bg_spatial_hash_query_noalloc, pts, qs, 1 << 18)
        L.check(cb, None)
        assert ta == tb > 0
        assert np.array_equal(oa, ob), f"offsets differ (n={n})"
        assert np.array_equal(ha.view(np.uint8), hb.view(np.uint8)), f"hits differ (n={n})"


def test_no_alloc_query_rejects_sizes_beyond_its_maximum():
    lib = L.load()
    buf, off = DevBuf(lib, 64), DevBuf(lib, 16)
    total = C.c_uint64()
    for n_points, n_queries in (((1 << 20) + 1, 1), (1, (1 << 20) + 1)):
        code = lib.bg_spatial_hash_query_noalloc(1.0, buf.p, n_points, buf.p, n_queries, off.p, buf.p, 1,
                                                 C.byref(total))
        assert code == L.BG_ERR_CAPACITY
