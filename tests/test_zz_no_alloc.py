"""godar_get_totals_noalloc (the totals through a fixed per-GPU array, no
device allocation) against godar_get_totals, through the C ABI: two calls in
a row, the second on a smaller field, so that the second must not see what
the first left in the shared array."""
import ctypes

import numpy as np
import pytest

from godar_amd import _lib
from godar_amd import generators as g
from workloads import config0_params, kick, packed_box, squeeze

pytestmark = pytest.mark.gpu


def _both(ctx):
    L = _lib.load()
    a = np.zeros(8, dtype=np.int64)
    b = np.zeros(8, dtype=np.int64)
    lp = ctypes.POINTER(ctypes.c_int64)
    _lib.check(ctx.h, L.godar_get_totals(ctx.h, a.ctypes.data_as(lp)))
    _lib.check(ctx.h, L.godar_get_totals_noalloc(ctx.h, b.ctypes.data_as(lp)))
    return a, b


def test_totals_noalloc_matches_allocating_entry_point():
    from godar_amd.sim import Context
    big = g.config0(0)
    small = kick(squeeze(packed_box(64, 3)), 1)
    c1 = Context(big.n, 1, "fp64")
    c1.add_field(config0_params(big), big)
    c1.step(20)
    a, b = _both(c1)
    print("[noalloc] 1,000 floes:", a.tolist(), flush=True)
    assert a[0] > 1000 and a[2] > 100
    assert np.array_equal(a, b)
    c2 = Context(small.n, 1, "mixed")
    c2.add_field(config0_params(small, sigmanc_crit=1e5), small)
    c2.step(5)
    a2, b2 = _both(c2)
    print("[noalloc] 64 floes:", a2.tolist(), flush=True)
    assert a2[0] < a[0] and a2[1] > 0 and a2[6] > 0      # the dense box ridges
    assert np.array_equal(a2, b2)
    # the larger context again, after the smaller one used the array
    a3, b3 = _both(c1)
    assert np.array_equal(a3, a) and np.array_equal(b3, a)
    c1.close()
    c2.close()
