"""Argument and failure handling of the C ABI on a GPU: capacity changes that
would cut live records, allocations that fail, checkpoint blocks that do not
fit the context, and parameters the step does not implement.  Each case
checks that the call is refused with the right code and that the context
goes on exactly as a twin that never saw the call."""
import ctypes
import re

import numpy as np
import pytest

from godar_amd import _lib
from godar_amd.params import GodarParams
from workloads import config0_params, kick, packed_box, squeeze

pytestmark = pytest.mark.gpu

STATE = ["x", "y", "u", "v", "omega", "r", "h", "theta"]


def _ctx(n, members=1, precision="fp64"):
    from godar_amd.sim import Context
    return Context(n, members, precision, record_forces=True)


def _box():
    f = kick(squeeze(packed_box(64, 3)), 1)
    return f, config0_params(f, sigmanc_crit=1e5)


def _same(a, b, ma, mb):
    fa, fb = a.floes(ma), b.floes(mb)
    for k in STATE:
        assert np.array_equal(fa[k], fb[k]), k
    assert a.pairs(ma, "contact")[0].tolist() == b.pairs(mb, "contact")[0].tolist()


def test_reserve_refuses_to_cut_live_records():
    """godar_reserve may not shrink the history list below the records a
    floe holds (the kernels walk every record up to the floe's count); a
    shrink that keeps them is applied, and the run stays bitwise equal to
    one that never changed its capacities."""
    f, p = _box()
    a, b = _ctx(f.n), _ctx(f.n)
    ma, mb = a.add_field(p, f), b.add_field(p, f)
    a.step(30)
    b.step(30)
    with pytest.raises(_lib.GodarError) as e:
        a.reserve(hist=1)
    assert e.value.code == _lib.GODAR_ERR_ARG
    live = int(re.search(r"histories (\d+)", str(e.value)).group(1))
    assert live > 1
    a.reserve(hist=live)          # down to the longest list: allowed
    a.step(25)                    # later steps may need more; capacity grows on its own
    b.step(25)
    _same(a, b, ma, mb)
    a.close()
    b.close()


def test_failed_reserve_keeps_the_context():
    """An allocation that fails (an outer list of 65,534 entries for 2^21
    floe slots needs about a terabyte) returns GODAR_ERR_NOMEM and leaves the
    context with its lists: the next steps run and equal a twin's."""
    f, p = _box()
    a, b = _ctx(1 << 21), _ctx(f.n)
    ma, mb = a.add_field(p, f), b.add_field(p, f)
    a.step(5)
    b.step(5)
    with pytest.raises(_lib.GodarError) as e:
        a.reserve(outer=65534)
    assert e.value.code == _lib.GODAR_ERR_NOMEM
    a.step(20)
    b.step(20)
    _same(a, b, ma, mb)
    a.close()
    b.close()


def test_member_add_refuses_stale_tree_cadence():
    """ntree != 1 (a neighbour tree of positions older than the step,
    stepper.f90:44-50) is refused rather than run with current positions."""
    f, p = _box()
    p.ntree = 5
    a = _ctx(f.n)
    with pytest.raises(_lib.GodarError) as e:
        a.add_field(p, f)
    assert e.value.code == _lib.GODAR_ERR_ARG
    assert "ntree" in str(e.value)
    p.ntree = 1
    a.add_field(p, f)
    a.close()


def _hist_offset(block):
    """Byte offset of the first contact-history record of a checkpoint block
    (header 40 B, parameters, 12 + 4 doubles per floe, n history counts)."""
    n = int(np.frombuffer(block[12:16], dtype=np.int32)[0])
    base = 40 + ctypes.sizeof(GodarParams)
    return base + n * (12 + 4) * 8 + n * 4, n, base


def test_restore_refuses_foreign_blocks():
    """A checkpoint whose history names a partner outside (i, n), or that was
    saved in a context of the other precision, is refused; the intact block
    restores and continues exactly."""
    f, p = _box()
    a = _ctx(f.n)
    m = a.add_field(p, f)
    a.step(30)
    block = a.checkpoint(m, p)
    a.step(10)
    off, n, base = _hist_offset(block)
    hc = np.frombuffer(block[off - 4 * n:off], dtype=np.int32)
    assert hc.sum() > 0
    for bad_j in (0, n, -1):
        bad = bytearray(block)
        bad[off:off + 4] = np.int32(bad_j).tobytes()
        b = _ctx(f.n)
        with pytest.raises(_lib.GodarError) as e:
            b.restore(bytes(bad))
        assert e.value.code == _lib.GODAR_ERR_ARG, bad_j
        b.close()
    mixed = _ctx(f.n, 1, "mixed")
    with pytest.raises(_lib.GodarError) as e:
        mixed.restore(block)
    assert e.value.code == _lib.GODAR_ERR_ARG
    assert "precision" in str(e.value)
    mixed.close()
    b = _ctx(f.n)
    mb = b.restore(block)
    b.step(10)
    _same(a, b, m, mb)
    a.close()
    b.close()
