"""ll_iterate_noalloc (the static-scratch variant of ll_iterate) against
ll_iterate, through the C-ABI on the GPU: identical per-iteration roundoff
and identical digits, on two calls in a row -- the second on a smaller
input -- and the size limit reported as an argument error."""
import numpy as np
import pytest

import cudalucas_amd as L
from cudalucas_amd import workloads as W
from conftest import announce

pytestmark = pytest.mark.gpu


def _pair(p, n, iters, offset, **kw):
    """(digits, offset, per-iteration errors) of ll_iterate_noalloc and of
    ll_iterate on two contexts loaded with the same state."""
    out = []
    for noalloc in (True, False):
        with L.LLContext(p, n, **kw) as c:
            c.load_initial(offset)
            c.iterate(3)  # leave the low-entropy start
            if noalloc:
                m, e = c.iterate_noalloc(iters, per_iter=True)
            else:
                m, e = c.iterate(iters, per_iter=True)
            d, off = c.read_digits()
            out.append((d, off, e, m))
    return out


@pytest.mark.parametrize("p,n,iters,kw", [
    (W.LARGE_P, W.LARGE_N, 300, {}),                                   # 4M words, two sweeps
    (2553659, 131072, 50, dict(flags=L.LL_OPT_FORCE_SWEEPS)),          # then a smaller input
])
def test_noalloc_matches_iterate(p, n, iters, kw):
    announce("noalloc vs iterate p=%d n=%d iters=%d" % (p, n, iters))
    (d1, o1, e1, m1), (d2, o2, e2, m2) = _pair(p, n, iters, 17, **kw)
    assert o1 == o2
    assert (d1 == d2).all()
    assert (e1 == e2).all() and m1 == m2
    assert m1 < 0.35


def test_noalloc_two_calls_in_a_row_then_smaller():
    """Two calls on one context, then a call on a smaller context, each
    against ll_iterate on a twin; the shared scratch keeps the previous
    call's carries, which must not leak into the next."""
    cases = [(W.LARGE_P, W.LARGE_N), (W.LARGE_P, W.LARGE_N), (1257787, 65536)]
    announce("noalloc consecutive calls")
    with L.LLContext(*cases[0]) as a, L.LLContext(*cases[0]) as b:
        a.load_initial(5)
        b.load_initial(5)
        for k in (40, 7):
            ma, ea = a.iterate_noalloc(k, per_iter=True)
            mb, eb = b.iterate(k, per_iter=True)
            assert (ea == eb).all()
            assert (a.read_digits()[0] == b.read_digits()[0]).all()
    p, n = cases[2]
    with L.LLContext(p, n, flags=L.LL_OPT_FORCE_SWEEPS) as a, \
            L.LLContext(p, n, flags=L.LL_OPT_FORCE_SWEEPS) as b:
        a.load_initial(9)
        b.load_initial(9)
        ma, ea = a.iterate_noalloc(33, per_iter=True)
        mb, eb = b.iterate(33, per_iter=True)
        assert (ea == eb).all()
        da, oa = a.read_digits()
        db, ob = b.read_digits()
        assert oa == ob and (da == db).all()


def test_noalloc_resident_path():
    p, n = W.SMALL_P, W.SMALL_N
    announce("noalloc resident p=%d n=%d" % (p, n))
    (d1, o1, e1, m1), (d2, o2, e2, m2) = _pair(p, n, 2000, 3)
    assert o1 == o2 and (d1 == d2).all() and (e1 == e2).all()


def test_noalloc_rejects_a_shuttle_beyond_the_scratch():
    """8M words with two-line carry blocks: (2048 / 2) * 2048 * 8 bytes of
    shuttle = 16 MiB, more than the scratch holds: an argument error, no
    kernel launched."""
    announce("noalloc size limit")
    with L.LLContext(149447533, 8388608, width=2048, height=2048, lines=2,
                     flags=L.LL_OPT_FORCE_SWEEPS) as c:
        c.load_initial(0)
        with pytest.raises(L.LLError) as ex:
            c.iterate_noalloc(1)
        assert ex.value.code == L.LL_ERR_ARG
        assert "no-allocation scratch" in str(ex.value)
        c.iterate(2)  # the context is still usable
