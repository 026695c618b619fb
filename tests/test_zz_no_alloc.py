"""who_owns without allocation: xkrt_hip_mcc_who_owns_static (staging in the
library's static device array) gives exactly xkrt_hip_mcc_who_owns's answers,
on two calls in a row (the second smaller), on two trees sharing the array."""
import numpy as np
import pytest

from xkaapi_amd import coherency as K
from xkaapi_amd import streams as S
from xkaapi_amd.resolver import ResolverError

pytestmark = pytest.mark.gpu

E2BIG, EINVAL = -7, -22


def queries(n, ld, cols, seed):
    """n query tiles with log-uniform extents in [1, 2048], clipped to the frame"""
    rng = np.random.default_rng(seed)
    row = rng.integers(0, ld, n).astype(np.uint64)
    m = np.minimum(np.exp2(rng.random(n) * 11).astype(np.uint64), ld - row)
    col = rng.integers(0, cols, n).astype(np.uint64)
    nn = np.minimum(np.exp2(rng.random(n) * 11).astype(np.uint64), cols - col)
    return (col * ld + row) * 8, m, nn


def planned_tree(n_stream, ld, s):
    st = S.coherency(n_stream)
    sub = st.take(np.flatnonzero((st.ld == ld) & (st.s == s)))
    t = K.MemoryTree(ld, s, (0, 8192))
    _, _, fail = t.plan_counts(K.to_mcc_array(sub.dev, sub.mode, sub.addr, sub.m, sub.n))
    assert fail < 0
    return t


def test_who_owns_static_equals_who_owns():
    with planned_tree(3000, 8192, 8) as t1, planned_tree(2000, 16384, 4) as t2:
        for t, ld, s in ((t1, 8192, 8), (t2, 16384, 4)):
            for n, seed in ((20000, 1), (700, 2)):  # two calls in a row, the second smaller
                addr, m, nn = queries(n, ld, 8192, seed)
                addr = addr // 8 * s
                want = t.who_owns(addr, m, nn)
                got = t.who_owns_static(addr, m, nn)
                assert np.array_equal(got, want), (ld, n)
                assert len(set(want.tolist())) > 1  # host, GPU and shared owners all appear
        # the largest call the array takes, then a refused one (checked on the host)
        addr, m, nn = queries(K.MemoryTree.WHO_OWNS_STATIC_MAX, 8192, 8192, 3)
        assert np.array_equal(t1.who_owns_static(addr, m, nn), t1.who_owns(addr, m, nn))
        addr, m, nn = queries(K.MemoryTree.WHO_OWNS_STATIC_MAX + 1, 8192, 8192, 4)
        with pytest.raises(ResolverError) as ei:
            t1.who_owns_static(addr, m, nn)
        assert ei.value.rc == E2BIG
        with pytest.raises(ResolverError) as ei:
            t1.who_owns_static(np.array([8], np.uint64), np.zeros(1, np.uint64), np.ones(1, np.uint64))
        assert ei.value.rc == EINVAL
