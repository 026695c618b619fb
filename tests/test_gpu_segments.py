"""Segment (1-D) accesses seen by a BLAS domain on the GPU (through the C ABI):
the golden fixture from the reference engine, random mixed streams against
the oracle, and the refusals where the reference marks whole nodes."""
import os

import numpy as np
import pytest

from oracle import khp
from oracle import segments as G
from xkaapi_amd import resolver as R

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gpu_edges(st, ld, s, tpb, **kw):
    acc = R.to_access_array(st["task"], st["mode"], st["addr"], st["m"], st["n"], kind=st["kind"])
    with R.DependencyDomain(ld, s, **kw) as dom:
        first = None
        for a, b in R.batches_by_tasks(st["task"], tpb):
            f = dom.submit(acc[a:b])
            first = f if first is None else first
        e = dom.edges().astype(np.int64) - (first or 0)
    return np.unique(e, axis=0) if len(e) else e.reshape(0, 2)


def test_golden_fixture():
    z = np.load(os.path.join(GOLDEN, "segments_mixed.npz"))
    st = {k: z[k] for k in ("task", "mode", "kind", "addr", "m", "n")}
    for bx, by, tpb, kc in ((32, 4, 50, 2), (16, 2, 7, 2), (64, 8, 2000, 2)):
        got = gpu_edges(st, int(z["ld"]), int(z["s"]), tpb, bin_x_bytes=bx, bin_y_cols=by, kernel=kc)
        assert np.array_equal(got, z["edges"]), (bx, by, tpb)


@pytest.mark.parametrize("seed", range(3))
def test_random_mixed_matches_oracle(seed):
    ld, cols, s = 16, 30, 2
    st = G.mixed_stream(1500, ld, cols, s, 40 + seed, pseg=0.4, maxm=14, maxn=6)
    rects, nrect, marks, nmark = G.stream_rects(st["kind"], st["addr"], st["m"], st["n"], ld, s)
    want = np.unique(khp.resolve_marks(st["task"], st["mode"], rects, nrect, marks, nmark).astype(np.int64), axis=0)
    for bx, by, tpb, kc in ((32, 8, 100, 1), (16, 4, 13, 2)):
        assert np.array_equal(gpu_edges(st, ld, s, tpb, bin_x_bytes=bx, bin_y_cols=by, kernel=kc), want)


def test_segment_refused_once_column_zero_is_touched():
    ld, s = 8, 1
    tile0 = R.to_access_array(np.array([0], np.uint32), np.array([2], np.uint8), np.array([0], np.uint64),
                              np.array([8], np.uint64), np.array([2], np.uint64))
    seg = R.to_access_array(np.array([1], np.uint32), np.array([1], np.uint8), np.array([42], np.uint64),
                            np.array([3], np.uint64), np.array([1], np.uint64), kind=np.array([1], np.uint8))
    with R.DependencyDomain(ld, s, bin_x_bytes=8, bin_y_cols=4) as dom:
        dom.submit(tile0)
        with pytest.raises(R.ResolverError) as ei:
            dom.submit(seg)
        assert ei.value.rc == -95  # -ENOTSUP
    low = R.to_access_array(np.array([0], np.uint32), np.array([1], np.uint8), np.array([3], np.uint64),
                            np.array([4], np.uint64), np.array([1], np.uint64), kind=np.array([1], np.uint8))
    with R.DependencyDomain(ld, s, bin_x_bytes=8, bin_y_cols=4) as dom:
        with pytest.raises(R.ResolverError) as ei:
            dom.submit(low)
        assert ei.value.rc == -95
