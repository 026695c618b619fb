"""Writes the golden fixtures under tests/golden/ from the reference's own KHP
engine (oracle/_ref/libkhpref.so, built by `make -C oracle` from
/root/reference).  Run in the build container:  python tests/golden/make_golden.py

Fixtures (npz, allow_pickle=False):
  cholesky_t2_edges.npz      T = 2 Cholesky (the closed-form 3-edge DAG)
  cholesky_4096_nb256.npz    tiled Cholesky 4096^2 fp64, 256^2 tiles (816 tasks)
  lu_1024_nb128.npz          tiled LU 1024^2 fp64, 128^2 tiles
  heat_1024_t128_s4.npz      heat halo stream 1024^2 fp32, 128^2 tiles, 4 steps
  random2d_4000.npz          first 4000 random tasks over 65536^2 (4 B elements)
  wrap_mixed.npz             wrapped 2-rect tiles, every access mode, seeded
  partition_coherency.npz    memory-tree partition geometry of 500 coherency accesses (view 0)
  segments_mixed.npz         tiles and byte segments (at or above the first column) in one BLAS
                             domain: (pred, succ) access-index edges, unique sorted rows
  intervals_k1.npz           byte segments at 64-bit addresses in one segment domain (the
                             reference's KHPTree<1> with the interval-tree hooks): (pred, succ) rows
  program_mixed.npz          one task domain over segments and tiles of three (ld, s) pairs (the
                             reference's dispatch over its own trees): (pred, succ) rows
Run with fixture names to write only those.
Each resolver fixture holds the stream (task, slot, mode, addr, ld, m, n, s),
the edges as (pred_task, pred_slot, succ_task, succ_slot) sorted rows, and
the per-access fragment counts F_i.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import refkhp  # noqa: E402
from oracle import segments as G  # noqa: E402
from xkaapi_amd import streams as S  # noqa: E402


def wrap_stream(n=1500, ld=16, cols=40, s=2, maxm=12, maxn=6, seed=7):
    rng = np.random.default_rng(seed)
    task = (np.arange(n) // 2).astype(np.uint32)
    slot = (np.arange(n) % 2).astype(np.uint16)
    mode = rng.choice(np.array([1, 2, 3, 4, 5, 6, 8, 9], np.uint8), n)
    m = rng.integers(1, maxm + 1, n).astype(np.uint64)
    nn = rng.integers(1, maxn + 1, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    col = rng.integers(0, cols, n).astype(np.uint64)
    return S.Stream(task, slot, mode, np.ones(n, np.uint8), (col * ld + row) * s, np.full(n, ld, np.uint64),
                    m, nn, np.full(n, s, np.uint64), name="wrap")


def edge_rows(st, e):
    rows = np.stack([st.task[e[:, 0]].astype(np.int64), st.slot[e[:, 0]].astype(np.int64),
                     st.task[e[:, 1]].astype(np.int64), st.slot[e[:, 1]].astype(np.int64)], axis=1)
    return rows[np.lexsort(rows.T[::-1])]


def save_resolver(name, st):
    r, nr = st.rects()
    e, f = refkhp.resolve(st.task, st.mode, r, nr, want_frags=True)
    np.savez_compressed(os.path.join(HERE, name), task=st.task, slot=st.slot, mode=st.mode, addr=st.addr,
                        ld=st.ld, m=st.m, n=st.n, s=st.s, edges=edge_rows(st, e), frag=f.astype(np.int64))
    print(name, len(st), "accesses", len(e), "edges")


def intervals_stream(n=3000, space=1 << 20, maxsz=1 << 14, base=0x7F12_3456_0000, seed=23):
    rng = np.random.default_rng(seed)
    task = np.repeat(np.arange(n, dtype=np.uint32), rng.integers(1, 4, n))[:n]
    mode = rng.choice(np.array([1, 2, 3, 4, 5, 6, 8, 9], np.uint8), n)
    a = rng.integers(0, space, n).astype(np.uint64)
    b = np.minimum(a + np.exp2(rng.random(n) * np.log2(maxsz)).astype(np.uint64) + np.uint64(1), np.uint64(space))
    return task, mode, np.stack([a, b], axis=1) + np.uint64(base)


def save_intervals():
    task, mode, seg = intervals_stream()
    e = refkhp.resolve_intervals(task, mode, seg)
    np.savez_compressed(os.path.join(HERE, "intervals_k1.npz"), task=task, mode=mode, seg=seg,
                        edges=np.unique(e.astype(np.int64), axis=0))
    print("intervals_k1.npz", len(task), "accesses", len(e), "edges")


def save_program():
    p = G.program_stream(3000, 21)
    e = refkhp.resolve_program(*[p[k] for k in ("task", "mode", "kind", "addr", "m", "n", "ld", "s")])
    np.savez_compressed(os.path.join(HERE, "program_mixed.npz"), **p, edges=np.unique(e.astype(np.int64), axis=0))
    print("program_mixed.npz", len(p["task"]), "accesses", len(e), "edges")


def main(names=()):
    if not refkhp.available():
        raise SystemExit("oracle/_ref/libkhpref.so missing: run make -C oracle")
    if names:
        for nm in names:
            {"intervals_k1.npz": save_intervals, "program_mixed.npz": save_program}[nm]()
        return
    save_intervals()
    save_program()
    save_resolver("cholesky_t2_edges.npz", S.cholesky(512, 256))
    save_resolver("cholesky_4096_nb256.npz", S.cholesky(4096, 256))
    save_resolver("lu_1024_nb128.npz", S.lu(1024, 128))
    save_resolver("heat_1024_t128_s4.npz", S.heat(1024, 128, 4, 4))
    save_resolver("random2d_4000.npz", S.random2d(4000))
    save_resolver("wrap_mixed.npz", wrap_stream())
    st = S.coherency(3000)
    sel = np.flatnonzero(st.ld == 8192)[:500]
    sub = st.take(sel)
    r, nr = sub.rects()
    off, parts = refkhp.partition(sub.dev, r, nr)
    np.savez_compressed(os.path.join(HERE, "partition_coherency.npz"), dev=sub.dev, mode=sub.mode, addr=sub.addr,
                        ld=sub.ld, m=sub.m, n=sub.n, s=sub.s, part_off=off, parts=parts.astype(np.int64))
    print("partition_coherency.npz", len(sub), "accesses", len(parts), "partites")
    ms = G.mixed_stream(2000, 32, 48, 4, seed=17, pseg=0.35, maxm=24, maxn=8)
    rects, nrect, marks, nmark = G.stream_rects(ms["kind"], ms["addr"], ms["m"], ms["n"], 32, 4)
    e = refkhp.resolve_marks(ms["task"], ms["mode"], rects, nrect, marks, nmark)
    np.savez_compressed(os.path.join(HERE, "segments_mixed.npz"), task=ms["task"], slot=ms["slot"], mode=ms["mode"],
                        kind=ms["kind"], addr=ms["addr"], m=ms["m"], n=ms["n"], ld=np.int64(32), s=np.int64(4),
                        edges=np.unique(e.astype(np.int64), axis=0))
    print("segments_mixed.npz", len(ms["task"]), "accesses", len(e), "edges")


if __name__ == "__main__":
    main(sys.argv[1:])
