"""Writes tests/golden/digests.json: edge-set digests of the full-size streams,
computed by the oracle's C restatement (oracle/liboracle.so) in the build
container, for the GPU parity tests at the workloads' real sizes.

    python tests/golden/make_digests.py [--ref]

Per stream: access count, edge count, and the SHA-256 of the sorted edge keys
(pred << 32 | succ, access indices of the stream, little-endian u64).  For
the random-stream prefix also the per-batch sums of F_i (canonical fragments
overlapped by each access after its insertion) and the per-batch edge counts
(by successor batch), the units of the resolver's algorithmic byte count.
With --ref the reference's own KHP engine (oracle/_ref) resolves the random
prefix too and must give the same digest.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import khp, refkhp  # noqa: E402
from xkaapi_amd import streams as S  # noqa: E402

RANDOM_PREFIX = 102400  # 25 batches of 4096 single-access tasks


def digest(e):
    """(count, sha256) of a (pred, succ) edge array, order-free."""
    e = np.asarray(e, np.uint64).reshape(-1, 2)
    k = np.sort((e[:, 0] << np.uint64(32)) | e[:, 1])
    if len(k) > 1:
        assert (np.diff(k) > 0).all(), "duplicate edges"
    return int(len(k)), hashlib.sha256(k.astype("<u8").tobytes()).hexdigest()


def main():
    use_ref = "--ref" in sys.argv
    out = {}
    t = time.time()
    st = S.heat(16384, 128, 4, 200)
    r, nr = st.rects()
    e = khp.resolve(st.task, st.mode, r, nr)
    n, h = digest(e)
    out["heat_16384_t128_s200"] = {"accesses": len(st), "tasks": st.ntasks, "edges": n, "sha256": h}
    print("heat", n, f"{time.time() - t:.1f}s", flush=True)
    del e

    t = time.time()
    st = S.cholesky(8192, 128)
    r, nr = st.rects()
    n, h = digest(khp.resolve(st.task, st.mode, r, nr))
    out["cholesky_8192_nb128"] = {"accesses": len(st), "tasks": st.ntasks, "edges": n, "sha256": h}
    print("cholesky", n, f"{time.time() - t:.1f}s", flush=True)

    t = time.time()
    st = S.random2d(RANDOM_PREFIX)
    r, nr = st.rects()
    e, f = khp.resolve(st.task, st.mode, r, nr, want_frags=True)
    n, h = digest(e)
    bsz = 4096
    nb = -(-len(st) // bsz)
    frag_b = [int(f[i * bsz:(i + 1) * bsz].sum()) for i in range(nb)]
    edge_b = np.bincount(np.asarray(e[:, 1], np.int64) // bsz, minlength=nb).tolist()
    out[f"random2d_{RANDOM_PREFIX}"] = {"accesses": len(st), "tasks": st.ntasks, "edges": n, "sha256": h,
                                        "batch_tasks": bsz, "sum_frags_per_batch": frag_b,
                                        "edges_per_batch": [int(x) for x in edge_b]}
    print("random", n, sum(frag_b), f"{time.time() - t:.1f}s", flush=True)
    if use_ref:
        if not refkhp.available():
            raise SystemExit("--ref: oracle/_ref/libkhpref.so missing (make -C oracle)")
        t = time.time()
        er = refkhp.resolve(st.task, st.mode, r, nr)
        assert digest(er) == (n, h), "oracle and reference engine differ on the random prefix"
        print("reference engine agrees on the random prefix", f"{time.time() - t:.1f}s", flush=True)
        out[f"random2d_{RANDOM_PREFIX}"]["reference_engine_agrees"] = True

    with open(os.path.join(HERE, "digests.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
