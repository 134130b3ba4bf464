"""Autotuner of the batched ensemble: picks the marching segment length of
the y pass for each size bucket of 2D problems by timing a few candidates on
that bucket's problems alone (a separate batch, discarded afterwards)."""
from __future__ import annotations

import time

BUCKETS = ((0, 128), (128, 256), (256, 1 << 30))   # 2D side ranges [lo, hi)
CANDIDATES = (4, 8, 16, 32, 64)


def bucket_of(pb):
    side = pb.n[1]
    for q, (lo, hi) in enumerate(BUCKETS):
        if lo <= side < hi:
            return q
    return len(BUCKETS) - 1


def tune_seglen(problems, prec: int = 0, steps: int = 6, warmup: int = 2, candidates=CANDIDATES,
                max_per_bucket: int = 16):
    """{bucket: best seglen} for the 2D problems of `problems` (time per step
    of the bucket's batch, median of the timed steps)."""
    from . import GpuBatch

    best = {}
    for q in range(len(BUCKETS)):
        group = [pb for pb in problems if pb.dim == 2 and bucket_of(pb) == q][:max_per_bucket]
        if not group:
            continue
        b = GpuBatch(group)
        timing = {}
        for cand in candidates:
            for p in range(len(group)):
                b.set_seglen(p, cand)
            b.init_from_prims()
            b.begin_run()
            b.advance(warmup, prec)
            b.sync()
            t = []
            for _ in range(steps):
                t0 = time.perf_counter()
                b.advance(1, prec)
                b.sync()
                t.append(time.perf_counter() - t0)
            timing[cand] = sorted(t)[len(t) // 2]
        b.close()
        best[q] = min(timing, key=timing.get)
    return best


def apply(batch, problems, table):
    """Sets each 2D problem's segment length from the bucket table."""
    for p, pb in enumerate(problems):
        if pb.dim == 2 and bucket_of(pb) in table:
            batch.set_seglen(p, table[bucket_of(pb)])
