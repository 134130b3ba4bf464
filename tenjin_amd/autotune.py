"""Autotuners of launch shapes: the y-tile shape of the stage kernels and the
marching segment length (rows per wave of the y / z passes of the flux-plane
form).  Every choice gives bit-identical results; it only moves time, and the
best one depends on where the work is (THINC's interface rows make some
waves much longer than others), which no static model sees.

- tune_tiles: the batched ensemble's stage kernels, per size bucket of 2D
  problems, timing each tile shape (BATCH_TILES) on that bucket's problems
  alone (a separate batch, discarded afterwards).
- tune_seglen: the same for the segment length of the flux-plane form.
- tune_level_seglen: one level, timing a few candidates (0 = the library's own
  pick) in turns on a scratch level advanced to a given step, discarded
  afterwards."""
from __future__ import annotations

import time

BUCKETS = ((0, 128), (128, 256), (256, 1 << 30))   # 2D side ranges [lo, hi)
CANDIDATES = (4, 8, 16, 32, 64)
LEVEL_CANDIDATES = (0, 8, 12)


def bucket_of(pb):
    side = pb.n[1]
    for q, (lo, hi) in enumerate(BUCKETS):
        if lo <= side < hi:
            return q
    return len(BUCKETS) - 1


def _tune_buckets(problems, prec, steps, warmup, candidates, max_per_bucket, setter, stage2d):
    """{bucket: best candidate}: each 2D bucket's problems alone in a batch,
    every candidate set on all of them by setter(batch, p, cand), timed as
    the median of `steps` single steps after `warmup` steps."""
    from . import GpuBatch

    best = {}
    for q in range(len(BUCKETS)):
        group = [pb for pb in problems if pb.dim == 2 and bucket_of(pb) == q][:max_per_bucket]
        if not group:
            continue
        b = GpuBatch(group)
        b.set_stage2d(stage2d)
        timing = {}
        for cand in candidates:
            for p in range(len(group)):
                setter(b, p, cand)
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


def tune_tiles(problems, prec: int = 0, steps: int = 6, warmup: int = 2, max_per_bucket: int = 16):
    """{bucket: best y-tile shape (index into BATCH_TILES)} for the stage
    kernels of the 2D problems of `problems`."""
    from . import BATCH_TILES

    return _tune_buckets(problems, prec, steps, warmup, range(len(BATCH_TILES)), max_per_bucket,
                         lambda b, p, c: b.set_tile(p, c), True)


def tune_seglen(problems, prec: int = 0, steps: int = 6, warmup: int = 2, candidates=CANDIDATES,
                max_per_bucket: int = 16):
    """{bucket: best seglen} for the flux-plane form of the 2D problems of
    `problems` (time per step of the bucket's batch, median of the timed
    steps)."""
    return _tune_buckets(problems, prec, steps, warmup, candidates, max_per_bucket,
                         lambda b, p, c: b.set_seglen(p, c), False)


def apply(batch, problems, table):
    """Sets each 2D problem's segment length from the bucket table."""
    for p, pb in enumerate(problems):
        if pb.dim == 2 and bucket_of(pb) in table:
            batch.set_seglen(p, table[bucket_of(pb)])


def apply_tiles(batch, problems, table):
    """Sets each 2D problem's y-tile shape from the bucket table."""
    for p, pb in enumerate(problems):
        if pb.dim == 2 and bucket_of(pb) in table:
            batch.set_tile(p, table[bucket_of(pb)])


def tune_level_seglen(pb, prec: int = 0, fuse: int = 0, at_step: int = 0, steps: int = 3, rounds: int = 3,
                      candidates=LEVEL_CANDIDATES, margin: float = 0.03, device: int = 0):
    """(seglen, {candidate: seconds for its rounds x steps loop bodies}) for
    the level of problem pb, timed on a scratch level (discarded afterwards).
    The scratch level first runs max(at_step, 1) loop bodies with the
    library's own pick, so that the trial sees the state the run will spend
    its time in (THINC's interface work moves as the flow develops); the
    candidates then take turns, `steps` loop bodies each, for `rounds`
    rounds, so that all of them see the same drift of the state.  The
    library's own pick (0) is kept unless another candidate is faster by more
    than `margin`, about the spread of so short a trial."""
    from . import GpuLevel

    g = GpuLevel(pb, device)
    timing = {c: 0.0 for c in candidates}
    try:
        g.set_tuning(0, fuse)
        g.init_from_prims(pb.prims, pb.shift_pressure)
        g.begin_run(float("inf"))
        g.advance(max(at_step, 1), prec)
        g.sync()
        for _ in range(rounds):
            for cand in candidates:
                g.set_tuning(cand, fuse)
                t0 = time.perf_counter()
                g.advance(steps, prec)
                g.sync()
                timing[cand] += time.perf_counter() - t0
    finally:
        g.close()
    return choose(timing, margin), timing


def choose(timing, margin: float = 0.03):
    """The fastest candidate of {candidate: seconds}, except that the library's
    own pick (0) stays unless the fastest beats it by more than `margin`."""
    best = min(timing, key=timing.get)
    if 0 in timing and timing[best] * (1.0 + margin) >= timing[0]:
        best = 0
    return best
