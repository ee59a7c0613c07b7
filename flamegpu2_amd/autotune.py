"""Launch-shape autotuner for batched simulations.

A batch's results do not depend on its launch shape (block size, agents per
thread, persistent blocks): every simulation gives the same bits under every
shape.  So the tuner measures candidate shapes on the real run's first steps
instead of on a copy, then keeps the fastest for the rest of the run.

Search space: block size {64, 128, 256, 512, 1024} x agents per thread {1, 2, 4},
plus persistent grids (blocks per CU) for the best one-block-per-tile shape.  Each
candidate runs two untimed steps (repack, graph capture of both buffer
arrangements) and `steps` timed steps.
"""

BLOCKS = (64, 128, 256, 512, 1024)
AGENTS_PER_THREAD = (1, 2, 4)
PERSISTENT = (2, 4, 8)


def candidates(blocks=BLOCKS, apts=AGENTS_PER_THREAD, persistent=PERSISTENT):
    out = [(b, a, 0) for b in blocks for a in apts]
    return out, list(persistent)


def steps_needed(steps=2, blocks=BLOCKS, apts=AGENTS_PER_THREAD, persistent=PERSISTENT):
    """Simulation steps one tuning pass consumes."""
    return (len(blocks) * len(apts) + len(persistent)) * (steps + 2)


def autotune(batch, steps=2, blocks=BLOCKS, apts=AGENTS_PER_THREAD, persistent=PERSISTENT):
    """Times every candidate on `batch` (a models.Batch) and leaves the fastest set.
    Returns (best (block, agents_per_thread, blocks_per_cu), {shape: ms per step})."""
    timings = {}
    shapes, pers = candidates(blocks, apts, persistent)
    for shape in shapes:
        batch.set_tuning(*shape)
        timings[shape] = batch.timed_steps(steps) / steps
    best = min(timings, key=timings.get)
    for bpc in pers:
        shape = (best[0], best[1], bpc)
        batch.set_tuning(*shape)
        timings[shape] = batch.timed_steps(steps) / steps
    best = min(timings, key=timings.get)
    batch.set_tuning(*best)
    return best, timings
