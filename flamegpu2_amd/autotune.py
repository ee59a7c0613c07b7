"""Launch-shape autotuner for batched simulations.

A batch's results do not depend on its launch shape (block size, agents per
thread, persistent blocks): every simulation gives the same bits under every
shape.  So the tuner measures candidate shapes on the real run's first steps
instead of on a copy, then keeps the fastest for the rest of the run.

Search space: block size {64, 128, 256, 512, 1024} x agents per thread {1, 2, 4},
plus persistent grids for the best one-block-per-tile shape.  A persistent grid is
sized in resident waves per CU (16, 24 or 32 of them: blocks per CU = waves x 64 /
block size), so that it holds as many waves as the one-block-per-tile grid keeps
resident.  Each candidate runs two untimed steps (repack, graph capture of both
buffer arrangements) and `steps` timed steps.

The bin-width multiplier (1 or 0.5: spatial bins as wide as the radius, or half
as wide and searched two bins either side) is the third dimension.  It changes the
neighbour visit order, so the FP32 summation order, so it cannot be switched inside
a run: choose_bin_width() times each multiplier on throwaway copies of the batch
before the run starts, and the run is then built with the winner.
"""

BLOCKS = (64, 128, 256, 512, 1024)
AGENTS_PER_THREAD = (1, 2, 4)
PERSISTENT_WAVES = (16, 24, 32)   # resident waves per CU of a persistent grid
BINS_PER_RADIUS = (1, 2)   # bin-width multipliers 1.0 and 0.5


def persistent_blocks_per_cu(block, waves):
    """Blocks per CU of a persistent grid holding `waves` wave64s per CU."""
    return max(1, (waves * 64) // block)


def candidates(blocks=BLOCKS, apts=AGENTS_PER_THREAD, persistent=PERSISTENT_WAVES):
    out = [(b, a, 0) for b in blocks for a in apts]
    return out, list(persistent)


def steps_needed(steps=2, blocks=BLOCKS, apts=AGENTS_PER_THREAD, persistent=PERSISTENT_WAVES):
    """Simulation steps one tuning pass consumes."""
    return (len(blocks) * len(apts) + len(persistent)) * (steps + 2)


def autotune(batch, steps=2, blocks=BLOCKS, apts=AGENTS_PER_THREAD, persistent=PERSISTENT_WAVES):
    """Times every candidate on `batch` (a models.Batch) and leaves the fastest set.
    Returns (best (block, agents_per_thread, blocks_per_cu), {shape: ms per step})."""
    timings = {}
    shapes, pers = candidates(blocks, apts, persistent)
    for shape in shapes:
        batch.set_tuning(*shape)
        timings[shape] = batch.timed_steps(steps) / steps
    best = min(timings, key=timings.get)
    for waves in pers:
        shape = (best[0], best[1], persistent_blocks_per_cu(best[0], waves))
        if shape in timings:
            continue
        batch.set_tuning(*shape)
        timings[shape] = batch.timed_steps(steps) / steps
    best = min(timings, key=timings.get)
    batch.set_tuning(*best)
    return best, timings


def choose_bin_width(make_batch, steps=2, bins_per_radius=BINS_PER_RADIUS):
    """make_batch(bins_per_radius) builds a models.Batch holding copies of the run's
    simulations.  Each copy runs two untimed steps and `steps` timed steps with the
    default launch shape, and is then closed.  Returns (best bins_per_radius,
    {bins_per_radius: ms per step})."""
    timings = {}
    for k in bins_per_radius:
        b = make_batch(k)
        try:
            timings[k] = b.timed_steps(steps) / steps
        finally:
            b.close()
    return min(timings, key=timings.get), timings
