"""Launch-shape autotuner.

Times a few steps of the actual batch under each candidate engine
configuration and keeps the fastest:

  kernel_mode       fused (one workgroup runs whole ticks, many per launch)
                    or split (batch-wide channel kernel + per-cell solve)
  block_threads     0 (per size bin: 64 / 128 / 256 threads by the bin's
                    largest cell) or a fixed workgroup size
  steps_per_launch  ticks fused into one launch (fused mode) or per launch
                    round (split mode)

Cells are binned by LDS need inside the engine (up to three bins launched
concurrently); the tuner picks the shape for the batch as given.  Candidates
are measured on the engine's own clock (HIP events around the launches) after
one warm-up round, so host-side setup is not timed.
"""
import itertools
import time

from .engine import Engine

DEFAULT_SPACE = {
    "kernel_mode": ("fused", "split"),
    "block_threads": (0, 128, 256),
    "steps_per_launch": (50,),
}


def tune(cells, precision="mixed", dt=25e-6, steps=100, space=None, log=None):
    """Returns (best_options, trials): trials is a list of
    (options, compartment-steps/s)."""
    space = dict(DEFAULT_SPACE if space is None else space)
    keys = sorted(space)
    ncomp = sum(c.n for c in cells)
    trials = []
    for combo in itertools.product(*(space[k] for k in keys)):
        opts = dict(zip(keys, combo))
        eng = Engine(precision, **opts)
        try:
            for c in cells:
                eng.add_cell(c, dt)
            eng.reinit(dt)
            spl = opts.get("steps_per_launch", 50) or 50
            eng.step(spl)                          # warm-up round
            eng.sync()
            t = time.perf_counter()
            eng.step(steps)
            eng.sync()
            wall = time.perf_counter() - t
        finally:
            eng.close()
        rate = ncomp * steps / wall
        trials.append((opts, rate))
        if log:
            log(f"[autotune] {opts}: {rate:.3e} compartment-steps/s")
    best = max(trials, key=lambda x: x[1])[0]
    return best, trials
