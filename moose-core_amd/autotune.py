"""Launch-shape autotuner.

`tune_bins` tunes the live engine's launch plan bin by bin, on the engine's
own clock (HIP events around the launches, hsg_last_step_timing):

  threads     threads per cell workgroup of the bin's per-cell kernel
              (64 / 128 / 192 / 256; the block-form split solve has a fixed shape)
  fast_pass   the fast channel pass (all loads of a compartment first) or
              the generic one, where the bin's cells allow both

The bins of one launch round run concurrently, so each candidate is judged
by the whole round's time per step with the other bins at their current
shapes: bins are visited longest first, each keeps its fastest candidate
(the engine's own choice unless another is faster by `min_gain`), and the
choice is stored in the engine (hsg_set_bin_shape) for every later plan.
Tuning steps the model, so a caller tunes on its batch and then calls
`reinit` before the run it keeps.

`tune` is the engine-level sweep (kernel mode, steps per launch) over fresh
engines, for options that are fixed when an engine is created.
"""
import itertools
import time

from .engine import Engine, HSolveError

DEFAULT_SPACE = {
    "kernel_mode": ("fused", "split"),
    "block_threads": (0, 128, 256),
    "steps_per_launch": (50,),
}

THREADS = (64, 128, 192, 256)


def _ms_per_step(eng, steps, repeats):
    best = float("inf")
    for _ in range(repeats):
        eng.step(steps)
        eng.sync()
        _, total_ms, _ = eng.last_timing()
        best = min(best, total_ms / steps)
    return best


def _candidates(info):
    """(threads, fast_pass) shapes to measure for one bin, the engine's own
    choice first."""
    fixed = info["split"] and not info["msg"] and "block" in info["kernel"]
    threads = (0,) if fixed else (0,) + THREADS
    out = []
    for t in threads:
        for f in (-1, 0, 1):
            out.append((t, f))
    return out


def tune_bins(eng, steps=50, repeats=2, min_gain=0.01, log=None):
    """Tunes every bin of `eng`'s plan in place.  Returns a dict: the
    chosen shape of each bin (with its kernel and cells), the time per step
    before and after, and every trial."""
    eng.set_bin_shape(-1)                        # start from the engine's choices
    eng.step(steps)                              # plan, and one timed round per bin
    eng.sync()
    plan = eng.plan_info()
    base = _ms_per_step(eng, steps, repeats)
    current = base
    trials = []
    chosen = {b: (0, -1) for b in range(len(plan))}
    for b in sorted(range(len(plan)), key=lambda k: -plan[k]["round_ms"]):
        best_shape, best_ms = (0, -1), current
        for shape in _candidates(plan[b]):
            if shape == (0, -1):
                continue
            try:
                eng.set_bin_shape(b, *shape)
            except HSolveError:
                continue                         # not available for this bin
            ms = _ms_per_step(eng, steps, repeats)
            trials.append({"bin": b, "threads": shape[0], "fast_pass": shape[1], "ms_per_step": ms})
            if log:
                log(f"[autotune] bin {b} ({plan[b]['kernel']}, {plan[b]['cells']} cells) "
                    f"threads {shape[0]} fast {shape[1]}: {ms:.4f} ms/step")
            if ms < best_ms * (1.0 - min_gain):
                best_shape, best_ms = shape, ms
        eng.set_bin_shape(b, *best_shape)
        chosen[b] = best_shape
        current = best_ms
    tuned = _ms_per_step(eng, steps, repeats)
    final = eng.plan_info()
    bins = []
    for b, info in enumerate(final):
        bins.append({"bin": b, "kernel": info["kernel"], "cells": info["cells"], "threads": info["threads"],
                     "fast_slots": info["fast_slots"], "chosen": {"threads": chosen[b][0], "fast_pass": chosen[b][1]}})
    return {"bins": bins, "ms_per_step_default": base, "ms_per_step_tuned": tuned, "trials": trials}


def tune(cells, precision="mixed", dt=25e-6, steps=100, space=None, log=None):
    """Engine-level sweep: returns (best_options, trials), trials a list of
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
