"""hsg_step_cells_noalloc (the entry point that allocates no device memory)
against hsg_step_cells, through the C ABI: the same cells stepped by the two
entry points give bit-identical traces and spike steps (FP64), on two calls
in a row, the second with fewer cells and fewer steps, and the no-allocation
calls leave the device's free memory unchanged.

The cells: small HH cells (fused bins), a Ca cell, and a cell of 4500
compartments, which the automatic mode runs split (batch-wide channel
kernel over its tiles, so the launch tables hold tiles as well as orders).
"""
import numpy as np
import pytest

from moose_core_amd import workloads

pytestmark = pytest.mark.gpu


def _free_device_memory():
    """hipMemGetInfo of the HIP runtime the engine library runs on."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _cells():
    rng = np.random.default_rng(5)
    specs = [workloads.random_hh_cell(rng, int(n), inject=float(rng.uniform(0.5e-9, 1.5e-9)))
             for n in (50, 400, 1300, 90)]
    specs.append(workloads.random_ca_cell(rng, 300, inject=0.8e-9))
    specs.append(workloads.random_hh_cell(rng, 4500, inject=3e-9))
    return specs


CALLS = [
    # (cell positions, steps): every cell with unequal counts, then fewer
    # cells for fewer steps
    ([0, 1, 2, 3, 4, 5], [120, 75, 133, 60, 98, 110]),
    ([1, 3, 5], [40, 17, 29]),
]


def _run(pkg, specs, noalloc):
    dt = 25e-6
    eng = pkg.Engine("fp64", steps_per_launch=32)
    ids = [eng.add_cell(s, dt) for s in specs]
    for cid, s in zip(ids, specs):
        eng.set_probes(cid, [0, s.n // 2], 400)
    eng.reinit(dt)
    free = []
    # an allocating call with one step on both engines: it sizes the
    # engine's tables for these cells and is the first work on the engine's
    # streams (the HIP runtime sets up a stream's queue on its first use)
    eng.step_cells(ids, 1)
    eng.sync()
    if noalloc:
        free.append(_free_device_memory())
    for pos, steps in CALLS:
        cells = [ids[p] for p in pos]
        print(f"call: cells {pos} steps {steps} {'noalloc' if noalloc else 'step_cells'}", flush=True)
        if noalloc:
            eng.step_cells_noalloc(cells, steps)
        else:
            eng.step_cells(cells, steps)
        eng.sync()
        if noalloc:
            free.append(_free_device_memory())
    traces = [eng.read_probes(cid) for cid in ids]
    spikes = eng.read_spikes()
    eng.close()
    return traces, spikes, free


def test_noalloc_equals_step_cells(pkg):
    print("case: hsg_step_cells_noalloc vs hsg_step_cells, 6 cells, two calls, fp64", flush=True)
    specs = _cells()
    want, want_spk, _ = _run(pkg, specs, noalloc=False)
    got, got_spk, free = _run(pkg, specs, noalloc=True)
    for k, (a, b) in enumerate(zip(got, want)):
        assert a.shape == b.shape, k
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got_spk, want_spk)
    assert len(want_spk) > 0
    # nothing allocated on the device by either no-allocation call
    assert free[1] == free[0] and free[2] == free[1], free
