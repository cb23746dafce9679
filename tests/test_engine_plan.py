"""The engine's launch plan: built when the live set changes, reused across
calls, freed with the engine.

* Stepping cells with unequal step counts again and again (what a serving
  loop does when requests finish in different quanta) keeps the device's
  free memory flat and the results equal to the oracle's (FP64, bit for bit).
* A cell listed twice in one hsg_step_cells call is an argument error (two
  workgroups would otherwise step the same cell at once).
* Reading probes between calls does not change the plan: the bins and their
  launch shapes stay the same, and traces read in pieces equal one long run.
"""
import numpy as np
import pytest

from moose_core_amd import workloads
from moose_core_amd.engine import HSolveError

pytestmark = pytest.mark.gpu


def _free_device_memory():
    """hipMemGetInfo of the HIP runtime the engine library runs on (the
    same library, so the same context)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _cells():
    rng = np.random.default_rng(3)
    return [workloads.random_hh_cell(rng, int(n), inject=float(rng.uniform(0.4e-9, 1.2e-9)))
            for n in (64, 300, 900, 150)]


def test_unequal_steps_reuse_plans_and_memory_stays_flat(pkg, oracle_mod):
    print("case: step_cells with unequal step counts x 40, fp64", flush=True)
    specs = _cells()
    dt = 25e-6
    eng = pkg.Engine("fp64", steps_per_launch=8)
    ids = [eng.add_cell(s, dt) for s in specs]
    total = np.zeros(len(ids), dtype=np.int64)
    for cid, s in zip(ids, specs):
        eng.set_probes(cid, [0], 40 * 13)
    eng.reinit(dt)
    free0 = None
    for it in range(40):
        steps = np.array([5 + (it + k) % 9 for k in range(len(ids))], dtype=np.int64)
        eng.step_cells(ids, steps)
        total += steps
        if it == 4:
            eng.sync()
            free0 = _free_device_memory()
    eng.sync()
    free1 = _free_device_memory()
    assert free1 >= free0 - (1 << 20), (free0, free1)
    for cid, spec, n in zip(ids, specs, total):
        gpu = eng.read_probes(cid)[:, 0]
        assert len(gpu) == n
        ref = oracle_mod.OracleCell(spec, dt)
        ref.reinit(dt)
        np.testing.assert_array_equal(gpu, ref.run(dt, int(n), [0])[:, 0])
    eng.close()


def test_duplicate_cell_in_step_cells_is_an_error(pkg):
    print("case: step_cells with a repeated cell id", flush=True)
    specs = _cells()
    eng = pkg.Engine("fp64")
    ids = [eng.add_cell(s, 25e-6) for s in specs]
    eng.reinit(25e-6)
    with pytest.raises(HSolveError) as ei:
        eng.step_cells([ids[0], ids[1], ids[0]], [3, 3, 3])
    assert "twice" in str(ei.value)
    # nothing was stepped
    assert eng.cell_info(ids[0])["steps_done"] == 0
    eng.close()


def test_probe_reads_keep_the_plan(pkg, oracle_mod):
    print("case: probes read between calls, plan kept, fp64", flush=True)
    specs = _cells()
    dt, chunks = 25e-6, [37, 50, 13, 100]
    eng = pkg.Engine("fp64", steps_per_launch=20)
    ids = [eng.add_cell(s, dt) for s in specs]
    for cid in ids:
        eng.set_probes(cid, [0, 1], 200)
    eng.reinit(dt)
    pieces = {cid: [] for cid in ids}
    plan0 = None
    for k in chunks:
        eng.step(k)
        for cid in ids:
            pieces[cid].append(eng.read_probes(cid))
        plan = [(b["kernel"], b["cells"], b["threads"], b["lds_bytes"]) for b in eng.plan_info()]
        plan0 = plan0 or plan
        assert plan == plan0
    for cid, spec in zip(ids, specs):
        gpu = np.concatenate(pieces[cid])
        ref = oracle_mod.OracleCell(spec, dt)
        ref.reinit(dt)
        np.testing.assert_array_equal(gpu, ref.run(dt, sum(chunks), [0, 1]))
    eng.close()
