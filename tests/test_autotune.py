"""The autotuners: the engine-level sweep returns one of its candidates; the
per-bin tuner sets measured shapes on the live plan, never ends slower than
the engine's own choices beyond timing noise, refuses shapes a bin cannot
take, and a run with the tuned shapes still equals the oracle bit for bit
(FP64)."""
import numpy as np
import pytest

from moose_core_amd import workloads

pytestmark = pytest.mark.gpu


def test_tuned_shape_is_a_candidate_and_exact(pkg, oracle_mod):
    print("case: autotune on a 12-cell batch, fp64", flush=True)
    from moose_core_amd import autotune
    rng = np.random.default_rng(41)
    specs = [workloads.random_hh_cell(rng, int(n), inject=1e-9) for n in rng.integers(40, 600, size=12)]
    packed = [s.pack() for s in specs]
    space = {"kernel_mode": ("fused", "split"), "block_threads": (0, 128), "steps_per_launch": (20,)}
    best, trials = autotune.tune(packed, "fp64", steps=40, space=space)
    assert len(trials) == 4 and all(r > 0 for _, r in trials)
    assert best in [t for t, _ in trials]
    eng = pkg.Engine("fp64", **best)
    ids = [eng.add_cell(p, 25e-6) for p in packed]
    for cid, s in zip(ids, specs):
        eng.set_probes(cid, [0], 120)
    eng.reinit(25e-6)
    eng.step(120)
    for cid, s in zip(ids, specs):
        ref = oracle_mod.OracleCell(s, 25e-6)
        ref.reinit(25e-6)
        np.testing.assert_array_equal(eng.read_probes(cid), ref.run(25e-6, 120, [0]))
    eng.close()


def test_tune_bins_sets_shapes_and_stays_exact(pkg, oracle_mod):
    print("case: tune_bins on a 3-bin ensemble batch, fp64", flush=True)
    from moose_core_amd import autotune
    from moose_core_amd.engine import HSolveError
    rng = np.random.default_rng(43)
    sizes = np.concatenate([rng.integers(50, 200, size=40), rng.integers(600, 900, size=24),
                            rng.integers(1500, 2000, size=16)])
    specs = [workloads.random_hh_cell(rng, int(n), inject=float(rng.uniform(0, 0.3e-9))) for n in sizes]
    dt, nsteps = 25e-6, 150
    eng = pkg.Engine("fp64", steps_per_launch=25)
    ids = [eng.add_cell(s, dt) for s in specs]
    eng.reinit(dt)
    res = autotune.tune_bins(eng, steps=25, repeats=2)
    print(res["ms_per_step_default"], res["ms_per_step_tuned"], [b["chosen"] for b in res["bins"]])
    assert len(res["bins"]) >= 2 and res["trials"]
    assert res["ms_per_step_tuned"] <= 1.15 * res["ms_per_step_default"]
    for b in res["bins"]:
        if b["chosen"]["threads"]:
            assert b["threads"] == b["chosen"]["threads"]
    # shapes a bin cannot take are refused
    with pytest.raises(HSolveError):
        eng.set_bin_shape(0, 96, -1)
    with pytest.raises(HSolveError):
        eng.set_bin_shape(len(res["bins"]), 0, -1)
    # the tuned plan from reinit equals the oracle bit for bit
    for cid in ids:
        eng.set_probes(cid, [0], nsteps)
    eng.reinit(dt)
    eng.step(nsteps)
    for cid, s in list(zip(ids, specs))[::7]:
        ref = oracle_mod.OracleCell(s, dt)
        ref.reinit(dt)
        np.testing.assert_array_equal(eng.read_probes(cid), ref.run(dt, nsteps, [0]))
    eng.set_bin_shape(-1)
    eng.close()
