"""The autotuner returns one of its candidates, and a batch run with the
chosen shape still equals the oracle bit for bit (FP64)."""
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
