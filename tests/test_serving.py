"""Serving loop: continuous batching with admission by device-memory
footprint.  A request's result is the same alone and in the mix (FP64 and
mixed forms, bit for bit, since cells do not interact), and FP64 results
equal the oracle's."""
import numpy as np
import pytest

from moose_core_amd import workloads

pytestmark = pytest.mark.gpu


def requests(serving):
    rng = np.random.default_rng(31)
    reqs = []
    sizes = [120, 700, 60, 1500, 300, 45, 900, 200]
    for k, n in enumerate(sizes):
        make = workloads.random_ca_cell if k % 3 == 1 else workloads.random_hh_cell
        spec = make(rng, n, inject=float(rng.uniform(0.3e-9, 1.2e-9)))
        reqs.append(serving.Request(k, spec.pack(), int(rng.integers(150, 500)), arrival=0.001 * k,
                                    probes=(0, n // 2)))
    return reqs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_alone_equals_in_mix(pkg, precision):
    print(f"case: serving alone vs mix, {precision}", flush=True)
    from moose_core_amd import serving
    reqs = requests(serving)
    # a small budget forces queueing and slot reuse
    srv = serving.Server(precision, dt=25e-6, mem_budget=3 << 19, quantum=64)
    mix = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert srv.max_active < len(reqs)
    srv.close()
    assert sorted(mix) == list(range(len(reqs)))
    for pick in (1, 3, 6):
        alone = serving.Server(precision, dt=25e-6, mem_budget=64 << 20, quantum=37)
        (res,) = alone.run([reqs[pick]], realtime=False)
        alone.close()
        np.testing.assert_array_equal(res.trace, mix[pick].trace)
        np.testing.assert_array_equal(res.spikes, mix[pick].spikes)


def test_fp64_results_equal_the_oracle(pkg, oracle_mod):
    print("case: serving fp64 vs oracle", flush=True)
    from moose_core_amd import serving
    reqs = requests(serving)[:5]
    srv = serving.Server("fp64", dt=25e-6, mem_budget=32 << 20, quantum=50)
    res = {r.rid: r for r in srv.run(reqs, realtime=True)}
    srv.close()
    for r in reqs:
        ref = oracle_mod.OracleCell(r.spec.spec, 25e-6)
        ref.reinit(25e-6)
        tr = ref.run(25e-6, r.nsteps, list(r.probes))
        np.testing.assert_array_equal(res[r.rid].trace, tr)
        g, steps = ref.read_spikes()
        np.testing.assert_array_equal(res[r.rid].spikes, np.array(steps, dtype=np.int64))
        assert res[r.rid].latency >= 0
