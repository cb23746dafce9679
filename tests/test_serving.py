"""Serving loop: continuous batching with admission by device-memory
footprint.  A request's result is the same alone and in the mix (FP64 and
mixed forms, bit for bit, since cells do not interact), and FP64 results
equal the oracle's.  A mix with a 5000-compartment cell and HHChannel2D
K-AHP channels is held to the oracle in both forms."""
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


def _large_2d_requests(serving):
    """A 5000-compartment Ca cell with the 2-D K-AHP (split mode, block
    form of the solve, a request of its own bin; one HHChannel2D per
    compartment, so the channel kernel advances it without the external
    kernel), a 400-compartment one, and an HH cell, with staggered
    arrivals."""
    rng = np.random.default_rng(33)
    specs = [workloads.random_ca_cell(rng, 5000, inject=6e-9, kahp2d=True),
             workloads.random_ca_cell(rng, 400, inject=3e-9, kahp2d=True),
             workloads.random_hh_cell(rng, 800, inject=1.0e-9)]
    steps = (600, 450, 500)
    return [serving.Request(k, s.pack(), n, arrival=0.002 * k, probes=(0, s.n // 2, s.n - 1))
            for k, (s, n) in enumerate(zip(specs, steps))], specs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_large_2d_requests_against_the_oracle(pkg, oracle_mod, precision):
    """Served in a mix, every request against the oracle: FP64 within 1e-6 V
    (the HHChannel2D gates use exp, so not bit for bit) with identical spike
    steps; mixed within 1e-4 V with spike steps within one step."""
    print(f"case: serving with a 5000-compartment 2-D K-AHP cell, {precision}", flush=True)
    from moose_core_amd import serving
    dt = 20e-6
    reqs, specs = _large_2d_requests(serving)
    srv = serving.Server(precision, dt=dt, mem_budget=1 << 30, quantum=64)
    res = {r.rid: r for r in srv.run(reqs, realtime=False)}
    srv.close()
    tol = 1e-6 if precision == "fp64" else 1e-4
    nspk = 0
    for r, s in zip(reqs, specs):
        ref = oracle_mod.OracleCell(s, dt)
        ref.reinit(dt)
        tr = ref.run(dt, r.nsteps, list(r.probes))
        _, steps = ref.read_spikes()
        err = float(np.abs(res[r.rid].trace - tr).max())
        print(f"request {r.rid}: n = {s.n}, max|dVm| = {err:.3e} V, spikes {list(steps)}", flush=True)
        assert err <= tol, (r.rid, err)
        mine = list(res[r.rid].spikes)
        assert len(mine) == len(steps), (r.rid, mine, list(steps))
        if precision == "fp64":
            assert mine == list(steps)
        else:
            assert all(abs(a - b) <= 1 for a, b in zip(mine, steps))
        nspk += len(steps)
    assert nspk >= 3
