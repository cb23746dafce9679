"""Reference branches off the main configurations, each pinned on the CPU and
run on the GPU against the oracle:

* non-integer gate powers, taken as exp(p log x) and 0 for x <= 0
  (ChannelStruct::powerN, hsolve/HSolveStruct.cpp:29-34);
* caAdvance = 0: the Ca pools driven by Vm of the new step rather than VMid
  (hsolve/HSolveActive.cpp:183-200);
* a Ca-dependent Z gate with no CaConc: the channel's external [Ca] through
  the Ca table when it is positive, else the V table at the gate's Ca-table
  column (hsolve/HSolveActive.cpp:296-312);
* SpikeGen with edgeTriggered = 0: a spike on every step above threshold
  once the refractory time has passed (biophysics/SpikeGen.cpp:202-216).

CPU pins: the oracle equals the reference's own hsolve sources on these cells
bit for bit (where the reference tree is present to build them), powerN
equals the closed form x**p, and the spike steps follow from the Vm trace by
the rule's definition.  GPU: FP64 within 1e-6 V of the oracle with identical
spike steps, mixed within 1e-4 V with spikes within one step.
"""
import os

import numpy as np
import pytest

from moose_core_amd import workloads
from moose_core_amd._abi import SPIKEGEN_DT
from moose_core_amd.cellspec import CellSpec, channel_records, interleave_channels, passive_compartments

HERE = os.path.dirname(os.path.abspath(__file__))
FP64_TOL = 1e-6
MIXED_TOL = 1e-4

needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "hsolve_ref")),
    reason="reference harness not built (needs the moose-core sources)")


def _tree(rng, n):
    parent, depth = workloads.random_tree_parents(rng, n, 2)
    maxd = max(1, int(depth.max()))
    length = rng.uniform(5e-6, 50e-6, size=n)
    dia = 2e-6 - 1.5e-6 * depth / maxd
    length[0] = 20e-6
    dia[0] = 20e-6
    comp, area = passive_compartments(length, dia)
    axial = np.stack([parent[1:], np.arange(1, n)], axis=1).astype(np.int32)
    return comp, area, axial


def power_cell(rng, n, inject):
    """Squid HH with non-integer powers: Na m^2.5 h^1.5, K n^3.7."""
    spec = workloads.random_hh_cell(rng, n, inject=inject)
    ch = spec.channels
    na = ch["Ypower"] > 0
    ch["Xpower"][na] = 2.5
    ch["Ypower"][na] = 1.5
    ch["Xpower"][~na] = 3.7
    return spec


def zgate_cell(rng, n, inject):
    """Squid HH plus a K-AHP-like channel whose Z gate is Ca-dependent
    (useConcentration) with no CaConc to read: its [Ca] is the channel's
    external concentration (addConc), and where that is not positive the
    reference reads the V table at the gate's Ca-table column."""
    comp, area, axial = _tree(rng, n)
    comp["inject"][0] = inject
    c = np.arange(n, dtype=np.int32)
    na = channel_records(n, c, 1200.0 * area, 0.050, Xpower=3.0, Ypower=1.0, gates=(0, 1, -1))
    k = channel_records(n, c, 360.0 * area, -0.077, Xpower=4.0, gates=(2, -1, -1))
    ahp = channel_records(n, c, 30.0 * area, -0.080, Zpower=2.0, gates=(-1, -1, 3), useConcentration=1)
    gates = workloads.hh_gates() + [workloads.ca_gates()[6]]
    spk = np.zeros(1, SPIKEGEN_DT)
    spk["compartment"] = 0
    spk["edgeTriggered"] = 1
    return CellSpec(comp, axial, seed=0, channels=interleave_channels([na, k, ahp]), gates=gates, spikegens=spk)


def edge0_cell(rng, n, inject, refract):
    """Squid HH driven to fire repetitively, with two SpikeGens on the soma:
    edgeTriggered = 0 (threshold 0, refractT `refract`) and edgeTriggered = 1
    (threshold 0, refractT 0)."""
    spec = workloads.random_hh_cell(rng, n, inject=inject)
    spk = np.zeros(2, SPIKEGEN_DT)
    spk["compartment"] = 0
    spk["edgeTriggered"] = [0, 1]
    spk["threshold"] = 0.0
    spk["refractT"] = [refract, 0.0]
    spec.spikegens = spk
    return spec


def ext_conc(spec):
    """addConc inputs for the Z-gate cell: every third AHP channel gets a
    positive external [Ca] (the Ca-table branch), the rest keep 0 (the
    V-table fallback).  Returns (channel index, conc) pairs."""
    ahp = np.nonzero(spec.channels["useConcentration"] == 1)[0]
    return [(int(ch), 0.05 + 0.3 * (k % 5) / 4) for k, ch in enumerate(ahp) if k % 3 == 0]


# ---------------------------------------------------------------------------
# CPU pins
# ---------------------------------------------------------------------------
@needs_ref
@pytest.mark.parametrize("kind", ["power", "zgate", "zgate_ext"])
def test_oracle_matches_reference_on_branches(oracle_mod, kind):
    """The oracle equals the reference's own HSolveActive step bit for bit on
    the non-integer-power cell and the Z-gate cell (external [Ca] on some
    channels, the V-table fallback on the rest)."""
    rng = np.random.default_rng(61)
    spec = power_cell(rng, 90, 1.5e-9) if kind == "power" else zgate_cell(rng, 70, 1.5e-9)
    dt, nsteps = 25e-6, 400
    ref = oracle_mod.OracleCell(spec, dt)
    if kind == "zgate_ext":
        for ch, conc in ext_conc(spec):
            ref.add_conc(ch, conc)
    V, VMid, state, ca, Gk = oracle_mod.run_reference(ref, dt, nsteps)
    cell = oracle_mod.OracleCell(spec, dt)
    if kind == "zgate_ext":
        for ch, conc in ext_conc(spec):
            cell.add_conc(ch, conc)
    cell.reinit(dt)
    order = cell.hines_order()
    tr = cell.run(dt, nsteps, np.arange(spec.n))
    np.testing.assert_array_equal(tr, V[:, order])
    np.testing.assert_array_equal(cell.state(), state)


def test_oracle_power_closed_form(oracle_mod):
    """Gk of a non-integer-power channel is Gbar * x^p * y^q from the gate
    states the oracle reports (x**p by numpy, relative 1e-14)."""
    rng = np.random.default_rng(62)
    spec = power_cell(rng, 30, 1.0e-9)
    dt = 25e-6
    cell = oracle_mod.OracleCell(spec, dt)
    cell.reinit(dt)
    cell.run(dt, 120)
    idx = np.arange(len(spec.channels))
    Gk = cell.get("Gk", idx)
    X, Y = cell.get("X", idx), cell.get("Y", idx)
    ch = spec.channels
    want = ch["Gbar"] * X ** ch["Xpower"] * np.where(ch["Ypower"] > 0, Y ** np.where(ch["Ypower"] > 0, ch["Ypower"], 1),
                                                     1.0)
    np.testing.assert_allclose(Gk, want, rtol=1e-14, atol=0)


def test_oracle_edge_rule_from_trace(oracle_mod):
    """SpikeGen::process on the oracle's own soma trace: with edgeTriggered
    = 1 and refractT = 0 a spike is the first step of each excursion above
    threshold; with edgeTriggered = 0 every step above threshold whose time
    is at least refractT after the last spike (half a step of slack) is one."""
    rng = np.random.default_rng(63)
    dt, N, refract = 25e-6, 3000, 0.3e-3
    spec = edge0_cell(rng, 40, 1.2e-9, refract)
    cell = oracle_mod.OracleCell(spec, dt)
    cell.reinit(dt)
    v = cell.run(dt, N, [0])[:, 0]
    gens, steps = cell.read_spikes()
    above = v > 0.0
    onset = [s + 1 for s in range(N) if above[s] and (s == 0 or not above[s - 1])]
    assert len(onset) >= 3
    assert list(steps[gens == 1]) == onset
    want, last = [], -refract
    for s in range(N):
        t = dt * (s + 1)
        if above[s] and t + dt / 2 >= last + refract:
            want.append(s + 1)
            last = t
    got = list(steps[gens == 0])
    assert got == want
    # edge 0 fires more than once per excursion when an excursion outlasts
    # refractT
    assert len(got) > len(onset)


# ---------------------------------------------------------------------------
# GPU against the oracle
# ---------------------------------------------------------------------------
def _gpu_vs_oracle(pkg, oracle_mod, specs, dt, nsteps, precision, mode, caAdvance=1, concs=None):
    eng = pkg.Engine(precision, kernel_mode=mode)
    ids = [eng.add_cell(s, dt, caAdvance) for s in specs]
    for cid, s in zip(ids, specs):
        eng.set_probes(cid, np.arange(s.n), nsteps)
    refs = []
    for k, (cid, s) in enumerate(zip(ids, specs)):
        r = oracle_mod.OracleCell(s, dt, caAdvance)
        for ch, conc in (concs[k] if concs else []):
            eng.add_conc(cid, ch, conc)
            r.add_conc(ch, conc)
        refs.append(r)
    eng.reinit(dt)
    eng.step(nsteps)
    spikes = eng.read_spikes()
    tol = FP64_TOL if precision == "fp64" else MIXED_TOL
    worst, nspk = 0.0, 0
    for cid, s, r in zip(ids, specs, refs):
        r.reinit(dt)
        ref = r.run(dt, nsteps, np.arange(s.n))
        got = eng.read_probes(cid)
        err = float(np.abs(got - ref).max())
        worst = max(worst, err)
        assert err <= tol, (cid, err)
        g, st = r.read_spikes()
        mine = spikes[spikes["cell"] == cid]
        nspk += len(st)
        for gen in set(g.tolist()) | set(mine["spikegen"].tolist()):
            a = list(mine[mine["spikegen"] == gen]["step"])
            b = list(st[g == gen])
            assert len(a) == len(b), (cid, gen)
            if precision == "fp64":
                assert a == b, (cid, gen)
            else:
                assert all(abs(x - y) <= 1 for x, y in zip(a, b)), (cid, gen)
    eng.close()
    return worst, nspk


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_non_integer_powers(pkg, oracle_mod, precision, mode):
    print(f"case: non-integer powers (3 cells), 800 steps, {precision}, {mode}", flush=True)
    rng = np.random.default_rng(64)
    specs = [power_cell(rng, n, float(rng.uniform(1.0e-9, 2.0e-9))) for n in (40, 300, 1100)]
    worst, nspk = _gpu_vs_oracle(pkg, oracle_mod, specs, 25e-6, 800, precision, mode)
    assert nspk > 0
    print(f"worst max|dVm| {worst:.3e} V, {nspk} spikes")


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_ca_advance_zero(pkg, oracle_mod, precision, mode):
    print(f"case: caAdvance = 0 (3 Ca cells), 600 steps, {precision}, {mode}", flush=True)
    rng = np.random.default_rng(65)
    specs = [workloads.random_ca_cell(rng, n, inject=float(rng.uniform(0.6e-9, 1.2e-9))) for n in (30, 200, 700)]
    worst, nspk = _gpu_vs_oracle(pkg, oracle_mod, specs, 20e-6, 600, precision, mode, caAdvance=0)
    print(f"worst max|dVm| {worst:.3e} V, {nspk} spikes")


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_z_gate_fallback(pkg, oracle_mod, precision, mode):
    print(f"case: Ca-dependent Z gate without a pool (3 cells), 600 steps, {precision}, {mode}", flush=True)
    rng = np.random.default_rng(66)
    specs = [zgate_cell(rng, n, float(rng.uniform(1.0e-9, 2.0e-9))) for n in (25, 160, 500)]
    concs = [ext_conc(s) for s in specs]
    worst, nspk = _gpu_vs_oracle(pkg, oracle_mod, specs, 25e-6, 600, precision, mode, concs=concs)
    print(f"worst max|dVm| {worst:.3e} V, {nspk} spikes")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_spikegen_edge_zero(pkg, oracle_mod, mode):
    print(f"case: SpikeGen edgeTriggered 0 and 1, 3000 steps, fp64, {mode}", flush=True)
    rng = np.random.default_rng(67)
    specs = [edge0_cell(rng, n, 1.2e-9, r) for n, r in ((40, 0.3e-3), (250, 0.0), (90, 0.2e-3))]
    worst, nspk = _gpu_vs_oracle(pkg, oracle_mod, specs, 25e-6, 3000, "fp64", mode)
    assert nspk > 20
