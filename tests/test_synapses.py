"""Objects that send (Gk, Ek) to a compartment on every tick, beside the
solver: Leakage (biophysics/Leakage.cpp:90-102, its Gk folded into EmByRm at
setup, hsolve/HSolvePassive.cpp:138-156) and SynChan (biophysics/SynChan.cpp)
fed by a SimpleSynHandler (synapse/SimpleSynHandler.cpp) whose synapses
(synapse/Synapse.cpp) receive spikes from the host or from SpikeGens of
cells in the same engine.

CPU tests pin the oracle against what the mathematics fixes: the passive
steady state of a compartment with a leak (fold plus per-tick send), the
closed-form sums of the SynChan recurrences after one event, the peak of a
unit event's conductance (the norm makes it Gbar as dt -> 0, for tau1 != tau2,
tau1 == tau2 and tau2 == 0), and the tick at which an event is first seen
(the first tick whose end time reaches it, never the tick that sent it).
GPU tests compare the HIP path (through the C ABI) with the oracle: FP64
within 1e-6 V with identical spike steps, mixed within 0.1 mV and one step;
networks against oracle cells stepped one tick at a time, each spike at
t = dt * step delivered to its targets as the event (t + delay, weight).
TimeTable sources (builtins/TimeTable.cpp:187-205): the host's send plan
against a tick-by-tick restatement on the CPU, and the object layer against
an oracle cell that receives each send after the tick that makes it.
"""
import math

import numpy as np
import pytest

from moose_core_amd import workloads
from moose_core_amd._abi import COMPARTMENT_DT, LEAK_DT, SYNAPSE_DT, SYNCHAN_DT
from moose_core_amd.cellspec import CellSpec

FP64_TOL = 1e-6
MIXED_TOL = 1e-4


def one_compartment(Rm=1e9, Cm=1e-11, Em=-0.065, V=-0.065):
    c = np.zeros(1, COMPARTMENT_DT)
    c["Ra"], c["Rm"], c["Cm"], c["Em"], c["initVm"], c["Vm"] = 1e6, Rm, Cm, Em, V, V
    return CellSpec(c, np.zeros((0, 2), np.int32))


def leak(comp, Gk=0.0, Ek=-0.07, Gbar=0.0, mod=1.0):
    r = np.zeros(1, LEAK_DT)[0]
    r["compartment"], r["Gk"], r["Ek"], r["Gbar"], r["modulation"] = comp, Gk, Ek, Gbar, mod
    return r


def synchan(comp, nsyn, Gbar=1e-9, Ek=0.0, tau1=1e-3, tau2=3e-3, mod=1.0):
    r = np.zeros(1, SYNCHAN_DT)[0]
    r["compartment"], r["n_synapses"] = comp, nsyn
    r["Gbar"], r["Ek"], r["tau1"], r["tau2"], r["modulation"] = Gbar, Ek, tau1, tau2, mod
    return r


def synapses(pairs):
    return np.array(pairs, SYNAPSE_DT).reshape(-1)


# ---------------------------------------------------------------------------
# Leakage
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("Gk_setup", [0.0, 2e-10])
def test_leak_steady_state(oracle_mod, Gk_setup):
    """Cm dV/dt = (Em - V)/Rm + Gk_setup*Ek + G*(Ek - V) with G = Gbar*mod:
    the fold adds to the right-hand side only, the per-tick send is a
    conductance.  The trace settles at the closed-form fixed point."""
    spec = one_compartment()
    G, Ek = 3e-9 * 0.5, -0.08
    spec.leaks = np.array([leak(0, Gk=Gk_setup, Ek=Ek, Gbar=3e-9, mod=0.5)], LEAK_DT)
    o = oracle_mod.OracleCell(spec, 50e-6)
    o.reinit(50e-6)
    v = o.run(50e-6, 4000, [0])[:, 0]
    Rm, Em = 1e9, -0.065
    vinf = (Em / Rm + Gk_setup * Ek + G * Ek) / (1.0 / Rm + G)
    assert abs(v[-1] - vinf) < 1e-9 * abs(vinf)
    # Im reads the fold and the last tick's send
    assert abs(o.get("Im", [0])[0]) < 1e-18


def test_leak_is_a_host_gkek_every_tick(oracle_mod):
    """A Leakage with Gk = 0 at setup is, bit for bit, a host addGkEk of
    (Gbar * modulation, Ek) before every tick."""
    rng = np.random.default_rng(3)
    base = workloads.random_hh_cell(rng, 25, inject=0.2e-9)
    spec = workloads.random_hh_cell(np.random.default_rng(3), 25, inject=0.2e-9)
    spec.leaks = np.array([leak(4, Ek=-0.07, Gbar=2e-9, mod=0.7), leak(11, Ek=-0.05, Gbar=1e-9)], LEAK_DT)
    a = oracle_mod.OracleCell(spec, 25e-6)
    a.reinit(25e-6)
    ta = a.run(25e-6, 400, np.arange(25))
    b = oracle_mod.OracleCell(base, 25e-6)
    b.reinit(25e-6)
    tb = []
    for _ in range(400):
        b.add_gkek(4, 2e-9 * 0.7, -0.07)
        b.add_gkek(11, 1e-9 * 1.0, -0.05)
        tb.append(b.run(25e-6, 1, np.arange(25))[0])
    np.testing.assert_array_equal(ta, np.array(tb))


# ---------------------------------------------------------------------------
# SynChan
# ---------------------------------------------------------------------------
def syn_cell(tau1, tau2, Gbar=1e-9, w=1.0, delay=0.0):
    spec = one_compartment()
    spec.synchans = np.array([synchan(0, 1, Gbar=Gbar, tau1=tau1, tau2=tau2)], SYNCHAN_DT)
    spec.synapses = synapses([(w, delay)])
    return spec


@pytest.mark.parametrize("tau1,tau2", [(1e-3, 3e-3), (2e-3, 2e-3), (1.5e-3, 0.0), (4e-3, 0.5e-3)])
def test_unit_event_peaks_at_gbar(oracle_mod, tau1, tau2):
    """A unit event (activation 1/dt for one tick) gives a conductance whose
    peak tends to Gbar as dt -> 0: the norm scales the dual exponential (or
    the alpha function, or the single exponential) to unit peak."""
    dt, Gbar = 1e-6, 2e-9
    o = oracle_mod.OracleCell(syn_cell(tau1, tau2, Gbar=Gbar), dt)
    o.reinit(dt)
    o.add_spike(0, 0, 0.0)
    gk = []
    for _ in range(int(12 * max(tau1, tau2) / dt)):
        o.run(dt, 1)
        gk.append(o.get("GkSyn", [0])[0])
    assert abs(max(gk) - Gbar) < 2e-3 * Gbar


def test_synchan_recurrence_closed_form(oracle_mod):
    """After one event of weight w seen on tick k0: X_k = (w/dt) tau1 (1-a)
    a^m, Y_k = (w/dt) tau1 (1-a) tau2 (1-b) (b^(m+1) - a^(m+1)) / (b - a),
    m = k - k0, a = exp(-dt/tau1), b = exp(-dt/tau2)."""
    dt, tau1, tau2, w = 25e-6, 1e-3, 3e-3, 0.7
    o = oracle_mod.OracleCell(syn_cell(tau1, tau2, w=w), dt)
    o.reinit(dt)
    o.add_spike(0, 0, 0.0)
    a, b = math.exp(-dt / tau1), math.exp(-dt / tau2)
    c = (w / dt) * tau1 * (1 - a)
    for m in range(300):
        o.run(dt, 1)
        X, Y = o.get("XSyn", [0])[0], o.get("YSyn", [0])[0]
        assert X == pytest.approx(c * a ** m, rel=1e-12)
        assert Y == pytest.approx(c * tau2 * (1 - b) * (b ** (m + 1) - a ** (m + 1)) / (b - a), rel=1e-10)


def test_event_seen_on_the_first_tick_reaching_it(oracle_mod):
    """An event at time tau is popped on the first tick whose end time
    dt * k reaches tau; one sent at a time already past is seen on the next
    tick.  Y is zero before and positive from that tick on."""
    dt = 50e-6
    for t_spike, delay, first in ((0.0, 0.0, 1), (3 * dt, 0.0, 3), (3 * dt, 1.5e-4, 6), (7.3 * dt, 1e-4, 10)):
        o = oracle_mod.OracleCell(syn_cell(1e-3, 2e-3, delay=delay), dt)
        o.reinit(dt)
        ys = []
        if t_spike == 0.0:
            o.add_spike(0, 0, t_spike)
        for k in range(1, 15):
            if t_spike > 0.0 and k == 1:
                o.add_spike(0, 0, t_spike)
            o.run(dt, 1)
            ys.append(o.get("YSyn", [0])[0])
        seen = next(k for k, y in enumerate(ys, 1) if y > 0.0)
        assert seen == first, (t_spike, delay)


def test_reinit_drops_queued_events(oracle_mod):
    o = oracle_mod.OracleCell(syn_cell(1e-3, 2e-3, delay=1e-3), 50e-6)
    o.reinit(50e-6)
    o.add_spike(0, 0, 0.0)
    o.reinit(50e-6)
    o.run(50e-6, 100)
    assert o.get("YSyn", [0])[0] == 0.0


# ---------------------------------------------------------------------------
# cells with synapses, networks
# ---------------------------------------------------------------------------
def syn_hh_cell(rng, n, nsyn_chans=3, nsyn=4, leaks=2):
    spec = workloads.random_hh_cell(rng, n, inject=0.0)
    comps = rng.choice(n, size=nsyn_chans, replace=False)
    recs, syn = [], []
    for c in comps:
        recs.append(synchan(int(c), nsyn, Gbar=float(rng.uniform(2e-9, 6e-9)), Ek=0.0,
                            tau1=float(rng.uniform(0.5e-3, 2e-3)), tau2=float(rng.uniform(2e-3, 6e-3))))
        syn += [(float(rng.uniform(0.5, 1.5)), float(rng.uniform(0.0, 2e-3))) for _ in range(nsyn)]
    spec.synchans = np.array(recs, SYNCHAN_DT)
    spec.synapses = synapses(syn)
    lc = rng.choice(n, size=leaks, replace=False)
    spec.leaks = np.array([leak(int(c), Gk=0.0, Ek=-0.07, Gbar=float(rng.uniform(0.5e-9, 2e-9))) for c in lc],
                          LEAK_DT)
    return spec


def poisson_inputs(rng, spec, rate, t_end):
    """(synchan, synapse, time) host spikes."""
    out = []
    for k, r in enumerate(spec.synchans):
        for j in range(int(r["n_synapses"])):
            t = rng.exponential(1.0 / rate)
            while t < t_end:
                out.append((k, j, float(t)))
                t += rng.exponential(1.0 / rate)
    return out


def run_network(pkg, oracle_mod, specs, conns, host, dt, nsteps, precision, mode, chunks=(1.0,)):
    """GPU engine against oracle cells stepped tick by tick, spikes routed
    as the reference's messages deliver them."""
    eng = pkg.Engine(precision, kernel_mode=mode)
    ids = [eng.add_cell(s, dt) for s in specs]
    for (a, g, b, k, j) in conns:
        eng.connect_spike(ids[a], g, ids[b], k, j)
    for cid, s in zip(ids, specs):
        eng.set_probes(cid, np.arange(s.n), nsteps)
    eng.reinit(dt)
    for (c, k, j, t) in host:
        eng.add_spike(ids[c], k, j, t)
    done = 0
    for f in chunks:
        m = int(round(f * nsteps)) - done
        eng.step(m)
        done += m
    spikes = eng.read_spikes()
    gpu = [eng.read_probes(cid) for cid in ids]
    gspk = [sorted(spikes[spikes["cell"] == cid]["step"].tolist()) for cid in ids]
    eng.close()

    refs = [oracle_mod.OracleCell(s, dt) for s in specs]
    for r in refs:
        r.reinit(dt)
    for (c, k, j, t) in host:
        refs[c].add_spike(k, j, t)
    traces = [[] for _ in specs]
    rspk = [[] for _ in specs]
    for step in range(1, nsteps + 1):
        for ci, r in enumerate(refs):
            traces[ci].append(r.run(dt, 1, np.arange(specs[ci].n))[0])
        # spikes of this tick, by cell then spike generator, to their targets
        for ci, r in enumerate(refs):
            gens, steps = r.read_spikes()
            for g, sstep in zip(gens, steps):
                rspk[ci].append(int(sstep))
                for (a, gg, b, k, j) in conns:
                    if a == ci and gg == g:
                        refs[b].add_spike(k, j, dt * float(sstep))
    return gpu, [np.array(t) for t in traces], [sorted(x) for x in gspk], [sorted(x) for x in rspk]


def test_oracle_network_harness_routes_by_message(oracle_mod):
    """The harness's routing on the oracle: a cell driven by a strong
    synapse from a spiking cell fires after the source, no earlier than the
    delay allows."""
    rng = np.random.default_rng(1)
    src = workloads.random_hh_cell(rng, 20, inject=1.5e-9)
    dst = workloads.random_hh_cell(rng, 20, inject=0.0)
    dst.synchans = np.array([synchan(0, 1, Gbar=40e-9, tau1=0.5e-3, tau2=1e-3)], SYNCHAN_DT)
    dst.synapses = synapses([(1.0, 2e-3)])
    dt = 25e-6
    refs = [oracle_mod.OracleCell(s, dt) for s in (src, dst)]
    for r in refs:
        r.reinit(dt)
    first_src, dsteps = None, []
    for step in range(1, 1201):
        for r in refs:
            r.run(dt, 1)
        for s in refs[0].read_spikes()[1]:
            first_src = first_src or int(s)
            refs[1].add_spike(0, 0, dt * float(s))
        dsteps += list(refs[1].read_spikes()[1])
    assert first_src is not None and len(dsteps) > 0
    assert dsteps[0] > first_src + 2e-3 / dt


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_fp64_synapses_host_spikes(pkg, oracle_mod, mode):
    print(f"case: 3 HH cells with SynChans + leaks, host Poisson spikes, 1000 steps, fp64, {mode}", flush=True)
    rng = np.random.default_rng(41)
    specs = [syn_hh_cell(rng, n) for n in (30, 120, 257)]
    specs.append(workloads.hh_ball_tree(inject=0.6e-9))
    host = []
    for c, s in enumerate(specs[:3]):
        host += [(c, k, j, t) for (k, j, t) in poisson_inputs(rng, s, 300.0, 0.025)]
    gpu, ref, gs, rs = run_network(pkg, oracle_mod, specs, [], host, 25e-6, 1000, "fp64", mode, (0.3, 1.0))
    for g, r in zip(gpu, ref):
        assert np.abs(g - r).max() <= FP64_TOL
    assert gs == rs
    assert sum(len(x) for x in rs[:3]) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_fp64_synapses_beside_chan2d(pkg, oracle_mod, mode):
    """Cells whose compartments carry HHChannel2D objects as well as
    SynChans and Leakage objects (the external kernel sums all three kinds),
    next to cells with only SynChans and leaks (summed in the step kernels)."""
    print(f"case: K-AHP 2-D cells + SynChans + leaks, host spikes, 800 steps, fp64, {mode}", flush=True)
    rng = np.random.default_rng(43)
    specs = []
    for n in (40, 150):
        s = workloads.random_ca_cell(rng, n, inject=0.3e-9, kahp2d=True)
        extra = syn_hh_cell(rng, n)
        s.synchans, s.synapses, s.leaks = extra.synchans, extra.synapses, extra.leaks
        # one SynChan and one leak on a compartment that has a 2-D channel
        s.synchans["compartment"][0] = s.chans2d["compartment"][0]
        s.leaks["compartment"][0] = s.chans2d["compartment"][0]
        specs.append(s)
    specs.append(syn_hh_cell(rng, 90))
    host = []
    for c, s in enumerate(specs):
        host += [(c, k, j, t) for (k, j, t) in poisson_inputs(rng, s, 300.0, 0.02)]
    gpu, ref, gs, rs = run_network(pkg, oracle_mod, specs, [], host, 25e-6, 800, "fp64", mode, (0.5, 1.0))
    for g, r in zip(gpu, ref):
        assert np.abs(g - r).max() <= FP64_TOL
    assert gs == rs


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_mixed_synapses_host_spikes(pkg, oracle_mod, mode):
    print(f"case: SynChan cells, host spikes, 1000 steps, mixed, {mode}", flush=True)
    rng = np.random.default_rng(41)
    specs = [syn_hh_cell(rng, n) for n in (30, 120, 257)]
    host = []
    for c, s in enumerate(specs):
        host += [(c, k, j, t) for (k, j, t) in poisson_inputs(rng, s, 300.0, 0.025)]
    gpu, ref, gs, rs = run_network(pkg, oracle_mod, specs, [], host, 25e-6, 1000, "mixed", mode)
    for g, r, a, b in zip(gpu, ref, gs, rs):
        assert np.abs(g - r).max() <= MIXED_TOL
        assert len(a) == len(b) and all(abs(x - y) <= 1 for x, y in zip(a, b))


def ring_network(rng, ncells=4):
    """Cells in a ring: cell i's SpikeGen drives a strong synapse on cell
    i+1 (delays 1-3 ms) and a weaker one on cell i+2; cell 0 is driven by a
    constant current."""
    specs = []
    for i in range(ncells):
        s = workloads.random_hh_cell(rng, int(rng.integers(20, 90)), inject=1.2e-9 if i == 0 else 0.0)
        s.synchans = np.array([synchan(0, 2, Gbar=30e-9, tau1=0.5e-3, tau2=1.5e-3),
                               synchan(int(s.n // 2), 1, Gbar=5e-9, tau1=1e-3, tau2=3e-3, Ek=-0.07)], SYNCHAN_DT)
        s.synapses = synapses([(1.0, float(rng.uniform(1e-3, 3e-3))), (0.8, float(rng.uniform(1e-3, 3e-3))),
                               (1.0, float(rng.uniform(1e-3, 3e-3)))])
        specs.append(s)
    conns = []
    for i in range(ncells):
        conns.append((i, 0, (i + 1) % ncells, 0, 0))
        conns.append((i, 0, (i + 2) % ncells, 0, 1))
        conns.append((i, 0, (i + 1) % ncells, 1, 0))
    return specs, conns


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_fp64_spike_network(pkg, oracle_mod, mode):
    print(f"case: 4-cell ring network (SpikeGen -> synapse, 1-3 ms delays), 1600 steps, fp64, {mode}", flush=True)
    specs, conns = ring_network(np.random.default_rng(7))
    gpu, ref, gs, rs = run_network(pkg, oracle_mod, specs, conns, [], 25e-6, 1600, "fp64", mode, (0.37, 1.0))
    for g, r in zip(gpu, ref):
        assert np.abs(g - r).max() <= FP64_TOL
    assert gs == rs
    # activity travelled round the ring
    assert all(len(x) > 0 for x in rs)


@pytest.mark.gpu
def test_gpu_synapse_fields(pkg, oracle_mod):
    """SynChan and synapse fields through the solver: tau1, Gbar and a
    weight written between runs, Gk/X/Y read back."""
    print("case: SynChan fields (tau1, Gbar, weight) between runs, fp64", flush=True)
    rng = np.random.default_rng(5)
    spec = syn_hh_cell(rng, 60)
    host = poisson_inputs(rng, spec, 400.0, 0.02)
    dt = 25e-6
    eng = pkg.Engine("fp64")
    cid = eng.add_cell(spec, dt)
    eng.reinit(dt)
    ref = oracle_mod.OracleCell(spec, dt)
    ref.reinit(dt)
    for (k, j, t) in host:
        eng.add_spike(cid, k, j, t)
        ref.add_spike(k, j, t)
    eng.step(300)
    ref.run(dt, 300)
    idx = np.arange(len(spec.synchans))
    for f in ("GkSyn", "XSyn", "YSyn"):
        np.testing.assert_allclose(eng.get(cid, f, idx), ref.get(f, idx), rtol=1e-9, atol=1e-20, err_msg=f)
    for f, v in (("tau1", 1.7e-3), ("GbarSyn", 9e-9)):
        eng.set(cid, f, [1], [v])
        ref.set(f, [1], [v])
    eng.set(cid, "weight", [2], [2.5])
    ref.set("weight", [2], [2.5])
    eng.set(cid, "GbarLeak", [0], [3e-9])
    ref.set("GbarLeak", [0], [3e-9])
    eng.step(500)
    ref.run(dt, 500)
    v_gpu = eng.get(cid, "Vm", np.arange(spec.n))
    v_ref = ref.get("Vm", np.arange(spec.n))
    assert np.abs(v_gpu - v_ref).max() <= FP64_TOL
    np.testing.assert_allclose(eng.get(cid, "GkSyn", idx), ref.get("GkSyn", idx), rtol=1e-6, atol=1e-20)
    np.testing.assert_allclose(eng.get(cid, "Im", [0, 5]), ref.get("Im", [0, 5]), rtol=1e-6, atol=1e-15)
    eng.close()


# ---------------------------------------------------------------------------
# object layer
# ---------------------------------------------------------------------------
def build_pair(moose):
    """Two three-compartment HH cells, each with a SynChan on its soma fed
    by a SimpleSynHandler; each cell's SpikeGen drives the other's synapse,
    and a Leakage sits on cell b's dendrite.  Cell a is driven by current."""
    moose.reset()
    moose.Neutral("/library")
    gates = workloads.hh_gates()
    na = moose.HHChannel("/library/Na")
    na.Ek, na.Xpower, na.Ypower = 0.050, 3.0, 1.0
    k = moose.HHChannel("/library/K")
    k.Ek, k.Xpower = -0.077, 4.0
    for gate, g in ((na.gateX, gates[0]), (na.gateY, gates[1]), (k.gateX, gates[2])):
        gate.min, gate.max = g.min, g.max
        gate.tableA, gate.tableB = g.tableA, g.tableB
    cells = {}
    for name, inject in (("a", 1.0e-9), ("b", 0.0)):
        moose.Neutral(f"/{name}")
        comps = []
        for i, (L, d) in enumerate(((20e-6, 20e-6), (50e-6, 2e-6), (50e-6, 2e-6))):
            c = moose.Compartment(f"/{name}/c{i}")
            area = math.pi * d * L
            c.Rm, c.Cm, c.Ra = 1.0 / area, 0.01 * area, 4.0 * L / (math.pi * d * d)
            c.Em = c.initVm = c.Vm = -0.065
            comps.append(c)
            for proto, gbar in ((na, 1200.0), (k, 360.0)):
                ch = moose.copy(proto, c, proto.name)
                ch.Gbar = gbar * area
                moose.connect(ch, "channel", c, "channel")
        comps[0].inject = inject
        moose.connect(comps[0], "axial", comps[1], "raxial")
        moose.connect(comps[1], "axial", comps[2], "raxial")
        sg = moose.SpikeGen(f"/{name}/spike")
        sg.threshold, sg.refractT = 0.0, 1e-3
        moose.connect(comps[0], "VmOut", sg, "Vm")
        syn = moose.SynChan(f"/{name}/c0/syn")
        syn.Gbar, syn.Ek, syn.tau1, syn.tau2 = 20e-9, 0.0, 0.5e-3, 2e-3
        moose.connect(syn, "channel", comps[0], "channel")
        h = moose.SimpleSynHandler(f"/{name}/c0/syn/handler")
        h.numSynapses = 2
        moose.connect(h, "activationOut", syn, "activation")
        cells[name] = (comps, sg, syn, h)
    lk = moose.Leakage("/b/c2/leak")
    lk.Gbar, lk.Ek = 1e-9, -0.07
    moose.connect(lk, "channel", cells["b"][0][2], "channel")
    ha, hb = cells["a"][3], cells["b"][3]
    hb.synapse[0].delay, hb.synapse[0].weight = 1.5e-3, 1.0
    ha.synapse[1].delay, ha.synapse[1].weight = 2.5e-3, 0.5
    moose.connect(cells["a"][1], "spikeOut", hb.synapse[0], "addSpike")
    moose.connect(cells["b"][1], "spikeOut", ha.synapse[1], "addSpike")
    return cells


def test_layer_collects_synchans(pkg):
    from moose_core_amd import moose
    cells = build_pair(moose)
    hs = moose.HSolve("/b/hsolve")
    hs.dt = 25e-6
    spec, objs = hs.build_spec(moose._deep_search_compartment(moose.element("/b")))
    assert objs["synchans"] == [cells["b"][2]]
    sc = spec.synchans[0]
    assert (sc["compartment"], sc["n_synapses"], sc["tau1"], sc["tau2"], sc["Gbar"]) == (0, 2, 0.5e-3, 2e-3, 20e-9)
    np.testing.assert_array_equal(spec.synapses["delay"], [1.5e-3, 0.0])
    np.testing.assert_array_equal(spec.synapses["weight"], [1.0, 1.0])
    lk = spec.leaks[0]
    assert (lk["compartment"], lk["Gbar"], lk["Ek"], lk["modulation"], lk["Gk"]) == (2, 1e-9, -0.07, 1.0, 0.0)
    moose.reset()


@pytest.mark.gpu
def test_gpu_layer_synapse_pair(pkg, oracle_mod):
    """Two solvers joined by spike messages through the object layer,
    against oracle cells on the layer's descriptions stepped tick by tick
    with the same routing."""
    print("case: object layer, 2 HH cells joined by SpikeGen -> synapse messages, fp64", flush=True)
    from moose_core_amd import moose
    cells = build_pair(moose)
    moose.setClock(8, 25e-6)
    tabs = {}
    for name in ("a", "b"):
        t = moose.Table(f"/{name}/vm")
        moose.connect(t, "requestOut", cells[name][0][0], "getVm")
        tabs[name] = t
    solvers = {}
    for name in ("a", "b"):
        hs = moose.HSolve(f"/{name}/hsolve")
        hs.dt = 25e-6
        hs.target = f"/{name}"
        solvers[name] = hs
    specs = [solvers["a"]._spec, solvers["b"]._spec]
    import copy
    specs = [copy.deepcopy(s) for s in specs]
    moose.reinit()
    moose.start(0.03)
    dt, nsteps = 25e-6, int(round(0.03 / 25e-6))
    refs = [oracle_mod.OracleCell(s, dt) for s in specs]
    for r in refs:
        r.reinit(dt)
    conns = [(0, 0, 1, 0, 0), (1, 0, 0, 0, 1)]
    tr = [[], []]
    for step in range(1, nsteps + 1):
        for ci, r in enumerate(refs):
            tr[ci].append(r.run(dt, 1, [0])[0, 0])
        for ci, r in enumerate(refs):
            gens, steps = r.read_spikes()
            for g, s in zip(gens, steps):
                for (a, gg, b, k, j) in conns:
                    if a == ci and gg == g:
                        refs[b].add_spike(k, j, dt * float(s))
    for ci, name in enumerate(("a", "b")):
        v = tabs[name].vector[1:]
        assert len(v) == nsteps
        assert np.abs(v - np.array(tr[ci])).max() <= FP64_TOL
    # b fired from a's input
    assert tabs["b"].vector.max() > 0.0
    moose.reset()


@pytest.mark.gpu
def test_gpu_step_cells_with_synapses(pkg, oracle_mod):
    """hsg_step_cells (the serving path) on SynChan cells advancing by
    different step counts: each cell's handler queue is popped over its own
    ticks.  Spike connections are refused there."""
    print("case: step_cells on 3 SynChan cells with different step counts, fp64", flush=True)
    rng = np.random.default_rng(61)
    specs = [syn_hh_cell(rng, n) for n in (40, 90, 160)]
    dt = 25e-6
    eng = pkg.Engine("fp64")
    ids = [eng.add_cell(s, dt) for s in specs]
    eng.reinit(dt)
    refs = [oracle_mod.OracleCell(s, dt) for s in specs]
    for r in refs:
        r.reinit(dt)
    for c, s in enumerate(specs):
        for (k, j, t) in poisson_inputs(rng, s, 400.0, 0.02):
            eng.add_spike(ids[c], k, j, t)
            refs[c].add_spike(k, j, t)
    for steps in ((100, 250, 40), (300, 0, 500)):
        eng.step_cells(ids, list(steps))
        for r, m in zip(refs, steps):
            if m:
                r.run(dt, m)
    for cid, s, r in zip(ids, specs, refs):
        v = eng.get(cid, "Vm", np.arange(s.n))
        assert np.abs(v - r.get("Vm", np.arange(s.n))).max() <= FP64_TOL
    eng.connect_spike(ids[0], 0, ids[1], 0, 0)
    with pytest.raises(pkg.HSolveError):
        eng.step_cells(ids, [1, 1, 1])
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_probes_of_other_fields(pkg, oracle_mod, mode):
    """Device probes on a CaConc's Ca, a SynChan's Gk and a VClamp's
    current beside Vm, every third step: equal to the oracle's values read
    after the same steps."""
    print(f"case: probes of Vm/Ca/GkSyn/current on a Ca cell, 300 steps every 3, fp64, {mode}", flush=True)
    from moose_core_amd._abi import VCLAMP_DT
    rng = np.random.default_rng(77)
    spec = workloads.random_ca_cell(rng, 70, inject=0.4e-9)
    extra = syn_hh_cell(rng, 70)
    spec.synchans, spec.synapses = extra.synchans, extra.synapses
    vc = np.zeros(1, VCLAMP_DT)
    vc[0] = (5, 0, 0.2e-3, 20e-6, 5e-6, spec.compartments["Cm"][5] * 1e4)
    spec.vclamps = vc
    dt, nsteps, every = 25e-6, 300, 3
    fields = ["Vm", "Ca", "GkSyn", "current", "Vm"]
    idx = [0, 3, 1, 0, 40]
    host = poisson_inputs(rng, spec, 500.0, nsteps * dt)
    eng = pkg.Engine("fp64", kernel_mode=mode)
    cid = eng.add_cell(spec, dt)
    eng.reinit(dt)
    eng.set(cid, "command", [0], [-0.05])
    for (k, j, t) in host:
        eng.add_spike(cid, k, j, t)
    eng.set_probe_fields(cid, fields, idx, nsteps // every, every)
    eng.step(nsteps)
    got = eng.read_probes(cid)
    eng.close()
    ref = oracle_mod.OracleCell(spec, dt)
    ref.reinit(dt)
    ref.set("command", [0], [-0.05])
    for (k, j, t) in host:
        ref.add_spike(k, j, t)
    want = []
    for step in range(1, nsteps + 1):
        ref.run(dt, 1)
        if step % every == 0:
            want.append([ref.get(f, [i])[0] for f, i in zip(fields, idx)])
    want = np.array(want)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-12)


# ---------------------------------------------------------------------------
# TimeTable sources (builtins/TimeTable.cpp)
# ---------------------------------------------------------------------------
def timetable_ticks(vec, every, dt, nsteps):
    """Brute force: TimeTable::process on every step whose time is a
    multiple of the tick's, from the start."""
    pos, sends, states = 0, [], []
    for u in range(1, nsteps + 1):
        if u % every:
            continue
        state = 0.0
        if pos < len(vec) and dt * u >= vec[pos]:
            sends.append((float(vec[pos]), u))
            pos += 1
            state = 1.0
        states.append(state)
    return sends, pos, (states[-1] if states else 0.0)


@pytest.mark.parametrize("every", [1, 4])
def test_timetable_sends_match_tick_by_tick(every):
    from moose_core_amd.moose import _timetable_sends
    dt = 25e-6
    # two times inside one tick (the second waits a tick), one on a tick,
    # time 0, and a late one
    vec = np.array([0.0, 1e-3, 1.01e-3, 1.02e-3, dt * 80, 3.3e-3, 9.0e-3])
    nsteps = 300
    want = timetable_ticks(vec, every, dt, nsteps)
    assert _timetable_sends(vec, 0, 0.0, every, dt, 0, nsteps) == want
    # the same run in pieces (a piece may hold no tick of the TimeTable)
    got, pos, state, s0 = [], 0, 0.0, 0
    for piece in (3, 1, 77, 2, 150, 67):
        sends, pos, state = _timetable_sends(vec, pos, state, every, dt, s0, piece)
        got += sends
        s0 += piece
    assert (got, pos, state) == want
    # time 0 goes out on the first tick; the three times inside one tick
    # leave on consecutive ticks
    assert want[0][0] == (0.0, every)
    u1 = [u for t, u in want[0] if t in (1e-3, 1.01e-3, 1.02e-3)]
    assert u1[1] - u1[0] == every and u1[2] - u1[1] == every
    # at most one send per tick, never before its listed time
    us = [u for _, u in want[0]]
    assert len(set(us)) == len(us)
    assert all(dt * u >= t for t, u in want[0])


@pytest.mark.gpu
@pytest.mark.parametrize("tt_dt", [25e-6, 1e-4])
def test_gpu_layer_timetable_to_synapses(pkg, oracle_mod, tt_dt):
    """TimeTable eventOut -> Synapse addSpike through the object layer
    (delays 0, 0.04 ms and 3 ms), against an oracle cell stepped one tick at
    a time, the TimeTable's sends applied after the tick that makes them."""
    print(f"case: object layer, TimeTable -> 3 synapses, TimeTable dt {tt_dt}, fp64", flush=True)
    from moose_core_amd import moose
    moose.reset()
    moose.Neutral("/m")
    c = moose.Compartment("/m/soma")
    c.Rm, c.Cm, c.Em, c.initVm, c.Vm = 1e9, 1e-11, -0.065, -0.065, -0.065
    syn = moose.SynChan("/m/soma/syn")
    syn.Gbar, syn.Ek, syn.tau1, syn.tau2 = 5e-9, 0.0, 0.5e-3, 2e-3
    moose.connect(syn, "channel", c, "channel")
    h = moose.SimpleSynHandler("/m/soma/syn/handler")
    h.numSynapses = 3
    moose.connect(h, "activationOut", syn, "activation")
    delays = (0.0, 0.04e-3, 3e-3)
    weights = (1.0, 0.5, 2.0)
    for y, dly, w in zip(h.synapse, delays, weights):
        y.delay, y.weight = dly, w
    tt = moose.TimeTable("/m/tt")
    times = [0.0, 2e-3, 2.01e-3, 2.02e-3, 5e-3, 5.5e-3, 11.3e-3]
    tt.vector = times
    moose.setClock(tt.tick, tt_dt)
    for y in h.synapse:
        moose.connect(tt, "eventOut", y, "addSpike")
    tab = moose.Table("/m/vm")
    moose.connect(tab, "requestOut", c, "getVm")
    moose.useClock(7, "/m/vm")
    dt = 25e-6
    moose.setClock(7, dt)
    hs = moose.HSolve("/m/hsolve")
    hs.dt = dt
    hs.target = "/m"
    import copy
    spec = copy.deepcopy(hs._spec)
    moose.reinit()
    moose.start(0.006)
    moose.start(0.014)
    nsteps = int(round(0.02 / dt))
    every = int(round(tt_dt / dt))
    ref = oracle_mod.OracleCell(spec, dt)
    ref.reinit(dt)
    pos, tr = 0, []
    for u in range(1, nsteps + 1):
        tr.append(ref.run(dt, 1, [0])[0, 0])
        if u % every == 0 and pos < len(times) and dt * u >= times[pos]:
            for j in range(3):
                ref.add_spike(0, j, times[pos])
            pos += 1
    v = tab.vector[1:]
    assert len(v) == nsteps
    assert np.abs(v - np.array(tr)).max() <= FP64_TOL
    assert np.ptp(tr) > 1e-3          # the synaptic input moved Vm
    assert pos == len(times)
    moose.reset()


def two_handler_model(moose):
    """One compartment whose SynChan is fed by two SimpleSynHandlers (2 and 1
    synapses); a TimeTable drives the second handler's synapse and the
    first handler's second synapse."""
    moose.reset()
    moose.Neutral("/m")
    c = moose.Compartment("/m/soma")
    c.Rm, c.Cm, c.Em, c.initVm, c.Vm = 1e9, 1e-11, -0.065, -0.065, -0.065
    syn = moose.SynChan("/m/soma/syn")
    syn.Gbar, syn.Ek, syn.tau1, syn.tau2 = 5e-9, 0.0, 0.5e-3, 2e-3
    moose.connect(syn, "channel", c, "channel")
    hs_ = []
    for name, n, pars in (("h1", 2, ((0.5e-3, 1.0), (1e-3, 0.5))), ("h2", 1, ((0.0, 2.0),))):
        h = moose.SimpleSynHandler(f"/m/soma/syn/{name}")
        h.numSynapses = n
        moose.connect(h, "activationOut", syn, "activation")
        for y, (dly, w) in zip(h.synapse, pars):
            y.delay, y.weight = dly, w
        hs_.append(h)
    tt = moose.TimeTable("/m/tt")
    tt.vector = [1e-3, 3e-3, 3.05e-3, 7e-3]
    moose.connect(tt, "eventOut", hs_[1].synapse[0], "addSpike")
    moose.connect(tt, "eventOut", hs_[0].synapse[1], "addSpike")
    return c, hs_, tt


def test_layer_synchan_with_two_handlers(pkg):
    from moose_core_amd import moose
    c, (h1, h2), _ = two_handler_model(moose)
    hs = moose.HSolve("/m/hsolve")
    hs.dt = 25e-6
    spec, objs = hs.build_spec(moose._deep_search_compartment(moose.element("/m")))
    assert spec.synchans[0]["n_synapses"] == 3
    np.testing.assert_array_equal(spec.synapses["weight"], [1.0, 0.5, 2.0])
    np.testing.assert_array_equal(spec.synapses["delay"], [0.5e-3, 1e-3, 0.0])
    moose.reset()


@pytest.mark.gpu
def test_gpu_layer_synchan_with_two_handlers(pkg, oracle_mod):
    """Two handlers on one SynChan through the object layer (a TimeTable and
    a host addSpike as sources) against an oracle cell holding the handlers'
    synapses in message order as one list.  The reference adds the two
    handlers' sums to the SynChan's activation one after the other; one
    queue sums all due events in time order, which can differ in the last
    bit of the activation, well inside the FP64 bound."""
    print("case: object layer, SynChan fed by 2 SimpleSynHandlers, fp64", flush=True)
    from moose_core_amd import moose
    c, (h1, h2), tt = two_handler_model(moose)
    tab = moose.Table("/m/vm")
    moose.connect(tab, "requestOut", c, "getVm")
    dt = 25e-6
    moose.setClock(8, dt)
    hs = moose.HSolve("/m/hsolve")
    hs.dt = dt
    hs.target = "/m"
    import copy
    spec = copy.deepcopy(hs._spec)
    moose.reinit()
    h1.synapse[0].addSpike(4e-3)
    moose.start(0.012)
    nsteps = int(round(0.012 / dt))
    ref = oracle_mod.OracleCell(spec, dt)
    ref.reinit(dt)
    ref.add_spike(0, 0, 4e-3)
    times, pos, tr = list(tt.vector), 0, []
    for u in range(1, nsteps + 1):
        tr.append(ref.run(dt, 1, [0])[0, 0])
        if pos < len(times) and dt * u >= times[pos]:
            ref.add_spike(0, 2, times[pos])     # h2.synapse[0]
            ref.add_spike(0, 1, times[pos])     # h1.synapse[1]
            pos += 1
    v = tab.vector[1:]
    assert len(v) == nsteps
    assert np.abs(v - np.array(tr)).max() <= FP64_TOL
    assert np.ptp(tr) > 1e-3
    moose.reset()
