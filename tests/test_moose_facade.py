"""The pymoose-style object layer (moose_core_amd.moose): a model built from
objects and messages reaches the solver exactly as the reference's HSolve
would read it, and runs through HSolve.target / reinit / start with zombie
field access.

CPU tests check the description the layer builds against the hand-built one
of the same model and run both through the oracle.  The GPU test runs the
model through the layer and compares with the oracle bit for bit (FP64).
"""
import numpy as np
import pytest

from moose_core_amd import workloads
from moose_core_amd.cellspec import passive_compartments


def build_ball_tree(moose, depth=7, inject=0.1e-9):
    """The ball-and-tree HH cell through the object API, in the creation
    and message order a MOOSE script would use (soma, then dendrites in
    heap order; channels copied from /library prototypes)."""
    moose.reset()
    moose.Neutral("/library")
    gates = workloads.hh_gates()
    na = moose.HHChannel("/library/Na")
    na.Ek = 0.050
    na.Xpower = 3.0
    na.Ypower = 1.0
    k = moose.HHChannel("/library/K")
    k.Ek = -0.077
    k.Xpower = 4.0
    for gate, g in ((na.gateX, gates[0]), (na.gateY, gates[1]), (k.gateX, gates[2])):
        gate.min, gate.max = g.min, g.max
        gate.tableA = g.tableA
        gate.tableB = g.tableB
    nd = 2 ** depth - 1
    n = nd + 1
    length = np.full(n, 10e-6)
    dia = np.full(n, 1e-6)
    length[0] = dia[0] = 20e-6
    comp, area = passive_compartments(length, dia)
    moose.Neutral("/cell")
    comps = []
    for i in range(n):
        c = moose.Compartment(f"/cell/c{i}")
        for f in ("Ra", "Rm", "Cm", "Em", "initVm", "Vm"):
            setattr(c, f, float(comp[f][i]))
        comps.append(c)
    comps[0].inject = inject
    moose.connect(comps[0], "axial", comps[1], "raxial")
    for kk in range(1, nd + 1):
        for ch in (2 * kk, 2 * kk + 1):
            if ch <= nd:
                moose.connect(comps[kk], "axial", comps[ch], "raxial")
    for i, c in enumerate(comps):
        for proto, gbar in ((na, 1200.0), (k, 360.0)):
            chan = moose.copy(proto, c, proto.name)
            chan.Gbar = gbar * area[i]
            moose.connect(chan, "channel", c, "channel")
    sg = moose.SpikeGen("/cell/spike")
    sg.threshold = 0.0
    sg.refractT = 0.0
    sg.edgeTriggered = 1
    moose.connect(comps[0], "VmOut", sg, "Vm")
    return comps, sg


def test_layer_builds_the_reference_description(pkg, oracle_mod):
    from moose_core_amd import moose
    comps, _ = build_ball_tree(moose)
    seed = moose._deep_search_compartment(moose.element("/cell"))
    # the last child first: the deepest search ends at the last compartment
    assert seed is comps[-1]
    hs = moose.HSolve("/cell/hsolve")
    spec, objs = hs.build_spec(seed)
    ref = workloads.hh_ball_tree()
    ref.seed = len(comps) - 1
    for f in ref.compartments.dtype.names:
        np.testing.assert_array_equal(spec.compartments[f], ref.compartments[f], err_msg=f)
    np.testing.assert_array_equal(spec.axial, ref.axial)
    for f in ref.channels.dtype.names:
        np.testing.assert_array_equal(spec.channels[f], ref.channels[f], err_msg=f)
    assert [g.id for g in spec.gates] == sorted(g.id for g in spec.gates)
    for g, r in zip(spec.gates, ref.gates):
        np.testing.assert_array_equal(g.tableA, r.tableA)
        np.testing.assert_array_equal(g.tableB, r.tableB)
    np.testing.assert_array_equal(spec.spikegens, ref.spikegens)
    # same model, same numbers through the oracle
    dt, n = 50e-6, 300
    a = oracle_mod.OracleCell(spec, dt)
    a.reinit(dt)
    b = oracle_mod.OracleCell(ref, dt)
    b.reinit(dt)
    np.testing.assert_array_equal(a.run(dt, n, np.arange(spec.n)), b.run(dt, n, np.arange(ref.n)))
    assert len(objs["channels"]) == 2 * len(comps)


def test_setup_rules(pkg):
    from moose_core_amd import moose
    build_ball_tree(moose, depth=2)
    hs = moose.HSolve("/cell/hsolve")
    with pytest.raises(ValueError, match="dt"):
        hs.target = "/cell"             # HSolve::setPath needs dt first
    with pytest.raises(ValueError):
        moose.Compartment("/nowhere/c")  # parent must exist
    na = moose.element("/library/Na")
    copy = moose.element("/cell/c1/Na")
    assert copy.gateX is na.gateX       # copies read the prototype gates
    moose.Neutral("/empty")
    hs.dt = 50e-6
    with pytest.raises(ValueError, match="No compartments"):
        hs.target = "/empty"


def test_caconc_shell_B(pkg):
    from moose_core_amd import moose
    moose.reset()
    p = moose.CaConc("/ca")
    p.B = 123.0
    assert p.shell_B() == 123.0          # no geometry: the model's B
    p.length, p.diameter, p.thick = 20e-6, 2e-6, 0.1e-6
    vol = np.pi * (2e-6) ** 2 * 20e-6 / 4 - np.pi * (0.9e-6) ** 2 * 20e-6
    assert p.shell_B() == pytest.approx(1.0 / (2 * 96485.3415 * vol), rel=1e-12)


@pytest.mark.gpu
def test_layer_runs_on_the_gpu_like_the_oracle(pkg, oracle_mod):
    print("case: object layer, ball tree, fp64, 400 + 200 steps", flush=True)
    from moose_core_amd import moose
    comps, sg = build_ball_tree(moose, inject=0.6e-9)
    vm = moose.Table("/cell/vm")
    moose.connect(vm, "requestOut", comps[0], "getVm")
    spk = moose.Table("/cell/spikes")
    moose.connect(sg, "spikeOut", spk, "spike")
    hs = moose.HSolve("/cell/hsolve")
    hs.dt = 50e-6
    hs.target = "/cell"
    spec = hs._spec
    moose.reinit()
    moose.start(400 * 50e-6)
    ref = oracle_mod.OracleCell(spec, 50e-6)
    ref.reinit(50e-6)
    tr = ref.run(50e-6, 400, np.arange(spec.n))
    soma = 0
    np.testing.assert_array_equal(vm.vector[1:], tr[:, soma])
    assert vm.vector[0] == spec.compartments["initVm"][soma]
    # zombie reads and writes go to the solver
    assert comps[5].Vm == tr[-1, 5]
    chan = moose.element("/cell/c3/Na")
    chan.Gbar = 2.0 * chan.Gbar
    ref.set("Gbar", [6], [chan.Gbar])           # channel 6 = Na of compartment 3
    comps[0].inject = 0.2e-9
    ref.set("inject", [0], [0.2e-9])
    moose.start(200 * 50e-6)
    tr2 = ref.run(50e-6, 200, np.arange(spec.n))
    np.testing.assert_array_equal(vm.vector[401:], tr2[:, soma])
    g, steps = ref.read_spikes()
    np.testing.assert_allclose(spk.vector, np.array(steps, dtype=np.float64) * 50e-6, rtol=0, atol=0)
    assert len(spk.vector) > 0
    moose.reset()
