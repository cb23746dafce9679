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


def test_setup_rules(pkg, capsys):
    from moose_core_amd import moose
    build_ball_tree(moose, depth=2)
    hs = moose.HSolve("/cell/hsolve")
    # HSolve::setPath needs dt first: it prints and returns (HSolve.cpp:317-321)
    hs.target = "/cell"
    assert "Must set 'dt' first" in capsys.readouterr().err
    assert hs.target == "" and hs._engine is None
    with pytest.raises(ValueError):
        moose.Compartment("/nowhere/c")  # parent must exist
    na = moose.element("/library/Na")
    copy = moose.element("/cell/c1/Na")
    assert copy.gateX is na.gateX       # copies read the prototype gates
    # a negative dt is refused with a message, the old dt kept (HSolve.cpp:409-418)
    hs.dt = -1.0
    assert "'dt' must be positive" in capsys.readouterr().err
    assert hs.dt == 0.0
    moose.Neutral("/empty")
    hs.dt = 50e-6
    # no compartment at or below the path: a warning, nothing set up
    for path in ("/empty", "/no/such/path"):
        hs.target = path
        assert "No compartments found" in capsys.readouterr().err
        assert hs.target == "" and hs.seed is None and hs._engine is None


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
    moose.setClock(8, 50e-6)                    # Tables every solver step
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


def add_ca_and_kc(moose, comps, kc_tables, gbar_kc=5e-9):
    """Per compartment: a CaConc fed by a CaL HHChannel (IkOut -> CaConc)
    and an HHChannel2D K(C) on Vm x [Ca] (Zindex VOLT_C1_INDEX) reading that
    pool through concen, all in a MOOSE script's creation/message order."""
    from moose_core_amd import workloads as W
    cag = W.ca_gates()
    cal = moose.HHChannel("/library/CaL")
    cal.Ek = 0.080
    cal.Xpower = 2.0
    g = cag[5]
    cal.gateX.min, cal.gateX.max = g.min, g.max
    cal.gateX.tableA = g.tableA
    cal.gateX.tableB = g.tableB
    kc = moose.HHChannel2D("/library/KC")
    kc.Ek = -0.080
    kc.Zpower = 1.0
    kc.Zindex = "VOLT_C1_INDEX"
    A, B = kc_tables
    for f in ("xmin", "xmax", "ymin", "ymax"):
        setattr(kc.gateZ, f, getattr(A, f))
    kc.gateZ.tableA = A.table
    kc.gateZ.tableB = B.table
    pools, kcs = [], []
    for i, c in enumerate(comps):
        p = moose.CaConc(c.path + "/ca")
        p.CaBasal = p.Ca = 1e-4
        p.tau = 0.020
        p.B = 5e12
        ch = moose.copy(cal, c, "CaL")
        ch.Gbar = 1e-9
        moose.connect(ch, "channel", c, "channel")
        moose.connect(ch, "IkOut", p, "current")
        k = moose.copy(kc, c, "KC")
        k.Gbar = gbar_kc
        moose.connect(k, "channel", c, "channel")
        moose.connect(p, "concOut", k, "concen")
        pools.append(p)
        kcs.append(k)
    return pools, kcs


def small_kc_tables():
    from moose_core_amd import workloads as W
    return W.kahp2d_tables(vdivs=60, cadivs=25)


def test_layer_builds_hhchannel2d(pkg, oracle_mod):
    from moose_core_amd import moose
    comps, _ = build_ball_tree(moose, depth=3)
    tabs = small_kc_tables()
    pools, kcs = add_ca_and_kc(moose, comps, tabs)
    hs = moose.HSolve("/cell/hsolve")
    spec, objs = hs.build_spec(moose._deep_search_compartment(moose.element("/cell")))
    assert len(spec.chans2d) == len(comps) and len(spec.tables2d) == 2
    np.testing.assert_array_equal(spec.tables2d[0].table, tabs[0].table)
    np.testing.assert_array_equal(spec.tables2d[1].table, tabs[1].table)
    cpos = {c.id: i for i, c in enumerate(objs["compartments"])}
    ppos = {p.id: i for i, p in enumerate(objs["pools"])}
    for k, ch in enumerate(objs["chans2d"]):
        r = spec.chans2d[k]
        comp = ch.path.rsplit("/", 1)[0]
        assert r["compartment"] == cpos[moose.element(comp).id]
        assert list(r["tableA"]) == [-1, -1, 0] and list(r["tableB"]) == [-1, -1, 1]
        assert list(r["dep"]) == [-1, -1, -1, -1, 0, 1]
        assert r["conc1"] == ppos[moose.element(comp + "/ca").id] and r["conc2"] == -1
        assert r["Gbar"] == 5e-9 and r["Zpower"] == 1.0
    # the oracle runs it and the 2-D channels conduct
    dt = 50e-6
    o = oracle_mod.OracleCell(spec, dt)
    o.reinit(dt)
    tr = o.run(dt, 200, np.arange(spec.n))
    assert np.isfinite(tr).all()
    assert (o.get("Gk2D", np.arange(len(kcs))) > 0).all()
    with pytest.raises(ValueError):
        kcs[0].Xindex = "NOT_AN_INDEX"


@pytest.mark.gpu
def test_layer_hhchannel2d_on_the_gpu(pkg, oracle_mod):
    print("case: object layer, ball tree depth 3 + CaL/CaConc/HHChannel2D, fp64, 300 steps", flush=True)
    from moose_core_amd import moose
    comps, _ = build_ball_tree(moose, depth=3, inject=0.8e-9)
    pools, kcs = add_ca_and_kc(moose, comps, small_kc_tables())
    moose.setClock(8, 50e-6)                    # Tables every solver step
    vm = moose.Table("/cell/vm")
    moose.connect(vm, "requestOut", comps[0], "getVm")
    hs = moose.HSolve("/cell/hsolve")
    hs.dt = 50e-6
    hs.target = "/cell"
    spec = hs._spec
    moose.reinit()
    moose.start(300 * 50e-6)
    ref = oracle_mod.OracleCell(spec, 50e-6)
    ref.reinit(50e-6)
    tr = ref.run(50e-6, 300, np.arange(spec.n))
    assert np.abs(vm.vector[1:] - tr[:, 0]).max() <= 1e-6
    k = hs._objs["chans2d"].index(kcs[3])
    assert kcs[3].Gk == pytest.approx(ref.get("Gk2D", [k])[0], rel=1e-6)
    kcs[3].Gbar = 2 * kcs[3].Gbar              # zombie write reaches the device
    assert kcs[3].Gbar == pytest.approx(1e-8)
    moose.reset()


# ---------------------------------------------------------------------------
# HHGate table generation (biophysics/HHGate.cpp:481-795)
# ---------------------------------------------------------------------------
EREST = -0.070
NA_M = [0.1e6 * (EREST + 0.025), -0.1e6, -1.0, -(EREST + 0.025), -0.010,    # alpha
        4e3, 0.0, 0.0, -EREST, 0.018,                                       # beta
        3000, -0.100, 0.050]


def _alpha(p, x):
    return (p[0] + p[1] * x) / (p[2] + np.exp((x + p[3]) / p[4]))


def test_setup_alpha_tables(pkg):
    from moose_core_amd import moose
    moose.reset()
    ch = moose.HHChannel("/na")
    ch.Xpower = 3
    g = ch.gateX
    g.setupAlpha(NA_M)
    A, B = g.tableA, g.tableB
    assert len(A) == 3001 and g.min == -0.1 and g.max == 0.05 and g.divs == 3000
    x = -0.1 + np.arange(3001) * (0.15 / 3000)
    alpha, beta = _alpha(NA_M[:5], x), _alpha(NA_M[5:10], x)
    # the alpha term has a 0/0 at x = -(EREST + 0.025) (x / (exp(x / k) - 1)):
    # away from it the tables are alpha and alpha + beta; at it, the average
    # of the two neighbours at +-dx/10, i.e. the limit 0.1e6 * 0.010 to O(dx^2)
    sing = np.abs(NA_M[2] + np.exp((x + NA_M[3]) / NA_M[4])) < 1e-6
    assert sing.sum() == 1
    np.testing.assert_allclose(A[~sing], alpha[~sing], rtol=1e-9)
    np.testing.assert_allclose(B[~sing] - A[~sing], beta[~sing], rtol=1e-9, atol=1e-9)
    assert A[sing][0] == pytest.approx(0.1e6 * 0.010, rel=1e-6)
    assert g.alpha == NA_M[:5] and g.beta == NA_M[5:10]
    assert g.alphaParms[-3:] == [3001.0, -0.1, 0.05]


def test_setup_tau_tables(pkg):
    from moose_core_amd import moose
    moose.reset()
    ch = moose.HHChannel("/ka")
    ch.Ypower = 1
    g = ch.gateY
    # tau = 0.02 (constant: B=0, C=1 with a huge F keeps exp() at 1 -> 0.04/2),
    # mInfinity = 1/(1 + exp((x + 0.078)/0.006))
    tau = [0.04, 0.0, 1.0, 0.0, 1e12]
    minf = [1.0, 0.0, 1.0, 0.078, 0.006]
    g.setupTau(tau + minf + [100, -0.1, 0.05])
    x = -0.1 + np.arange(101) * 0.0015
    t = _alpha(tau, x)
    m = _alpha(minf, x)
    np.testing.assert_allclose(g.tableA, m / t, rtol=1e-12)
    np.testing.assert_allclose(g.tableB, 1.0 / t, rtol=1e-12)


def test_gate_range_changes_and_field_forms(pkg):
    from moose_core_amd import moose
    moose.reset()
    ch = moose.HHChannel("/k")
    ch.Xpower = 4
    g = ch.gateX
    g.min, g.max = -0.1, 0.05          # before any table: nothing to rebuild
    tab = np.linspace(0.0, 30.0, 301)
    g.tableA = tab
    g.tableB = 2 * tab
    # a new range on a direct table refills it by lookups in itself at the
    # new grid, which lands on its own entries: values stay, range moves
    g.min = -0.12
    assert g.min == -0.12
    np.testing.assert_allclose(g.tableA, tab, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g.tableB, 2 * tab, rtol=1e-12, atol=1e-12)
    g.tableB = np.zeros(7)              # size mismatch: ignored
    assert len(g.tableB) == 301
    # the alpha/beta fields rebuild from the functional form with the
    # current table size as the number of divisions
    moose.reset()
    ch = moose.HHChannel("/k2")
    ch.Xpower = 1
    g = ch.gateX
    g.setupAlpha(NA_M[:10] + [50, -0.1, 0.05])
    g.alpha = NA_M[:5]
    assert g.divs == 51                 # HHGate::updateTables passes A_.size() as divs
    x = -0.1 + np.arange(52) * (0.15 / 51)
    ok = np.abs(NA_M[2] + np.exp((x + NA_M[3]) / NA_M[4])) >= 1e-6
    np.testing.assert_allclose(g.tableA[ok], _alpha(NA_M[:5], x)[ok], rtol=1e-9)


def test_setup_gate_form(pkg):
    from moose_core_amd import moose
    moose.reset()
    ch = moose.HHChannel("/g")
    ch.Xpower = 1
    g = ch.gateX
    g.min, g.max = -0.1, 0.05
    g.setupGate([2.0, 0.0, 1.0, 0.0, 1e12, 10, -0.1, 0.05, 0])     # alpha = 1
    g.setupGate([6.0, 0.0, 1.0, 0.0, 1e12, 10, -0.1, 0.05, 1])     # beta = 3
    # A = alpha, B = alpha + beta after the beta call
    np.testing.assert_allclose(g.tableA, np.ones(11))
    np.testing.assert_allclose(g.tableB, np.full(11, 4.0))


@pytest.mark.gpu
def test_table_default_clock_on_device(pkg, oracle_mod):
    """A Table on tick 8 (default 0.1 ms) with a 50 us solver: the device
    samples every second step (probe interval 2)."""
    print("case: object layer, Table on tick 8 default (every 2nd step), fp64", flush=True)
    from moose_core_amd import moose
    comps, _ = build_ball_tree(moose, depth=3, inject=0.8e-9)
    t8 = moose.Table("/cell/vm8")
    moose.connect(t8, "requestOut", comps[2], "getVm")
    hs = moose.HSolve("/cell/hsolve")
    hs.dt = 50e-6
    hs.target = "/cell"
    spec = hs._spec
    moose.reinit()
    moose.start(61 * 50e-6)
    moose.start(40 * 50e-6)
    ref = oracle_mod.OracleCell(spec, 50e-6)
    ref.reinit(50e-6)
    tr = ref.run(50e-6, 101, np.arange(spec.n))
    steps = np.arange(1, 102)
    np.testing.assert_array_equal(t8.vector[1:], tr[steps % 2 == 0, 2])
    assert len(t8.vector) == 1 + 50
    moose.reset()


@pytest.mark.gpu
def test_tables_sample_on_their_clock(pkg, oracle_mod):
    """Tables on tick 8 (default 0.1 ms) sample every second 50 us step,
    after the solver's step, plus the reinit sample; a Table moved to
    another tick samples at its own rate, across start() calls."""
    print("case: object layer, Table clocks (tick 8 default, tick 9 at 0.25 ms), fp64", flush=True)
    from moose_core_amd import moose
    comps, _ = build_ball_tree(moose, depth=4, inject=0.8e-9)
    t8 = moose.Table("/cell/vm8")
    moose.connect(t8, "requestOut", comps[0], "getVm")
    t9 = moose.Table("/cell/vm9")
    moose.connect(t9, "requestOut", comps[3], "getVm")
    moose.setClock(9, 250e-6)
    moose.useClock(9, "/cell/vm9", "process")
    hs = moose.HSolve("/cell/hsolve")
    hs.dt = 50e-6
    hs.target = "/cell"
    spec = hs._spec
    moose.reinit()
    moose.start(101 * 50e-6)
    moose.start(99 * 50e-6)
    ref = oracle_mod.OracleCell(spec, 50e-6)
    ref.reinit(50e-6)
    tr = ref.run(50e-6, 200, np.arange(spec.n))
    steps = np.arange(1, 201)
    np.testing.assert_array_equal(t8.vector[1:], tr[steps % 2 == 0, 0])
    np.testing.assert_array_equal(t9.vector[1:], tr[steps % 5 == 0, 3])
    assert t8.vector[0] == spec.compartments["initVm"][0] and len(t8.vector) == 101
    moose.reset()


# ---------------------------------------------------------------------------
# SWC morphologies (biophysics/ReadSwc.cpp)
# ---------------------------------------------------------------------------
TINY_SWC = """# soma with a dendrite that forks (one side with a zero-length point) and an axon
1 1 0 0 0 5.0 -1
2 3 10 0 0 1.0 1
3 3 20 0 0 1.0 2
4 3 30 5 0 0.8 3
5 3 30 -5 0 0.8 3
6 3 30 -5 0.001 0.8 5
7 3 40 -5 0 0.6 6
8 2 -10 0 0 0.5 1
"""


def test_swc_reader_structure(pkg, tmp_path):
    """Hand-traced ReadSwc on a small tree: the zero-length point 6 (1 nm from
    5) is dropped and 7 reparented to 5; branches are the runs ending at each
    fork or end, in segment order; names are <type>_<branch>_<position>;
    the soma is a cylinder of length 2r; compartments and axial messages are
    created branch by branch."""
    from moose_core_amd import moose
    moose.reset()
    f = tmp_path / "tiny.swc"
    f.write_text(TINY_SWC)
    segs, branches = moose._read_swc(str(f))
    assert not segs[5]["ok"] and segs[6]["pa"] == 5 and segs[4]["kids"] == [7]
    assert branches == [[1], [2, 3], [4], [5, 7], [8]]
    moose.loadModel(str(f), "/cell")
    names = [c.name for c in moose.element("/cell").children]
    assert names == ["soma", "dend_1_0", "dend_f_1_1", "dend_e_2_0", "dend_3_0", "dend_e_3_1", "axon_e_4_0"]
    soma = moose.element("/cell/soma")
    assert soma.length == pytest.approx(10e-6) and soma.diameter == pytest.approx(10e-6)
    # cylinder formulas with RM = 1, RA = 1, CM = 0.01 (LoadModels.cpp:165)
    d = moose.element("/cell/dend_e_3_1")      # segment 7, 10 um from 5
    L, dia = 10e-6, 1.2e-6
    assert d.Rm == pytest.approx(1.0 / (L * dia * np.pi), rel=1e-12)
    assert d.Ra == pytest.approx(L / (np.pi * (dia / 2) ** 2), rel=1e-12)
    assert d.Cm == pytest.approx(0.01 * L * dia * np.pi, rel=1e-12)
    axial = [(s.name, t.name) for s, sf, t, tf in moose._msgs if sf == "axial"]
    assert axial == [("soma", "dend_1_0"), ("dend_1_0", "dend_f_1_1"), ("dend_f_1_1", "dend_e_2_0"),
                     ("dend_f_1_1", "dend_3_0"), ("dend_3_0", "dend_e_3_1"), ("soma", "axon_e_4_0")]
    one_kid = tmp_path / "one_kid.swc"         # the soma ends no branch: no parent compartment
    one_kid.write_text("1 1 0 0 0 5 -1\n2 3 10 0 0 1 1\n3 3 20 0 0 1 2\n")
    moose.reset()
    with pytest.raises(ValueError, match="no parent compartment"):
        moose.loadModel(str(one_kid), "/c1")
    bad = tmp_path / "two_roots.swc"
    bad.write_text("1 1 0 0 0 5 -1\n2 3 10 0 0 1 -1\n")
    with pytest.raises(ValueError):
        moose._read_swc(str(bad))


def test_swc_h10_under_hsolve(pkg, oracle_mod):
    """The reference's h10 test morphology: every valid point a compartment,
    one tree, taken over by an HSolve (passive, Em = -65 mV) and run by the
    oracle to the passive steady state of a soma injection."""
    import os
    from moose_core_amd import moose
    moose.reset()
    path = os.path.join(os.path.dirname(__file__), "golden", "h10.CNG.swc")
    segs, branches = moose._read_swc(path)
    cell = moose.loadModel(path, "/h10")
    comps = [c for c in cell.children if isinstance(c, moose.Compartment)]
    assert len(comps) == sum(1 for s in segs if s["ok"])
    assert sum(len(b) for b in branches) == len(comps)
    for c in comps:
        c.Em = c.initVm = c.Vm = -0.065
    comps[0].inject = 0.1e-9
    hs = moose.HSolve("/h10/hsolve")
    hs.dt = 50e-6
    spec, _ = hs.build_spec(moose._deep_search_compartment(cell))
    assert spec.n == len(comps) and len(spec.axial) == len(comps) - 1
    o = oracle_mod.OracleCell(spec, 50e-6)
    o.reinit(50e-6)
    tr = o.run(50e-6, 4000, [0])
    assert -0.065 < tr[-1, 0] < 0.0 and abs(tr[-1, 0] - tr[-2, 0]) < 1e-9
    moose.reset()


@pytest.mark.gpu
def test_swc_h10_hh_on_the_gpu(pkg, oracle_mod):
    """h10 with squid HH channels copied onto every compartment (Gbar by
    area), run through HSolve on the GPU: bit-identical to the oracle."""
    print("case: h10.CNG.swc with HH channels, fp64, 400 steps", flush=True)
    import os
    from moose_core_amd import moose
    moose.reset()
    moose.Neutral("/library")
    gates = workloads.hh_gates()
    na = moose.HHChannel("/library/Na")
    na.Ek, na.Xpower, na.Ypower = 0.050, 3.0, 1.0
    k = moose.HHChannel("/library/K")
    k.Ek, k.Xpower = -0.077, 4.0
    for gate, g in ((na.gateX, gates[0]), (na.gateY, gates[1]), (k.gateX, gates[2])):
        gate.min, gate.max = g.min, g.max
        gate.tableA = g.tableA
        gate.tableB = g.tableB
    cell = moose.loadModel(os.path.join(os.path.dirname(__file__), "golden", "h10.CNG.swc"), "/h10")
    comps = [c for c in cell.children if isinstance(c, moose.Compartment)]
    for c in comps:
        c.Em = c.initVm = c.Vm = -0.065
        area = np.pi * c.diameter * c.length
        for proto, gbar in ((na, 1200.0), (k, 360.0)):
            ch = moose.copy(proto, c, proto.name)
            ch.Gbar = gbar * area
            moose.connect(ch, "channel", c, "channel")
    comps[0].inject = 2e-9
    moose.setClock(8, 50e-6)
    vm = moose.Table("/h10/vm")
    moose.connect(vm, "requestOut", comps[0], "getVm")
    hs = moose.HSolve("/h10/hsolve")
    hs.dt = 50e-6
    hs.target = "/h10"
    moose.reinit()
    moose.start(400 * 50e-6)
    ref = oracle_mod.OracleCell(hs._spec, 50e-6)
    ref.reinit(50e-6)
    tr = ref.run(50e-6, 400, np.arange(hs._spec.n))
    soma = hs._objs["compartments"].index(comps[0])
    np.testing.assert_array_equal(vm.vector[1:], tr[:, soma])
    assert vm.vector.max() > 0.0                # it fires
    moose.reset()


# ---------------------------------------------------------------------------
# GENESIS .p cell files (biophysics/ReadCell.cpp)
# ---------------------------------------------------------------------------
def _golden(name):
    import os
    return os.path.join(os.path.dirname(__file__), "golden", name)


def test_readcell_soma_p(pkg):
    """The reference's tests/core/soma.p: globals RA 4, CM 0.03, RM 2.8,
    EREST_ACT -80 mV, ELEAK -50 mV; one 16 x 16 um root compartment."""
    from moose_core_amd import moose
    moose.reset()
    moose.Neutral("/library")
    cell = moose.loadModel(_golden("soma.p"), "/neuron")
    (soma,) = cell.children
    L = d = 16e-6
    surf = L * d * np.pi
    assert soma.name == "soma" and soma.length == pytest.approx(L) and soma.diameter == pytest.approx(d)
    assert soma.Cm == pytest.approx(0.03 * surf, rel=1e-12)
    assert soma.Rm == pytest.approx(2.8 / surf, rel=1e-12)
    assert soma.Ra == pytest.approx(4.0 * L * 4.0 / (d * d * np.pi), rel=1e-12)
    assert soma.initVm == -0.080 and soma.Vm == -0.080 and soma.Em == -0.050


def test_readcell_ca1_minimal(pkg):
    """rdesigneur's ca1_minimal.p: *relative *cartesian coordinates, the '.'
    parent, EREST_ACT only (so Em = initVm = -60 mV), lengths from the
    relative offsets."""
    from moose_core_amd import moose
    moose.reset()
    moose.Neutral("/library")
    cell = moose.loadModel(_golden("ca1_minimal.p"), "/ca1")
    names = [c.name for c in cell.children]
    assert names[:6] == ["dend_3", "dend_5", "dend_6", "dend_8", "soma", "apical_10"]
    soma = moose.element("/ca1/soma")
    assert soma.y == pytest.approx((3 * 220 + 110 + 125) * 1e-6)     # chain of relative offsets
    assert soma.length == pytest.approx(125e-6) and soma.diameter == pytest.approx(8.46e-6)
    a14_2 = moose.element("/ca1/apical_14_2")                      # parent '.' = apical_14_1
    assert a14_2.length == pytest.approx(10e-6)
    for c in cell.children:
        assert c.Em == -0.06 and c.initVm == -0.06
        assert c.Cm == pytest.approx(0.03 * c.length * c.diameter * np.pi, rel=1e-12)
    axial = [(s.name, t.name) for s, sf, t, tf in moose._msgs if sf == "axial"]
    assert ("dend_8", "soma") in axial and ("apical_13_2", "apical_14_1") in axial
    assert len(axial) == len(names) - 1
    with pytest.raises(ValueError):
        moose.loadModel(_golden("h10.CNG.swc")[:-4] + ".txt", "/x")


P_FILE = """// a soma with HH channels and two dendrites
*cartesian
*relative
*asymmetric
*set_global RM 1.0
*set_global RA 1.0
*set_global CM 0.01
*set_global EREST_ACT -0.065
soma none 20 0 0 20 Na 1200 K 360 spike 0.0
dend1 soma 50 0 0 2 Na 120 K 36
dend2 . 50 0 0 1.5 RA 2.0 Na -1e-9 K 36
dend3 soma 0 40 0 2 K 36
"""


def build_p_cell(moose, tmp_path, inject=1e-9):
    moose.reset()
    moose.Neutral("/library")
    gates = workloads.hh_gates()
    na = moose.HHChannel("/library/Na")
    na.Ek, na.Xpower, na.Ypower = 0.050, 3.0, 1.0
    k = moose.HHChannel("/library/K")
    k.Ek, k.Xpower = -0.077, 4.0
    for gate, g in ((na.gateX, gates[0]), (na.gateY, gates[1]), (k.gateX, gates[2])):
        gate.min, gate.max = g.min, g.max
        gate.tableA = g.tableA
        gate.tableB = g.tableB
    sg = moose.SpikeGen("/library/spike")
    sg.refractT = 0.0
    f = tmp_path / "cell.p"
    f.write_text(P_FILE)
    cell = moose.loadModel(str(f), "/pcell")
    moose.element("/pcell/soma").inject = inject
    return cell


def test_readcell_channels(pkg, tmp_path):
    from moose_core_amd import moose
    cell = build_p_cell(moose, tmp_path)
    soma = moose.element("/pcell/soma")
    surf = 20e-6 * 20e-6 * np.pi
    assert moose.element("/pcell/soma/Na").Gbar == pytest.approx(1200 * surf, rel=1e-12)
    assert moose.element("/pcell/soma/spike").threshold == 0.0
    d2 = moose.element("/pcell/dend2")
    assert moose.element("/pcell/dend2/Na").Gbar == 1e-9               # negative: absolute
    assert d2.Ra == pytest.approx(4.0 * 2.0 * 50e-6 / (1.5e-6 ** 2 * np.pi), rel=1e-12)
    hs = moose.HSolve("/pcell/hsolve")
    hs.dt = 50e-6
    spec, objs = hs.build_spec(moose._deep_search_compartment(cell))
    assert spec.n == 4 and len(spec.channels) == 7 and len(spec.spikegens) == 1


@pytest.mark.gpu
def test_readcell_cell_on_the_gpu(pkg, oracle_mod, tmp_path):
    print("case: .p cell with HH channels and a SpikeGen, fp64, 600 steps", flush=True)
    from moose_core_amd import moose
    cell = build_p_cell(moose, tmp_path, inject=1.5e-10)
    moose.setClock(8, 50e-6)
    vm = moose.Table("/pcell/vm")
    moose.connect(vm, "requestOut", moose.element("/pcell/soma"), "getVm")
    spk = moose.Table("/pcell/spk")
    moose.connect(moose.element("/pcell/soma/spike"), "spikeOut", spk, "spike")
    hs = moose.HSolve("/pcell/hsolve")
    hs.dt = 50e-6
    hs.target = "/pcell"
    moose.reinit()
    moose.start(600 * 50e-6)
    ref = oracle_mod.OracleCell(hs._spec, 50e-6)
    ref.reinit(50e-6)
    tr = ref.run(50e-6, 600, np.arange(hs._spec.n))
    soma = hs._objs["compartments"].index(moose.element("/pcell/soma"))
    np.testing.assert_array_equal(vm.vector[1:], tr[:, soma])
    _, steps = ref.read_spikes()
    assert len(steps) > 0
    np.testing.assert_array_equal(spk.vector, np.array(steps, dtype=np.float64) * 50e-6)
    moose.reset()
