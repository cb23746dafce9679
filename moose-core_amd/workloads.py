"""Synthetic, seeded cell models for the benchmark configurations.

  hh_ball_tree()      one cell: a soma of 20 um plus a full binary dendritic tree
                      of depth 7 (127 compartments, 10 um x 1 um), squid HH Na/K
                      everywhere, 0.1 nA into the soma (run at dt = 50 us).
  random_hh_cell()    a random binary tree (recursive uniform splitting), branch
                      lengths U[5, 50] um, diameters tapering 2 -> 0.5 um with
                      depth, HH Na/K with Gbar scaled by U[0.8, 1.2], soma inject
                      U[0, 0.3] nA.  4096 of them with n ~ U{50..2000} is the
                      batched ensemble (dt = 25 us).
  random_ca_cell()    a random tree with 2-4 children per junction and six
                      channel types (Na, Kdr, KA, CaL, Ca-dependent K-AHP, plus
                      leak through Rm) and one CaConc pool per compartment;
                      with kahp2d the K-AHP is an HHChannel2D on a 3000 x 500
                      Vm x [Ca] table.

Passive parameters follow rdesigneur's buildCompt (RM = 1 ohm m^2,
RA = 1 ohm m, CM = 0.01 F/m^2, Em = initVm = -65 mV); channel kinetics follow
python/rdesigneur/rdesigneurProtos.py:65-114 (squid HH, EREST_ACT_HH = -70 mV).
"""
import numpy as np

from .cellspec import CellSpec, Gate, Table2D, channel_records, interleave_channels, passive_compartments
from ._abi import CACONC_DT, CHAN2D_DT, SPIKEGEN_DT

PER_MS = 1e3
EREST_ACT_HH = -0.070
PI = 3.14159265359
FARADAY = 96485.3365


def _xexp(x, k):
    """x / (exp(x/k) - 1) with its limit k at x = 0."""
    u = np.asarray(x, dtype=np.float64) / k
    small = np.abs(u) < 1e-9
    safe = np.where(small, 1.0, u)
    return np.where(small, k * (1.0 - u / 2.0), k * safe / np.expm1(safe))


def hh_gates(vmin=-0.100, vmax=0.050, divs=3000, id_base=1):
    """Squid HH gates (Na m, Na h, K n) on one grid: make_HH_Na / make_HH_K
    of python/rdesigneur/rdesigneurProtos.py:65-114."""
    v = np.linspace(vmin, vmax, divs + 1) - EREST_ACT_HH
    mv = v * 1e3
    m_a = PER_MS * 0.1 * _xexp(25.0 - mv, 10.0)
    m_b = PER_MS * 4.0 * np.exp(-mv / 18.0)
    h_a = PER_MS * 0.07 * np.exp(-v / 20e-3)
    h_b = PER_MS * 1.0 / (np.exp((30e-3 - v) / 10e-3) + 1.0)
    n_a = PER_MS * 0.01 * _xexp(10.0 - mv, 10.0)
    n_b = PER_MS * 0.125 * np.exp(-mv / 80.0)
    return [Gate(id_base + 0, vmin, vmax, divs, m_a, m_a + m_b),
            Gate(id_base + 1, vmin, vmax, divs, h_a, h_a + h_b),
            Gate(id_base + 2, vmin, vmax, divs, n_a, n_a + n_b)]


_HH = None


def _hh():
    global _HH
    if _HH is None:
        _HH = hh_gates()
    return _HH


def hh_channels(n, area, scale=1.0, gna=1200.0, gk=360.0):
    """Na (m^3 h, Ek 50 mV) and K (n^4, Ek -77 mV) on compartments 0..n-1
    (gate indices 0, 1, 2 refer to hh_gates())."""
    comp = np.arange(n, dtype=np.int32)
    na = channel_records(n, comp, gna * area * scale, 0.050, Xpower=3.0, Ypower=1.0,
                         gates=(0, 1, -1))
    k = channel_records(n, comp, gk * area * scale, -0.077, Xpower=4.0, gates=(2, -1, -1))
    return interleave_channels([na, k])


def hh_ball_tree(depth=7, inject=0.1e-9, soma_dia=20e-6, dend_len=10e-6, dend_dia=1e-6):
    """Soma + full binary tree of `depth` levels (2**depth - 1 dendrites)."""
    nd = 2 ** depth - 1
    n = nd + 1
    length = np.full(n, dend_len)
    dia = np.full(n, dend_dia)
    length[0] = soma_dia
    dia[0] = soma_dia
    comp, area = passive_compartments(length, dia)
    comp["inject"][0] = inject
    # dendrite k (1-based in heap order) -> compartment k; soma -> dendrite 1
    axial = [(0, 1)]
    for k in range(1, nd + 1):
        for c in (2 * k, 2 * k + 1):
            if c <= nd:
                axial.append((k, c))
    spk = np.zeros(1, SPIKEGEN_DT)
    spk["compartment"] = 0
    spk["threshold"] = 0.0
    spk["refractT"] = 0.0
    spk["edgeTriggered"] = 1
    return CellSpec(comp, np.array(axial, dtype=np.int32), seed=0,
                    channels=hh_channels(n, area), gates=_hh(), spikegens=spk)


def random_tree_parents(rng, n, max_children=2):
    """Parent of each node 1..n-1 (node 0 is the root): a random tree by
    recursive uniform splitting of the remaining nodes among k subtrees
    (k = 2, or k ~ U{2..max_children}); empty subtrees are allowed, so chains
    occur.  Random draws are taken from pre-drawn uniform arrays."""
    parent = [-1] * n
    depth = [0] * n
    ku = rng.random(n).tolist()                    # children-count draws
    cu = rng.random(n * max(1, max_children - 1)).tolist()   # cut draws
    ci = 0
    stack = [(0, n - 1)]
    nxt = 1
    while stack:
        node, m = stack.pop()
        if m <= 0:
            continue
        k = 2 if max_children <= 2 else 2 + int(ku[node] * (max_children - 1))
        cuts = sorted(int(cu[ci + j] * m) for j in range(k - 1))
        ci += k - 1
        prev = 0
        for cut in cuts + [m]:
            size = cut - prev
            prev = cut
            if size <= 0:
                continue
            c = nxt
            nxt += 1
            parent[c] = node
            depth[c] = depth[node] + 1
            stack.append((c, size - 1))
    assert nxt == n
    return np.array(parent, dtype=np.int64), np.array(depth, dtype=np.int64)


def random_hh_cell(rng, n, inject=None, spikegen=True):
    """One cell of the batched HH ensemble."""
    parent, depth = random_tree_parents(rng, n, 2)
    maxd = max(1, int(depth.max()))
    length = rng.uniform(5e-6, 50e-6, size=n)
    dia = 2e-6 - 1.5e-6 * depth / maxd
    length[0] = 20e-6
    dia[0] = 20e-6
    comp, area = passive_compartments(length, dia)
    scale = rng.uniform(0.8, 1.2)
    comp["inject"][0] = rng.uniform(0.0, 0.3e-9) if inject is None else inject
    axial = np.stack([parent[1:], np.arange(1, n)], axis=1).astype(np.int32)
    spk = np.zeros(1 if spikegen else 0, SPIKEGEN_DT)
    if spikegen:
        spk["compartment"] = 0
        spk["edgeTriggered"] = 1
    return CellSpec(comp, axial, seed=0, channels=hh_channels(n, area, scale), gates=_hh(),
                    spikegens=spk)


def hh_ensemble(ncells=4096, nmin=50, nmax=2000, seed=1):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=ncells)
    return [random_hh_cell(rng, int(s)) for s in sizes]


# ---------------------------------------------------------------------------
# six channel types with calcium (Traub-style kinetics, MOOSE conventions:
# tables A = alpha, B = alpha + beta; Ca-dependent K-AHP reads the pool)
# ---------------------------------------------------------------------------
def _abtables(alpha, beta):
    return alpha, alpha + beta


def ca_gates(vmin=-0.100, vmax=0.050, divs=3000, camin=0.0, camax=1.0, cadivs=500, id_base=100):
    v = np.linspace(vmin, vmax, divs + 1)
    mv = (v - EREST_ACT_HH) * 1e3
    g = []
    # Na m, h (squid)
    m_a = PER_MS * 0.1 * _xexp(25.0 - mv, 10.0)
    m_b = PER_MS * 4.0 * np.exp(-mv / 18.0)
    h_a = PER_MS * 0.07 * np.exp(-mv / 20.0)
    h_b = PER_MS * 1.0 / (np.exp((30.0 - mv) / 10.0) + 1.0)
    # Kdr n (squid)
    n_a = PER_MS * 0.01 * _xexp(10.0 - mv, 10.0)
    n_b = PER_MS * 0.125 * np.exp(-mv / 80.0)
    # KA a (activation) and b (inactivation), Connor-Stevens-like
    a_inf = 1.0 / (1.0 + np.exp(-(v + 0.060) / 0.0085))
    a_tau = 1e-3 * (0.37 + 1.0 / (np.exp((v + 0.0355) / 0.0165) + np.exp(-(v + 0.072) / 0.0155)))
    b_inf = 1.0 / (1.0 + np.exp((v + 0.078) / 0.006))
    b_tau = 20e-3 * np.ones_like(v)
    # CaL s (Traub 1991)
    s_a = 1.6e3 / (1.0 + np.exp(-0.072e3 * (v - 0.005)))
    s_b = 20.0 * _xexp(v * 1e3 + 8.9, 5.0)
    tabs = [_abtables(m_a, m_b), _abtables(h_a, h_b), _abtables(n_a, n_b),
            (a_inf / a_tau, 1.0 / a_tau), (b_inf / b_tau, 1.0 / b_tau), _abtables(s_a, s_b)]
    for k, (A, B) in enumerate(tabs):
        g.append(Gate(id_base + k, vmin, vmax, divs, A, B))
    # K-AHP q on [Ca] (mM): alpha = min(0.2e-4*Ca, 0.01) per ms (Traub), beta = 0.001/ms
    ca = np.linspace(camin, camax, cadivs + 1)
    q_a = np.minimum(20.0 * ca, 10.0)
    q_b = np.full_like(ca, 1.0)
    g.append(Gate(id_base + 6, camin, camax, cadivs, q_a, q_a + q_b))
    return g


def kahp2d_tables(vmin=-0.100, vmax=0.050, vdivs=3000, camin=0.0, camax=1.0, cadivs=500):
    """The K-AHP gate as an HHChannel2D Z gate over Vm x [Ca]: the Ca-driven
    alpha of ca_gates() scaled by a sigmoid in Vm (BK-like), beta 1/ms.
    Returns the (A, B) Interpol2D tables, x = Vm, y = [Ca]."""
    v = np.linspace(vmin, vmax, vdivs + 1)[:, None]
    ca = np.linspace(camin, camax, cadivs + 1)[None, :]
    a = np.minimum(20.0 * ca, 10.0) * (0.5 + 1.0 / (1.0 + np.exp(-(v + 0.030) / 0.008)))
    b = np.ones_like(a)
    mk = lambda t: Table2D(vdivs, cadivs, vmin, vmax, camin, camax, np.ascontiguousarray(t))
    return [mk(a), mk(a + b)]


_CA = None
_KAHP2D = None


def _kahp2d():
    global _KAHP2D
    if _KAHP2D is None:
        _KAHP2D = kahp2d_tables()
    return _KAHP2D


def _ca():
    global _CA
    if _CA is None:
        _CA = ca_gates()
    return _CA


def random_ca_cell(rng, n, inject=None, kahp2d=False):
    """A random tree with branching 2-4, six channel types and one Ca pool per
    compartment (tau 20 ms, B from the shell volume, 1/(2 F vol)).  kahp2d:
    the K-AHP channel is an HHChannel2D over Vm x [Ca] (kahp2d_tables, 3000 x
    500 divisions) instead of an HHChannel on [Ca]."""
    parent, depth = random_tree_parents(rng, n, 4)
    maxd = max(1, int(depth.max()))
    length = rng.uniform(5e-6, 50e-6, size=n)
    dia = 2e-6 - 1.5e-6 * depth / maxd
    length[0] = 20e-6
    dia[0] = 20e-6
    comp, area = passive_compartments(length, dia)
    comp["inject"][0] = rng.uniform(0.0, 0.3e-9) if inject is None else inject
    axial = np.stack([parent[1:], np.arange(1, n)], axis=1).astype(np.int32)
    c = np.arange(n, dtype=np.int32)
    scale = rng.uniform(0.8, 1.2)
    na = channel_records(n, c, 1200.0 * area * scale, 0.050, Xpower=3.0, Ypower=1.0, gates=(0, 1, -1))
    kdr = channel_records(n, c, 360.0 * area * scale, -0.077, Xpower=4.0, gates=(2, -1, -1))
    ka = channel_records(n, c, 50.0 * area * scale, -0.080, Xpower=3.0, Ypower=1.0, gates=(3, 4, -1))
    cal = channel_records(n, c, 10.0 * area * scale, 0.080, Xpower=2.0, gates=(5, -1, -1),
                          caTarget=c)
    kahp = channel_records(n, c, 20.0 * area * scale, -0.080, Zpower=1.0, gates=(-1, -1, 6),
                           useConcentration=1, caDepend=c)
    chans2d = np.zeros(0, CHAN2D_DT)
    tables2d = []
    if kahp2d:
        chans = interleave_channels([na, kdr, ka, cal])
        tables2d = _kahp2d()
        chans2d = np.zeros(n, CHAN2D_DT)
        chans2d["compartment"] = c
        chans2d["tableA"] = -1
        chans2d["tableB"] = -1
        chans2d["tableA"][:, 2] = 0
        chans2d["tableB"][:, 2] = 1
        chans2d["dep"] = -1
        chans2d["dep"][:, 4] = 0                    # Z table x: Vm
        chans2d["dep"][:, 5] = 1                    # Z table y: concen
        chans2d["conc1"] = c
        chans2d["conc2"] = -1
        chans2d["Gbar"] = 20.0 * area * scale
        chans2d["Ek"] = -0.080
        chans2d["Zpower"] = 1.0
        chans2d["modulation"] = 1.0
    else:
        chans = interleave_channels([na, kdr, ka, cal, kahp])
    pools = np.zeros(n, CACONC_DT)
    vol = PI * dia * dia * length / 4.0
    thick = np.minimum(0.1e-6, dia / 2.0)
    core = np.maximum(dia / 2.0 - thick, 0.0)
    vol = vol - PI * core * core * length
    pools["B"] = 1.0 / (2.0 * FARADAY * vol)          # mM per coulomb
    pools["tau"] = 0.020
    pools["CaBasal"] = 1e-4
    pools["Ca"] = 1e-4
    pools["ceiling"] = 0.0
    pools["floor"] = 0.0
    spk = np.zeros(1, SPIKEGEN_DT)
    spk["compartment"] = 0
    spk["edgeTriggered"] = 1
    return CellSpec(comp, axial, seed=0, channels=chans, gates=_ca(), pools=pools, spikegens=spk,
                    tables2d=tables2d, chans2d=chans2d)
