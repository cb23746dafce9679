"""HHChannel2D objects beside the solver (biophysics/HHChannel2D.cpp): their
Interpol2D lookup (builtins/Interpol2D.cpp:508-590), the exponential-Euler
gate step (biophysics/HHChannelBase.cpp:268-274), and their current reaching
the solver through HSolve::addGkEk (hsolve/HSolveInterface.cpp:246-252) from
the Vm and [Ca] the solver sent at the end of the previous tick.

CPU tests pin the oracle against what the mathematics fixes: bilinear
functions are reproduced exactly, the lookup agrees with scipy's
RegularGridInterpolator on clamped points, the gate step is the closed-form
solution of dx/dt = A - B x, a gate-less channel is the same current as a
host addGkEk, and instant gates read the previous tick's Vm / [Ca].  GPU
tests compare the HIP path (called through the C ABI) with the oracle on
cells whose K-AHP is a 2-D channel: FP64 within 1e-6 V (the GPU's exp
and glibc's may differ in the last bit, so not bit for bit) with identical
spike steps, mixed within 0.1 mV and one step.
"""
import ctypes as C

import numpy as np
import pytest

from moose_core_amd import workloads
from moose_core_amd._abi import CHAN2D_DT, Table2DDesc
from moose_core_amd.cellspec import Table2D

FP64_TOL = 1e-6
MIXED_TOL = 1e-4


def table_desc(t):
    arr = np.ascontiguousarray(t.table, dtype=np.float64).reshape(-1)
    d = Table2DDesc()
    d.xdivs, d.ydivs = t.xdivs, t.ydivs
    d.xmin, d.xmax, d.ymin, d.ymax = t.xmin, t.xmax, t.ymin, t.ymax
    d.table = arr.ctypes.data_as(C.POINTER(C.c_double))
    return d, arr


def lookup(oracle_mod, t, xs, ys):
    d, keep = table_desc(t)
    L = oracle_mod.lib()
    return np.array([L.or_interp2d(C.byref(d), float(x), float(y)) for x, y in zip(xs, ys)])


def test_interp2d_reproduces_bilinear_functions(oracle_mod):
    # nx != ny and an asymmetric function: a transposed index or a swapped
    # fraction changes the result
    xdivs, ydivs, x0, x1, y0, y1 = 7, 4, -0.1, 0.05, 0.0, 2.0
    f = lambda x, y: 0.3 + 2.0 * x - 0.7 * y + 5.0 * x * y
    xg = np.linspace(x0, x1, xdivs + 1)[:, None]
    yg = np.linspace(y0, y1, ydivs + 1)[None, :]
    t = Table2D(xdivs, ydivs, x0, x1, y0, y1, f(xg, yg))
    rng = np.random.default_rng(3)
    xs = rng.uniform(x0, x1, 400)
    ys = rng.uniform(y0, y1, 400)
    np.testing.assert_allclose(lookup(oracle_mod, t, xs, ys), f(xs, ys), rtol=0, atol=1e-12)
    # outside the table: the value at the clamped point, including the last
    # row and column (where the lookup reads no neighbour beyond the edge)
    xo = np.array([-1.0, 1.0, x1, x1, -1.0, 0.0, 0.0, 1.0])
    yo = np.array([1.0, 1.0, 5.0, -5.0, 5.0, y1, 9.0, 0.5])
    np.testing.assert_allclose(lookup(oracle_mod, t, xo, yo),
                               f(np.clip(xo, x0, x1), np.clip(yo, y0, y1)), rtol=0, atol=1e-12)
    # the corner is the corner value
    assert lookup(oracle_mod, t, [x1], [y1])[0] == t.table[-1, -1]


def test_interp2d_matches_scipy(oracle_mod):
    from scipy.interpolate import RegularGridInterpolator
    rng = np.random.default_rng(5)
    xdivs, ydivs = 30, 11
    tab = rng.normal(size=(xdivs + 1, ydivs + 1))
    t = Table2D(xdivs, ydivs, -0.08, 0.04, 1e-4, 3.0, tab)
    xs = rng.uniform(-0.1, 0.06, 1000)
    ys = rng.uniform(-1.0, 4.0, 1000)
    ref = RegularGridInterpolator((np.linspace(-0.08, 0.04, xdivs + 1), np.linspace(1e-4, 3.0, ydivs + 1)),
                                  tab, method="linear")
    want = ref(np.stack([np.clip(xs, -0.08, 0.04), np.clip(ys, 1e-4, 3.0)], axis=1))
    np.testing.assert_allclose(lookup(oracle_mod, t, xs, ys), want, rtol=0, atol=1e-12)


def test_integrate_ee_closed_form(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(9)
    for _ in range(200):
        x0, A, B, dt = rng.uniform(0, 1), rng.uniform(0, 50), rng.uniform(1e-3, 2e3), rng.uniform(1e-6, 1e-3)
        exact = A / B + (x0 - A / B) * np.exp(-B * dt)     # solution of dx/dt = A - B x
        assert L.or_integrate_ee(x0, dt, A, B) == pytest.approx(exact, rel=1e-13, abs=1e-15)
    # B below EPSILON (1e-10): forward Euler
    assert L.or_integrate_ee(0.25, 1e-4, 3.0, 1e-12) == 0.25 + 3.0 * 1e-4


# ---------------------------------------------------------------------------
# cells with 2-D channels
# ---------------------------------------------------------------------------
def small_cell(n=12, inject=0.5e-9):
    """An HH chain-with-branches cell (random_hh_cell) of n compartments."""
    rng = np.random.default_rng(11)
    return workloads.random_hh_cell(rng, n, inject=inject)


def chan2d_records(comps, tabA=(-1, -1, -1), tabB=(-1, -1, -1), dep=(-1,) * 6, powers=(0, 0, 0),
                   Gbar=1e-9, Ek=-0.09, instant=0, conc1=-1):
    r = np.zeros(len(comps), CHAN2D_DT)
    r["compartment"] = comps
    r["instant"] = instant
    r["tableA"] = tabA
    r["tableB"] = tabB
    r["dep"] = dep
    r["conc1"] = conc1
    r["conc2"] = -1
    r["Gbar"] = Gbar
    r["Ek"] = Ek
    r["Xpower"], r["Ypower"], r["Zpower"] = powers
    r["modulation"] = 1.0
    return r


def test_gateless_channel_is_a_host_gkek(oracle_mod):
    """A 2-D channel with no gates is a fixed conductance: stepping it is
    the same, bit for bit, as a host addGkEk of (Gbar, Ek) before every
    tick."""
    base = small_cell()
    dt, steps, G, E = 25e-6, 120, 3e-9, -0.02
    spec = small_cell()
    spec.chans2d = chan2d_records([4], Gbar=G, Ek=E)
    a = oracle_mod.OracleCell(spec, dt)
    a.reinit(dt)
    ta = a.run(dt, steps, np.arange(spec.n))
    b = oracle_mod.OracleCell(base, dt)
    b.reinit(dt)
    tb = []
    for k in range(steps):
        b.add_gkek(4, G, E)
        tb.append(b.run(dt, 1, np.arange(base.n))[0])
    np.testing.assert_array_equal(ta, np.array(tb))
    assert a.get("Gk2D", [0])[0] == G


def linear_tables(n_x=3000, n_y=500):
    """A(x, y) = 2 + 30 x + 4 y, B = 8: an instant gate reads A/B."""
    x0, x1, y0, y1 = -0.1, 0.05, 0.0, 1.0
    xg = np.linspace(x0, x1, n_x + 1)[:, None]
    yg = np.linspace(y0, y1, n_y + 1)[None, :]
    A = 2.0 + 30.0 * xg + 4.0 * yg
    return [Table2D(n_x, n_y, x0, x1, y0, y1, A), Table2D(n_x, n_y, x0, x1, y0, y1, np.full_like(A, 8.0))]


def test_instant_gate_reads_previous_tick(oracle_mod):
    """An instant X gate on (Vm, [Ca]): after tick k its Gk is
    Gbar * A/B of the Vm and [Ca] at the end of tick k-1; at reinit (before
    the solver first sends them) of (0, 0)."""
    rng = np.random.default_rng(4)
    spec = workloads.random_ca_cell(rng, 30, inject=0.8e-9)
    spec.tables2d = linear_tables(300, 50)
    spec.chans2d = chan2d_records([0, 7], tabA=(0, -1, -1), tabB=(1, -1, -1), dep=(0, 1, -1, -1, -1, -1),
                                  powers=(1, 0, 0), instant=1, Gbar=2e-9)
    spec.chans2d["conc1"] = [0, 7]
    dt = 25e-6
    c = oracle_mod.OracleCell(spec, dt)
    c.reinit(dt)
    G = 2e-9
    np.testing.assert_allclose(c.get("Gk2D", [0, 1]), G * 2.0 / 8.0, rtol=1e-13)
    pools = c.get("Ca", [0, 7])
    assert pools[0] > 0
    for _ in range(40):
        V = c.get("Vm", [0, 7])
        ca = c.get("Ca", [0, 7])
        c.run(dt, 1)
        want = G * (2.0 + 30.0 * V + 4.0 * ca) / 8.0
        np.testing.assert_allclose(c.get("Gk2D", [0, 1]), want, rtol=1e-12)
        np.testing.assert_allclose(c.get("X2D", [0, 1]), want / G, rtol=1e-12)


def test_gate_relaxes_in_closed_form(oracle_mod):
    """Constant tables: Z relaxes from the value set after reinit as
    A/B + (z0 - A/B) exp(-k B dt)."""
    spec = small_cell()
    n_x, n_y = 4, 3
    A = np.full((n_x + 1, n_y + 1), 30.0)
    B = np.full((n_x + 1, n_y + 1), 120.0)
    spec.tables2d = [Table2D(n_x, n_y, -0.1, 0.05, 0.0, 1.0, A), Table2D(n_x, n_y, -0.1, 0.05, 0.0, 1.0, B)]
    spec.chans2d = chan2d_records([2], tabA=(-1, -1, 0), tabB=(-1, -1, 1), dep=(-1, -1, -1, -1, 0, -1),
                                  powers=(0, 0, 2), Gbar=1e-9)
    dt = 50e-6
    c = oracle_mod.OracleCell(spec, dt)
    c.reinit(dt)
    assert c.get("Z2D", [0])[0] == 0.25
    c.set("Z2D", [0], [0.9])
    for k in range(1, 60):
        c.run(dt, 1)
        z = 0.25 + (0.9 - 0.25) * np.exp(-k * 120.0 * dt)
        assert c.get("Z2D", [0])[0] == pytest.approx(z, rel=1e-12)
        assert c.get("Gk2D", [0])[0] == pytest.approx(1e-9 * z * z, rel=1e-12)


def test_kahp2d_cell_differs_from_1d_only_by_the_gate(oracle_mod):
    """The 2-D K-AHP workload runs in the oracle and starts out close to the
    1-D K-AHP cell of the same seed (the 2-D gate scales alpha by a Vm
    factor and starts from [Ca] = 0, not the basal level); both spike."""
    dt, n = 25e-6, 2000
    a = workloads.random_ca_cell(np.random.default_rng(21), 60, inject=0.6e-9, kahp2d=True)
    b = workloads.random_ca_cell(np.random.default_rng(21), 60, inject=0.6e-9)
    assert len(a.chans2d) == 60 and len(a.channels) == 4 * 60
    ra = oracle_mod.OracleCell(a, dt)
    ra.reinit(dt)
    ta = ra.run(dt, n, [0])
    rb = oracle_mod.OracleCell(b, dt)
    rb.reinit(dt)
    tb = rb.run(dt, n, [0])
    assert np.isfinite(ta).all()
    assert len(ra.read_spikes()[1]) > 0 and len(rb.read_spikes()[1]) > 0
    assert np.abs(ta[:20] - tb[:20]).max() < 1e-4


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------
def kahp2d_batch():
    rng = np.random.default_rng(17)
    specs = [workloads.random_ca_cell(rng, int(n), inject=float(rng.uniform(0.3e-9, 1.0e-9)), kahp2d=True)
             for n in (40, 150, 333)]
    # cells without 2-D channels in the same engine
    specs += [workloads.random_ca_cell(rng, 90, inject=0.7e-9), workloads.hh_ball_tree(inject=0.6e-9)]
    # a cell with the 2-D channel on every other compartment only (the
    # external kernel's general path; the cells above, with one on every
    # compartment, take its direct path)
    half = workloads.random_ca_cell(rng, 120, inject=0.8e-9, kahp2d=True)
    half.chans2d = half.chans2d[::2]
    specs.append(half)
    return specs


def run_both(pkg, oracle_mod, specs, dt, nsteps, precision, mode):
    eng = pkg.Engine(precision, kernel_mode=mode)
    ids = [eng.add_cell(s, dt) for s in specs]
    for cid, s in zip(ids, specs):
        eng.set_probes(cid, np.arange(s.n), nsteps)
    eng.reinit(dt)
    eng.step(nsteps // 2)
    eng.step(nsteps - nsteps // 2)
    out = []
    spikes = eng.read_spikes()
    for cid, s in zip(ids, specs):
        r = oracle_mod.OracleCell(s, dt)
        r.reinit(dt)
        ref = r.run(dt, nsteps, np.arange(s.n))
        out.append((eng.read_probes(cid), ref, list(spikes[spikes["cell"] == cid]["step"]), list(r.read_spikes()[1])))
    eng.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_fp64_kahp2d_batch(pkg, oracle_mod, mode):
    print(f"case: K-AHP 2-D batch (3 cells + 2 plain), 600 steps, fp64, {mode}", flush=True)
    for gpu, ref, s_gpu, s_ref in run_both(pkg, oracle_mod, kahp2d_batch(), 25e-6, 600, "fp64", mode):
        assert np.abs(gpu - ref).max() <= FP64_TOL
        assert s_gpu == s_ref


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_gpu_mixed_kahp2d_batch(pkg, oracle_mod, mode):
    print(f"case: K-AHP 2-D batch, 600 steps, mixed, {mode}", flush=True)
    for gpu, ref, s_gpu, s_ref in run_both(pkg, oracle_mod, kahp2d_batch(), 25e-6, 600, "mixed", mode):
        assert np.abs(gpu - ref).max() <= MIXED_TOL
        assert len(s_gpu) == len(s_ref)
        assert all(abs(a - b) <= 1 for a, b in zip(s_gpu, s_ref))


@pytest.mark.gpu
def test_gpu_chan2d_fields_and_external_current(pkg, oracle_mod):
    """2-D channel fields through the solver (Gk2D, Z2D, Gbar2D), and a host
    addGkEk on a compartment that also has a 2-D channel (one external sum
    per compartment, host entries first)."""
    print("case: K-AHP 2-D cell 120 compartments, fields + addGkEk, fp64", flush=True)
    dt = 25e-6
    spec = workloads.random_ca_cell(np.random.default_rng(8), 120, inject=0.9e-9, kahp2d=True)
    eng = pkg.Engine("fp64")
    cid = eng.add_cell(spec, dt)
    eng.reinit(dt)
    ref = oracle_mod.OracleCell(spec, dt)
    ref.reinit(dt)
    idx = np.arange(0, 120, 7)
    np.testing.assert_allclose(eng.get(cid, "Z2D", idx), ref.get("Z2D", idx), rtol=1e-12)
    eng.step(150)
    ref.run(dt, 150)
    for f in ("Gk2D", "Z2D"):
        np.testing.assert_allclose(eng.get(cid, f, idx), ref.get(f, idx), rtol=1e-9, atol=1e-18, err_msg=f)
    eng.set(cid, "Gbar2D", [3], [5e-9])
    ref.set("Gbar2D", [3], [5e-9])
    eng.add_gkek(cid, 3, 2e-9, -0.03)
    ref.add_gkek(3, 2e-9, -0.03)
    eng.step(1)
    ref.run(dt, 1)
    eng.step(100)
    ref.run(dt, 100)
    v_gpu = eng.get(cid, "Vm", np.arange(120))
    v_ref = ref.get("Vm", np.arange(120))
    assert np.abs(v_gpu - v_ref).max() <= FP64_TOL
    np.testing.assert_allclose(eng.get(cid, "Im", idx), ref.get("Im", idx), rtol=1e-6, atol=1e-15)
    eng.close()
