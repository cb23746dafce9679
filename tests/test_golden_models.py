"""End-to-end golden vectors from the reference's own test suite.

tests/rdesigneur/test_32_squid_axon_propgn.py (moose-core) builds a squid
axon with rdesigneur and checks 41 soma Vm samples (numpy.allclose).  The
model is rebuilt here through the object layer as rdesigneur would build it:

* compartments as rdesigneur's buildCompt makes them
  (python/rdesigneur/rdesigneurProtos.py:620-638: length from dx, dy; Ra, Rm,
  Cm from RA, RM, CM with that file's PI = 3.14159265359; Em = initVm =
  -65 mV), a 10 um soma and 200 axon segments of 10 um x 1 um bending as the
  test's makeAxonProto does;
* HH Na and K channels with tables as make_HH_Na / make_HH_K fill them
  (rdesigneurProtos.py:61-108: 3000 divisions over -110..50 mV and
  -120..40 mV, the squid rate functions of V - EREST_ACT_HH), distributed to
  every compartment with Gbar = density * len * dia * PI
  (biophysics/Neuron.cpp:703-712, PI = 3.141592653589793), Na before K;
* an HSolve at dt = 50 us on the soma, a Table on the 0.1 ms plot clock
  (tick 8), and the stimulus "(t>0.005 && t<0.2) * 2e-11" into the soma's
  inject: a Function on tick 12 at 0.1 ms (rdesigneur.py:1252-1330,
  scheduling/Clock.cpp:928), which runs after the solver's tick 6 at
  t = dt * step for even steps, so the injection starts with step 103 (the
  first even step with t > 5 ms is 102; 50e-6 * 100 == 0.005 exactly).

The oracle (CPU) and the GPU path (FP64 and mixed) are each checked against
the reference's 41 values with the reference test's own numpy.allclose.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "squid_axon_propagation.json")))

PI_PROTOS = 3.14159265359        # rdesigneurProtos.py's PI (buildCompt)
PI = 3.141592653589793           # basecode/header.h (channel distribution)
EREST_ACT_HH = -0.070
DT = 50e-6
ONSET_STEP = 102                 # the Function's first firing with t > 5 ms
NSTEPS = 800                     # 40 ms


def squid_tables(kind):
    """Gate tables (A = alpha, B = alpha + beta, per second) of the squid
    channels over the test model's voltage grids, in the arithmetic order of
    the model's rate functions."""
    if kind == "Na":
        v = np.linspace(-110e-3, 50e-3, 3001) - EREST_ACT_HH
        m_alpha = 1e3 * (25 - v * 1e3) / (10 * (np.exp((25 - v * 1e3) / 10) - 1))
        m_beta = 1e3 * 4 * np.exp(- v * 1e3 / 18)
        h_alpha = 1e3 * 0.07 * np.exp(-v / 20e-3)
        h_beta = 1e3 * 1 / (np.exp((30e-3 - v) / 10e-3) + 1)
        return [(-110e-3, 50e-3, m_alpha, m_alpha + m_beta), (-110e-3, 50e-3, h_alpha, h_alpha + h_beta)]
    v = np.linspace(-120e-3, 40e-3, 3001) - EREST_ACT_HH
    n_alpha = 1e3 * (10 - v * 1e3) / (100 * (np.exp((10 - v * 1e3) / 10) - 1))
    n_beta = 1e3 * 0.125 * np.exp(- v * 1e3 / 80)
    return [(-120e-3, 40e-3, n_alpha, n_alpha + n_beta)]


def build_squid_axon(moose, nseg=200, comptLen=10e-6, comptDia=1e-6, RM=1.0, RA=10.0, CM=0.01):
    moose.reset()
    moose.Neutral("/library")
    protos = []
    for name, ek, xp, yp in (("Na", 50e-3, 3, 1), ("K", -77e-3, 4, 0)):
        ch = moose.HHChannel(f"/library/{name}")
        ch.Ek = ek
        ch.Xpower = xp
        if yp:
            ch.Ypower = yp
        for gate, (vmin, vmax, A, B) in zip((ch.gateX, ch.gateY), squid_tables(name)):
            gate.min, gate.max, gate.divs = vmin, vmax, 3000
            gate.tableA = A
            gate.tableB = B
        protos.append(ch)
    moose.Neutral("/model")
    moose.Neutral("/model/elec")

    def build_compt(name, dia, x, y, dx, dy):
        length = np.sqrt(dx * dx + dy * dy)
        c = moose.Compartment(f"/model/elec/{name}")
        c.diameter, c.length = dia, length
        xa = dia * dia * PI_PROTOS / 4.0
        sa = length * dia * PI_PROTOS
        c.Ra = length * RA / xa
        c.Rm = RM / sa
        c.Cm = CM * sa
        c.Em = c.initVm = -0.065
        return c

    comps = [build_compt("soma", 10e-6, 0.0, 0.0, comptLen, 0.0)]
    theta, x, y = 0.0, comptLen, 0.0
    for i in range(nseg):
        dx = comptLen * np.cos(theta)
        dy = comptLen * np.sin(theta)
        r = np.sqrt(x * x + y * y)
        theta += comptLen / r
        c = build_compt(f"axon{i}", comptDia, x, y, dx, dy)
        moose.connect(comps[-1], "axial", c, "raxial")
        comps.append(c)
        x += dx
        y += dy
    for proto, dens in zip(protos, (1200.0, 360.0)):
        for c in comps:
            ch = moose.copy(proto, c, proto.name)
            ch.Gbar = dens * c.length * c.diameter * PI
            moose.connect(ch, "channel", c, "channel")
    return comps


def samples_ok(v):
    return np.allclose(v, np.array(GOLD["Vm"]))


def test_squid_axon_oracle(pkg, oracle_mod):
    """The oracle on the rebuilt model reproduces the reference's 41 samples."""
    from moose_core_amd import moose
    comps = build_squid_axon(moose)
    hs = moose.HSolve("/model/elec/hsolve")
    hs.dt = DT
    spec, objs = hs.build_spec(comps[0])
    soma = objs["compartments"].index(comps[0])
    o = oracle_mod.OracleCell(spec, DT)
    o.reinit(DT)
    v = [o.get("Vm", [soma])[0]]
    tr = o.run(DT, ONSET_STEP, [soma])[:, 0]
    o.set("inject", [soma], [2e-11])
    tr = np.concatenate([tr, o.run(DT, NSTEPS - ONSET_STEP, [soma])[:, 0]])
    v = np.array(v + list(tr[19::20]))
    assert len(v) == 41
    print("max |dV| vs the reference's samples:", np.abs(v - np.array(GOLD["Vm"])).max())
    assert samples_ok(v)
    moose.reset()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_squid_axon_gpu(pkg, precision):
    """The same model through the object layer on the GPU: Table on the
    0.1 ms plot clock, the stimulus written at the Function's firing."""
    print(f"case: squid axon (rdesigneur test_32), 201 compartments, {precision}", flush=True)
    from moose_core_amd import moose
    moose.set_precision(precision)
    try:
        comps = build_squid_axon(moose)
        moose.Neutral("/model/graphs")
        tab = moose.Table("/model/graphs/plot0")
        moose.connect(tab, "requestOut", comps[0], "getVm")
        moose.setClock(8, 0.1e-3)
        hs = moose.HSolve("/model/elec/hsolve")
        hs.dt = DT
        hs.target = "/model/elec/soma"
        moose.reinit()
        moose.start(ONSET_STEP * DT)
        comps[0].inject = 2e-11
        moose.start((NSTEPS - ONSET_STEP) * DT)
        v = tab.vector
        assert len(v) == 401
        print("max |dV| vs the reference's samples:", np.abs(v[::10] - np.array(GOLD["Vm"])).max(), flush=True)
        assert samples_ok(v[::10])
    finally:
        moose.set_precision("fp64")
        moose.reset()
