"""Why the mixed form keeps gate states in FP64 (CPU, the oracle only).

Cell 3622 of the batched HH ensemble (seed 1, n ~ U{50..2000}; this one has
68 compartments and 0.298 nA at the soma) fires every ~700 steps just above
rheobase, so its spike times are the most sensitive in the ensemble.  On the
oracle's FP64 solve:

* rounding Gbar and Ek to FP32 once (the mixed form stores them so) leaves
  every spike step unchanged and moves Vm by well under 1 mV;
* rounding every gate state to FP32 after every step (what an FP32 state
  array does) moves the spikes by several steps: near rest a gate's per-step
  increment is below half its ulp, so the rounding drops it, and the bias
  builds up over the ~900 steps to the first spike.

The GPU's mixed form therefore keeps the states and their update in FP64,
with FP32 tables, parameters, powers and currents.
"""
import numpy as np
import pytest

from moose_core_amd import workloads

DT, NSTEPS = 25e-6, 5000


def _ensemble_cell(k, ncells=4096, nmin=50, nmax=2000, seed=1):
    """Cell k of workloads.hh_ensemble(ncells, nmin, nmax, seed), built
    without the cells after it."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=ncells)
    for i in range(k + 1):
        cell = workloads.random_hh_cell(rng, int(sizes[i]))
    return cell


def _run(oracle_mod, spec, params32=False, states32=False):
    o = oracle_mod.OracleCell(spec, DT)
    o.reinit(DT)
    ch = np.arange(len(spec.channels))
    withY = ch[spec.channels["Ypower"] > 0]
    if params32:
        for f in ("Gbar", "Ek"):
            o.set(f, ch, o.get(f, ch).astype(np.float32).astype(np.float64))
    if not states32:
        tr = o.run(DT, NSTEPS, [0])[:, 0]
    else:
        tr = np.zeros(NSTEPS)
        for k in range(NSTEPS):
            tr[k] = o.run(DT, 1, [0])[0, 0]
            for f, sel in (("X", ch), ("Y", withY)):
                o.set(f, sel, o.get(f, sel).astype(np.float32).astype(np.float64))
    return tr, [int(s) for s in o.read_spikes()[1]]


@pytest.fixture(scope="module")
def cell_and_base(oracle_mod):
    spec = _ensemble_cell(3622)
    assert spec.n == 68
    base, spikes = _run(oracle_mod, spec)
    assert len(spikes) >= 5
    return spec, base, spikes


def test_fp32_parameters_keep_the_spike_steps(oracle_mod, cell_and_base):
    spec, base, spikes = cell_and_base
    tr, s = _run(oracle_mod, spec, params32=True)
    assert s == spikes
    assert np.abs(tr - base).max() < 1e-3


def test_fp32_gate_states_move_the_spikes(oracle_mod, cell_and_base):
    spec, base, spikes = cell_and_base
    tr, s = _run(oracle_mod, spec, states32=True)
    assert len(s) == len(spikes)
    shift = max(abs(a - b) for a, b in zip(s, spikes))
    print(f"spike steps {spikes} -> {s}, max|dVm| {np.abs(tr - base).max():.3e} V")
    assert shift >= 5
