"""The oracle against the reference: bit-exact against the reference's own
hsolve sources (oracle/_ref/hsolve_ref) and against the committed golden
fixtures written from that build (tests/golden/make_golden.py)."""
import glob
import os

import numpy as np
import pytest

from moose_core_amd import workloads
from moose_core_amd.cellspec import CellSpec, Gate

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def spec_from_fixture(z):
    gates = []
    for i in range(int(z["ngates"])):
        gid, divs = z[f"gate{i}_hdr"]
        mn, mx = z[f"gate{i}_rng"]
        gates.append(Gate(int(gid), float(mn), float(mx), int(divs), z[f"gate{i}_A"], z[f"gate{i}_B"]))
    return CellSpec(z["compartments"], z["axial"], int(z["seed"]), z["channels"], gates,
                    z["pools"], z["spikegens"], z["leaks"])


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_golden(oracle_mod, path):
    z = np.load(path)
    spec = spec_from_fixture(z)
    dt, nsteps, stride = float(z["dt"]), int(z["nsteps"]), int(z["stride"])
    cell = oracle_mod.OracleCell(spec, dt)
    assert np.array_equal(cell.hines_order(), z["hines_order"])
    cell.reinit(dt)
    tr = cell.run(dt, nsteps, np.arange(spec.n))
    np.testing.assert_array_equal(tr[stride - 1::stride], z["V"])
    np.testing.assert_array_equal(cell.state(), z["state"])
    np.testing.assert_array_equal(cell.ca(), z["ca"])
    Gk = cell.get("Gk", np.arange(len(spec.channels)))
    # fixture Gk is in the oracle's channel order (Hines order of compartments)
    assert np.array_equal(np.sort(Gk), np.sort(z["Gk"]))


needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "hsolve_ref")),
    reason="reference harness not built (needs the moose-core sources)")


@needs_ref
@pytest.mark.parametrize("seed,n,kind", [(11, 120, "hh"), (12, 33, "hh"), (13, 90, "ca"), (14, 1, "hh")])
def test_oracle_matches_reference_build(oracle_mod, seed, n, kind):
    rng = np.random.default_rng(seed)
    spec = (workloads.random_hh_cell(rng, n, inject=1.2e-9) if kind == "hh"
            else workloads.random_ca_cell(rng, n, inject=0.7e-9))
    dt, nsteps = 25e-6, 400
    ref = oracle_mod.OracleCell(spec, dt)
    V, VMid, state, ca, Gk = oracle_mod.run_reference(ref, dt, nsteps)
    cell = oracle_mod.OracleCell(spec, dt)
    cell.reinit(dt)
    order = cell.hines_order()
    tr = cell.run(dt, nsteps, np.arange(spec.n))
    np.testing.assert_array_equal(tr, V[:, order])
    np.testing.assert_array_equal(cell.state(), state)
    np.testing.assert_array_equal(cell.ca(), ca)


@needs_ref
def test_oracle_matches_reference_with_inputs(oracle_mod):
    """caAdvance 0, a basal inject changed between runs, and no reinit."""
    spec = workloads.random_ca_cell(np.random.default_rng(21), 50, inject=0.2e-9)
    dt = 20e-6
    ref = oracle_mod.OracleCell(spec, dt, caAdvance=0)
    V, VMid, state, ca, Gk = oracle_mod.run_reference(ref, dt, 300, reinit=False)
    cell = oracle_mod.OracleCell(spec, dt, caAdvance=0)
    order = cell.hines_order()
    tr = cell.run(dt, 300, np.arange(spec.n))
    np.testing.assert_array_equal(tr, V[:, order])
    np.testing.assert_array_equal(cell.ca(), ca)


@needs_ref
def test_reference_timing_mode(oracle_mod, tmp_path):
    """The reference harness's CPU-baseline mode (bench.py's cpu_baseline):
    it loads every cell given, steps them and reports their compartments and
    a positive loop time."""
    import subprocess
    rng = np.random.default_rng(8)
    cells = [workloads.random_hh_cell(rng, n, inject=1e-10) for n in (60, 250)]
    flats = []
    for k, c in enumerate(cells):
        o = oracle_mod.OracleCell(c, 25e-6)
        flats.append(str(tmp_path / f"c{k}.bin"))
        o.dump_flat(flats[-1])
        o.close()
    out = subprocess.run([oracle_mod.REF_BIN, "--time", "50", repr(25e-6)] + flats, capture_output=True,
                         text=True, check=True).stdout.split()
    assert int(out[0]) == sum(c.n for c in cells)
    assert float(out[1]) > 0.0
