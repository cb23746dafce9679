"""Writes the golden fixtures of tests/golden/ from the reference's own hsolve
sources (oracle/_ref/hsolve_ref, built by `make -C oracle ref` from the
moose-core tree).  Each fixture holds the model (as the C-ABI description
arrays) and what the reference computed for it: V of every compartment every
`stride` steps, and the final VMid, gate states, Ca and Gk.

The reference harness is fed the oracle's setup of the model (Hines numbering,
tables, channel bookkeeping); see oracle/ref_harness.cpp.

usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import __graft_entry__  # noqa: E402

__graft_entry__.import_package()
import oracle  # noqa: E402
from moose_core_amd import workloads  # noqa: E402

CASES = {
    # name: (builder, dt, nsteps, stride)
    "ball_tree_hh": (lambda: workloads.hh_ball_tree(inject=1e-9), 50e-6, 600, 5),
    "random_hh_60": (lambda: workloads.random_hh_cell(np.random.default_rng(3), 60, inject=0.8e-9),
                     25e-6, 800, 5),
    "random_ca_40": (lambda: workloads.random_ca_cell(np.random.default_rng(5), 40, inject=0.5e-9),
                     20e-6, 600, 5),
}


def spec_arrays(spec):
    d = {"compartments": spec.compartments, "axial": spec.axial, "seed": np.int32(spec.seed),
         "channels": spec.channels, "pools": spec.pools, "spikegens": spec.spikegens,
         "leaks": spec.leaks}
    for i, g in enumerate(spec.gates):
        d[f"gate{i}_hdr"] = np.array([g.id, g.divs], dtype=np.int64)
        d[f"gate{i}_rng"] = np.array([g.min, g.max])
        d[f"gate{i}_A"] = np.asarray(g.tableA)
        d[f"gate{i}_B"] = np.asarray(g.tableB)
    d["ngates"] = np.int32(len(spec.gates))
    return d


def main():
    if not oracle.ref_available():
        raise SystemExit("build the reference harness first: make -C oracle ref")
    for name, (build, dt, nsteps, stride) in CASES.items():
        spec = build()
        cell = oracle.OracleCell(spec, dt)
        V, VMid, state, ca, Gk = oracle.run_reference(cell, dt, nsteps)
        order = cell.hines_order()
        out = spec_arrays(spec)
        out.update({"dt": np.float64(dt), "nsteps": np.int64(nsteps), "stride": np.int64(stride),
                    "hines_order": order,
                    # V in the caller's (desc) numbering, every `stride` steps
                    "V": V[stride - 1::stride][:, order],
                    "VMid": VMid, "state": state, "ca": ca, "Gk": Gk})
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{path}: {spec.n} compartments, {nsteps} steps, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
