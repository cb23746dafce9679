"""Benchmark: compartment-steps/s of the batched HSolve ensemble on one MI355X.

Workload (the batched-ensemble configuration): 4096 independent cells, each a
random binary tree (recursive uniform splitting) with n ~ U{50..2000}
compartments, branch lengths U[5, 50] um, diameters tapering 2 -> 0.5 um,
squid HH Na/K in every compartment with Gbar scaled by U[0.8, 1.2] per cell,
soma inject U[0, 0.3] nA, dt = 25 us.  All cells advance together in one
persistent fused launch per `steps_per_launch` ticks.

Timed: `--steps` ticks of the whole ensemble (default 5000, the full run),
after `--warmup` untimed ticks, bracketed by device synchronisation.  The
headline `value` is the mixed-precision form (FP32 gates/tables/currents, FP64
Hines solve and Vm); the FP64 form (bit-identical to the oracle) is run the
same way and reported beside it.

Accuracy of the mixed form against the FP64 form comes from a separate,
untimed run of `--accuracy-steps` ticks (default: the workload's full length)
after a fresh reinit of both engines, whatever `--steps` is: the soma Vm of
every cell at every step and every spike step.

The default line also carries a short serving leg (Poisson arrivals of Ca
cells with the 2-D K-AHP, continuous batching): p50/p99 request latency and
sustained compartment-steps/s (`--serving-requests 0` leaves it out).

The CPU baseline is the reference's own HSolve step (hsolve/HSolveActive.cpp
and the files it calls, compiled from the reference's sources into
oracle/_ref/hsolve_ref, on the oracle's setup of each cell) on a bounded
sample of the same cells, one cell after another on one core and sharded
over processes on the GPU's share of host cores.  Where that build is absent,
or the workload has objects beside the solver (HHChannel2D, SynChan), which
the reference runs outside HSolve, the oracle (the bit-exact C restatement)
is timed instead.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

pkg = __graft_entry__.import_package()
from moose_core_amd import autotune, workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0


def bytes_per_cs(w_chan, w_2d, gates, chans, gates2d=0, chans2d=0, pools=0):
    """Algorithmic bytes per compartment-step: the compulsory traffic with
    the state resident between steps, lookup tables and elimination
    temporaries excluded (they stay in LDS/L2).  V read+write 16; every gate
    state read+written (FP64 in both forms) 16; Gbar and Ek of every channel
    read (w each); CmByDt, EmByRm, the diagonal base and the off-diagonal
    (FP64) 32; per Ca pool its state read+written, f1 and f2 read (FP64) 32.
    w = 8 in the FP64 form and 4 for the HHChannels in the mixed form; the
    HHChannel2D objects stay FP64 in both (w_2d = 8)."""
    return 16 + 16 * gates + 2 * w_chan * chans + 16 * gates2d + 2 * w_2d * chans2d + 32 + 32 * pools


# squid HH cells: Na m^3 h and K n^4 (3 gates, 2 channels) -> 128 / 112 B
BYTES_PER_CS = {"fp64": bytes_per_cs(8, 8, 3, 2), "mixed": bytes_per_cs(4, 8, 3, 2)}
# the large Ca cells: Na m^3 h, Kdr n^4, KA a^3 b, CaL s^2 (6 gates, 4
# channels), the 2-D K-AHP (1 gate, 1 channel, FP64 in both forms) and one
# CaConc per compartment -> 272 / 240 B
BYTES_PER_CS_LARGE = {"fp64": bytes_per_cs(8, 8, 6, 4, 1, 1, 1), "mixed": bytes_per_cs(4, 8, 6, 4, 1, 1, 1)}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_cells(ncells, seed):
    t = time.time()
    cells = workloads.hh_ensemble(ncells, 50, 2000, seed)
    log(f"[bench] built {ncells} cells ({sum(c.n for c in cells)} compartments) in {time.time() - t:.1f}s")
    return cells


def build_large_ca(ncells, seed, nmin=5000, nmax=30000):
    """Large uneven cells with calcium: branching 2-4, Na/Kdr/KA/CaL, K-AHP as
    an HHChannel2D on a 3000 x 500 Vm x [Ca] table, one CaConc per
    compartment, n ~ U{5000..30000}; the soma injected with U[0.8, 1.2] pA
    per compartment, so that the cells fire (the spike-time accuracy of the
    mixed form is then measured on them)."""
    t = time.time()
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=ncells)
    cells = [workloads.random_ca_cell(rng, int(n), inject=float(rng.uniform(0.8e-12, 1.2e-12)) * int(n),
                                      kahp2d=True) for n in sizes]
    log(f"[bench] built {ncells} large Ca cells ({sum(c.n for c in cells)} compartments) in {time.time() - t:.1f}s")
    return cells


def build_uniform_hh(ncells, seed, n=2048):
    """ncells random binary-tree HH cells of exactly n compartments."""
    rng = np.random.default_rng(seed)
    cells = [workloads.random_hh_cell(rng, n) for _ in range(ncells)]
    log(f"[bench] built {ncells} HH cells of {n} compartments")
    return cells


def build_network(ncells, seed, fan_in=10):
    """A spiking network: random binary-tree HH cells (n ~ U{50..500}, soma
    inject U[0, 0.3] nA), each with a SynChan on its soma (tau 1/3 ms, Ek 0)
    whose fan_in synapses (weight U[0.5, 1.5], delay U[1, 5] ms) receive the
    SpikeGens of random other cells.  Returns the cells and the connections
    (source cell, spikegen, target cell, synchan, synapse)."""
    from moose_core_amd._abi import SYNAPSE_DT, SYNCHAN_DT
    rng = np.random.default_rng(seed)
    cells, conns = [], []
    for c in range(ncells):
        s = workloads.random_hh_cell(rng, int(rng.integers(50, 501)))
        sc = np.zeros(1, SYNCHAN_DT)
        sc["compartment"], sc["n_synapses"] = 0, fan_in
        sc["Gbar"], sc["Ek"], sc["tau1"], sc["tau2"], sc["modulation"] = 3e-9, 0.0, 1e-3, 3e-3, 1.0
        s.synchans = sc
        syn = np.zeros(fan_in, SYNAPSE_DT)
        syn["weight"] = rng.uniform(0.5, 1.5, fan_in)
        syn["delay"] = rng.uniform(1e-3, 5e-3, fan_in)
        s.synapses = syn
        cells.append(s)
        for j, src in enumerate(rng.choice(ncells, size=fan_in, replace=False)):
            conns.append((int(src), 0, c, 0, j))
    log(f"[bench] built a {ncells}-cell network ({sum(c.n for c in cells)} compartments, "
        f"{len(conns)} connections)")
    return cells, conns


# workload name -> (description, default cells, dt, default steps, steps per
# launch, graph chunk)
WORKLOADS = {
    "ensemble": ("batched ensemble: 4096 random binary-tree HH cells, n~U{50..2000}, dt=25us, 5000 steps",
                 4096, 25e-6, 5000, 50, 0),
    "large_ca": ("large uneven Ca cells: 256 random trees (branching 2-4), n~U{5000..30000}, Na/Kdr/KA/CaL + "
                 "2-D K-AHP (Vm x [Ca], 3000 x 500), one CaConc per compartment, soma injected with "
                 "U[0.8,1.2] pA per compartment (the cells fire), dt=20us, 10000 steps, graphs of 100 steps",
                 256, 20e-6, 10000, 100, 1),
    "hh_2048": ("uniform HH cells: 1953 random binary trees of 2048 compartments (4.0 M), dt=25us, "
                "1000 steps", 1953, 25e-6, 1000, 50, 0),
    "network": ("spiking network: 4096 random binary-tree HH cells, n~U{50..500}, a somatic SynChan each "
                "with 10 synapses from random cells' SpikeGens (delays 1-5 ms), dt=25us, 4000 steps",
                4096, 25e-6, 4000, 50, 0),
}


def run_form(precision, packed, dt, steps, warmup, spl, graph, threads, mode="fused", conns=(), acc_steps=0,
             plans=None, tune=True, tuned=None):
    """Tunes the launch plan bin by bin (autotune.tune_bins, untimed) and
    reinits; times `steps` ticks after `warmup` untimed ones; then, untimed,
    a fresh reinit and `acc_steps` ticks recording the soma Vm of every cell
    and the spikes (the accuracy leg)."""
    eng = pkg.Engine(precision, steps_per_launch=spl, graph_chunk=graph, block_threads=threads,
                     kernel_mode=mode)
    t = time.time()
    ids = [eng.add_cell(p, dt) for p in packed]
    for (a, g, b, k, j) in conns:
        eng.connect_spike(ids[a], g, ids[b], k, j)
    log(f"[bench] {precision}: {len(ids)} cells set up in {time.time() - t:.1f}s")
    if tune and threads == 0:
        eng.reinit(dt)
        t = time.time()
        res = autotune.tune_bins(eng, steps=max(10, spl), repeats=2)
        log(f"[bench] {precision}: autotuned {len(res['bins'])} bins in {time.time() - t:.1f}s: "
            f"{res['ms_per_step_default']:.4f} -> {res['ms_per_step_tuned']:.4f} ms/step")
        if tuned is not None:
            tuned[precision] = res
    for cid in ids:
        eng.set_probes(cid, [0], max(steps, acc_steps))
    eng.reinit(dt)
    if warmup:
        eng.step(warmup)
        for cid in ids:
            eng.read_probes(cid)
        eng.read_spikes()
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    wall = time.perf_counter() - t0
    launches, total_ms, mean_ms = eng.last_timing()
    if plans is not None:
        plans[precision] = eng.plan_info()
    log(f"[bench] {precision}: {steps} steps in {wall:.3f}s, {launches} launches, mean launch {mean_ms:.3f} ms")
    soma = spikes = None
    if acc_steps:
        for cid in ids:
            eng.read_probes(cid)
        eng.read_spikes()
        eng.reinit(dt)
        t1 = time.perf_counter()
        eng.step(acc_steps)
        eng.sync()
        log(f"[bench] {precision}: accuracy leg, {acc_steps} steps from reinit in {time.perf_counter() - t1:.3f}s")
        soma = np.stack([eng.read_probes(cid)[:, 0] for cid in ids], axis=1)
        spikes = eng.read_spikes()
    eng.close()
    return wall, launches, total_ms, mean_ms, soma, spikes


def measured_traffic(key, compartment_steps):
    """HBM bytes per launch round from the committed FETCH_SIZE/WRITE_SIZE
    profile of this workload and form (profiles/traffic.json: rocprofv3 PMC
    runs, scaled by a plain device copy of known size), and the profile it
    comes from; (None, None) when no profile of this workload exists."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)[key]
    except (OSError, KeyError, ValueError):
        return None, None
    per_cs = t["fetch_bytes_per_compartment_step"] + t["write_bytes_per_compartment_step"]
    return per_cs * compartment_steps, f"{t['source']} ({per_cs:.1f} B per compartment-step)"


def spike_steps(spikes, ncells):
    out = [[] for _ in range(ncells)]
    for r in spikes:
        out[int(r["cell"])].append(int(r["step"]))
    return out


# host cores timed beside one GPU: the box's CPU share of one GPU
HOST_CORES_PER_GPU = 16


def cpu_baseline(cells, dt, budget_s=15.0):
    """Oracle on the host: single core (cells one after another) and the
    GPU's share of host cores (cells sharded over threads; ctypes releases
    the GIL).  The sample is every 16th cell of the workload (sizes spread
    like the whole), stepped for as many ticks as ~budget_s seconds of one
    core's work allow."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    sample = cells[::16]
    probe_steps = 20
    # calibrate: compartment-steps/s on one core
    c0 = oracle.OracleCell(sample[0], dt)
    c0.reinit(dt)
    t = time.perf_counter()
    c0.run(dt, probe_steps)
    rate1 = sample[0].n * probe_steps / max(time.perf_counter() - t, 1e-9)
    total_cs = budget_s * rate1 / 2
    nsteps = int(max(10, min(5000, total_cs / sum(c.n for c in sample))))
    objs = []
    for c in sample:
        o = oracle.OracleCell(c, dt)
        o.reinit(dt)
        objs.append(o)
    t = time.perf_counter()
    for o in objs:
        o.run(dt, nsteps)
    single = sum(c.n for c in sample) * nsteps / (time.perf_counter() - t)
    ncores = max(1, min(HOST_CORES_PER_GPU, os.cpu_count() or 1))
    for o in objs:
        o.reinit(dt)
    chunks = [objs[i::ncores] for i in range(ncores)]
    ths = [threading.Thread(target=lambda ch: [o.run(dt, nsteps) for o in ch], args=(ch,)) for ch in chunks]
    t = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    multi = sum(c.n for c in sample) * nsteps / (time.perf_counter() - t)
    return {
        "value": multi, "unit": "compartment-steps/s", "cores": ncores, "kind": "port",
        "single_core_value": single,
        "sample": f"every 16th cell of the workload ({len(sample)} cells, {sum(c.n for c in sample)} "
                  f"compartments) x {nsteps} steps, oracle (bit-exact C restatement of HSolve), "
                  f"-O3 -ffp-contract=off, {ncores} threads (one GPU's share of the host)",
    }


def _ref_time(binary, nsteps, dt, flats):
    """One hsolve_ref process timing its cells: (compartments, seconds)."""
    out = subprocess.run([binary, "--time", str(int(nsteps)), repr(float(dt))] + flats, capture_output=True,
                         text=True, check=True).stdout.split()
    return int(out[0]), float(out[1])


def cpu_baseline_reference(cells, dt, budget_s=15.0):
    """The reference's own HSolveActive::step (oracle/_ref/hsolve_ref, built
    from hsolve/*.cpp by oracle/Makefile) on the oracle's setup of every 16th
    cell: one process stepping the sample on one core, then the sample
    sharded over processes on the GPU's share of host cores (wall time from
    their start to the last one's end).  None when the build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not oracle.ref_available():
        return None
    sample = cells[::16]
    ncomp = sum(c.n for c in sample)
    ncores = max(1, min(HOST_CORES_PER_GPU, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as td:
        flats = []
        for k, c in enumerate(sample):
            o = oracle.OracleCell(c, dt)
            path = os.path.join(td, f"cell{k}.bin")
            o.dump_flat(path)
            o.close()
            flats.append(path)
        n0, s0 = _ref_time(oracle.REF_BIN, 20, dt, flats[:1])
        rate1 = n0 * 20 / max(s0, 1e-9)
        nsteps = int(max(10, min(5000, budget_s * rate1 / 2 / ncomp)))
        n1, s1 = _ref_time(oracle.REF_BIN, nsteps, dt, flats)
        single = n1 * nsteps / s1
        shards = [flats[i::ncores] for i in range(ncores) if flats[i::ncores]]
        t = time.perf_counter()
        procs = [subprocess.Popen([oracle.REF_BIN, "--time", str(nsteps), repr(float(dt))] + sh,
                                  stdout=subprocess.PIPE, text=True) for sh in shards]
        for p in procs:
            p.communicate()
        wall = time.perf_counter() - t
        if any(p.returncode != 0 for p in procs):
            raise RuntimeError("hsolve_ref failed")
    return {
        "value": ncomp * nsteps / wall, "unit": "compartment-steps/s", "cores": len(shards), "kind": "reference",
        "single_core_value": single,
        "sample": f"every 16th cell of the workload ({len(sample)} cells, {ncomp} compartments) x {nsteps} steps, "
                  f"the reference's own HSolveActive::step (hsolve/*.cpp built from its sources, -O3 "
                  f"-ffp-contract=off) on the oracle's setup of each cell, {len(shards)} processes (one GPU's "
                  f"share of the host), wall time including process start",
    }


def serving_requests(rng, count, rate, dt):
    """`count` requests with Poisson arrivals at `rate` per second: a seeded
    random Ca cell with the 2-D K-AHP of n ~ log-U[100, 30000] compartments
    for a duration log-U[10 ms, 1 s] at dt."""
    from moose_core_amd import serving
    reqs, arrival = [], 0.0
    for k in range(count):
        arrival += rng.exponential(1.0 / rate)
        n = int(round(np.exp(rng.uniform(np.log(100), np.log(30000)))))
        steps = int(round(np.exp(rng.uniform(np.log(0.010 / dt), np.log(1.0 / dt)))))
        cell = workloads.random_ca_cell(rng, n, kahp2d=True).pack()
        reqs.append(serving.Request(k, cell, steps, arrival))
    return reqs


def serving_leg(count, rate, seed, quantum=100, mem_budget_gib=64, kernel_mode="auto"):
    """The serving loop (continuous batching, admission by device-memory
    footprint) over `count` Poisson requests, mixed precision, dt = 25 us.
    Returns the summary (p50/p99 latency, sustained compartment-steps/s)."""
    from moose_core_amd import serving
    dt = 25e-6
    rng = np.random.default_rng(seed)
    t = time.time()
    reqs = serving_requests(rng, count, rate, dt)
    log(f"[bench] serving: {count} requests built in {time.time() - t:.1f}s "
        f"({sum(r.spec.n for r in reqs)} compartments, {sum(r.nsteps for r in reqs)} request-steps)")
    srv = serving.Server("mixed", dt, mem_budget=mem_budget_gib << 30, quantum=quantum, kernel_mode=kernel_mode)
    # warm-up: one small request through the loop (kernels loaded, arena grown)
    srv.run([serving.Request(-1, workloads.random_ca_cell(rng, 200, kahp2d=True).pack(), 200, 0.0)],
            realtime=False)
    res = srv.run(reqs, realtime=True)
    summ = serving.summarize(res, srv.wall, srv.cs)
    summ.update({"rate_per_s": rate, "quantum_steps": quantum, "mem_budget_gib": mem_budget_gib,
                 "max_active": srv.max_active, "quanta": srv.quanta})
    srv.close()
    log(f"[bench] serving: p50 {summ['p50_latency_s']:.2f}s p99 {summ['p99_latency_s']:.2f}s, "
        f"{summ['sustained_cs_per_s']:.3e} compartment-steps/s")
    return summ


def run_serving(args):
    """The serving workload on its own: one JSON line with the p50/p99
    request latency (arrival to result) and the sustained throughput."""
    summ = serving_leg(args.cells, args.rate, args.seed, args.quantum, args.mem_budget, args.kernel_mode)
    out = {
        "metric": "serving: sustained compartment-steps/s with p50/p99 request latency",
        "value": summ["sustained_cs_per_s"], "unit": "compartment-steps/s",
        "n_gpus": 1, "steps": None, "warmup": 1, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-channels/f64-hines", "data": "synthetic (seeded random morphologies, Poisson arrivals)",
        "config": {"workload": "serving loop: Poisson arrivals, n~logU[100,30000] Ca cells with 2-D K-AHP, "
                               "duration logU[10ms,1s], dt=25us, mixed precision",
                   "requests": args.cells, "rate_per_s": args.rate, "quantum_steps": args.quantum,
                   "mem_budget_gib": args.mem_budget, "max_active": summ["max_active"], "quanta": summ["quanta"]},
        "p50_latency_s": summ["p50_latency_s"], "p99_latency_s": summ["p99_latency_s"],
        "wall_s": summ["wall_s"], "compartment_steps": summ["compartment_steps"],
    }
    print(json.dumps(out))


def copy_probe(mib):
    """Plain device-to-device copy of a known size, for counter calibration."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 5}))


def accuracy_block(soma, soma64, spikes, spk64, ncells, steps):
    """Mixed form against the FP64 form over the accuracy leg: per cell the
    max |dVm| of the soma trace, and spike steps within one step."""
    d = soma - soma64
    s_mixed = spike_steps(spikes, ncells)
    s_64 = spike_steps(spk64, ncells)
    worst = 0
    mismatched = 0
    for a, b in zip(s_mixed, s_64):
        if len(a) != len(b):
            mismatched += 1
            continue
        if a:
            worst = max(worst, int(np.max(np.abs(np.array(a) - np.array(b)))))
    per_cell = np.abs(d).max(axis=0)
    spike_ok = np.array([len(a) == len(b) and (not a or np.max(np.abs(np.array(a) - np.array(b))) <= 1)
                         for a, b in zip(s_mixed, s_64)])
    bad = np.nonzero((per_cell > 1e-4) | ~spike_ok)[0]
    return {
        "steps": steps, "max_abs_dVm": float(per_cell.max()), "rms_dVm": float(np.sqrt(np.mean(d * d))),
        "per_cell_max_abs_dVm_p50": float(np.percentile(per_cell, 50)),
        "per_cell_max_abs_dVm_p99": float(np.percentile(per_cell, 99)),
        "cells_within_0.1mV": int(np.sum(per_cell <= 1e-4)),
        "cells_within_0.1mV_and_spikes_within_1dt": int(np.sum((per_cell <= 1e-4) & spike_ok)),
        "cells": ncells, "cells_outside_bound": [int(c) for c in bad[:16]],
        "spike_step_max_err": worst, "cells_with_spike_count_mismatch": mismatched,
        "spikes": int(sum(len(a) for a in s_64)),
        "probe": "soma of every cell, every step, from reinit (untimed run)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="ensemble", choices=sorted(WORKLOADS) + ["serving"],
                    help="ensemble: the headline batched-ensemble line; large_ca: the large Ca cells "
                         "with 2-D K-AHP channels; hh_2048: uniform 2048-compartment HH cells; network: "
                         "4096 HH cells joined by SpikeGen -> SynChan synapses; serving: the request loop "
                         "(p50/p99 latency)")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: request arrivals per second")
    ap.add_argument("--quantum", type=int, default=100, help="serving: steps per batch quantum")
    ap.add_argument("--mem-budget", type=int, default=64, help="serving: device memory budget, GiB")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--accuracy-steps", type=int, default=None,
                    help="ticks of the untimed accuracy leg (default: the workload's full length; 0: none)")
    ap.add_argument("--serving-requests", type=int, default=None,
                    help="requests of the serving leg added to the line (default 24 for the ensemble, "
                         "0 otherwise)")
    ap.add_argument("--cells", type=int, default=None)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--steps-per-launch", type=int, default=None)
    ap.add_argument("--graph-chunk", type=int, default=None)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--kernel-mode", default="auto", choices=["fused", "split", "auto"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the engine's own launch shapes")
    ap.add_argument("--forms", default="mixed,fp64")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(cells are independent and would shard by bin-packing across GPUs)"}))
        sys.exit(2)
    if args.workload == "serving":
        args.cells = 200 if args.cells is None else args.cells
        run_serving(args)
        return
    desc, dcells, dt, dsteps, dspl, dgraph = WORKLOADS[args.workload]
    args.cells = dcells if args.cells is None else args.cells
    args.steps = dsteps if args.steps is None else args.steps
    args.steps_per_launch = dspl if args.steps_per_launch is None else args.steps_per_launch
    args.graph_chunk = dgraph if args.graph_chunk is None else args.graph_chunk
    if args.accuracy_steps is None:
        args.accuracy_steps = dsteps if "mixed" in args.forms and "fp64" in args.forms else 0
    if args.serving_requests is None:
        args.serving_requests = 24 if args.workload == "ensemble" else 0
    conns = ()
    if args.workload == "ensemble":
        cells = build_cells(args.cells, args.seed)
    elif args.workload == "network":
        cells, conns = build_network(args.cells, args.seed)
    elif args.workload == "hh_2048":
        cells = build_uniform_hh(args.cells, args.seed)
    else:
        cells = build_large_ca(args.cells, args.seed)
    packed = [c.pack() for c in cells]
    ncomp = sum(c.n for c in cells)
    res = {}
    plans = {}
    tuned = {}
    for form in args.forms.split(","):
        res[form] = run_form(form, packed, dt, args.steps, args.warmup, args.steps_per_launch,
                             args.graph_chunk, args.threads, args.kernel_mode, conns, args.accuracy_steps,
                             plans, not args.no_autotune, tuned)
    head = "mixed" if "mixed" in res else list(res)[0]
    wall, launches, total_ms, mean_ms, soma, spikes = res[head]
    value = ncomp * args.steps / wall
    # one launch round: every size bin's dispatch, concurrently on its own
    # stream; algorithmic bytes of the round over its duration (HIP events
    # on the engine stream, which forks to and joins the bin streams)
    bpcs = BYTES_PER_CS_LARGE if args.workload == "large_ca" else BYTES_PER_CS
    steps_per_launch = args.steps / max(1, launches)
    bytes_per_launch = bpcs[head] * ncomp * steps_per_launch
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9
    wl_key = {"ensemble": f"hh_ensemble_{args.cells}", "large_ca": f"large_ca_{args.cells}",
              "hh_2048": f"hh_2048_{args.cells}", "network": f"network_{args.cells}"}[args.workload]
    traffic, traffic_src = measured_traffic(f"{head}/{wl_key}", ncomp * steps_per_launch)
    bins = plans[head]
    dom = max(bins, key=lambda b: b["round_ms"]) if bins else None
    out = {
        "metric": "compartment-steps/s (batched cells, FP32-channel mixed form)",
        "value": value,
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-channels/f64-hines" if head == "mixed" else "f64",
        "data": "synthetic (seeded random morphologies)",
        "config": {"workload": desc, "cells": args.cells, "compartments": ncomp, "dt": dt,
                   "steps_per_launch": args.steps_per_launch, "graph_chunk": args.graph_chunk,
                   "kernel_mode": args.kernel_mode,
                   "bins": [{k: b[k] for k in ("kernel", "cells", "compartments", "threads", "solve_threads",
                                               "lds_bytes", "split", "fast_slots", "round_ms")}
                            | ({"tuned": tuned[head]["bins"][i]["chosen"]} if head in tuned else {})
                            for i, b in enumerate(bins)]}
                  | ({"connections": len(conns), "spikes_routed": int(len(spikes))} if conns else {})
                  | ({"autotune": {"ms_per_step_default": tuned[head]["ms_per_step_default"],
                                   "ms_per_step_tuned": tuned[head]["ms_per_step_tuned"],
                                   "trials": len(tuned[head]["trials"]),
                                   "method": "per-bin threads and channel pass, each bin's candidates timed on the "
                                             "whole round with HIP events (tune_bins), then reinit"}}
                     if head in tuned else {}),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": (f"{dom['kernel']} (the longest of the {len(bins)} concurrent size-bin "
                                f"dispatches of one launch round)") if dom else None,
                     "launch": "one launch round: every size bin's dispatch on its own stream",
                     "bytes_per_compartment_step": bpcs[head], "mean_launch_ms": mean_ms},
    }
    if "fp64" in res and head != "fp64":
        w64, l64, t64, m64, soma64, spk64 = res["fp64"]
        out["fp64_value"] = ncomp * args.steps / w64
        out["mixed_over_fp64"] = value / out["fp64_value"]
        out["fp64_roofline_frac"] = (bpcs["fp64"] * ncomp * (args.steps / max(1, l64))
                                     / (m64 * 1e-3) / 1e9) / HBM_PEAK_GBS
        if soma is not None and soma64 is not None:
            out["accuracy_mixed_vs_fp64"] = accuracy_block(soma, soma64, spikes, spk64, len(cells),
                                                           args.accuracy_steps)
    if args.serving_requests > 0:
        out["serving"] = serving_leg(args.serving_requests, args.rate, args.seed, args.quantum, args.mem_budget)
        out["serving"]["requests_workload"] = ("Poisson arrivals, n~logU[100,30000] Ca cells with 2-D K-AHP, "
                                               "duration logU[10ms,1s], dt=25us, mixed precision")
    if not args.no_cpu:
        # objects beside the solver run outside the reference's HSolve: the
        # oracle covers them
        beside = any(len(c.chans2d) or len(c.synchans) for c in cells)
        ref = None if beside else cpu_baseline_reference(cells, dt)
        if ref is not None:
            log("[bench] CPU baseline: the reference's HSolve step")
            out["cpu_baseline"] = ref
        else:
            log("[bench] CPU baseline (oracle) ...")
            out["cpu_baseline"] = cpu_baseline(cells, dt)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
