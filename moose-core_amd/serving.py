"""Continuous-batching request loop over one engine.

Simulation requests (a cell, a number of steps, an arrival time) arrive as
they come.  The loop admits waiting requests into the running batch whenever
the engine's device-memory budget has room (admission by footprint: the
engine's arena is sized to the budget and `add_cell` fails with HSG_E_NOMEM
when a cell does not fit), keeps every admitted cell's Vm / gate / Ca state
resident in HBM, and advances the whole batch in quanta of `quantum` steps
with one `step_cells` call per quantum (cells that need fewer steps get
fewer).  A request finishes when its steps are done: its soma trace and
spikes are read back, its slot is released, and the next waiting request can
be admitted before the next quantum.  Requests that arrive while others run
share those launches.

Each cell's arithmetic does not depend on its batch-mates, so a request's
result is the same alone and in any mix (bit-identical; checked by the GPU
tests).
"""
import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .engine import Engine, HSolveError


@dataclass
class Request:
    rid: int
    spec: object            # CellSpec or PackedCell
    nsteps: int
    arrival: float = 0.0    # seconds after the loop starts
    probes: tuple = (0,)    # compartments whose Vm is returned (desc indices)
    caAdvance: int = 1


@dataclass
class Result:
    rid: int
    n: int
    nsteps: int
    arrival: float
    admitted: float
    finished: float
    trace: np.ndarray = field(repr=False)   # [nsteps, len(probes)]
    spikes: np.ndarray = field(repr=False)  # step indices (1-based) per spike

    @property
    def latency(self):
        return self.finished - self.arrival


class Server:
    """One engine serving requests at a fixed dt."""

    def __init__(self, precision="mixed", dt=25e-6, mem_budget=8 << 30, quantum=100, **engine_opts):
        self.dt = float(dt)
        self.quantum = int(quantum)
        self.eng = Engine(precision, mem_budget=int(mem_budget), **engine_opts)
        self.eng.reinit(self.dt)

    def close(self):
        self.eng.close()

    def run(self, requests, realtime=True):
        """Serves `requests` (any order) and returns their Results in
        completion order.  realtime: arrivals follow the wall clock (a
        request is visible once its arrival time has passed); otherwise all
        are visible at once and admitted by footprint."""
        eng = self.eng
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        active = {}          # cell -> [request, steps left, admitted, spike steps]
        results = []
        t0 = time.perf_counter()
        self.cs = 0          # compartment-steps done
        self.quanta = 0
        self.max_active = 0

        def now():
            return time.perf_counter() - t0

        while pending or waiting or active:
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                waiting.append(pending.popleft())
            # admission by footprint, first come first served
            while waiting:
                r = waiting[0]
                cell = None
                try:
                    cell = eng.add_cell(r.spec, self.dt, r.caAdvance)
                    eng.set_probes(cell, list(r.probes), r.nsteps)
                except HSolveError as e:
                    if cell is not None:
                        eng.remove_cell(cell)
                    if e.code == _abi.HSG_E_NOMEM and active:
                        break           # wait for a slot to free up
                    raise
                eng.reinit_cell(cell)
                active[cell] = [r, r.nsteps, now(), []]
                waiting.popleft()
            self.max_active = max(self.max_active, len(active))
            if not active:
                if pending and realtime:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
                continue
            cells = list(active)
            steps = [min(self.quantum, active[c][1]) for c in cells]
            eng.step_cells(cells, steps)
            self.quanta += 1
            for c, s in zip(cells, steps):
                st = active[c]
                st[1] -= s
                self.cs += s * _ncomp(st[0].spec)
            spk = eng.read_spikes()
            for rec in spk:
                if int(rec["cell"]) in active:
                    active[int(rec["cell"])][3].append(int(rec["step"]))
            done = [c for c in cells if active[c][1] == 0]
            if done:
                eng.sync()
                tf = now()
                for c in done:
                    r, _, ta, sp = active.pop(c)
                    tr = eng.read_probes(c)
                    eng.remove_cell(c)
                    results.append(Result(r.rid, _ncomp(r.spec), r.nsteps, r.arrival, ta, tf, tr,
                                          np.array(sp, dtype=np.int64)))
        self.wall = now()
        return results


def _ncomp(spec):
    return spec.n


def poisson_requests(rng, count, rate, make_cell, steps_range, dt=25e-6):
    """`count` requests with exponential inter-arrival times at `rate` per
    second; make_cell(rng) -> CellSpec; durations uniform in steps_range."""
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        out.append(Request(k, make_cell(rng).pack(), int(rng.integers(steps_range[0], steps_range[1] + 1)), t))
    return out


def summarize(results, wall, cs):
    lat = np.array([r.latency for r in results])
    return {"requests": len(results), "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "wall_s": wall,
            "compartment_steps": int(cs), "sustained_cs_per_s": cs / wall if wall > 0 else 0.0}
