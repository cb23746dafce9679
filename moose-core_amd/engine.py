"""ctypes binding of libhsolve_amd.so (include/hsolve_amd.h).

The library is built in-tree by csrc/build.sh (or __graft_entry__.build()).
There is no CPU fallback: if the library or a GPU is missing, the engine
raises.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import FIELDS, CellInfo, Options, Spike

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhsolve_amd.so")
_LIB = None

# exported entry points of include/hsolve_amd.h
EXPORTS = [
    "hsg_create", "hsg_destroy", "hsg_add_cell", "hsg_remove_cell", "hsg_reinit",
    "hsg_reinit_cell", "hsg_step", "hsg_step_cells", "hsg_step_cells_noalloc", "hsg_get_field", "hsg_set_field",
    "hsg_add_inject", "hsg_add_gkek", "hsg_add_conc", "hsg_add_ica", "hsg_set_pulsegen", "hsg_add_spike", "hsg_connect_spike", "hsg_read_spikes",
    "hsg_set_probes", "hsg_set_probes_every", "hsg_set_probe_fields", "hsg_read_probes", "hsg_cell_info_get", "hsg_hines_order",
    "hsg_sync", "hsg_last_step_timing", "hsg_plan_info", "hsg_set_bin_shape", "hsg_last_error",
]


class HSolveError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def load_library(path=None):
    """Loads (once) and returns the HIP engine library; raises if absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise HSolveError(_abi.HSG_E_STATE,
                          f"{p} is not built: run moose-core_amd/csrc/build.sh "
                          "(or __graft_entry__.build())")
    lib = C.CDLL(p)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    P = C.POINTER
    sig = {
        "hsg_create": ([P(Options), P(vp)], C.c_int),
        "hsg_destroy": ([vp], C.c_int),
        "hsg_add_cell": ([vp, P(_abi.CellDesc), dbl, i32, P(i32)], C.c_int),
        "hsg_remove_cell": ([vp, i32], C.c_int),
        "hsg_reinit": ([vp, dbl], C.c_int),
        "hsg_reinit_cell": ([vp, i32], C.c_int),
        "hsg_step": ([vp, i64], C.c_int),
        "hsg_step_cells": ([vp, P(i32), P(i64), i32], C.c_int),
        "hsg_step_cells_noalloc": ([vp, P(i32), P(i64), i32], C.c_int),
        "hsg_get_field": ([vp, i32, i32, P(i32), i32, P(dbl)], C.c_int),
        "hsg_set_field": ([vp, i32, i32, P(i32), i32, P(dbl)], C.c_int),
        "hsg_add_inject": ([vp, i32, i32, dbl], C.c_int),
        "hsg_add_gkek": ([vp, i32, i32, dbl, dbl], C.c_int),
        "hsg_add_conc": ([vp, i32, i32, dbl], C.c_int),
        "hsg_add_ica": ([vp, i32, i32, dbl], C.c_int),
        "hsg_set_pulsegen": ([vp, i32, i32, vp], C.c_int),
        "hsg_add_spike": ([vp, i32, i32, i32, dbl], C.c_int),
        "hsg_connect_spike": ([vp, i32, i32, i32, i32, i32], C.c_int),
        "hsg_read_spikes": ([vp, P(Spike), i32, P(i32), P(i64)], C.c_int),
        "hsg_set_probes": ([vp, i32, P(i32), i32, i64], C.c_int),
        "hsg_set_probes_every": ([vp, i32, P(i32), i32, i64, i32], C.c_int),
        "hsg_set_probe_fields": ([vp, i32, P(i32), P(i32), i32, i64, i32], C.c_int),
        "hsg_read_probes": ([vp, i32, P(dbl), i64, P(i64)], C.c_int),
        "hsg_cell_info_get": ([vp, i32, P(CellInfo)], C.c_int),
        "hsg_hines_order": ([vp, i32, P(i32)], C.c_int),
        "hsg_sync": ([vp], C.c_int),
        "hsg_last_step_timing": ([vp, P(i32), P(dbl), P(dbl)], C.c_int),
        "hsg_plan_info": ([vp, P(_abi.BinInfo), i32, P(i32)], C.c_int),
        "hsg_set_bin_shape": ([vp, i32, i32, i32], C.c_int),
        "hsg_last_error": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    if path is None:
        _LIB = lib
    return lib


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


SPIKE_DT = np.dtype([("cell", "i4"), ("spikegen", "i4"), ("step", "i8")])
assert SPIKE_DT.itemsize == C.sizeof(Spike)


class Engine:
    """A batch of HSolve cells on one GPU."""

    def __init__(self, precision="fp64", device=0, mem_budget=0, steps_per_launch=0,
                 graph_chunk=0, block_threads=0, kernel_mode="auto"):
        self.lib = load_library()
        o = Options()
        o.precision = _abi.HSG_MIXED if precision == "mixed" else _abi.HSG_FP64
        o.device = device
        o.mem_budget = int(mem_budget)
        o.steps_per_launch = int(steps_per_launch)
        o.graph_chunk = int(graph_chunk)
        o.block_threads = int(block_threads)
        if kernel_mode not in _abi.KERNEL_MODES:
            raise ValueError(f"kernel_mode must be one of {sorted(_abi.KERNEL_MODES)}")
        o.kernel_mode = _abi.KERNEL_MODES[kernel_mode]
        self.precision = precision
        h = C.c_void_p()
        self._check(self.lib.hsg_create(C.byref(o), C.byref(h)))
        self.h = h
        self._packed = {}

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.hsg_last_error().decode(errors="replace")
            raise HSolveError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.hsg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- cells
    def add_cell(self, spec, dt, caAdvance=1):
        packed = spec if isinstance(spec, _abi.PackedCell) else spec.pack()
        cid = C.c_int32()
        self._check(self.lib.hsg_add_cell(self.h, packed.ref(), float(dt), int(caAdvance),
                                          C.byref(cid)))
        return cid.value

    def remove_cell(self, cell):
        self._check(self.lib.hsg_remove_cell(self.h, int(cell)))

    def reinit(self, dt):
        self._check(self.lib.hsg_reinit(self.h, float(dt)))

    def reinit_cell(self, cell):
        self._check(self.lib.hsg_reinit_cell(self.h, int(cell)))

    def step(self, nsteps=1):
        self._check(self.lib.hsg_step(self.h, int(nsteps)))

    def step_cells(self, cells, nsteps):
        c, cp = _i32(cells)
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(nsteps, dtype=np.int64), c.shape))
        self._check(self.lib.hsg_step_cells(self.h, cp, s.ctypes.data_as(C.POINTER(C.c_int64)),
                                            len(c)))

    def step_cells_noalloc(self, cells, nsteps):
        """step_cells without device allocations (hsg_step_cells_noalloc):
        launch tables in the library's static per-GPU area."""
        c, cp = _i32(cells)
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(nsteps, dtype=np.int64), c.shape))
        self._check(self.lib.hsg_step_cells_noalloc(self.h, cp, s.ctypes.data_as(C.POINTER(C.c_int64)),
                                                    len(c)))

    def sync(self):
        self._check(self.lib.hsg_sync(self.h))

    # ---- fields
    def get(self, cell, field, idx):
        code = FIELDS[field] if isinstance(field, str) else int(field)
        i, ip = _i32(np.atleast_1d(idx))
        out = np.zeros(len(i), dtype=np.float64)
        self._check(self.lib.hsg_get_field(self.h, int(cell), code, ip, len(i),
                                           out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set(self, cell, field, idx, vals):
        code = FIELDS[field] if isinstance(field, str) else int(field)
        i, ip = _i32(np.atleast_1d(idx))
        v, vp = _f64(np.broadcast_to(np.asarray(vals, dtype=np.float64), i.shape))
        self._check(self.lib.hsg_set_field(self.h, int(cell), code, ip, len(i), vp))

    def add_inject(self, cell, compartment, value):
        self._check(self.lib.hsg_add_inject(self.h, int(cell), int(compartment), float(value)))

    def add_gkek(self, cell, compartment, Gk, Ek):
        self._check(self.lib.hsg_add_gkek(self.h, int(cell), int(compartment), float(Gk), float(Ek)))

    def add_conc(self, cell, channel, conc):
        self._check(self.lib.hsg_add_conc(self.h, int(cell), int(channel), float(conc)))

    def add_ica(self, cell, pool, iCa):
        self._check(self.lib.hsg_add_ica(self.h, int(cell), int(pool), float(iCa)))

    def add_spike(self, cell, synchan, synapse, time):
        """Synapse::addSpike: a spike at `time` through the synapse."""
        self._check(self.lib.hsg_add_spike(self.h, int(cell), int(synchan), int(synapse), float(time)))

    def connect_spike(self, src_cell, spikegen, dst_cell, synchan, synapse):
        """spikeOut of a SpikeGen -> addSpike of a synapse (any two cells)."""
        self._check(self.lib.hsg_connect_spike(self.h, int(src_cell), int(spikegen), int(dst_cell),
                                               int(synchan), int(synapse)))

    def set_pulsegen(self, cell, index, rec):
        """New parameters for the cell's index-th PulseGen (one PULSEGEN_DT
        record); its output level carries over."""
        r = np.ascontiguousarray(np.asarray(rec, _abi.PULSEGEN_DT).reshape(1))
        self._check(self.lib.hsg_set_pulsegen(self.h, int(cell), int(index), r.ctypes.data))

    # ---- outputs
    def read_spikes(self, cap=1 << 16):
        """Spikes since the last call (and clears them): records of
        (cell, spikegen, step)."""
        out = np.zeros(cap, dtype=SPIKE_DT)
        n = C.c_int32()
        dropped = C.c_int64()
        self._check(self.lib.hsg_read_spikes(self.h, out.ctypes.data_as(C.POINTER(Spike)), cap,
                                             C.byref(n), C.byref(dropped)))
        self.spikes_dropped = dropped.value
        return out[: n.value].copy()

    def set_probes(self, cell, compartments, capacity, every=1):
        """Vm of `compartments` sampled every `every` steps (after the step),
        up to `capacity` samples until read_probes."""
        c, cp = _i32(compartments)
        if every == 1:
            self._check(self.lib.hsg_set_probes(self.h, int(cell), cp, len(c), int(capacity)))
        else:
            self._check(self.lib.hsg_set_probes_every(self.h, int(cell), cp, len(c), int(capacity), int(every)))
        self._probe_n = getattr(self, "_probe_n", {})
        self._probe_n[int(cell)] = (len(c), int(capacity))

    def set_probe_fields(self, cell, fields, idx, capacity, every=1):
        """Probes on fields ("Vm", "Ca", "current", "GkSyn") of objects
        `idx`, sampled on the device every `every` steps; read_probes
        returns one column per probe."""
        f, fp = _i32([FIELDS[x] for x in fields])
        i, ip = _i32(idx)
        if len(f) != len(i):
            raise ValueError("one index per field")
        self._check(self.lib.hsg_set_probe_fields(self.h, int(cell), fp, ip, len(i), int(capacity), int(every)))
        self._probe_n = getattr(self, "_probe_n", {})
        self._probe_n[int(cell)] = (len(i), int(capacity))

    def read_probes(self, cell):
        n, cap = self._probe_n[int(cell)]
        out = np.zeros(max(1, n * cap), dtype=np.float64)
        ns = C.c_int64()
        self._check(self.lib.hsg_read_probes(self.h, int(cell),
                                             out.ctypes.data_as(C.POINTER(C.c_double)), cap,
                                             C.byref(ns)))
        return out[: ns.value * n].reshape(ns.value, n)

    def cell_info(self, cell):
        ci = CellInfo()
        self._check(self.lib.hsg_cell_info_get(self.h, int(cell), C.byref(ci)))
        return {k: getattr(ci, k) for k, _ in CellInfo._fields_}

    def hines_order(self, cell, ncomp):
        out = np.zeros(ncomp, dtype=np.int32)
        self._check(self.lib.hsg_hines_order(self.h, int(cell),
                                             out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def plan_info(self):
        """The launch plan of the live set: one dict per size bin (launch
        shape, kernel, and its dispatches' duration in the last timed
        launch round)."""
        cap = 64
        arr = (_abi.BinInfo * cap)()
        n = C.c_int32()
        self._check(self.lib.hsg_plan_info(self.h, arr, cap, C.byref(n)))
        out = []
        for b in arr[: min(n.value, cap)]:
            d = {k: getattr(b, k) for k, _ in _abi.BinInfo._fields_}
            d["kernel"] = b.kernel.decode()
            out.append(d)
        return out

    def set_bin_shape(self, bin, threads=0, fast_pass=-1):
        """Launch shape of plan bin `bin` (hsg_set_bin_shape): threads per
        cell workgroup (0: the engine's choice) and the fast channel pass
        (-1: the engine's choice).  bin = -1 clears every choice."""
        self._check(self.lib.hsg_set_bin_shape(self.h, int(bin), int(threads), int(fast_pass)))

    def last_timing(self):
        n = C.c_int32()
        tot = C.c_double()
        mean = C.c_double()
        self._check(self.lib.hsg_last_step_timing(self.h, C.byref(n), C.byref(tot), C.byref(mean)))
        return n.value, tot.value, mean.value
