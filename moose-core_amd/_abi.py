"""ctypes mirror of include/hsolve_amd.h (structures and constants).

The model description structures are the drop-in boundary: they carry a cell
in MOOSE object terms (compartments in creation order, axial messages in
creation order, channels in channel-message order, gate prototypes with their
object ids), which is what HSolve reads at setup
(hsolve/HSolvePassive.cpp:45-205, hsolve/HSolveActiveSetup.cpp:162-526).
"""
import ctypes as C

import numpy as np

HSG_OK = 0
HSG_E_ARG = -1
HSG_E_MODEL = -2
HSG_E_STATE = -3
HSG_E_DEVICE = -4
HSG_E_NOMEM = -5
HSG_E_UNSUPPORTED = -6

HSG_FP64 = 0
HSG_MIXED = 1

HSG_KERNEL_AUTO = 0
HSG_KERNEL_FUSED = 1
HSG_KERNEL_SPLIT = 2
KERNEL_MODES = {"auto": HSG_KERNEL_AUTO, "fused": HSG_KERNEL_FUSED, "split": HSG_KERNEL_SPLIT}

FIELDS = {
    "Vm": 1, "Im": 2, "inject": 3, "initVm": 4, "Cm": 5, "Em": 6, "Rm": 7, "Ra": 8,
    "Gbar": 20, "Ek": 21, "Gk": 22, "Ik": 23, "X": 24, "Y": 25, "Z": 26,
    "modulation": 27, "instant": 28,
    "Ca": 40, "CaBasal": 41, "ceiling": 42, "floor": 43,
    "Gbar2D": 60, "Ek2D": 61, "Gk2D": 62, "X2D": 64, "Y2D": 65, "Z2D": 66, "modulation2D": 67,
    "GbarSyn": 80, "EkSyn": 81, "GkSyn": 82, "tau1": 83, "tau2": 84, "modulationSyn": 85,
    "XSyn": 86, "YSyn": 87, "weight": 90, "delay": 91,
    "GbarLeak": 95, "EkLeak": 96, "modulationLeak": 97,
    "command": 100, "current": 101,
    "pulseOutput": 110, "pulseInput": 111,
}
COMPARTMENT_FIELDS = {k for k, v in FIELDS.items() if v < 20}
CHANNEL_FIELDS = {k for k, v in FIELDS.items() if 20 <= v < 40}
POOL_FIELDS = {k for k, v in FIELDS.items() if 40 <= v < 60}
CHAN2D_FIELDS = {k for k, v in FIELDS.items() if 60 <= v < 80}
SYNCHAN_FIELDS = {k for k, v in FIELDS.items() if 80 <= v < 90}
SYNAPSE_FIELDS = {k for k, v in FIELDS.items() if 90 <= v < 95}
LEAK_FIELDS = {k for k, v in FIELDS.items() if 95 <= v < 100}
VCLAMP_FIELDS = {k for k, v in FIELDS.items() if 100 <= v < 110}
PULSE_FIELDS = {k for k, v in FIELDS.items() if v >= 110}


class GateDesc(C.Structure):
    _fields_ = [("id", C.c_int64), ("min", C.c_double), ("max", C.c_double),
                ("divs", C.c_uint32), ("tableA", C.POINTER(C.c_double)),
                ("tableB", C.POINTER(C.c_double))]


class Table2DDesc(C.Structure):
    _fields_ = [("xdivs", C.c_uint32), ("ydivs", C.c_uint32), ("xmin", C.c_double),
                ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
                ("table", C.POINTER(C.c_double))]


class CellDesc(C.Structure):
    _fields_ = [("n_compartments", C.c_int32), ("compartments", C.c_void_p),
                ("n_axial", C.c_int32), ("axial", C.c_void_p),
                ("seed", C.c_int32),
                ("n_channels", C.c_int32), ("channels", C.c_void_p),
                ("n_gates", C.c_int32), ("gates", C.POINTER(GateDesc)),
                ("n_pools", C.c_int32), ("pools", C.c_void_p),
                ("n_spikegens", C.c_int32), ("spikegens", C.c_void_p),
                ("n_leaks", C.c_int32), ("leaks", C.c_void_p),
                ("n_tables2d", C.c_int32), ("tables2d", C.POINTER(Table2DDesc)),
                ("n_chans2d", C.c_int32), ("chans2d", C.c_void_p),
                ("n_pulsegens", C.c_int32), ("pulsegens", C.c_void_p),
                ("n_synchans", C.c_int32), ("synchans", C.c_void_p), ("synapses", C.c_void_p),
                ("n_vclamps", C.c_int32), ("vclamps", C.c_void_p)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("device", C.c_int32),
                ("mem_budget", C.c_int64), ("steps_per_launch", C.c_int32),
                ("graph_chunk", C.c_int32), ("block_threads", C.c_int32),
                ("kernel_mode", C.c_int32)]


class Spike(C.Structure):
    _fields_ = [("cell", C.c_int32), ("spikegen", C.c_int32), ("step", C.c_int64)]


class CellInfo(C.Structure):
    _fields_ = [("n_compartments", C.c_int32), ("n_channels", C.c_int32),
                ("n_gates", C.c_int32), ("n_pools", C.c_int32),
                ("n_spikegens", C.c_int32), ("n_levels", C.c_int32),
                ("max_group", C.c_int32), ("device_bytes", C.c_int64),
                ("steps_done", C.c_int64), ("vMin", C.c_double), ("vMax", C.c_double),
                ("caMin", C.c_double), ("caMax", C.c_double), ("vDiv", C.c_int32), ("caDiv", C.c_int32)]


class BinInfo(C.Structure):
    _fields_ = [("cells", C.c_int32), ("threads", C.c_int32), ("solve_threads", C.c_int32),
                ("lds_bytes", C.c_int32), ("split", C.c_int32), ("msg", C.c_int32),
                ("fast_slots", C.c_int32), ("slow_groups", C.c_int32), ("ext", C.c_int32),
                ("round_steps", C.c_int32), ("compartments", C.c_int64), ("round_ms", C.c_double),
                ("kernel", C.c_char * 64)]


# numpy record layouts identical to the C structs (natural alignment)
COMPARTMENT_DT = np.dtype([(k, "f8") for k in ("Ra", "Rm", "Cm", "Em", "initVm", "Vm", "inject")])
CHANNEL_DT = np.dtype([("compartment", "i4"), ("instant", "i4"), ("useConcentration", "i4"),
                       ("gate", "i4", (3,)), ("caTarget", "i4"), ("caDepend", "i4")]
                      + [(k, "f8") for k in ("Gbar", "Ek", "Xpower", "Ypower", "Zpower",
                                             "X", "Y", "Z", "modulation")], align=True)
CACONC_DT = np.dtype([(k, "f8") for k in ("Ca", "CaBasal", "tau", "B", "ceiling", "floor")])
SPIKEGEN_DT = np.dtype([("compartment", "i4"), ("edgeTriggered", "i4"),
                        ("threshold", "f8"), ("refractT", "f8")], align=True)
LEAK_DT = np.dtype([("compartment", "i4"), ("Gk", "f8"), ("Ek", "f8"), ("Gbar", "f8"),
                    ("modulation", "f8")], align=True)
SYNCHAN_DT = np.dtype([("compartment", "i4"), ("n_synapses", "i4")]
                      + [(k, "f8") for k in ("Gbar", "Ek", "tau1", "tau2", "modulation")], align=True)
SYNAPSE_DT = np.dtype([("weight", "f8"), ("delay", "f8")])
VCLAMP_DT = np.dtype([("compartment", "i4"), ("mode", "i4")] + [(k, "f8") for k in ("tau", "ti", "td", "gain")])
CHAN2D_DT = np.dtype([("compartment", "i4"), ("instant", "i4"), ("tableA", "i4", (3,)),
                      ("tableB", "i4", (3,)), ("dep", "i4", (6,)), ("conc1", "i4"), ("conc2", "i4")]
                     + [(k, "f8") for k in ("Gbar", "Ek", "Xpower", "Ypower", "Zpower",
                                            "X", "Y", "Z", "modulation")], align=True)

assert CHANNEL_DT.itemsize == 104 and SPIKEGEN_DT.itemsize == 24 and LEAK_DT.itemsize == 40
assert SYNCHAN_DT.itemsize == 48 and SYNAPSE_DT.itemsize == 16 and VCLAMP_DT.itemsize == 40
MAX_PULSES = 4
PULSEGEN_DT = np.dtype([("compartment", "i4"), ("n_pulses", "i4"), ("baseLevel", "f8"),
                        ("delay", "f8", (MAX_PULSES,)), ("width", "f8", (MAX_PULSES,)),
                        ("level", "f8", (MAX_PULSES,)), ("trigMode", "i4"), ("input_src", "i4")], align=True)

assert CHAN2D_DT.itemsize == 136
assert PULSEGEN_DT.itemsize == 120


def _ptr(a):
    return a.ctypes.data if a is not None and len(a) else None


def as_records(a, dt):
    """a as a contiguous array of dt; a structured array of another layout
    (an older fixture) is copied field by field, missing fields zero."""
    a = np.asarray(a)
    if a.dtype == dt:
        return np.ascontiguousarray(a)
    if a.dtype.names is None:
        return np.ascontiguousarray(a, dtype=dt)
    out = np.zeros(a.shape, dt)
    for name in a.dtype.names:
        if name in dt.names:
            out[name] = a[name]
    return out


class PackedCell:
    """A cell description packed into C layout; keeps its arrays alive."""

    def __init__(self, spec):
        self.spec = spec
        self.compartments = np.ascontiguousarray(spec.compartments, dtype=COMPARTMENT_DT)
        self.axial = np.ascontiguousarray(spec.axial, dtype=np.int32).reshape(-1)
        self.channels = np.ascontiguousarray(spec.channels, dtype=CHANNEL_DT)
        self.pools = np.ascontiguousarray(spec.pools, dtype=CACONC_DT)
        self.spikegens = np.ascontiguousarray(spec.spikegens, dtype=SPIKEGEN_DT)
        self.leaks = as_records(spec.leaks, LEAK_DT)
        self.synchans = as_records(getattr(spec, "synchans", np.zeros(0, SYNCHAN_DT)), SYNCHAN_DT)
        self.synapses = as_records(getattr(spec, "synapses", np.zeros(0, SYNAPSE_DT)), SYNAPSE_DT)
        self.vclamps = as_records(getattr(spec, "vclamps", np.zeros(0, VCLAMP_DT)), VCLAMP_DT)
        if len(self.synapses) != int(self.synchans["n_synapses"].sum()):
            raise ValueError("synapses: need the sum of the synchans' n_synapses entries")
        self.chans2d = np.ascontiguousarray(spec.chans2d, dtype=CHAN2D_DT)
        self.pulsegens = np.ascontiguousarray(getattr(spec, "pulsegens", np.zeros(0, PULSEGEN_DT)),
                                              dtype=PULSEGEN_DT)
        self._tables = []           # every table array the descriptors point into
        t2 = (Table2DDesc * max(1, len(spec.tables2d)))()
        for i, t in enumerate(spec.tables2d):
            arr = np.ascontiguousarray(t.table, dtype=np.float64).reshape(-1)
            if arr.size != (t.xdivs + 1) * (t.ydivs + 1):
                raise ValueError(f"2-D table {i}: needs (xdivs+1)*(ydivs+1) values")
            self._tables.append(arr)
            t2[i].xdivs, t2[i].ydivs = t.xdivs, t.ydivs
            t2[i].xmin, t2[i].xmax, t2[i].ymin, t2[i].ymax = t.xmin, t.xmax, t.ymin, t.ymax
            t2[i].table = arr.ctypes.data_as(C.POINTER(C.c_double))
        self._t2 = t2
        gates = (GateDesc * max(1, len(spec.gates)))()
        for i, g in enumerate(spec.gates):
            a = np.ascontiguousarray(g.tableA, dtype=np.float64)
            b = np.ascontiguousarray(g.tableB, dtype=np.float64)
            if len(a) != g.divs + 1 or len(b) != g.divs + 1:
                raise ValueError(f"gate {g.id}: tables need divs+1={g.divs + 1} entries")
            self._tables += [a, b]
            gates[i].id = g.id
            gates[i].min = g.min
            gates[i].max = g.max
            gates[i].divs = g.divs
            gates[i].tableA = a.ctypes.data_as(C.POINTER(C.c_double))
            gates[i].tableB = b.ctypes.data_as(C.POINTER(C.c_double))
        self._gates = gates
        d = CellDesc()
        d.n_compartments = len(self.compartments)
        d.compartments = _ptr(self.compartments)
        d.n_axial = len(self.axial) // 2
        d.axial = _ptr(self.axial)
        d.seed = spec.seed
        d.n_channels = len(self.channels)
        d.channels = _ptr(self.channels)
        d.n_gates = len(spec.gates)
        d.gates = gates
        d.n_pools = len(self.pools)
        d.pools = _ptr(self.pools)
        d.n_spikegens = len(self.spikegens)
        d.spikegens = _ptr(self.spikegens)
        d.n_leaks = len(self.leaks)
        d.leaks = _ptr(self.leaks)
        d.n_tables2d = len(spec.tables2d)
        d.tables2d = self._t2
        d.n_chans2d = len(self.chans2d)
        d.chans2d = _ptr(self.chans2d)
        d.n_pulsegens = len(self.pulsegens)
        d.pulsegens = _ptr(self.pulsegens)
        d.n_synchans = len(self.synchans)
        d.synchans = _ptr(self.synchans)
        d.synapses = _ptr(self.synapses)
        d.n_vclamps = len(self.vclamps)
        d.vclamps = _ptr(self.vclamps)
        self.desc = d

    @property
    def n(self):
        return len(self.compartments)

    def ref(self):
        return C.byref(self.desc)
