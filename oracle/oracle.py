"""Python binding of the oracle (liboracle.so) and of the reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check of
__graft_entry__.py and by the cpu_baseline leg of bench.py; never by the
product.  The oracle restates MOOSE's HSolve on the CPU (hsolve_oracle.c);
the harness (_ref/hsolve_ref) runs the reference's own hsolve sources.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "hsolve_ref")
_LIB = None


def build(ref=False):
    """Builds liboracle.so (and the reference harness if asked and possible)."""
    targets = ["liboracle.so"] + (["_ref/hsolve_ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, dbl, P = C.c_void_p, C.c_int, C.c_double, C.POINTER
        sig = {
            "or_create": ([vp, dbl, i32, P(vp)], i32),
            "or_destroy": ([vp], None),
            "or_reinit": ([vp, dbl], None),
            "or_step": ([vp, dbl, dbl], None),
            "or_run": ([vp, dbl, C.c_long, P(i32), i32, P(dbl)], None),
            "or_n_compartments": ([vp], i32),
            "or_n_channels": ([vp], i32),
            "or_n_states": ([vp], i32),
            "or_n_pools": ([vp], i32),
            "or_hines_order": ([vp, P(i32)], None),
            "or_get_field": ([vp, i32, P(i32), i32, P(dbl)], i32),
            "or_set_field": ([vp, i32, P(i32), i32, P(dbl)], i32),
            "or_add_inject": ([vp, i32, dbl], i32),
            "or_add_gkek": ([vp, i32, dbl, dbl], i32),
            "or_add_spike": ([vp, i32, i32, dbl], i32),
            "or_add_conc": ([vp, i32, dbl], i32),
            "or_add_ica": ([vp, i32, dbl], i32),
            "or_read_spikes": ([vp, P(i32), P(C.c_long), i32], i32),
            "or_get_V": ([vp, P(dbl)], None),
            "or_get_VMid": ([vp, P(dbl)], None),
            "or_get_state": ([vp, P(dbl)], None),
            "or_get_ca": ([vp, P(dbl)], None),
            "or_dump_flat": ([vp, C.c_char_p], i32),
            "or_matrix_create": ([i32, P(i32), P(i32), P(dbl), P(dbl), P(dbl), P(dbl), P(dbl), dbl], vp),
            "or_matrix_destroy": ([vp], None),
            "or_matrix_getA": ([vp, i32, i32], dbl),
            "or_matrix_getB": ([vp, i32], dbl),
            "or_matrix_getVMid": ([vp, i32], dbl),
            "or_matrix_getV": ([vp, i32], dbl),
            "or_matrix_update": ([vp], None),
            "or_matrix_forward": ([vp], None),
            "or_matrix_backward": ([vp], None),
            "or_lookup_row": ([dbl, dbl, C.c_uint, dbl, P(C.c_uint)], dbl),
            "or_interp2d": ([vp, dbl, dbl], dbl),
            "or_integrate_ee": ([dbl, dbl, dbl, dbl], dbl),
        }
        for name, (a, r) in sig.items():
            f = getattr(L, name)
            f.argtypes = a
            f.restype = r
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


FIELD = {"Vm": 1, "Im": 2, "inject": 3, "initVm": 4, "Cm": 5, "Em": 6, "Rm": 7, "Ra": 8,
         "Gbar": 20, "Ek": 21, "Gk": 22, "Ik": 23, "X": 24, "Y": 25, "Z": 26,
         "modulation": 27, "instant": 28, "Ca": 40, "CaBasal": 41, "ceiling": 42, "floor": 43,
         "Gbar2D": 60, "Ek2D": 61, "Gk2D": 62, "X2D": 64, "Y2D": 65, "Z2D": 66, "modulation2D": 67,
         "GbarSyn": 80, "EkSyn": 81, "GkSyn": 82, "tau1": 83, "tau2": 84, "modulationSyn": 85,
         "XSyn": 86, "YSyn": 87, "weight": 90, "delay": 91,
         "GbarLeak": 95, "EkLeak": 96, "modulationLeak": 97, "command": 100, "current": 101,
         "pulseOutput": 110, "pulseInput": 111}


class OracleCell:
    """One cell solved by the CPU oracle."""

    def __init__(self, spec, dt, caAdvance=1):
        L = lib()
        self.L = L
        self.packed = spec.pack() if hasattr(spec, "pack") else spec
        h = C.c_void_p()
        rc = L.or_create(C.cast(self.packed.ref(), C.c_void_p), float(dt), int(caAdvance), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"oracle setup failed ({rc})")
        self.h = h
        self.ncomp = len(self.packed.compartments)
        self.n = L.or_n_compartments(h)

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reinit(self, dt):
        self.L.or_reinit(self.h, float(dt))

    def run(self, dt, nsteps, probes=None):
        probes = np.asarray([] if probes is None else probes, dtype=np.int32)
        tr = np.zeros(max(1, nsteps * len(probes)), dtype=np.float64)
        self.L.or_run(self.h, float(dt), int(nsteps), _ip(probes), len(probes),
                      _dp(tr) if len(probes) else None)
        return tr[: nsteps * len(probes)].reshape(nsteps, len(probes))

    def hines_order(self):
        out = np.zeros(self.ncomp, dtype=np.int32)
        self.L.or_hines_order(self.h, _ip(out))
        return out

    def get(self, field, idx):
        idx = np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int32)
        out = np.zeros(len(idx))
        rc = self.L.or_get_field(self.h, FIELD[field], _ip(idx), len(idx), _dp(out))
        if rc:
            raise IndexError(f"oracle get {field} failed ({rc})")
        return out

    def set(self, field, idx, vals):
        idx = np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int32)
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(vals, dtype=np.float64), idx.shape))
        rc = self.L.or_set_field(self.h, FIELD[field], _ip(idx), len(idx), _dp(v))
        if rc:
            raise IndexError(f"oracle set {field} failed ({rc})")

    def add_inject(self, c, v):
        self.L.or_add_inject(self.h, int(c), float(v))

    def add_spike(self, synchan, synapse, time):
        if self.L.or_add_spike(self.h, int(synchan), int(synapse), float(time)):
            raise IndexError("oracle add_spike: no such synchan or synapse")

    def add_gkek(self, c, g, e):
        self.L.or_add_gkek(self.h, int(c), float(g), float(e))

    def add_conc(self, ch, conc):
        self.L.or_add_conc(self.h, int(ch), float(conc))

    def add_ica(self, p, v):
        self.L.or_add_ica(self.h, int(p), float(v))

    def read_spikes(self, cap=1 << 16):
        g = np.zeros(cap, dtype=np.int32)
        s = np.zeros(cap, dtype=np.int64)
        n = self.L.or_read_spikes(self.h, _ip(g), s.ctypes.data_as(C.POINTER(C.c_long)), cap)
        return g[:n], s[:n]

    def V(self):
        out = np.zeros(self.n)
        self.L.or_get_V(self.h, _dp(out))
        return out

    def state(self):
        out = np.zeros(self.L.or_n_states(self.h))
        self.L.or_get_state(self.h, _dp(out))
        return out

    def ca(self):
        out = np.zeros(self.L.or_n_pools(self.h))
        self.L.or_get_ca(self.h, _dp(out))
        return out

    def dump_flat(self, path):
        rc = self.L.or_dump_flat(self.h, path.encode())
        if rc:
            raise RuntimeError("dump failed")


def ref_available():
    return os.path.exists(REF_BIN)


def run_reference(cell, dt, nsteps, reinit=True):
    """Runs the reference's own hsolve step on the oracle's setup of `cell`
    (an OracleCell right after creation).  Returns V after every step
    (nsteps, n) in Hines order and the final VMid, state, ca, Gk."""
    with tempfile.TemporaryDirectory() as td:
        flat = os.path.join(td, "flat.bin")
        out = os.path.join(td, "out.bin")
        cell.dump_flat(flat)
        subprocess.run([REF_BIN, flat, str(int(nsteps)), repr(float(dt)), out, "1" if reinit else "0"],
                       check=True)
        raw = np.fromfile(out, dtype=np.uint8)
    hdr = raw[:16].view(np.int64)
    n, ns = int(hdr[0]), int(hdr[1])
    body = raw[16:].view(np.float64)
    V = body[: n * ns].reshape(ns, n)
    rest = body[n * ns:]
    nst = cell.L.or_n_states(cell.h)
    npool = cell.L.or_n_pools(cell.h)
    nch = cell.L.or_n_channels(cell.h)
    VMid = rest[:n]
    state = rest[n:n + nst]
    ca = rest[n + nst:n + nst + npool]
    Gk = rest[n + nst + npool:n + nst + npool + nch]
    return V, VMid, state, ca, Gk


class OracleMatrix:
    """A bare Hines matrix (HinesMatrix::setup on a tree in Hines indices)."""

    def __init__(self, children, Ra, Rm, Cm, Em, V, dt):
        n = len(children)
        cstart = np.zeros(n + 1, dtype=np.int32)
        for i, c in enumerate(children):
            cstart[i + 1] = cstart[i] + len(c)
        ch = np.array([x for c in children for x in c] or [0], dtype=np.int32)
        self.arrs = [cstart, ch] + [np.ascontiguousarray(a, dtype=np.float64) for a in (Ra, Rm, Cm, Em, V)]
        L = lib()
        self.L = L
        self.n = n
        self.h = L.or_matrix_create(n, _ip(cstart), _ip(ch), *[_dp(a) for a in self.arrs[2:]], float(dt))

    def __del__(self):
        try:
            self.L.or_matrix_destroy(self.h)
        except Exception:
            pass

    def A(self):
        return np.array([[self.L.or_matrix_getA(self.h, i, j) for j in range(self.n)] for i in range(self.n)])

    def B(self):
        return np.array([self.L.or_matrix_getB(self.h, i) for i in range(self.n)])

    def VMid(self):
        return np.array([self.L.or_matrix_getVMid(self.h, i) for i in range(self.n)])

    def V(self):
        return np.array([self.L.or_matrix_getV(self.h, i) for i in range(self.n)])

    def update(self):
        self.L.or_matrix_update(self.h)

    def forward(self):
        self.L.or_matrix_forward(self.h)

    def backward(self):
        self.L.or_matrix_backward(self.h)
