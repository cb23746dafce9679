"""A pymoose-style object layer over the engine: build a model from
Compartment, HHChannel, HHGate, HHChannel2D, HHGate2D, CaConc, SpikeGen,
Leakage, PulseGen, SynChan, SimpleSynHandler (with its synapses) and Table
objects
connected by messages, hand it to an HSolve through `target`, then
`reinit()` / `start(t)` and read or write the (zombified) fields.

This mirrors the part of the pymoose surface that an HSolve-managed model
uses (python/moose/__init__.py:57-66 generates the classes from the Cinfos):

  objects and paths        moose.Neutral/Compartment/...('/path'), element(),
                           connect(src, 'srcField', dest, 'destField'), delete()
  HSolve                   fields dt, target, caAdvance, seed, vDiv, vMin,
                           vMax, caDiv, caMin, caMax (hsolve/HSolve.cpp:37-195);
                           setting target runs setup (HSolve::setPath,
                           hsolve/HSolve.cpp:313-334), which needs dt first
  clock                    reinit(), start(runtime), setClock()
  zombie fields            Compartment (biophysics/CompartmentBase.cpp:191-235),
                           HHChannel (biophysics/ChanBase.cpp:90-117,
                           HHChannelBase.cpp:31-76), CaConc
                           (biophysics/CaConcBase.cpp:58-140): reads and writes
                           go to the solver once the model is taken over
                           (hsolve/HSolveInterface.cpp)

Message order is creation order, and object ids grow with creation, so the
solver sees the model exactly as MOOSE would (Hines numbering, channel order,
gate order).  Only models run by an HSolve advance: `start` steps every
HSolve's cells on the GPU; there is no per-object CPU integration here.
"""
import math
import sys

import numpy as np

from . import _abi
from .cellspec import CellSpec, Gate, Table2D

FARADAY = 96485.3415      # basecode/header.h:30
CA_VALENCE = 2            # biophysics/CaConcBase.cpp:15

_objects = {}             # path -> element
_next_id = [1]
_msgs = []                # (src, srcField, dest, destField) in creation order
_solvers = []             # HSolve elements with a target
_engines = {}             # (precision, dt) -> Engine
_precision = ["fp64"]
# Clock tick dts (scheduling/Clock.cpp:999-1016); Tables run on tick 8
_CLOCK_DT = {t: 50e-6 for t in range(8)}
_CLOCK_DT.update({8: 1e-4, 9: 0.0, 10: 0.01})
_CLOCK_DT.update({t: 0.1 for t in range(11, 20)})
_clock_dt = dict(_CLOCK_DT)


def _norm(path):
    if not path.startswith("/"):
        path = "/" + path
    return path.rstrip("/") or "/"


class Neutral:
    """A MOOSE element: a path, a unique id in creation order, children."""
    _zfields = ()

    def __init__(self, path="/"):
        path = _norm(path)
        if path in _objects:
            old = _objects[path]
            if type(old) is not type(self) and type(self) is not Neutral:
                raise ValueError(f"{path} already exists as a {type(old).__name__}")
            self.__dict__ = old.__dict__
            return
        parent = path.rsplit("/", 1)[0] or "/"
        if path != "/" and parent not in _objects:
            raise ValueError(f"cannot create {path}: parent {parent} does not exist")
        d = self.__dict__
        d["path"] = path
        d["name"] = path.rsplit("/", 1)[1] if path != "/" else "/"
        d["id"] = _next_id[0]
        _next_id[0] += 1
        d["children"] = []
        d["_solver"] = None       # (HSolve, index) once zombified
        if path != "/":
            _objects[parent].children.append(self)
        _objects[path] = self
        self._init_fields()

    def _init_fields(self):
        pass

    @property
    def className(self):
        return type(self).__name__

    def isA(self, name):
        return any(c.__name__ == name for c in type(self).__mro__)

    def __repr__(self):
        return f"<moose.{self.className}: id={self.id}, path={self.path}>"

    # zombie field access: a taken-over object reads and writes the solver
    def __getattr__(self, name):
        d = self.__dict__
        if name in type(self)._zfields and d.get("_solver") is not None:
            hs, idx = d["_solver"]
            return hs._get(self, name, idx)
        raise AttributeError(f"{type(self).__name__} has no field {name}")

    def __setattr__(self, name, value):
        d = self.__dict__
        if name in type(self)._zfields and d.get("_solver") is not None:
            hs, idx = d["_solver"]
            hs._set(self, name, idx, value)
            return
        self._set_local(name, value)

    def _set_local(self, name, value):
        object.__setattr__(self, name, value)


class Compartment(Neutral):
    _zfields = ("Vm", "Im", "inject", "initVm", "Cm", "Em", "Rm", "Ra")

    def _init_fields(self):
        d = self.__dict__
        # biophysics/CompartmentBase.cpp defaults
        d.update(Vm=-0.06, Cm=1.0, Em=-0.06, initVm=-0.06, Rm=1.0, Ra=1.0, inject=0.0, Im=0.0,
                 length=0.0, diameter=0.0, x0=0.0, y0=0.0, z0=0.0, x=0.0, y=0.0, z=0.0)


class SymCompartment(Compartment):
    pass


SINGULARITY = 1.0e-6      # biophysics/HHGate.cpp:16


class HHGate(Neutral):
    """An HHGate (biophysics/HHGate.cpp): tables A (alpha) and B (alpha +
    beta) over [min, max], set directly (tableA/tableB) or generated from
    the GENESIS functional form (setupAlpha, setupTau, setupGate, the
    alpha/beta/tau/mInfinity fields).  The table generation and the
    range/size changes restate the reference's arithmetic step by step,
    including its quirks, so a gate built either way holds the reference's
    numbers."""

    def _init_fields(self):
        d = self.__dict__
        # HHGate::HHGate(Id, Id) (HHGate.cpp:214-225)
        d.update(_A=[0.0], _B=[0.0], _xmin=0.0, _xmax=1.0, _invDx=1.0, useInterpolation=False,
                 _direct=False, _alpha=[], _beta=[], _tau=[], _minf=[])

    # -- fields --------------------------------------------------------
    @property
    def tableA(self):
        return np.array(self._A, dtype=np.float64)

    @property
    def tableB(self):
        return np.array(self._B, dtype=np.float64)

    @property
    def min(self):
        return self._xmin

    @property
    def max(self):
        return self._xmax

    @property
    def divs(self):
        return len(self._A) - 1

    @property
    def alpha(self):
        return list(self._alpha)

    @property
    def beta(self):
        return list(self._beta)

    @property
    def tau(self):
        return list(self._tau)

    @property
    def mInfinity(self):
        return list(self._minf)

    @property
    def alphaParms(self):
        # HHGate::getAlphaParms (HHGate.cpp:498-508)
        return list(self._alpha) + list(self._beta) + [float(len(self._A)), self._xmin, self._xmax]

    def _set_local(self, name, value):
        d = self.__dict__
        if name == "tableA":            # HHGate::setTableA (:436-448)
            v = [float(x) for x in np.asarray(value, dtype=np.float64).reshape(-1)]
            if len(v) < 2:
                return
            d["_direct"] = True
            d["_A"] = v
            d["_invDx"] = (len(v) - 1) / (d["_xmax"] - d["_xmin"])
        elif name == "tableB":          # HHGate::setTableB (:456-468)
            v = [float(x) for x in np.asarray(value, dtype=np.float64).reshape(-1)]
            d["_direct"] = True
            if len(v) != len(d["_A"]):
                return
            d["_B"] = v
        elif name == "min":             # HHGate::setMin (:368-383)
            d["_xmin"] = float(value)
            xdivs = len(d["_A"]) - 1
            if d["_direct"] and xdivs > 0:
                d["_invDx"] = xdivs / (d["_xmax"] - d["_xmin"])
                self._tab_fill("_A", xdivs, d["_xmin"], d["_xmax"])
                self._tab_fill("_B", xdivs, d["_xmin"], d["_xmax"])
            else:
                self._update_tables()
        elif name == "max":             # HHGate::setMax (:390-406)
            d["_xmax"] = float(value)
            xdivs = len(d["_A"]) - 1
            if d["_direct"] and xdivs > 0:
                d["_invDx"] = xdivs / (d["_xmax"] - d["_xmin"])
                self._tab_fill("_A", xdivs, d["_xmin"], d["_xmax"])
                self._tab_fill("_B", xdivs, d["_xmin"], d["_xmax"])
            else:
                self._update_tables()
        elif name == "divs":            # HHGate::setDivs (:413-429)
            val = int(value)
            if d["_direct"]:
                d["_invDx"] = val / (d["_xmax"] - d["_xmin"])
                self._tab_fill("_A", val, d["_xmin"], d["_xmax"])
                self._tab_fill("_B", val, d["_xmin"], d["_xmax"])
            else:
                d["_A"] = (d["_A"] + [0.0] * (val + 1))[:val + 1]
                d["_B"] = (d["_B"] + [0.0] * (val + 1))[:val + 1]
                d["_invDx"] = val / (d["_xmax"] - d["_xmin"])
                self._update_tables()
        elif name in ("alpha", "beta", "tau", "mInfinity"):   # :288-361
            v = [float(x) for x in value]
            if len(v) != 5:
                return
            key = {"alpha": "_alpha", "beta": "_beta", "tau": "_tau", "mInfinity": "_minf"}[name]
            d[key] = v
            self._update_tables()
        else:
            object.__setattr__(self, name, value)

    # -- table generation (HHGate.cpp:481-795) ---------------------------
    def setupAlpha(self, parms):
        """HHGate::setupAlpha: 13 numbers, alpha A B C D F, beta A B C D F,
        divs, min, max; tables A = alpha, B = alpha + beta."""
        parms = [float(x) for x in parms]
        if len(parms) != 13:
            return
        self._setup_tables(parms, False)
        self.__dict__["_alpha"] = parms[0:5]
        self.__dict__["_beta"] = parms[5:10]

    def setupTau(self, parms):
        """HHGate::setupTau: the same 13 numbers read as tau and mInfinity;
        tables A = minf/tau, B = 1/tau."""
        parms = [float(x) for x in parms]
        if len(parms) != 13:
            return
        self._setup_tables(parms, True)

    def tweakAlpha(self):
        """HHGate::tweakAlpha is a no-op in this MOOSE version (:524-527)."""

    def tweakTau(self):
        """HHGate::tweakTau is a no-op in this MOOSE version (:529-532)."""

    def setupGate(self, parms):
        """HHGate::setupGate (:670-752): A B C D F size min max isbeta.  As
        in the reference, x stays at min + dx/2 for every entry, and an
        isbeta call converts the pair to (alpha, alpha + beta)."""
        if len(parms) != 9:
            return
        A, B, C, D, F = (float(x) for x in parms[:5])
        size, mn, mx, isBeta = int(parms[5]), float(parms[6]), float(parms[7]), bool(parms[8])
        d = self.__dict__
        key = "_B" if isBeta else "_A"
        ip = d[key]
        if size <= 0:
            size = len(ip) - 1
            if size <= 0:
                return
        else:
            ip = (ip + [0.0] * (size + 1))[:size + 1]
        dx = (mx - mn) / float(size)
        x = mn + dx / 2.0
        for i in range(size + 1):
            if abs(F) < SINGULARITY:
                ip[i] = 0.0
            else:
                temp2 = C + math.exp((x + D) / F)
                if abs(temp2) < SINGULARITY:
                    if i > 0:               # the reference reads ip[i - 1]
                        ip[i] = ip[i - 1]
                else:
                    ip[i] = (A + B * x) / temp2
        d[key] = ip
        if isBeta:
            if len(d["_A"]) != len(d["_B"]):
                if len(d["_A"]) > len(d["_B"]):
                    self._tab_fill("_B", len(d["_A"]) - 1, d["_xmin"], d["_xmax"])
                else:
                    self._tab_fill("_A", len(d["_B"]) - 1, d["_xmin"], d["_xmax"])
            d["_B"] = [a + b for a, b in zip(d["_A"], d["_B"])]      # tweakTables(0)

    def _update_tables(self):
        # HHGate::updateTables (:784-795): the current table size as divs
        d = self.__dict__
        if not d["_alpha"] or not d["_beta"]:
            return
        self._setup_tables(d["_alpha"] + d["_beta"] + [float(len(d["_A"])), d["_xmin"], d["_xmax"]], False)

    def _setup_tables(self, parms, doTau):
        """HHGate::setupTables (:539-641)."""
        if parms[10] < 1:
            return
        xdivs = int(parms[10])
        d = self.__dict__
        A = [0.0] * (xdivs + 1)
        B = [0.0] * (xdivs + 1)
        d["_xmin"], d["_xmax"] = parms[11], parms[12]
        d["_invDx"] = xdivs / (d["_xmax"] - d["_xmin"])
        dx = (d["_xmax"] - d["_xmin"]) / xdivs
        x = d["_xmin"]
        temp2 = 0.0
        for i in range(xdivs + 1):
            if abs(parms[4]) < SINGULARITY:
                temp = 0.0
                A[i] = temp
            else:
                temp2 = parms[2] + math.exp((x + parms[3]) / parms[4])
                if abs(temp2) < SINGULARITY:
                    temp2 = parms[2] + math.exp((x + dx / 10.0 + parms[3]) / parms[4])
                    temp = (parms[0] + parms[1] * (x + dx / 10)) / temp2
                    temp2 = parms[2] + math.exp((x - dx / 10.0 + parms[3]) / parms[4])
                    temp += (parms[0] + parms[1] * (x - dx / 10)) / temp2
                    temp /= 2.0
                    A[i] = temp
                else:
                    temp = (parms[0] + parms[1] * x) / temp2
                    A[i] = temp
            if abs(parms[9]) < SINGULARITY:
                B[i] = 0.0
            else:
                temp2 = parms[7] + math.exp((x + parms[8]) / parms[9])
                if abs(temp2) < SINGULARITY:
                    temp2 = parms[7] + math.exp((x + dx / 10.0 + parms[8]) / parms[9])
                    temp = (parms[5] + parms[6] * (x + dx / 10)) / temp2
                    temp2 = parms[7] + math.exp((x - dx / 10.0 + parms[8]) / parms[9])
                    temp += (parms[5] + parms[6] * (x - dx / 10)) / temp2
                    temp /= 2.0
                    B[i] = temp
                else:
                    B[i] = (parms[5] + parms[6] * x) / temp2
            if not doTau and abs(temp2) > SINGULARITY:
                B[i] += temp
            x += dx
        if doTau:
            pa = pb = 0.0
            for i in range(xdivs + 1):
                temp, temp2 = A[i], B[i]
                if abs(temp) < SINGULARITY:
                    A[i], B[i] = pa, pb
                else:
                    A[i] = temp2 / temp
                    B[i] = 1.0 / temp
                pa, pb = A[i], B[i]
        d["_A"], d["_B"] = A, B

    def _lookup_table(self, tab, v, interp):
        # HHGate::lookupTable (:231-246)
        d = self.__dict__
        if v <= d["_xmin"]:
            return tab[0]
        if v >= d["_xmax"]:
            return tab[-1]
        index = int((v - d["_xmin"]) * d["_invDx"])
        if interp:
            frac = (v - d["_xmin"] - index / d["_invDx"]) * d["_invDx"]
            # a point just below max can round onto the last entry; the
            # reference then reads one past the end with frac ~ 0
            nxt = tab[index + 1] if index + 1 < len(tab) else tab[index]
            return tab[index] * (1 - frac) + nxt * frac
        return tab[index]

    def _tab_fill(self, key, newXdivs, newXmin, newXmax):
        """HHGate::tabFill (:754-774): resizes, then refills each
        entry by an interpolated lookup in the table being refilled, with
        the gate's current range (as the reference does)."""
        if newXdivs < 3:
            return
        d = self.__dict__
        table = list(d[key])
        newDx = (newXmax - newXmin) / newXdivs
        table = (table + [0.0] * (newXdivs + 1))[:newXdivs + 1]
        for i in range(newXdivs + 1):
            table[i] = self._lookup_table(table, newXmin + i * newDx, True)
        d[key] = table


class HHChannel(Neutral):
    _zfields = ("Gbar", "Ek", "Gk", "Ik", "X", "Y", "Z", "modulation", "instant")

    def _init_fields(self):
        d = self.__dict__
        d.update(Gbar=0.0, Ek=0.0, Gk=0.0, Ik=0.0, Xpower=0.0, Ypower=0.0, Zpower=0.0,
                 X=0.0, Y=0.0, Z=0.0, instant=0, useConcentration=0, modulation=1.0)

    def _set_local(self, name, value):
        object.__setattr__(self, name, value)
        # setting a power creates the gate child (HHChannelBase::setXpower)
        for p, g in (("Xpower", "gateX"), ("Ypower", "gateY"), ("Zpower", "gateZ")):
            if name == p and value > 0 and _norm(self.path + "/" + g) not in _objects:
                HHGate(self.path + "/" + g)

    def gate(self, which):
        """The gate prototype: a copied channel uses its original's gates
        (HSolveUtils::gates with getOriginals, hsolve/HSolveUtils.cpp:75-120)."""
        src = self.__dict__.get("_original") or self
        return _objects.get(_norm(src.path + "/gate" + which))

    @property
    def gateX(self):
        return self.gate("X")

    @property
    def gateY(self):
        return self.gate("Y")

    @property
    def gateZ(self):
        return self.gate("Z")


class HHGate2D(Neutral):
    """An HHGate2D (biophysics/HHGate2D.cpp): two Interpol2D tables, A and B,
    indexed [x][y] over [xmin, xmax] x [ymin, ymax]; setting a table sets
    its divisions (Interpol2D::setTableVector)."""

    def _init_fields(self):
        d = self.__dict__
        d.update(xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, xdivs=0, ydivs=0,
                 tableA=np.zeros((1, 1)), tableB=np.zeros((1, 1)))

    def _set_local(self, name, value):
        if name in ("tableA", "tableB"):
            value = np.array(value, dtype=np.float64)
            if value.ndim != 2:
                raise ValueError("HHGate2D tables are 2-D ([x][y])")
            object.__setattr__(self, "xdivs", value.shape[0] - 1)
            object.__setattr__(self, "ydivs", value.shape[1] - 1)
        object.__setattr__(self, name, value)


# HHChannel2D::dependency (biophysics/HHChannel2D.cpp:273-303): what the x and
# y of a gate's tables read (0 Vm, 1 concen, 2 concen2, -1 nothing)
_INDEX_DEPS = {"VOLT_INDEX": (0, -1), "C1_INDEX": (1, -1), "C2_INDEX": (2, -1),
               "VOLT_C1_INDEX": (0, 1), "VOLT_C2_INDEX": (0, 2), "C1_C2_INDEX": (1, 2)}


class HHChannel2D(Neutral):
    """An HHChannel2D (biophysics/HHChannel2D.cpp): gates on 2-D tables of
    Vm and up to two concentrations (Xindex/Yindex/Zindex), integrated by
    exponential Euler; its current reaches the compartment through the
    "channel" message.  Taken over with its cell, it runs on the GPU beside
    the solver and its fields read and write there."""
    _zfields = ("Gbar", "Ek", "Gk", "X", "Y", "Z", "modulation")
    _zmap = {f: f + "2D" for f in _zfields}

    def _init_fields(self):
        d = self.__dict__
        d.update(Gbar=0.0, Ek=0.0, Gk=0.0, Xpower=0.0, Ypower=0.0, Zpower=0.0,
                 X=0.0, Y=0.0, Z=0.0, instant=0, modulation=1.0,
                 Xindex="", Yindex="", Zindex="")

    def _set_local(self, name, value):
        if name in ("Xindex", "Yindex", "Zindex") and value not in _INDEX_DEPS and value != "":
            raise ValueError(f"HHChannel2D: unknown index '{value}'")
        object.__setattr__(self, name, value)
        for p, g in (("Xpower", "gateX"), ("Ypower", "gateY"), ("Zpower", "gateZ")):
            if name == p and value > 0 and _norm(self.path + "/" + g) not in _objects:
                HHGate2D(self.path + "/" + g)

    def gate(self, which):
        src = self.__dict__.get("_original") or self
        return _objects.get(_norm(src.path + "/gate" + which))

    @property
    def gateX(self):
        return self.gate("X")

    @property
    def gateY(self):
        return self.gate("Y")

    @property
    def gateZ(self):
        return self.gate("Z")


class CaConc(Neutral):
    _zfields = ("Ca", "CaBasal", "ceiling", "floor")

    def _init_fields(self):
        d = self.__dict__
        d.update(Ca=0.0, CaBasal=0.0, tau=1.0, B=1.0, thick=0.0, ceiling=1e9, floor=0.0,
                 length=0.0, diameter=0.0)

    def shell_B(self):
        """B as the solver takes it over.  Zombifying a CaConc re-applies
        its geometry through CaConcBase::updateDimensions
        (biophysics/CaConcBase.cpp:259-269, :385-388): B = 1/(2 F vol) of
        the shell of thickness `thick`.  When length or diameter was never
        set the reference reads uninitialised members there; this layer then
        keeps the B the model set."""
        if not (self.length > 0 and self.diameter > 0):
            return self.B
        vol = math.pi * self.diameter * self.diameter * self.length * 0.25
        if 0 < self.thick < self.diameter / 2.0:
            core = self.diameter / 2.0 - self.thick
            vol -= math.pi * core * core * self.length
        return 1.0 / (CA_VALENCE * FARADAY * vol)


class SpikeGen(Neutral):
    def _init_fields(self):
        # biophysics/SpikeGen.cpp:127-134
        self.__dict__.update(threshold=0.0, refractT=0.0, edgeTriggered=1)


class _PulseLookup:
    """PulseGen's level/width/delay LookupValueFinfos: pg.delay[i] reads or
    writes entry i (device/PulseGen.cpp:300-355; an index at or past count
    is refused, as the reference warns and ignores)."""

    def __init__(self, pg, key):
        self._pg, self._key = pg, key

    def __getitem__(self, i):
        vals = self._pg.__dict__[self._key]
        return vals[i] if 0 <= i < len(vals) else 0.0

    def __setitem__(self, i, value):
        vals = self._pg.__dict__[self._key]
        if not 0 <= i < len(vals):
            raise IndexError(f"PulseGen: index {i} past count {len(vals)}; set 'count' first")
        vals[i] = float(value)

    def __len__(self):
        return len(self._pg.__dict__[self._key])


class PulseGen(Neutral):
    """A pulse train (device/PulseGen.cpp:359-424) sent through 'output' to a
    compartment's 'injectMsg': free running (trigMode 0), or started by its
    'input' (1, external trigger) or gated by it (2, external gate).  The
    input is the 'input' field, or another PulseGen's 'output' connected to
    it (which must run free).  Under an HSolve it is evaluated on the GPU
    before each tick; 'output' and 'input' read from there."""
    _zfields = ("output", "input")
    _zmap = {"output": "pulseOutput", "input": "pulseInput"}

    def _init_fields(self):
        # device/PulseGen.cpp:148-165: two pulses, all zero
        self.__dict__.update(_level=[0.0, 0.0], _width=[0.0, 0.0], _delay=[0.0, 0.0],
                             baseLevel=0.0, trigMode=0, output=0.0, input=0.0)

    level = property(lambda self: _PulseLookup(self, "_level"))
    width = property(lambda self: _PulseLookup(self, "_width"))
    delay = property(lambda self: _PulseLookup(self, "_delay"))

    @property
    def count(self):
        return len(self._level)

    def _set_local(self, name, value):
        d = self.__dict__
        if name == "count":
            n = int(value)
            if n <= 0:
                raise ValueError("PulseGen: invalid pulse count")
            for k in ("_level", "_width", "_delay"):       # PulseGen::setCount
                d[k] = (d[k] + [0.0] * n)[:n]
            return
        first = {"firstLevel": ("_level", 0), "firstWidth": ("_width", 0), "firstDelay": ("_delay", 0),
                 "secondLevel": ("_level", 1), "secondWidth": ("_width", 1),
                 "secondDelay": ("_delay", 1)}
        if name in first:
            key, i = first[name]
            if i < len(d[key]):             # the second* setters need count >= 2
                d[key][i] = float(value)
            return
        if name == "trigMode" and int(value) not in (0, 1, 2):
            # PulseGen::process's default branch: the mode is kept but no
            # pulse is produced; this layer refuses it instead
            raise ValueError("PulseGen: trigMode is 0, 1 or 2")
        object.__setattr__(self, name, value)

    def __getattr__(self, name):
        first = {"firstLevel": ("_level", 0), "firstWidth": ("_width", 0), "firstDelay": ("_delay", 0),
                 "secondLevel": ("_level", 1), "secondWidth": ("_width", 1),
                 "secondDelay": ("_delay", 1)}
        if name in first:
            key, i = first[name]
            vals = self.__dict__[key]
            return vals[i] if i < len(vals) else 0.0
        return Neutral.__getattr__(self, name)

    def _record(self, compartment, input_src=-1):
        n = self.count
        if n > _abi.MAX_PULSES:
            raise ValueError(f"{self.path}: at most {_abi.MAX_PULSES} pulses under HSolve")
        r = np.zeros(1, _abi.PULSEGEN_DT)[0]
        r["compartment"], r["n_pulses"], r["baseLevel"] = compartment, n, self.baseLevel
        r["delay"][:n], r["width"][:n], r["level"][:n] = self._delay, self._width, self._level
        r["trigMode"], r["input_src"] = int(self.trigMode), input_src
        return r


def _pulse_list(cidx):
    """The PulseGens a cell's solver runs, in creation order (the order clock
    tick 0 processes them): those feeding its compartments' injectMsg, with
    their compartment index, and the PulseGens whose 'output' drives one of
    those's 'input', with -1 (they feed no compartment); and each one's
    input source (an index in the list, or -1)."""
    feeds = {s.id: (s, cidx[d.id]) for s, sf, d, df in _msgs
             if sf == "output" and df == "injectMsg" and isinstance(s, PulseGen) and d.id in cidx}
    src = {}
    for pg, _ in list(feeds.values()):
        drv = [s for s, sf, d, df in _msgs if d is pg and df == "input" and sf == "output" and isinstance(s, PulseGen)]
        if len(drv) > 1:
            raise ValueError(f"{pg.path}: more than one PulseGen drives its input under HSolve")
        if drv:
            src[pg.id] = drv[0]
            if drv[0].id not in feeds:
                feeds[drv[0].id] = (drv[0], -1)
    order = sorted(feeds.values(), key=lambda x: x[0].id)
    pos = {pg.id: k for k, (pg, _) in enumerate(order)}
    return [(pg, ci, pos[src[pg.id].id] if pg.id in src else -1) for pg, ci in order]


class VClamp(Neutral):
    """A voltage clamp (biophysics/VClamp.cpp): senses its compartment's Vm
    (VmOut -> sensedIn), low-pass filters the command (commandIn or the
    'command' field) and sends a PID current to the compartment's injectMsg
    on clock tick 0.  Under an HSolve it runs on the GPU; 'command' writes
    and 'current' reads go there."""
    _zfields = ("command", "current")

    def _init_fields(self):
        # biophysics/VClamp.cpp:196-205
        self.__dict__.update(mode=0, tau=0.0, ti=0.0, td=-1.0, gain=0.0, command=0.0, current=0.0)

    def _set_local(self, name, value):
        if name in ("mode", "tau", "ti", "td", "gain") and self.__dict__.get("_solver") is not None:
            raise NotImplementedError(f"{self.path}: set VClamp {name} before HSolve.target")
        object.__setattr__(self, name, value)


class Leakage(Neutral):
    """A Leakage channel (biophysics/Leakage.cpp): not taken over by the
    solver.  Its Gk field (Gbar * modulation once reinit has run) is folded
    into the compartment's EmByRm when the solver reads the model, and on
    every tick it sends (Gbar * modulation, Ek) to the compartment, which
    under the solver becomes external current (run on the GPU)."""

    def _init_fields(self):
        self.__dict__.update(Gk=0.0, Ek=0.0, Gbar=0.0, modulation=1.0)


class Synapse:
    """One synapse of a SimpleSynHandler (synapse/Synapse.cpp): weight,
    delay; a spike arriving on addSpike at time t queues the event
    (t + delay, weight)."""

    def __init__(self, handler, index):
        self.handler, self.index = handler, index
        self.weight, self.delay = 1.0, 0.0
        self.id = _next_id[0]
        _next_id[0] += 1
        self.path = f"{handler.path}/synapse[{index}]"

    def isA(self, name):
        return name == "Synapse"

    def _target(self):
        """(solver, SynChan index, synapse index within the SynChan): the
        SynChan's synapses are those of its handlers in message order."""
        hs, k = self.handler._solver_of()
        sc = hs._objs["synchans"][k]
        off = 0
        for h in _sources(sc, "activation", "SimpleSynHandler"):
            if h is self.handler:
                return hs, k, off + self.index
            off += len(h.synapse)
        raise RuntimeError(f"{self.path}: its handler does not feed the SynChan")

    def addSpike(self, time):
        """Synapse::addSpike from the host (a spike at `time` seconds)."""
        hs, k, j = self._target()
        hs._engine.add_spike(hs._cell, k, j, float(time))


class SimpleSynHandler(Neutral):
    """A SimpleSynHandler (synapse/SimpleSynHandler.cpp): synapses as field
    elements (numSynapses, synapse[i]); on each tick it sends the activation
    sum(weight / dt) of the events due to the SynChan on its activationOut.
    Under a solver its queue and the SynChan run beside it (queue on the
    host, SynChan on the GPU)."""

    def _init_fields(self):
        self.__dict__.update(synapse=[])

    @property
    def numSynapses(self):
        return len(self.synapse)

    def _set_local(self, name, value):
        if name == "numSynapses":
            n = int(value)
            syn = self.__dict__["synapse"]
            while len(syn) < n:
                syn.append(Synapse(self, len(syn)))
            del syn[n:]
            return
        object.__setattr__(self, name, value)

    def _solver_of(self):
        for chan in _targets(self, "activationOut", "SynChan"):
            z = chan.__dict__.get("_solver")
            if z is not None:
                return z
        raise RuntimeError(f"{self.path}: its SynChan is not under an HSolve")


class TimeTable(Neutral):
    """A TimeTable (builtins/TimeTable.cpp): a list of spike times sent on
    eventOut, typically to a Synapse's addSpike.  On each of its clock ticks
    (8 by default, TimeTable.cpp:181-205) it sends the next listed time when
    the tick's time has reached it, at most one time per tick, stamped with
    the listed time (not the tick's).  Under a solver the sends are worked
    out on the host before each run and queued on the synapses' handlers."""

    def _init_fields(self):
        self.__dict__.update(vector=np.zeros(0), filename="", method=4, state=0.0, tick=8, _pos=0)

    def _set_local(self, name, value):
        d = self.__dict__
        if name == "vector":
            d["vector"] = np.asarray(value, dtype=np.float64).reshape(-1).copy()
        elif name == "filename":
            # TimeTable::setFilename: whitespace-separated numbers up to the
            # first token that is not one
            vals = []
            with open(value) as f:
                for tok in f.read().split():
                    try:
                        vals.append(float(tok))
                    except ValueError:
                        break
            d["filename"] = str(value)
            d["vector"] = np.array(vals, dtype=np.float64)
        elif name == "method":
            if int(value) != 4:
                raise ValueError("TimeTable: only method 4 (loading from file) is supported")
            d["method"] = 4
        elif name == "state":
            raise AttributeError("TimeTable.state is read-only")
        else:
            object.__setattr__(self, name, value)

    def clearVec(self):
        self.__dict__["vector"] = np.zeros(0)


def _timetable_sends(vec, pos, state, every, dt, s0, nsteps):
    """TimeTable::process (builtins/TimeTable.cpp:187-205) on the ticks of
    one run: the tick whose time is dt * u runs when u is a multiple of
    `every`, for s0 < u <= s0 + nsteps.  Returns the sends as (listed time,
    u), the position after the run and the state after its last tick."""
    sends = []
    u = (s0 // every + 1) * every
    last = (s0 + nsteps) // every * every
    if u > last:
        return sends, pos, state
    while pos < len(vec) and u <= last:
        tau = float(vec[pos])
        if tau > dt * u:
            # the first tick whose time reaches the listed time
            k = max(u, int(math.ceil(tau / dt / every)) * every)
            while k - every >= u and dt * (k - every) >= tau:
                k -= every
            while dt * k < tau:
                k += every
            u = k
            if u > last:
                break
        sends.append((tau, u))
        pos += 1
        u += every
    state = 1.0 if sends and sends[-1][1] == last else 0.0
    return sends, pos, state


def _feed_timetables(group, dt, nsteps):
    """TimeTable eventOut -> Synapse addSpike messages into the solvers of
    one engine: the events the TimeTables send during the next `nsteps`
    steps, queued on the handlers before the run.  A send on the tick of
    step u is seen by the handler from step u + 1 on.  Queued early, an event
    (t + delay) later than the tick that sends it pops on the same tick as in
    the reference; an earlier one is given the time of the middle of step
    u + 1, where the reference's handler first sees it."""
    ours = {id(hs) for hs in group}
    for tt in [o for o in _objects.values() if isinstance(o, TimeTable)]:
        dests = []
        for syn in _targets(tt, "eventOut", "Synapse"):
            try:
                hs, k, j = syn._target()
            except RuntimeError:
                continue
            if id(hs) in ours:
                dests.append((hs, k, j, syn))
        if not dests:
            continue
        tdt = _clock_dt.get(tt.tick, 0.0)
        every = int(round(tdt / dt)) if tdt > 0 else 0
        if every < 1 or abs(every * dt - tdt) > 1e-9 * tdt:
            raise ValueError(f"{tt.path}: its tick's dt ({tdt}) is not a multiple of the solver dt ({dt})")
        s0 = dests[0][0].__dict__.get("_steps", 0)
        d = tt.__dict__
        sends, d["_pos"], d["state"] = _timetable_sends(d["vector"], d["_pos"], d["state"], every, dt, s0, nsteps)
        for tau, u in sends:
            for hs, k, j, syn in dests:
                when = tau if tau + syn.delay > dt * u else dt * (u + 0.5) - syn.delay
                hs._engine.add_spike(hs._cell, k, j, when)


class SynChan(Neutral):
    """A SynChan (biophysics/SynChan.cpp): dual-exponential conductance
    (tau1, tau2; normalised so the peak of a unit activation is Gbar),
    driven by its SimpleSynHandler's activation.  Taken over with its cell,
    it runs on the GPU beside the solver and its fields read and write
    there."""
    _zfields = ("Gbar", "Ek", "Gk", "tau1", "tau2", "modulation", "X", "Y")
    _zmap = {"Gbar": "GbarSyn", "Ek": "EkSyn", "Gk": "GkSyn", "modulation": "modulationSyn",
             "X": "XSyn", "Y": "YSyn"}

    def _init_fields(self):
        # biophysics/SynChan.cpp:95-108, ChanCommon.cpp:13-20
        self.__dict__.update(Gbar=0.0, Ek=0.0, Gk=0.0, tau1=1.0e-3, tau2=1.0e-3, modulation=1.0,
                             X=0.0, Y=0.0, normalizeWeights=0)


class Table(Neutral):
    """Records a compartment's Vm (requestOut -> getVm) on its clock tick
    (8 unless moved with useClock), after the solver's tick of the same step
    (Table::process, builtins/Table.cpp:254-285), plus one sample at reinit;
    or a SpikeGen's spike times (spikeOut -> spike)."""

    def _init_fields(self):
        self.__dict__.update(_data=[], tick=8)

    @property
    def vector(self):
        return np.array(self._data, dtype=np.float64)

    def clearVec(self):
        self._data.clear()


def element(path):
    if isinstance(path, (Neutral, Synapse)):
        return path
    p = _norm(path)
    if p not in _objects:
        raise ValueError(f"no element at {p}")
    return _objects[p]


def exists(path):
    return _norm(path) in _objects


def connect(src, srcField, dest, destField, msgType="Single"):
    src, dest = element(src), element(dest)
    _msgs.append((src, srcField, dest, destField))
    return len(_msgs) - 1


def copy(orig, parent, name=None):
    """moose.copy of one element (fields only; a copied HHChannel keeps
    reading its original's gates, as MOOSE's copies share the prototype
    gates)."""
    orig, parent = element(orig), element(parent)
    name = name or orig.name
    new = type(orig)(parent.path.rstrip("/") + "/" + name)
    for k, v in orig.__dict__.items():
        if k in ("path", "name", "id", "children", "_solver"):
            continue
        new.__dict__[k] = v.copy() if isinstance(v, (list, np.ndarray)) else v
    if isinstance(orig, (HHChannel, HHChannel2D)):
        new.__dict__["_original"] = orig.__dict__.get("_original") or orig
    return new


def delete(path):
    obj = element(path)
    for c in list(obj.children):
        delete(c)
    parent = _objects[obj.path.rsplit("/", 1)[0] or "/"]
    parent.children.remove(obj)
    del _objects[obj.path]
    _msgs[:] = [m for m in _msgs if m[0] is not obj and m[2] is not obj]


def reset():
    """Forgets every element and solver (a fresh MOOSE session)."""
    for eng in _engines.values():
        eng.close()
    _engines.clear()
    _objects.clear()
    _msgs.clear()
    _solvers.clear()
    _clock_dt.clear()
    _clock_dt.update(_CLOCK_DT)
    _next_id[0] = 1
    Neutral("/")


def set_precision(precision):
    """'fp64' (the reference's arithmetic) or 'mixed' (FP32 channels) for
    solvers set up after this call."""
    if precision not in ("fp64", "mixed"):
        raise ValueError(precision)
    _precision[0] = precision


def _targets(obj, field, cls=None):
    """Elements at the far end of messages leaving `obj` from `field`."""
    return [d for s, sf, d, df in _msgs if s is obj and sf == field and (cls is None or d.isA(cls))]


def _sources(obj, field, cls=None):
    """Elements sending messages into `obj`'s `field`."""
    return [s for s, sf, d, df in _msgs if d is obj and df == field and (cls is None or s.isA(cls))]


def _adjacent(comp):
    """HSolveUtils::adjacent (hsolve/HSolveUtils.cpp:36-45): axial targets,
    then raxial targets (the parent, reached through the same message)."""
    out = [d for s, sf, d, df in _msgs if s is comp and sf == "axial" and d.isA("Compartment")]
    out += [s for s, sf, d, df in _msgs if d is comp and df == "raxial" and sf == "axial" and s.isA("Compartment")]
    return out


def _channels(comp, cls):
    """Channels on a compartment's "channel" messages, in message order."""
    out = []
    for s, sf, d, df in _msgs:
        if d is comp and df == "channel" and sf == "channel" and s.isA(cls):
            out.append(s)
        elif s is comp and sf == "channel" and df == "channel" and d.isA(cls):
            out.append(d)
    return out


def _deep_search_compartment(base):
    """HSolve::deepSearchForCompartment (hsolve/HSolve.cpp:346-393): depth
    first, the last child of each level first."""
    cstack = [[base]]
    while cstack:
        if not cstack[-1]:
            cstack.pop()
            if cstack:
                cstack[-1].pop()
            continue
        cur = cstack[-1][-1]
        if cur.isA("Compartment"):
            return cur
        cstack.append(list(cur.children))
    return None


class HSolve(Neutral):
    """moose.HSolve: set dt, then target (a path at or above the cell's
    compartments); the model is read, taken over and advanced on the GPU."""

    def _init_fields(self):
        self.__dict__.update(dt=0.0, caAdvance=1, vDiv=3000, vMin=-0.100, vMax=0.050,
                             caDiv=3000, caMin=0.0, caMax=1000.0, _target="", seed=None,
                             _cell=None, _engine=None, _objs=None, _tables=[], _spk=[])

    @property
    def target(self):
        return self._target

    @target.setter
    def target(self, path):
        # HSolve::setPath (hsolve/HSolve.cpp:315-334): without dt, or with no
        # compartment at or below the path, it prints and returns, leaving
        # the solver as it was (on no compartment, the seed cleared)
        if self.dt == 0.0:
            print("Error: HSolve::setPath(): Must set 'dt' first.", file=sys.stderr)
            return
        try:
            seed = _deep_search_compartment(element(path))
        except ValueError:   # no element there: no compartment found
            seed = None
        if seed is None:
            object.__setattr__(self, "seed", None)
            print(f"Warning: HSolve::setPath(): No compartments found at or below '{path}'.", file=sys.stderr)
            return
        object.__setattr__(self, "_target", path)
        object.__setattr__(self, "seed", seed)
        self._setup(seed)

    def _set_local(self, name, value):
        if name == "dt" and value < 0:
            # HSolve::setDt (hsolve/HSolve.cpp:409-418): prints, keeps the old dt
            print("Error: HSolve: 'dt' must be positive.", file=sys.stderr)
            return
        if name == "target":
            HSolve.target.fset(self, value)
            return
        object.__setattr__(self, name, value)

    # -- setup: the model in MOOSE object order --------------------------
    def build_spec(self, seed=None):
        """The cell description of the model reachable from `seed`, in
        creation/message order (a CellSpec), plus the objects behind each
        index."""
        seed = seed or self.seed
        comps, seen, todo = [], {seed.id}, [seed]
        while todo:                           # everything connected to the seed
            c = todo.pop()
            comps.append(c)
            for a in _adjacent(c):
                if a.id not in seen:
                    seen.add(a.id)
                    todo.append(a)
        comps.sort(key=lambda o: o.id)        # creation order
        cidx = {c.id: i for i, c in enumerate(comps)}
        ct = np.zeros(len(comps), _abi.COMPARTMENT_DT)
        for i, c in enumerate(comps):
            d = c.__dict__
            ct[i] = (d["Ra"], d["Rm"], d["Cm"], d["Em"], d["initVm"], d["Vm"], d["inject"])
        axial = [(cidx[s.id], cidx[d.id]) for s, sf, d, df in _msgs
                 if sf == "axial" and df == "raxial" and s.id in cidx and d.id in cidx]
        chans, gates, pools, leaks, spikes = [], {}, [], [], []
        pidx = {}
        for c in comps:
            for ch in _channels(c, "HHChannel"):
                chans.append((cidx[c.id], ch))
            for lk in _channels(c, "Leakage"):
                d = lk.__dict__
                leaks.append((cidx[c.id], d["Gk"], d["Ek"], d["Gbar"], d["modulation"]))
            for sg in _targets(c, "VmOut", "SpikeGen"):
                spikes.append((cidx[c.id], sg))
        crec = np.zeros(len(chans), _abi.CHANNEL_DT)
        for k, (ci, ch) in enumerate(chans):
            d = ch.__dict__
            g = []
            for w in "XYZ":
                gt = ch.gate(w) if d[w + "power"] > 0 else None
                if gt is not None and gt.id not in gates:
                    gates[gt.id] = gt
                g.append(gt.id if gt is not None else -1)
            tgt = _targets(ch, "IkOut", "CaConc")
            dep = _targets(ch, "concen", "CaConc") + _sources(ch, "concen", "CaConc")
            for p in tgt + dep:
                if p.id not in pidx:
                    pidx[p.id] = len(pools)
                    pools.append(p)
            crec[k] = (ci, d["instant"], d["useConcentration"], g,
                       pidx[tgt[0].id] if tgt else -1, pidx[dep[-1].id] if dep else -1,
                       d["Gbar"], d["Ek"], d["Xpower"], d["Ypower"], d["Zpower"],
                       d["X"], d["Y"], d["Z"], d["modulation"])
        gid = sorted(gates)
        gpos = {g: i for i, g in enumerate(gid)}
        for k in range(len(crec)):
            crec[k]["gate"] = [gpos[g] if g >= 0 else -1 for g in crec[k]["gate"]]
        glist = []
        for g in gid:
            gt = gates[g]
            glist.append(Gate(g, gt.min, gt.max, int(gt.divs), np.asarray(gt.tableA, np.float64),
                              np.asarray(gt.tableB, np.float64)))
        prec = np.zeros(len(pools), _abi.CACONC_DT)
        for q, p in enumerate(pools):
            d = p.__dict__
            prec[q] = (d["Ca"], d["CaBasal"], d["tau"], p.shell_B(), d["ceiling"], d["floor"])
        # HHChannel2D objects on the compartments, in message order; each
        # gate's A and B tables become Interpol2D entries (once per gate)
        ch2, tabs, tpos = [], [], {}
        for c in comps:
            for ch in _channels(c, "HHChannel2D"):
                ch2.append((cidx[c.id], ch))
        c2rec = np.zeros(len(ch2), _abi.CHAN2D_DT)
        for k, (ci, ch) in enumerate(ch2):
            d = ch.__dict__
            ta, tb, dep = [-1, -1, -1], [-1, -1, -1], [-1] * 6
            for a, w in enumerate("XYZ"):
                if not d[w + "power"] > 0:
                    continue
                gt = ch.gate(w)
                if gt.id not in tpos:
                    tpos[gt.id] = len(tabs)
                    g = gt.__dict__
                    for t in (g["tableA"], g["tableB"]):
                        tabs.append(Table2D(int(g["xdivs"]), int(g["ydivs"]), g["xmin"], g["xmax"],
                                            g["ymin"], g["ymax"], np.asarray(t, np.float64)))
                ta[a], tb[a] = 2 * tpos[gt.id], 2 * tpos[gt.id] + 1
                dep[2 * a], dep[2 * a + 1] = _INDEX_DEPS.get(d[w + "index"], (-1, -1))
            concs = []
            for field in ("concen", "concen2"):
                src = _sources(ch, field, "CaConc")
                if src and src[-1].id not in pidx:
                    raise ValueError(f"{ch.path}: {field} from a CaConc the solver does not run")
                concs.append(pidx[src[-1].id] if src else -1)
            c2rec[k] = (ci, d["instant"], ta, tb, dep, concs[0], concs[1], d["Gbar"], d["Ek"],
                        d["Xpower"], d["Ypower"], d["Zpower"], d["X"], d["Y"], d["Z"], d["modulation"])
        srec = np.zeros(len(spikes), _abi.SPIKEGEN_DT)
        for k, (ci, sg) in enumerate(spikes):
            d = sg.__dict__
            srec[k] = (ci, int(d["edgeTriggered"]), d["threshold"], d["refractT"])
        lrec = np.zeros(len(leaks), _abi.LEAK_DT)
        for k, rec in enumerate(leaks):
            lrec[k] = rec
        # SynChans on the compartments, in message order, each with the
        # synapses of the handlers feeding it (handlers in message order: one
        # queue holds their events, so several handlers' activations reach
        # the SynChan as one sum per tick)
        syns, synrec, synapses = [], [], []
        for c in comps:
            for sc in _channels(c, "SynChan"):
                syn = [y for h in _sources(sc, "activation", "SimpleSynHandler") for y in h.synapse]
                d = sc.__dict__
                synrec.append((cidx[c.id], len(syn), d["Gbar"], d["Ek"], d["tau1"], d["tau2"], d["modulation"]))
                synapses += [(y.weight, y.delay) for y in syn]
                syns.append(sc)
        screc = np.array(synrec, _abi.SYNCHAN_DT) if synrec else np.zeros(0, _abi.SYNCHAN_DT)
        sprec = np.array(synapses, _abi.SYNAPSE_DT) if synapses else np.zeros(0, _abi.SYNAPSE_DT)
        # VClamps sending to the compartments' injectMsg, in creation order;
        # each senses its own compartment
        vcs = sorted(((src, cidx[d.id]) for src, sf, d, df in _msgs
                      if sf == "currentOut" and df == "injectMsg" and isinstance(src, VClamp)
                      and d.id in cidx), key=lambda x: x[0].id)
        vrec = np.zeros(len(vcs), _abi.VCLAMP_DT)
        for k, (vc, ci) in enumerate(vcs):
            sensed = _sources(vc, "sensedIn")
            if [x.id for x in sensed] != [comps[ci].id]:
                raise ValueError(f"{vc.path}: a VClamp under HSolve senses the compartment it drives")
            d = vc.__dict__
            vrec[k] = (ci, int(d["mode"]), d["tau"], d["ti"], d["td"], d["gain"])
        # PulseGens feeding the compartments' injectMsg and those driving
        # their inputs, in creation order (the order clock tick 0 runs them)
        pgs = _pulse_list(cidx)
        grec = np.zeros(len(pgs), _abi.PULSEGEN_DT)
        for k, (pg, ci, src) in enumerate(pgs):
            grec[k] = pg._record(ci, src)
        spec = CellSpec(ct, np.array(axial, dtype=np.int32).reshape(-1, 2), seed=cidx[seed.id],
                        channels=crec, gates=glist, pools=prec, spikegens=srec, leaks=lrec,
                        tables2d=tabs, chans2d=c2rec, pulsegens=grec, synchans=screc, synapses=sprec,
                        vclamps=vrec)
        objs = {"compartments": comps, "channels": [ch for _, ch in chans], "pools": pools,
                "spikegens": [sg for _, sg in spikes], "chans2d": [ch for _, ch in ch2],
                "pulsegens": [pg for pg, _, _ in pgs], "pulse_links": [(ci, src) for _, ci, src in pgs],
                "synchans": syns, "vclamps": [vc for vc, _ in vcs],
                "leaks": [lk for c in comps for lk in _channels(c, "Leakage")]}
        return spec, objs

    def _setup(self, seed):
        from .engine import Engine
        spec, objs = self.build_spec(seed)
        key = (_precision[0], float(self.dt))
        if key not in _engines:
            _engines[key] = Engine(_precision[0])
        eng = _engines[key]
        cell = eng.add_cell(spec, float(self.dt), int(self.caAdvance))
        object.__setattr__(self, "_engine", eng)
        object.__setattr__(self, "_cell", cell)
        # setup overwrites the lookup-grid fields with the grids it derived
        # from the gates (hsolve/HSolveActiveSetup.cpp:366-409)
        info = eng.cell_info(cell)
        for f in ("vMin", "vMax", "vDiv", "caMin", "caMax", "caDiv"):
            object.__setattr__(self, f, info[f])
        object.__setattr__(self, "_objs", objs)
        object.__setattr__(self, "_spec", spec)
        # zombify (hsolve/HSolve.cpp:235-262): fields now live in the solver
        for kind in ("compartments", "channels", "pools", "chans2d", "synchans", "vclamps", "pulsegens"):
            for i, o in enumerate(objs[kind]):
                o.__dict__["_solver"] = (self, i)
                for f in type(o)._zfields:       # reads now reach the solver
                    o.__dict__.pop(f, None)
        _solvers.append(self)

    # -- zombie field access (hsolve/HSolveInterface.cpp) ----------------
    def _get(self, obj, field, idx):
        field = getattr(type(obj), "_zmap", {}).get(field, field)
        return float(self._engine.get(self._cell, field, [idx])[0])

    def _set(self, obj, field, idx, value):
        field = getattr(type(obj), "_zmap", {}).get(field, field)
        self._engine.set(self._cell, field, [idx], [float(value)])

    def _push_leaks(self):
        """Leakage fields written since setup reach the device's per-tick
        send (the object itself keeps running beside the solver)."""
        sent = self._spec.leaks
        for k, lk in enumerate(self._objs["leaks"]):
            d = lk.__dict__
            for f, name in (("Gbar", "GbarLeak"), ("Ek", "EkLeak"), ("modulation", "modulationLeak")):
                if d[f] != sent[k][f]:
                    self._engine.set(self._cell, name, [k], [d[f]])
                    sent[k][f] = d[f]

    def _push_synapses(self):
        """Synapse weights and delays written since setup."""
        sent = self._spec.synapses
        j = 0
        for sc in self._objs["synchans"]:
            for y in [y for h in _sources(sc, "activation", "SimpleSynHandler") for y in h.synapse]:
                for f, name in (("weight", "weight"), ("delay", "delay")):
                    v = getattr(y, f)
                    if v != sent[j][f]:
                        self._engine.set(self._cell, name, [j], [v])
                        sent[j][f] = v
                j += 1

    def _push_pulsegens(self):
        """PulseGen fields written since setup reach the device before the
        next step, as the reference's PulseGen reads them every tick."""
        sent = self._spec.pulsegens
        cidx = {c.id: i for i, c in enumerate(self._objs["compartments"])}
        now = [(pg.id, ci, src) for pg, ci, src in _pulse_list(cidx)]
        if now != [(pg.id, ci, src) for pg, (ci, src) in zip(self._objs["pulsegens"], self._objs["pulse_links"])]:
            raise ValueError("PulseGen messages to the solver's compartments or PulseGen inputs changed "
                             "after HSolve.target was set; connect them first")
        for k, (pg, (ci, src)) in enumerate(zip(self._objs["pulsegens"], self._objs["pulse_links"])):
            r = pg._record(ci, src)
            if r.tobytes() != sent[k].tobytes():
                self._engine.set_pulsegen(self._cell, k, r)
                sent[k] = r

    # -- recording -------------------------------------------------------
    # Table sources the device samples: getter -> (object kind, engine field)
    _DEVICE_PROBES = {"getVm": ("compartments", "Vm"), "getCa": ("pools", "Ca"),
                      "getCurrent": ("vclamps", "current"), "getGk": ("synchans", "GkSyn")}

    def _recorders(self):
        """Tables the device samples: (table, (engine field, index)); and
        Tables of spike times: (table, spikegen index)."""
        pos = {kind: {o.id: i for i, o in enumerate(self._objs[kind])}
               for kind, _ in self._DEVICE_PROBES.values()}
        vm = []
        for s, sf, d, df in _msgs:
            if sf == "requestOut" and isinstance(s, Table) and df in self._DEVICE_PROBES:
                kind, field = self._DEVICE_PROBES[df]
                if isinstance(d, Neutral) and d.id in pos[kind]:
                    vm.append((s, (field, pos[kind][d.id])))
        spk = []
        sgs = self._objs["spikegens"]
        spos = {g.id: i for i, g in enumerate(sgs)}
        for s, sf, d, df in _msgs:
            if sf == "spikeOut" and isinstance(d, Table) and s.id in spos:
                spk.append((d, spos[s.id]))
        return vm, spk

    def _field_recorders(self, dt):
        """Tables reading another field of an object this solver runs (a
        VClamp's current, a channel's Gk, ...): (table, object, field,
        sampling interval in steps).  The host reads these after the ticks
        they sample."""
        out = []
        for s, sf, d, df in _msgs:
            if sf != "requestOut" or not isinstance(s, Table) or not df.startswith("get"):
                continue
            z = d.__dict__.get("_solver") if isinstance(d, Neutral) else None
            if z is None or z[0] is not self:
                continue
            if df in self._DEVICE_PROBES and d in self._objs[self._DEVICE_PROBES[df][0]]:
                continue
            for f in (df[3:], df[3].lower() + df[4:]):
                if f in type(d)._zfields:
                    out.append((s, d, f, _table_every(s, dt)))
                    break
            else:
                raise ValueError(f"{s.path}: cannot record {df} of {d.path} under HSolve")
        return out


# ---------------------------------------------------------------------------
# SWC morphologies (biophysics/ReadSwc.cpp, biophysics/SwcSegment.cpp,
# shell/LoadModels.cpp:156-166)
# ---------------------------------------------------------------------------
_SWC_TYPE_NAME = ["undef", "soma", "axon", "dend", "apical", "dend_f", "dend_e", "custom", "bad",
                  "undef", "axon_f", "axon_e", "apical_f", "apical_e"]
_SWC_MIN_RADIUS = 0.04        # microns (ReadSwc.cpp:23)
_SWC_ZERO_LENGTH = 1e-2       # microns (ReadSwc::cleanZeroLength)


def _read_swc(fname):
    """ReadSwc::ReadSwc: the segments (index, type, x, y, z, radius, parent,
    kids) after validation, kid assignment, type refinement and the removal
    of zero-length segments; then the branches (runs between forks/ends) in
    the reference's order.  Returns (segs, branches) or raises ValueError."""
    segs = []
    with open(fname) as f:
        for line in f:
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            a = t.split()
            if len(a) != 7:
                continue                     # a bad segment is skipped
            pa = int(a[6])
            segs.append({"i": int(a[0]), "type": int(a[1]),
                         "v": (float(a[2]), float(a[3]), float(a[4])),
                         "r": float(a[5]), "pa": pa if pa > 0 else None, "kids": [], "bad": False})
    # validate (ReadSwc::validate)
    starts = sum(1 for s in segs if s["pa"] is None)
    orphans = sum(1 for k, s in enumerate(segs) if s["pa"] is not None and s["pa"] > k)
    badr = sum(1 for s in segs if s["r"] < _SWC_MIN_RADIUS)
    if not (starts == 1 and orphans == 0 and badr == 0):
        raise ValueError(f"ReadSwc: {fname} failed validation (starts {starts}, orphans {orphans}, "
                         f"radius < {_SWC_MIN_RADIUS}: {badr})")
    # assignKids + figureOutType
    for k, s in enumerate(segs):
        if s["pa"] is not None:
            segs[s["pa"] - 1]["kids"].append(k + 1)
    for s in segs:
        n = len(s["kids"])
        for base, fork, end in ((3, 5, 6), (4, 12, 13), (2, 10, 11)):
            if s["type"] == base:
                s["type"] = fork if n > 1 else end if n == 0 else base
                break

    def dist(a, b):
        return math.sqrt(sum((x - y) ** 2 for x, y in zip(a["v"], b["v"])))

    # cleanZeroLength
    for k in range(1, len(segs)):
        s = segs[k]
        pa = segs[s["pa"] - 1]
        if dist(s, pa) < _SWC_ZERO_LENGTH:
            temp = [c for c in pa["kids"] if c != s["i"]]
            for c in s["kids"]:
                segs[c - 1]["pa"] = pa["i"]
                temp.append(c)
            pa["kids"] = temp
            s["bad"] = True
    # SwcSegment::OK (SwcSegment.h:34-37): not bad, not undef, not custom
    for s in segs:
        s["ok"] = not s["bad"] and s["type"] not in (0, 7)
    # parseBranches: one branch per fork or end, back to the previous fork
    branches = []
    for s in segs:
        if s["ok"] and len(s["kids"]) != 1:
            cable = [s["i"]]
            if s["pa"] is not None:
                prev = s
                while True:
                    pa = segs[prev["pa"] - 1]
                    cable.append(pa["i"])
                    prev = pa
                    if not (prev["pa"] is not None and len(prev["kids"]) == 1):
                        break
                cable.pop()
            branches.append(cable[::-1])
    return segs, branches


def loadModel(filename, modelpath, solverClass="ee"):
    """moose.loadModel for cell morphologies (Shell::doLoadModel,
    shell/LoadModels.cpp:150-166): a Neuron at `modelpath`, filled from a
    GENESIS .p file (ReadCell) or an SWC file (ReadSwc: one Compartment per
    point, lambda 0.5 mm, RM 1, RA 1, CM 0.01), in the reference's creation
    and message order."""
    low = filename.lower()
    if low.endswith(".p"):
        return _ReadCell().read(filename, modelpath)
    if not low.endswith(".swc"):
        raise ValueError("loadModel: this layer reads .p and .swc cell files")
    segs, branches = _read_swc(filename)
    parent = Neutral(modelpath)
    RM, RA, CM = 1.0, 1.0, 0.01
    compts = {}
    for bi, cable in enumerate(branches):
        for j, si in enumerate(cable):
            seg = segs[si - 1]
            if seg["pa"] is not None and seg["pa"] not in compts:
                # the reference asserts here: the parent point lies on no
                # branch (a soma with exactly one child ends no branch)
                raise ValueError(f"ReadSwc::build: point {si} has no parent compartment "
                                 f"(point {seg['pa']} ends no branch)")
            pa = segs[seg["pa"] - 1] if seg["pa"] is not None else seg
            compts[si] = _make_swc_compartment(parent, seg, pa, RM, RA, CM, bi, j)
            if seg["pa"] is not None:
                connect(compts[seg["pa"]], "axial", compts[si], "raxial")
    return parent


def _make_swc_compartment(parent, seg, pa, RM, RA, CM, i, j):
    # makeCompt (ReadSwc.cpp:234-280)
    length = seg["r"] * 2.0
    name = "soma"
    if seg["pa"] is not None:
        length = math.sqrt(sum((x - y) ** 2 for x, y in zip(seg["v"], pa["v"])))
        name = f"{_SWC_TYPE_NAME[seg['type']]}_{i}_{j}"
        x0 = pa["v"]
    else:
        x0 = (seg["v"][0] - length, seg["v"][1], seg["v"][2])
    c = Compartment(parent.path.rstrip("/") + "/" + name)
    xa = seg["r"] * seg["r"] * math.pi * 1e-12
    length *= 1e-6
    dia = seg["r"] * 2.0e-6
    d = c.__dict__
    d["Rm"] = RM / (length * dia * math.pi)
    d["Ra"] = RA * length / xa
    d["Cm"] = CM * (length * dia * math.pi)
    d.update(diameter=dia, length=length, x0=x0[0] * 1e-6, y0=x0[1] * 1e-6, z0=x0[2] * 1e-6,
             x=seg["v"][0] * 1e-6, y=seg["v"][1] * 1e-6, z=seg["v"][2] * 1e-6)
    return c


# ---------------------------------------------------------------------------
# GENESIS .p cell files (biophysics/ReadCell.cpp)
# ---------------------------------------------------------------------------
def _calc_surf(length, dia):
    # calcSurf (ReadCell.cpp:617-626): a sphere when the length is zero
    return dia * dia * math.pi if length == 0.0 else length * dia * math.pi


class _ReadCell:
    """ReadCell (biophysics/ReadCell.cpp): the GENESIS cell reader.  Supported:
    comments (// and /* */), *cartesian/*polar, *relative/*absolute,
    *asymmetric, *set_global/*set_compt_param (RM RA CM EREST_ACT ELEAK),
    *start_cell (no graft), *double_endpoint[_off], *makeproto; compartment
    lines `name parent x y z d [chan dens]...` with the parent '.', 'none'
    or 'nil', the RA/RM/CM/Ra/Rm/Cm overrides, and channels, SpikeGens and
    CaConcs copied from /library with their densities scaled as the
    reference scales them.  Not supported (error): *symmetric
    (SymCompartment messages), grafted cells, *compt prototypes; the
    GENESIS 'addmsg' fields of prototypes are not applied."""

    def __init__(self):
        # ReadCell::ReadCell (ReadCell.cpp:22-105)
        self.RM, self.CM, self.RA = 10.0, 0.01, 1.0
        self.EREST_ACT, self.ELEAK = -0.065, -0.065
        self.erestFlag = self.eleakFlag = False
        self.polar = self.relative = self.doubleEnd = self.symmetric = False
        self.last = None
        lib = _objects.get("/library")
        self.protos = {c.name: c for c in lib.children} if lib is not None else {}

    def read(self, fname, path):
        # ReadCell::read / innerRead (ReadCell.cpp:107-219)
        self.cell = Neutral(path)
        mode = "data"
        with open(fname) as f:
            for raw in f:
                line = raw.strip()
                if not line:
                    continue
                if line.startswith("//"):
                    continue
                if "//" in line:
                    line = line[:line.index("//")]
                if line.startswith("/*"):
                    mode = "comment"
                elif "*/" in line:
                    mode = "data"
                    continue
                elif line[0] == "*":
                    mode = "script"
                if mode == "comment":
                    pos = line.find("*/")
                    if pos >= 0:
                        mode = "data"
                        if len(line) > pos + 2:
                            line = line[pos + 2:]
                if mode == "data":
                    self._data(line.split())
                elif mode == "script":
                    self._script(line.split())
                    mode = "data"
        return self.cell

    def _script(self, argv):
        # ReadCell::readScript (ReadCell.cpp:220-306)
        cmd = argv[0]
        if cmd == "*cartesian":
            self.polar = False
        elif cmd == "*polar":
            self.polar = True
        elif cmd == "*relative":
            self.relative = True
        elif cmd == "*absolute":
            self.relative = False
        elif cmd == "*symmetric":
            raise ValueError("ReadCell: *symmetric (SymCompartment) cells are not supported by this layer")
        elif cmd == "*asymmetric":
            self.symmetric = False
        elif cmd in ("*set_global", "*set_compt_param"):
            if len(argv) != 3:
                return
            v = float(argv[2])
            if argv[1] == "RM":
                self.RM = v
            if argv[1] == "RA":
                self.RA = v
            if argv[1] == "CM":
                self.CM = v
            if argv[1] == "EREST_ACT":
                self.EREST_ACT, self.erestFlag = v, True
            if argv[1] == "ELEAK":
                self.ELEAK, self.eleakFlag = v, True
        elif cmd == "*start_cell":
            if len(argv) != 1:
                raise ValueError("ReadCell: grafted cells (*start_cell <path>) are not supported by this layer")
        elif cmd == "*compt":
            raise ValueError("ReadCell: *compt prototypes are not supported by this layer")
        elif cmd == "*double_endpoint":
            self.doubleEnd = True
        elif cmd == "*double_endpoint_off":
            self.doubleEnd = False

    def _coords(self, a, b, c):
        x, y, z = 1.0e-6 * float(a), float(b), float(c)
        if self.polar:
            r, th, ph = x, y * math.pi / 180.0, z * math.pi / 180.0
            return r * math.sin(ph) * math.cos(th), r * math.sin(ph) * math.sin(th), r * math.cos(ph)
        return x, y * 1.0e-6, z * 1.0e-6

    def _data(self, argv):
        # ReadCell::readData (ReadCell.cpp:308-375)
        if len(argv) < 6:
            return
        off = 3 if self.doubleEnd else 0
        x0 = y0 = z0 = 0.0
        if self.doubleEnd:
            x0, y0, z0 = self._coords(argv[2], argv[3], argv[4])
        x, y, z = self._coords(argv[off + 2], argv[off + 3], argv[off + 4])
        d = 1.0e-6 * float(argv[off + 5])
        compt, length = self._compartment(argv[0], argv[1], x0, y0, z0, x, y, z, d)
        self._channels(compt, argv, d, length)

    def _compartment(self, name, parent, x0, y0, z0, x, y, z, d):
        # ReadCell::buildCompartment (ReadCell.cpp:377-555)
        if parent == ".":
            pa = self.last
        elif parent in ("none", "nil"):
            pa = None
        else:
            pp = _norm(self.cell.path + "/" + parent)
            if pp not in _objects:
                raise ValueError(f"ReadCell: could not find parent compt '{parent}' for child '{name}'")
            pa = _objects[pp]
        c = Compartment(self.cell.path.rstrip("/") + "/" + name)
        self.last = c
        if pa is not None:
            px, py, pz = pa.__dict__["x"], pa.__dict__["y"], pa.__dict__["z"]
            if not self.doubleEnd:
                x0, y0, z0 = px, py, pz
            if self.relative:
                x, y, z = x + px, y + py, z + pz
                if self.doubleEnd:
                    x0, y0, z0 = x0 + px, y0 + py, z0 + pz
            dx, dy, dz = x - x0, y - y0, z - z0
            length = math.sqrt(dx * dx + dy * dy + dz * dz)
            connect(pa, "axial", c, "raxial")
        else:
            length = math.sqrt(x * x + y * y + z * z)
        surf = _calc_surf(length, d)
        Ra = self.RA * length * 4.0 / (d * d * math.pi) if length > 0 else self.RA * 8.0 / (d * math.pi)
        eleak = self.EREST_ACT if (self.erestFlag and not self.eleakFlag) else self.ELEAK
        erest = self.ELEAK if (not self.erestFlag and self.eleakFlag) else self.EREST_ACT
        c.__dict__.update(x0=x0, y0=y0, z0=z0, x=x, y=y, z=z, diameter=d, length=length,
                          Rm=self.RM / surf, Ra=Ra, Cm=self.CM * surf, initVm=erest, Em=eleak, Vm=erest)
        return c, length

    def _channels(self, compt, argv, dia, length):
        # ReadCell::buildChannels / addChannel / addCanonicalChannel /
        # addSpikeGen / addCaConc (ReadCell.cpp:628-890)
        start = 9 if self.doubleEnd else 6
        if (len(argv) % 2) != (1 if self.doubleEnd else 0):
            return
        j = start
        while j < len(argv):
            chan = argv[j]
            j += 1
            value = float(argv[j])
            cd = compt.__dict__
            if chan == "RA":
                cd["Ra"] = (8.0 * value / (dia * math.pi) if length == 0.0
                            else 4.0 * value * length / (dia * dia * math.pi))
            elif chan == "RM":
                cd["Rm"] = value * _calc_surf(length, dia)
            elif chan == "CM":
                cd["Cm"] = value * _calc_surf(length, dia)
            elif chan in ("Rm", "Ra", "Cm"):
                cd[chan] = value
            elif chan == "kinModel":
                j += 2
            elif chan in ("m2c", "c2m"):
                j += 4
            elif chan in self.protos:
                proto = self.protos[chan]
                cp = copy(proto, compt, proto.name)
                if isinstance(proto, (HHChannel, HHChannel2D)):
                    connect(compt, "channel", cp, "channel")
                    cp.__dict__["Gbar"] = value * _calc_surf(length, dia) if value > 0 else -value
                elif isinstance(proto, SpikeGen):
                    connect(compt, "VmOut", cp, "Vm")
                    cp.__dict__["threshold"] = value
                elif isinstance(proto, CaConc):
                    thick = cp.__dict__["thick"]
                    if thick > dia / 2.0:
                        thick = 0.0
                    if value > 0.0:
                        if length > 0.0:
                            vol = (math.pi * length * (dia - thick) * thick if thick > 0.0
                                   else dia * dia * math.pi * length / 4.0)
                        else:
                            inner = dia - 2 * thick
                            vol = (math.pi * (dia ** 3 - inner ** 3) / 6.0 if thick > 0.0
                                   else math.pi * dia ** 3 / 6.0)
                        if vol > 0.0:
                            value /= vol
                    else:
                        value = -value
                    cp.__dict__["B"] = value
            j += 1


def setClock(tick, dt):
    """Sets a clock tick's dt.  Each HSolve advances at its own dt; Tables
    sample every round(tick dt / solver dt) solver steps."""
    _clock_dt[int(tick)] = float(dt)


def getClock(tick):
    return _clock_dt[int(tick)]


def useClock(tick, path, field="process"):
    """Puts the Table or TimeTable at `path` on clock `tick`."""
    obj = element(path)
    if isinstance(obj, (Table, TimeTable)):
        obj.__dict__["tick"] = int(tick)


def _table_every(table, dt):
    tdt = _clock_dt.get(table.__dict__["tick"], dt)
    return max(1, int(round(tdt / dt))) if tdt > 0 else 1


def _wire_spikes():
    """SpikeGen spikeOut -> synapse addSpike messages between the solvers'
    cells become engine connections (routed on the host between stretches
    no longer than the shortest delay)."""
    where = {}
    for hs in _solvers:
        for g, sg in enumerate(hs._objs["spikegens"]):
            where.setdefault(sg.id, []).append((hs, g))
    for s, sf, d, df in _msgs:
        if sf != "spikeOut" or df != "addSpike" or not isinstance(d, Synapse):
            continue
        try:
            hs_dst, k, j = d._target()
        except RuntimeError:
            continue
        srcs = where.get(s.id)
        if not srcs:
            raise ValueError(f"{s.path} -> {d.path}: the SpikeGen is not under an HSolve")
        for hs_src, g in srcs:
            if hs_src._engine is not hs_dst._engine:
                raise ValueError(f"{s.path} -> {d.path}: solvers with different dt or precision")
            key = (s.id, d.id, hs_src._cell, hs_dst._cell)
            wired = hs_src._engine.__dict__.setdefault("_wired", set())
            if key not in wired:
                hs_src._engine.connect_spike(hs_src._cell, g, hs_dst._cell, k, j)
                wired.add(key)


def reinit():
    """HSolve::reinit on every solver (hsolve/HSolveActiveSetup.cpp:43-52);
    Tables record their t = 0 sample; Leakage objects set Gk = Gbar *
    modulation (Leakage::vReinit)."""
    for o in list(_objects.values()):
        if isinstance(o, Leakage):
            o.__dict__["Gk"] = o.__dict__["Gbar"] * o.__dict__["modulation"]
        if isinstance(o, TimeTable):
            o.__dict__.update(_pos=0, state=0.0)
    _wire_spikes()
    done = set()
    for hs in _solvers:
        hs._push_leaks()
        hs._push_synapses()
        hs._push_pulsegens()
    for hs in _solvers:
        eng = hs._engine
        if id(eng) not in done:
            eng.reinit(float(hs.dt))
            eng.read_spikes()
            done.add(id(eng))
    for hs in _solvers:
        vm, spk = hs._recorders()
        for t, (field, i) in vm:
            t.clearVec()
            t._data.append(float(hs._engine.get(hs._cell, field, [i])[0]))
        for t, _ in spk:
            t.clearVec()
        for t, obj, f, _ in hs._field_recorders(float(hs.dt)):
            t.clearVec()
            t._data.append(float(getattr(obj, f)))
        hs.__dict__["_time"] = 0.0
        hs.__dict__["_steps"] = 0


def start(runtime):
    """Advances every HSolve by round(runtime / dt) steps on the GPU and
    fills the Tables connected to its compartments and spike generators."""
    byeng = {}
    for hs in _solvers:
        byeng.setdefault(id(hs._engine), []).append(hs)
    _wire_spikes()
    for group in byeng.values():
        eng = group[0]._engine
        dt = float(group[0].dt)
        nsteps = int(round(runtime / dt))
        if nsteps <= 0:
            continue
        rec = []
        for hs in group:
            vm, spk = hs._recorders()
            comps = sorted({key for _, key in vm})
            # one sampling interval per cell on the device: the Tables'
            # common one, or every step with the Tables decimated here
            evs = {_table_every(t, dt) for t, _ in vm}
            every = evs.pop() if len(evs) == 1 else 1
            if comps:
                eng.set_probe_fields(hs._cell, [f for f, _ in comps], [i for _, i in comps],
                                     nsteps // every + 1, every=every)
            rec.append((hs, vm, spk, comps, every, hs.__dict__.get("_steps", 0)))
            hs._push_pulsegens()
            hs._push_leaks()
            hs._push_synapses()
        _feed_timetables(group, dt, nsteps)
        # tables of other fields: stepped in stretches ending on their
        # sampling steps, read by the host
        ftabs = [(hs, r) for hs in group for r in hs._field_recorders(dt)]
        if ftabs:
            grid = 0
            for _, (_, _, _, e) in ftabs:
                grid = math.gcd(grid, e)
            s0 = group[0].__dict__.get("_steps", 0)
            done = 0
            while done < nsteps:
                m = min(grid - (s0 + done) % grid, nsteps - done)
                eng.step(m)
                done += m
                for _, (t, obj, f, e) in ftabs:
                    if (s0 + done) % e == 0:
                        t._data.append(float(getattr(obj, f)))
        else:
            eng.step(nsteps)
        spikes = eng.read_spikes()
        for hs, vm, spk, comps, every, s0 in rec:
            if comps:
                tr = eng.read_probes(hs._cell)
                col = {c: j for j, c in enumerate(comps)}
                # row r is step (s0 // every + 1 + r) * every
                steps = (s0 // every + 1 + np.arange(len(tr))) * every
                for t, i in vm:
                    te = _table_every(t, dt)
                    keep = steps % te == 0
                    t._data.extend(tr[keep, col[i]].tolist())
            mine = spikes[spikes["cell"] == hs._cell]
            for t, g in spk:
                t._data.extend((mine[mine["spikegen"] == g]["step"] * dt).tolist())
            hs.__dict__["_steps"] = hs.__dict__.get("_steps", 0) + nsteps


reset()
