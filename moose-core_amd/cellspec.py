"""Cell descriptions in MOOSE object terms.

A CellSpec is what one moose.HSolve object takes over: compartments in
creation order with their passive fields, the axial messages in creation
order, HHChannels in channel-message order with their HHGate prototypes, CaConc
pools, SpikeGens and Leakage channels (hsolve/HSolvePassive.cpp:45-205,
hsolve/HSolveActiveSetup.cpp:162-578).  It packs into the C structures of
include/hsolve_amd.h (see _abi.PackedCell).
"""
from dataclasses import dataclass, field

import numpy as np

from ._abi import (CACONC_DT, CHAN2D_DT, CHANNEL_DT, COMPARTMENT_DT, LEAK_DT, PULSEGEN_DT, SPIKEGEN_DT,
                   SYNAPSE_DT, SYNCHAN_DT, VCLAMP_DT, PackedCell)


@dataclass
class Gate:
    """An HHGate prototype: A = alpha, B = alpha + beta over [min, max]."""
    id: int
    min: float
    max: float
    divs: int
    tableA: np.ndarray
    tableB: np.ndarray


@dataclass
class Table2D:
    """An Interpol2D: (xdivs+1) x (ydivs+1) values over [xmin,xmax] x [ymin,ymax]."""
    xdivs: int
    ydivs: int
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    table: np.ndarray


@dataclass
class CellSpec:
    compartments: np.ndarray                       # COMPARTMENT_DT
    axial: np.ndarray                              # int32 (m, 2): (parent, child)
    seed: int = 0
    channels: np.ndarray = field(default_factory=lambda: np.zeros(0, CHANNEL_DT))
    gates: list = field(default_factory=list)
    pools: np.ndarray = field(default_factory=lambda: np.zeros(0, CACONC_DT))
    spikegens: np.ndarray = field(default_factory=lambda: np.zeros(0, SPIKEGEN_DT))
    leaks: np.ndarray = field(default_factory=lambda: np.zeros(0, LEAK_DT))
    tables2d: list = field(default_factory=list)
    chans2d: np.ndarray = field(default_factory=lambda: np.zeros(0, CHAN2D_DT))
    pulsegens: np.ndarray = field(default_factory=lambda: np.zeros(0, PULSEGEN_DT))
    synchans: np.ndarray = field(default_factory=lambda: np.zeros(0, SYNCHAN_DT))
    synapses: np.ndarray = field(default_factory=lambda: np.zeros(0, SYNAPSE_DT))
    vclamps: np.ndarray = field(default_factory=lambda: np.zeros(0, VCLAMP_DT))

    @property
    def n(self):
        return len(self.compartments)

    def pack(self):
        return PackedCell(self)


def passive_compartments(length, dia, RM=1.0, RA=1.0, CM=0.01, Em=-0.065, initVm=-0.065):
    """Cylinder passives as rdesigneur's buildCompt computes them
    (python/rdesigneur/rdesigneurProtos.py:620-638): Ra = RA*L/(pi d^2/4),
    Rm = RM/(pi d L), Cm = CM*pi d L."""
    length = np.asarray(length, dtype=np.float64)
    dia = np.asarray(dia, dtype=np.float64)
    PI = 3.14159265359
    xa = dia * dia * PI / 4.0
    sa = length * dia * PI
    c = np.zeros(len(length), COMPARTMENT_DT)
    c["Ra"] = length * RA / xa
    c["Rm"] = RM / sa
    c["Cm"] = CM * sa
    c["Em"] = Em
    c["initVm"] = initVm
    c["Vm"] = initVm
    c["inject"] = 0.0
    return c, sa


def channel_records(n, compartment, Gbar, Ek, Xpower=0.0, Ypower=0.0, Zpower=0.0,
                    gates=(-1, -1, -1), X=0.0, Y=0.0, Z=0.0, instant=0,
                    useConcentration=0, caTarget=-1, caDepend=-1, modulation=1.0):
    """n channel records (vectorised over compartment/Gbar/...)."""
    r = np.zeros(n, CHANNEL_DT)
    r["compartment"] = compartment
    r["instant"] = instant
    r["useConcentration"] = useConcentration
    r["gate"] = np.asarray(gates, dtype=np.int32)
    r["caTarget"] = caTarget
    r["caDepend"] = caDepend
    r["Gbar"] = Gbar
    r["Ek"] = Ek
    r["Xpower"] = Xpower
    r["Ypower"] = Ypower
    r["Zpower"] = Zpower
    r["X"] = X
    r["Y"] = Y
    r["Z"] = Z
    r["modulation"] = modulation
    return r


def interleave_channels(per_type):
    """Channel records of several types over the same compartments, ordered the
    way MOOSE lists a compartment's channel messages: compartment by
    compartment, types in the given order."""
    n = len(per_type[0])
    out = np.zeros(n * len(per_type), CHANNEL_DT)
    for t, r in enumerate(per_type):
        out[t::len(per_type)] = r
    order = np.argsort(out["compartment"], kind="stable")
    return out[order]
