"""MI355X-native HSolve: MOOSE's Hines compartmental solver on CDNA4.

Layers:
  cellspec / workloads ... cells in MOOSE object terms (what HSolve takes over)
  engine ................. ctypes binding of the C ABI (include/hsolve_amd.h)
  moose .................. the moose.HSolve object contract (target, dt,
                           reinit/start, zombie field reads and writes)
  serving ................ continuous-batching request loop
  autotune ............... per-size-bin launch-shape tuner
"""
from ._abi import FIELDS, HSG_FP64, HSG_MIXED
from .cellspec import CellSpec, Gate
from .engine import Engine, HSolveError, load_library, EXPORTS

__all__ = ["CellSpec", "Gate", "Engine", "HSolveError", "load_library", "EXPORTS", "FIELDS",
           "HSG_FP64", "HSG_MIXED"]
