"""The C-ABI library: it loads and exports every entry point that
include/hsolve_amd.h declares, and the ctypes mirror of its structures has the
C layout (checked against a C program compiled from the header).  No GPU
work is done here."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from moose_core_amd import _abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hsolve_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(hsg_\w+)\s*\(", text, re.M)))


def test_header_lists_python_exports():
    assert declared_functions() == sorted(engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = engine.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_last_error_without_device_work():
    lib = engine.load_library()
    assert isinstance(lib.hsg_last_error(), bytes)


def test_struct_layouts_match_c():
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "hsolve_amd.h"
#define P(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(T, f))
int main(void) {
  printf("hsg_gate_desc %zu\n", sizeof(hsg_gate_desc));
  printf("hsg_compartment_desc %zu\n", sizeof(hsg_compartment_desc));
  printf("hsg_channel_desc %zu\n", sizeof(hsg_channel_desc));
  printf("hsg_caconc_desc %zu\n", sizeof(hsg_caconc_desc));
  printf("hsg_spikegen_desc %zu\n", sizeof(hsg_spikegen_desc));
  printf("hsg_leak_desc %zu\n", sizeof(hsg_leak_desc));
  printf("hsg_table2d_desc %zu\n", sizeof(hsg_table2d_desc));
  printf("hsg_chan2d_desc %zu\n", sizeof(hsg_chan2d_desc));
  printf("hsg_cell_desc %zu\n", sizeof(hsg_cell_desc));
  printf("hsg_options %zu\n", sizeof(hsg_options));
  printf("hsg_spike %zu\n", sizeof(hsg_spike));
  printf("hsg_cell_info %zu\n", sizeof(hsg_cell_info));
  P(hsg_channel_desc, Gbar); P(hsg_channel_desc, caDepend); P(hsg_channel_desc, modulation);
  P(hsg_leak_desc, Gk); P(hsg_spikegen_desc, threshold); P(hsg_cell_desc, leaks);
  P(hsg_gate_desc, tableA); P(hsg_cell_desc, chans2d); P(hsg_chan2d_desc, Gbar);
  P(hsg_chan2d_desc, conc2); P(hsg_table2d_desc, table);
  printf("hsg_pulsegen_desc %zu\n", sizeof(hsg_pulsegen_desc));
  P(hsg_pulsegen_desc, baseLevel); P(hsg_pulsegen_desc, width); P(hsg_pulsegen_desc, level);
  P(hsg_cell_desc, pulsegens);
  printf("hsg_synchan_desc %zu\n", sizeof(hsg_synchan_desc));
  printf("hsg_synapse_desc %zu\n", sizeof(hsg_synapse_desc));
  printf("hsg_vclamp_desc %zu\n", sizeof(hsg_vclamp_desc));
  P(hsg_leak_desc, Gbar); P(hsg_leak_desc, modulation);
  P(hsg_synchan_desc, Gbar); P(hsg_synchan_desc, tau2); P(hsg_synchan_desc, modulation);
  P(hsg_vclamp_desc, tau); P(hsg_vclamp_desc, gain);
  P(hsg_cell_desc, synchans); P(hsg_cell_desc, synapses); P(hsg_cell_desc, vclamps);
  printf("hsg_bin_info %zu\n", sizeof(hsg_bin_info));
  P(hsg_bin_info, compartments); P(hsg_bin_info, round_ms); P(hsg_bin_info, kernel);
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "l")
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), "-o", exe, c], check=True)
        out = dict(line.rsplit(" ", 1) for line in subprocess.run([exe], capture_output=True, text=True,
                                                                  check=True).stdout.split("\n") if line)
    sizes = {k: int(v) for k, v in out.items()}
    assert sizes["hsg_gate_desc"] == ctypes.sizeof(_abi.GateDesc)
    assert sizes["hsg_compartment_desc"] == _abi.COMPARTMENT_DT.itemsize
    assert sizes["hsg_channel_desc"] == _abi.CHANNEL_DT.itemsize
    assert sizes["hsg_caconc_desc"] == _abi.CACONC_DT.itemsize
    assert sizes["hsg_spikegen_desc"] == _abi.SPIKEGEN_DT.itemsize
    assert sizes["hsg_leak_desc"] == _abi.LEAK_DT.itemsize
    assert sizes["hsg_cell_desc"] == ctypes.sizeof(_abi.CellDesc)
    assert sizes["hsg_options"] == ctypes.sizeof(_abi.Options)
    assert sizes["hsg_spike"] == ctypes.sizeof(_abi.Spike)
    assert sizes["hsg_cell_info"] == ctypes.sizeof(_abi.CellInfo)
    assert sizes["hsg_channel_desc.Gbar"] == _abi.CHANNEL_DT.fields["Gbar"][1]
    assert sizes["hsg_channel_desc.caDepend"] == _abi.CHANNEL_DT.fields["caDepend"][1]
    assert sizes["hsg_channel_desc.modulation"] == _abi.CHANNEL_DT.fields["modulation"][1]
    assert sizes["hsg_leak_desc.Gk"] == _abi.LEAK_DT.fields["Gk"][1]
    assert sizes["hsg_spikegen_desc.threshold"] == _abi.SPIKEGEN_DT.fields["threshold"][1]
    assert sizes["hsg_cell_desc.leaks"] == _abi.CellDesc.leaks.offset
    assert sizes["hsg_gate_desc.tableA"] == _abi.GateDesc.tableA.offset
    assert sizes["hsg_table2d_desc"] == ctypes.sizeof(_abi.Table2DDesc)
    assert sizes["hsg_chan2d_desc"] == _abi.CHAN2D_DT.itemsize
    assert sizes["hsg_cell_desc.chans2d"] == _abi.CellDesc.chans2d.offset
    assert sizes["hsg_chan2d_desc.Gbar"] == _abi.CHAN2D_DT.fields["Gbar"][1]
    assert sizes["hsg_chan2d_desc.conc2"] == _abi.CHAN2D_DT.fields["conc2"][1]
    assert sizes["hsg_table2d_desc.table"] == _abi.Table2DDesc.table.offset
    assert sizes["hsg_pulsegen_desc"] == _abi.PULSEGEN_DT.itemsize
    for f in ("baseLevel", "width", "level"):
        assert sizes["hsg_pulsegen_desc." + f] == _abi.PULSEGEN_DT.fields[f][1]
    assert sizes["hsg_cell_desc.pulsegens"] == _abi.CellDesc.pulsegens.offset
    assert sizes["hsg_synchan_desc"] == _abi.SYNCHAN_DT.itemsize
    assert sizes["hsg_synapse_desc"] == _abi.SYNAPSE_DT.itemsize
    assert sizes["hsg_vclamp_desc"] == _abi.VCLAMP_DT.itemsize
    for f in ("Gbar", "modulation"):
        assert sizes["hsg_leak_desc." + f] == _abi.LEAK_DT.fields[f][1]
    for f in ("Gbar", "tau2", "modulation"):
        assert sizes["hsg_synchan_desc." + f] == _abi.SYNCHAN_DT.fields[f][1]
    for f in ("tau", "gain"):
        assert sizes["hsg_vclamp_desc." + f] == _abi.VCLAMP_DT.fields[f][1]
    for f in ("synchans", "synapses", "vclamps"):
        assert sizes["hsg_cell_desc." + f] == getattr(_abi.CellDesc, f).offset
    assert sizes["hsg_bin_info"] == ctypes.sizeof(_abi.BinInfo)
    for f in ("compartments", "round_ms", "kernel"):
        assert sizes["hsg_bin_info." + f] == getattr(_abi.BinInfo, f).offset


def test_engine_without_gpu_raises_cleanly():
    """On a machine without a HIP device the engine refuses loudly (no CPU
    fallback)."""
    import torch  # noqa: F401  (device count only)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(engine.HSolveError):
        engine.Engine("fp64")
