"""CPU-side checks of the product library: it loads without a GPU and exports every entry
point include/cleo_hip.h declares; the ctypes mirror matches the header's struct sizes."""
import ctypes as C
import os
import re

import pytest

from cleo_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cleo_hip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cleo_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from cleo_amd import lib
    if not os.path.exists(lib.LIB_PATH):
        pytest.fail("libcleo_hip.so not built (run __graft_entry__.build())")
    so = C.CDLL(lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert len(declared()) >= 40


def test_struct_layouts_match_header():
    # sizes of the POD structs as gcc lays out include/cleo_hip.h on x86-64
    sizes = [C.sizeof(x) for x in (abi.DomainDesc, abi.SupersSoA, abi.GbxState, abi.CollParams, abi.CondParams,
                                   abi.SDMParams, abi.Error, abi.Stats, abi.Process)]
    assert sizes == [56, 64, 88, 64, 48, 136, 16, 72, 120]


def test_oracle_exports_mirror(oracle):
    names = [n.replace("cleo_", "orc_") for n in declared()
             if not n.startswith(("cleo_mlp", "cleo_event", "cleo_autotune", "cleo_time_kernel", "cleo_set_tiles",
                                  "cleo_get_tiles",
                                  "cleo_device_info", "cleo_upload_supers_aos", "cleo_download_supers_aos",
                                  "cleo_upload_supers_device", "cleo_run_steps_graph", "cleo_synchronize",
                                  "cleo_record_guess", "cleo_download_guess", "cleo_ndomain",
                                  # no context: the sort cleo_upload_supers_device's parity covers
                                  "cleo_sort_supers_noalloc",
                                  # device functions the oracle exports one by one (orc_saturation_pressure, ...)
                                  "cleo_thermo_eval",
                                  # host-only file readers: they compute nothing the oracle checks
                                  "cleo_read_"))]
    missing = [n for n in names if not hasattr(oracle.lib, n)]
    assert not missing, missing
