"""The C-ABI library loads and exports every entry point include/pcc.h
declares (no compute call: runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

from pcc import engine

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "pcc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcc_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_expected_surface():
    names = declared_functions()
    for must in ("pcc_create", "pcc_destroy", "pcc_set_state", "pcc_get_state", "pcc_step",
                 "pcc_get_constraints", "pcc_last_error", "pcc_detect"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(engine.LIB_PATH), reason="libpcc.so not built")
def test_library_exports_every_declared_symbol():
    L = C.CDLL(engine.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(engine.EXPORTED) <= set(declared_functions())


@pytest.mark.skipif(not os.path.exists(engine.LIB_PATH), reason="libpcc.so not built")
def test_pure_host_entry_points():
    p = engine.default_params()
    assert p.mode == 0 and p.linked_cell_size == 2.2 and p.warm_start_factor == 500.0
    L = engine.lib()
    assert L.pcc_status_string(1).decode() == "BBPGD did not converge"
    assert L.pcc_status_string(2).decode() == "Maximum number of constraint iterations reached"
    assert L.pcc_status_string(3).decode() == "Invalid mode"
    assert L.pcc_kernel_name(0).decode() == "body_apply"
