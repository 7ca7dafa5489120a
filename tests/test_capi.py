"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU call is made)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(cgfdm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("cgfdm_create", "cgfdm_step", "cgfdm_get_stations", "cgfdm_destroy"):
        assert must in names


def test_library_exports_all_declared_symbols():
    from cgfdm_amd import solver
    if not os.path.exists(solver.LIB_PATH):
        pytest.fail("libcgfdm_amd.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(solver.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(solver.EXPORTS) <= set(declared_symbols())


def test_error_message_without_gpu_work():
    from cgfdm_amd import solver
    lib = solver.lib()
    rc = lib.cgfdm_create(None, None, None, None)
    assert rc == 2
    assert b"null" in lib.cgfdm_last_error()
