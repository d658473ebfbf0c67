"""The C-ABI library loads and exports every function include/exadg_amd.h declares (no GPU)."""
import ctypes
import os
import re

import exadg_amd


def declared_functions():
    src = open(exadg_amd.HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(exadg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "exadg_laplace_vmult" in names and "exadg_laplace_create" in names
    assert len(names) >= 10


def test_library_exports_all_declared_symbols():
    assert os.path.exists(exadg_amd.LIB_PATH), "build with python -m exadg_amd.build"
    L = ctypes.CDLL(exadg_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_error_path_without_gpu():
    """Invalid arguments are rejected before any device work, with a message."""
    L = exadg_amd.lib()
    h = ctypes.c_void_p()
    m = exadg_amd.mesh_desc(2)
    b = exadg_amd.bc_desc()
    assert L.exadg_laplace_create(ctypes.byref(m), ctypes.byref(b), 9, 0, 1.0, ctypes.byref(h)) != 0
    assert b"degree" in L.exadg_last_error()
    assert L.exadg_laplace_create(None, ctypes.byref(b), 2, 0, 1.0, ctypes.byref(h)) == 1
    bad = exadg_amd.mesh_desc((0, 1, 1))
    assert L.exadg_laplace_create(ctypes.byref(bad), ctypes.byref(b), 2, 0, 1.0, ctypes.byref(h)) == 1
