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


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the built HIP library the package raises; there is no CPU fallback to route to."""
    import pytest
    monkeypatch.setattr(exadg_amd, "_lib", None)
    monkeypatch.setattr(exadg_amd, "LIB_PATH", str(tmp_path / "libexadg_amd.so"))
    with pytest.raises(ImportError, match="build it"):
        exadg_amd.lib()


def test_product_does_not_use_the_oracle():
    """The product package and its native sources neither import nor link the oracle."""
    root = os.path.dirname(os.path.abspath(exadg_amd.__file__))
    offenders = []
    for dirpath, _, files in os.walk(root):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                if re.search(r"^\s*(import|from)\s+oracle\b|liboracle|oracle_capi|oracle_core",
                             text, flags=re.M):
                    offenders.append(f)
    assert not offenders, offenders
