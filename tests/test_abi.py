"""The C-ABI library builds for gfx950, loads, and exports every function that
include/wcfe.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(REPO, "include", "wcfe.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wcfe_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_hot_path_entry_points():
    fns = header_functions()
    for name in ["wcfe_calculate_res", "wcfe_force_fa", "wcfe_update_fa", "wcfe_newton_step", "wcfe_dense_solve"]:
        assert name in fns


def test_library_exports_every_declared_symbol():
    import wholecellfe_amd as W
    if not os.path.exists(W.LIB_PATH):
        from wholecellfe_amd import build
        build.build()
    lib = ctypes.CDLL(W.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(W.EXPORTS) == header_functions()


def test_code_objects_target_gfx950():
    import wholecellfe_amd as W
    data = open(W.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_python_binding_refuses_without_library(tmp_path, monkeypatch):
    import wholecellfe_amd as W
    monkeypatch.setattr(W, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(W, "_lib", None)
    with pytest.raises(ImportError):
        W.lib()
