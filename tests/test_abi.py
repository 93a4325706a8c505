"""The C-ABI library loads and exports every entry point include/egt.h declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "egt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(egt_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "egt_grid_create" in syms and "egt_smooth" in syms and len(syms) >= 18


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "multigpuegtsmoothing_amd", "libegt.so")
    assert os.path.exists(path), "libegt.so not built"
    lib = ctypes.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_mirror_loads_without_gpu_calls():
    from multigpuegtsmoothing_amd import egt as E

    assert E.lib().egt_last_error_string() is not None
    assert E.lib().egt_destroy(None) == 0
