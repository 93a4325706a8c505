"""C layout of egt_stats against the ctypes mirror, and a reference-style C++ call site of the query search
compiled against include/egt.hpp (CPU only, nothing runs on a GPU)."""
import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_stats_layout_matches_ctypes(tmp_path):
    from multigpuegtsmoothing_amd import egt as E

    exe = str(tmp_path / "stats_layout")
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "stats_layout.c"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    c = dict(line.split() for line in out if line)
    assert int(c.pop("sizeof")) == ctypes.sizeof(E.Stats)
    for name, off in c.items():
        assert getattr(E.Stats, name).offset == int(off), name
    assert len(c) == len(E.Stats._fields_)


def test_query_call_site_compiles():
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "callsite_query.cpp")], check=True)
