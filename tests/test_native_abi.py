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


def test_index_space_guards(tmp_path):
    """The CSR and mesh corner-list limits (int32 offsets) accept counts up to the limit and refuse one past it,
    including totals past 2^32 that a 32-bit sum would wrap."""
    exe = str(tmp_path / "limits_check")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "multigpuegtsmoothing_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "limits_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout


def test_mesh_triangle_limit_refused_before_any_gpu_call():
    """egt_mesh_create refuses a triangle count whose corner lists would overflow int32 offsets with
    EGT_ERR_UNSUPPORTED, before touching the device (so this runs without a GPU)."""
    import numpy as np
    from multigpuegtsmoothing_amd import egt as E

    L = E.lib()
    x = np.zeros(3, np.float32)
    tri = np.zeros(3, np.uint32)
    h = ctypes.c_void_p()
    st = L.egt_mesh_create(x, x, x, 3, tri, (2 ** 31 - 1) // 6 + 1, 1, 1e-6, 1024, ctypes.byref(h))
    assert st == 4 and not h.value  # EGT_ERR_UNSUPPORTED
    assert b"triangles" in L.egt_last_error_string()
