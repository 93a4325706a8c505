"""Command line and PLY I/O on the CPU: the reader over ascii / little- / big-endian inputs with extra
properties, the writer's byte layout (the reference's WritePLY, MeshPLY.cpp:67-127), and argument errors of
KNNCUDA that are reported before any GPU work."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "multigpuegtsmoothing_amd", "KNNCUDA")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ply") / "ply_roundtrip")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "multigpuegtsmoothing_amd", "cli"),
                    os.path.join(ROOT, "tests", "native", "ply_roundtrip.cpp"), "-o", out], check=True)
    return out


def _mesh(seed=0, nv=7, nf=4):
    r = np.random.default_rng(seed)
    v = r.standard_normal((nv, 3)).astype(np.float32)
    nrm = r.standard_normal((nv, 3)).astype(np.float32)
    col = r.integers(0, 256, (nv, 3)).astype(np.uint8)
    f = r.integers(0, nv, (nf, 3)).astype(np.uint32)
    return v, nrm, col, f


def _header(fmt, nv, nf, extra=True):
    h = ["ply", f"format {fmt} 1.0", "comment made by a test", f"element vertex {nv}", "property float x",
         "property float y", "property float z"]
    if extra:
        h.append("property double quality")
    h += ["property float nx", "property float ny", "property float nz", "property uchar red",
          "property uchar green", "property uchar blue", f"element face {nf}",
          "property list uchar int vertex_indices", "element edge 1", "property int vertex1", "property int vertex2",
          "end_header"]
    return ("\n".join(h) + "\n").encode()


def _write_input(path, fmt, v, nrm, col, f):
    nv, nf = len(v), len(f)
    with open(path, "wb") as fh:
        fh.write(_header(fmt, nv, nf))
        if fmt == "ascii":
            for i in range(nv):
                fh.write(("%r %r %r 0.5 %r %r %r %d %d %d\n" % (float(v[i, 0]), float(v[i, 1]), float(v[i, 2]),
                                                               float(nrm[i, 0]), float(nrm[i, 1]), float(nrm[i, 2]),
                                                               col[i, 0], col[i, 1], col[i, 2])).encode())
            for t in f:
                fh.write(("3 %d %d %d\n" % tuple(int(a) for a in t)).encode())
            fh.write(b"0 1\n")
        else:
            e = "<" if fmt == "binary_little_endian" else ">"
            for i in range(nv):
                fh.write(struct.pack(e + "fffdfffBBB", *v[i].tolist(), 0.5, *nrm[i].tolist(), *col[i].tolist()))
            for t in f:
                fh.write(struct.pack(e + "Biii", 3, *[int(a) for a in t]))
            fh.write(struct.pack(e + "ii", 0, 1))


def _parse_output(path):
    data = open(path, "rb").read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    return data[:end].decode(), data[end:]


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_roundtrip_layout(harness, tmp_path, fmt):
    v, nrm, col, f = _mesh(1)
    src, dst = str(tmp_path / "in.ply"), str(tmp_path / "out.ply")
    _write_input(src, fmt, v, nrm, col, f)
    r = subprocess.run([harness, src, dst], capture_output=True, text=True, check=True)
    assert r.stdout.split() == [str(len(v)), str(len(f)), str(len(v)), str(len(v))]
    header, body = _parse_output(dst)
    assert header == ("ply\nformat binary_little_endian 1.0\nelement vertex 7\nproperty float x\n"
                      "property float y\nproperty float z\nproperty float nx\nproperty float ny\n"
                      "property float nz\nproperty uchar red\nproperty uchar green\nproperty uchar blue\n"
                      "element face 4\nproperty list uchar uint vertex_indices\nend_header\n")
    vt = np.dtype([("p", "<f4", 3), ("n", "<f4", 3), ("c", "u1", 3)])
    ft = np.dtype([("cnt", "u1"), ("i", "<u4", 3)])
    assert len(body) == len(v) * vt.itemsize + len(f) * ft.itemsize
    vv = np.frombuffer(body[:len(v) * vt.itemsize], vt)
    ff = np.frombuffer(body[len(v) * vt.itemsize:], ft)
    np.testing.assert_array_equal(vv["p"], v)
    np.testing.assert_array_equal(vv["n"], nrm)
    np.testing.assert_array_equal(vv["c"], col)
    assert np.all(ff["cnt"] == 3)
    np.testing.assert_array_equal(ff["i"], f)


def test_ply_points_only(harness, tmp_path):
    src, dst = str(tmp_path / "p.ply"), str(tmp_path / "o.ply")
    pts = np.arange(12, dtype=np.float32).reshape(4, 3)
    with open(src, "wb") as fh:
        fh.write(b"ply\nformat binary_little_endian 1.0\nelement vertex 4\nproperty float x\nproperty float y\n"
                 b"property float z\nend_header\n" + pts.tobytes())
    subprocess.run([harness, src, dst], check=True, capture_output=True)
    header, body = _parse_output(dst)
    assert "element face" not in header and "nx" not in header
    np.testing.assert_array_equal(np.frombuffer(body, "<f4").reshape(4, 3), pts)


@pytest.mark.parametrize("bad", ["quad", "noxyz", "truncated"])
def test_ply_rejects_bad_input(harness, tmp_path, bad):
    src, dst = str(tmp_path / "b.ply"), str(tmp_path / "o.ply")
    if bad == "quad":
        txt = "ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n" \
              "element face 1\nproperty list uchar int vertex_indices\nend_header\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n" \
              "4 0 1 2 3\n"
    elif bad == "noxyz":
        txt = "ply\nformat ascii 1.0\nelement vertex 1\nproperty float u\nend_header\n0\n"
    else:
        txt = "ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n" \
              "end_header\n0 0 0\n1 1\n"
    open(src, "w").write(txt)
    r = subprocess.run([harness, src, dst], capture_output=True, text=True)
    assert r.returncode == 1 and "PLY" in r.stderr


def test_cli_usage_and_card_count(tmp_path):
    if not os.path.exists(CLI):
        pytest.fail("KNNCUDA is not built (run __graft_entry__.build())")
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    src = str(tmp_path / "p.ply")
    pts = np.random.default_rng(0).standard_normal((50, 3)).astype(np.float32)
    with open(src, "wb") as fh:
        fh.write(b"ply\nformat binary_little_endian 1.0\nelement vertex 50\nproperty float x\nproperty float y\n"
                 b"property float z\nend_header\n" + pts.tobytes())
    r = subprocess.run([CLI, src, "8", "2", "0.65", "0", "2"], capture_output=True, text=True, cwd=str(tmp_path))
    assert r.returncode == 1 and "num_cards must be 1" in r.stderr
    assert not os.path.exists(tmp_path / "output.ply")
