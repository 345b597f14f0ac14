"""CPU checks of the STL readers: the library's axsq_stl_read (host code, no
GPU needed) and the Python reader in axom_amd/stl.py, against
quest::STLReader's rules (src/axom/quest/readers/STLReader.cpp:40-254) and the
reference's own test file (quest/tests/quest_stl_reader.cpp:33-52,99-140)."""
import ctypes
import struct

import numpy as np
import pytest

from axom_amd import _lib, stl


def lib_read(path):
    """(nodes (n,3) float64, nfaces) through the library's reader."""
    L = _lib.lib()
    p, nn, nf = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int64()
    _lib.check(L.axsq_stl_read(str(path).encode(), ctypes.byref(p), ctypes.byref(nn),
                               ctypes.byref(nf)))
    try:
        buf = (ctypes.c_double * (3 * nn.value)).from_address(p.value) if nn.value else []
        nodes = np.array(buf, dtype=np.float64).reshape(-1, 3)
    finally:
        L.axsq_host_free(p)
    return nodes, nf.value


def _reference_triangle_file(path):
    # the file quest_stl_reader.cpp:33-52 generates
    path.write_text("solid triangle\n\t facet normal 0.0 0.0 1.0\n\t\t outer loop\n"
                    "\t\t\t vertex 0.0 0.0 0.0\n\t\t\t vertex 1.0 0.0 0.0\n"
                    "\t\t\t vertex 0.0 1.0 0.0\n\t\t endloop\n\t endfacet\n"
                    "endsolid triangle\n")


def test_reference_single_triangle(tmp_path):
    f = tmp_path / "triangle.stl"
    _reference_triangle_file(f)
    want = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    nodes, nf = lib_read(f)
    assert nf == 1 and np.array_equal(nodes, want)
    v, t = stl.read_stl(str(f))
    assert np.array_equal(v, want) and t.tolist() == [[0, 1, 2]]


def test_missing_file(tmp_path):
    L = _lib.lib()
    p, nn, nf = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int64()
    rc = L.axsq_stl_read(str(tmp_path / "foo.stl").encode(), ctypes.byref(p), ctypes.byref(nn),
                         ctypes.byref(nf))
    assert rc == _lib.ERRORS_BY_NAME["AXSQ_ERR_INVALID_ARG"]
    assert b"foo.stl" in L.axsq_last_error()
    with pytest.raises(OSError):
        stl.read_stl(str(tmp_path / "foo.stl"))


def _mesh():
    """Two triangles whose coordinates are not float32-exact."""
    v = np.array([[0.1, 1.0 / 3.0, -2.5e-3], [1.7, 0.2, 0.3], [0.45, 2.0 / 3.0, 1e-7],
                  [-0.1, 0.123456789012345, 5.5]])
    t = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return v, t


def test_ascii_coordinates_are_doubles(tmp_path):
    """ASCII coordinates are read as double (STLReader.cpp:101-116): 0.1 stays
    0.1, not its float32 neighbour."""
    v, t = _mesh()
    f = tmp_path / "m.stl"
    stl.write_ascii_stl(str(f), v, t)
    want = v[t.reshape(-1)]
    assert not np.array_equal(want, want.astype(np.float32).astype(np.float64))
    nodes, nf = lib_read(f)
    assert nf == 2 and np.array_equal(nodes, want)
    pv, pt = stl.read_stl(str(f))
    assert np.array_equal(pv, want) and np.array_equal(pt, [[0, 1, 2], [3, 4, 5]])


def test_ascii_tokens_parse_like_strtod(tmp_path):
    """Each coordinate token is parsed as a whole decimal number; the two
    readers agree bit for bit with float() (correctly rounded) on awkward
    spellings."""
    toks = ["1e-3", "-0", "+.5", "7.", "1E+2", "0.30000000000000004", "123456789.123456789",
            "-4.9e-324", "2.2250738585072014e-308"]
    while len(toks) % 9:
        toks.append("0.25")
    lines = ["solid x"]
    for k in range(0, len(toks), 3):
        if k % 9 == 0:
            lines.append("facet normal 0 0 1\nouter loop")
        lines.append("vertex " + " ".join(toks[k:k + 3]))
        if k % 9 == 6:
            lines.append("endloop\nendfacet")
    lines.append("endsolid x")
    f = tmp_path / "t.stl"
    f.write_text("\n".join(lines) + "\n")
    want = np.array([float(x) for x in toks]).reshape(-1, 3)
    nodes, _ = lib_read(f)
    assert np.array_equal(nodes.view(np.int64), want.view(np.int64))
    pv, _ = stl.read_stl(str(f))
    assert np.array_equal(pv.view(np.int64), want.view(np.int64))


def test_ascii_reading_stops_at_a_bad_coordinate(tmp_path):
    f = tmp_path / "bad.stl"
    f.write_text("solid s\nvertex 0 0 0\nvertex 1 0 0\nvertex 0 1 0\n"
                 "vertex 0 0 1\nvertex nan 0 0\nvertex 1 1 1\nendsolid s\n")
    nodes, nf = lib_read(f)
    assert nodes.shape == (4, 3) and nf == 1  # the node after the last face is kept
    pv, pt = stl.read_stl(str(f))
    assert np.array_equal(pv, nodes) and pt.shape == (1, 3)


def test_binary_is_float_widened(tmp_path):
    """Binary coordinates are float32 widened to double (STLReader.cpp:131-190),
    and the format is recognised by its size (:40-86)."""
    v, t = _mesh()
    f = tmp_path / "m_bin.stl"
    stl.write_binary_stl(str(f), v, t)
    assert f.stat().st_size == 84 + 50 * 2
    want = v[t.reshape(-1)].astype(np.float32).astype(np.float64)
    nodes, nf = lib_read(f)
    assert nf == 2 and np.array_equal(nodes, want)
    pv, _ = stl.read_stl(str(f))
    assert np.array_equal(pv, want)


def test_binary_size_mismatch_is_read_as_ascii(tmp_path):
    """A file whose size does not match its binary triangle count is ASCII."""
    v, t = _mesh()
    f = tmp_path / "m_bin.stl"
    stl.write_binary_stl(str(f), v, t)
    data = f.read_bytes() + b"\n"
    f.write_bytes(data)
    nodes, nf = lib_read(f)
    assert nf == 0  # no "vertex" token in the binary bytes
    # and a header whose count is negative is ASCII too
    g = tmp_path / "neg.stl"
    g.write_bytes(b"vertex 1 2 3 ".ljust(80, b" ") + struct.pack("<i", -1) + b" vertex 4 5 6")
    nodes, nf = lib_read(g)
    assert nodes.tolist() == [[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]] and nf == 0
