"""STL reader feeding the signed-distance mesh.

Follows quest::STLReader (src/axom/quest/readers/STLReader.cpp:89-254):
  - format: binary when the file is at least 84 bytes long and its size is
    84 + 50 * (the int32 triangle count at byte 80), else ASCII (:40-86);
  - ASCII: every token "vertex" is followed by three coordinates parsed as
    double (:89-124), so a coordinate such as 0.1 keeps its double value;
  - binary: each triangle's nine float32 coordinates are widened to double
    (:127-190);
  - one node per facet corner (a triangle soup, no welding), faces =
    nodes // 3 with implicit connectivity (3i, 3i+1, 3i+2) (:207-254).
Pseudo-normals then rely on the closest-point clustering of SignedDistance.
The library's C reader (axsq_stl_read) follows the same rules.
"""
import os
import re
import struct

import numpy as np


def _is_binary(data):
    if len(data) < 84:
        return False
    ntris = struct.unpack_from("<i", data, 80)[0]
    return len(data) == 84 + 50 * ntris


# a decimal number as std::istream's double extraction reads one whole token
_NUMBER = re.compile(rb"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def _read_ascii(data):
    coords = []
    tok = data.split()
    k = 0
    while k < len(tok):
        if tok[k] != b"vertex":
            k += 1
            continue
        xyz = tok[k + 1:k + 4]
        if len(xyz) < 3 or not all(_NUMBER.fullmatch(t) for t in xyz):
            break  # the reference's stream extraction fails here and reading stops
        xyz = [float(t) for t in xyz]
        coords.append(xyz)
        k += 4
    return np.asarray(coords, dtype=np.float64).reshape(-1, 3)


def read_stl(path):
    """Returns (vertices (N,3) float64, triangles (N//3,3) int32)."""
    if not os.path.exists(path):
        raise OSError(f"STL file not found: {path}")
    with open(path, "rb") as f:
        data = f.read()
    if _is_binary(data):
        ntris = struct.unpack_from("<i", data, 80)[0]
        rec = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")])
        tris = np.frombuffer(data, dtype=rec, count=ntris, offset=84)
        verts = tris["v"].reshape(-1, 3).astype(np.float64)
    else:
        verts = _read_ascii(data)
    if verts.shape[0] < 3:
        raise ValueError(f"no triangles in STL file: {path}")
    nf = verts.shape[0] // 3
    t = np.arange(3 * nf, dtype=np.int32).reshape(-1, 3)
    return np.ascontiguousarray(verts), t


def write_ascii_stl(path, verts, tris, name="mesh"):
    """ASCII STL with every coordinate written with 17 significant digits, so
    that reading it back gives the same doubles."""
    with open(path, "w") as f:
        f.write(f"solid {name}\n")
        for a, b, c in tris:
            f.write("  facet normal 0 0 0\n    outer loop\n")
            for v in (a, b, c):
                f.write("      vertex {:.17g} {:.17g} {:.17g}\n".format(*verts[v]))
            f.write("    endloop\n  endfacet\n")
        f.write(f"endsolid {name}\n")


def write_binary_stl(path, verts, tris):
    """Binary STL (80-byte header, int32 count, 50 bytes per triangle); the
    coordinates are stored as float32."""
    tris = np.asarray(tris)
    rec = np.zeros(tris.shape[0], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)),
                                                  ("attr", "<u2")]))
    rec["v"] = np.asarray(verts, np.float64)[tris].astype(np.float32)
    with open(path, "wb") as f:
        f.write(b"binary stl".ljust(80, b" "))
        f.write(struct.pack("<i", tris.shape[0]))
        f.write(rec.tobytes())
