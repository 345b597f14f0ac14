"""STL reader feeding the signed-distance mesh.

Follows quest::STLReader (src/axom/quest/readers/STLReader.cpp:89-254): ASCII
or binary STL, vertex coordinates read as float32 and widened to float64,
one node per facet corner (a triangle soup, no welding; pseudo-normals
then rely on the closest-point clustering of SignedDistance).
"""
import os
import struct

import numpy as np


def _is_binary(data):
    if len(data) < 84:
        return False
    ntris = struct.unpack_from("<I", data, 80)[0]
    return len(data) == 84 + 50 * ntris


def read_stl(path):
    """Returns (vertices (3T,3) float64, triangles (T,3) int32)."""
    if not os.path.exists(path):
        raise OSError(f"STL file not found: {path}")
    data = open(path, "rb").read()
    if _is_binary(data):
        ntris = struct.unpack_from("<I", data, 80)[0]
        rec = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")])
        tris = np.frombuffer(data, dtype=rec, count=ntris, offset=84)
        verts = tris["v"].reshape(-1, 3).astype(np.float32).astype(np.float64)
    else:
        coords = []
        for line in data.decode("ascii", errors="replace").splitlines():
            tok = line.split()
            if len(tok) == 4 and tok[0] == "vertex":
                coords.append([np.float32(float(t)) for t in tok[1:]])
        if len(coords) % 3 != 0 or not coords:
            raise ValueError(f"malformed ASCII STL: {path}")
        verts = np.asarray(coords, dtype=np.float32).astype(np.float64)
    t = np.arange(verts.shape[0], dtype=np.int32).reshape(-1, 3)
    return np.ascontiguousarray(verts), t


def write_ascii_stl(path, verts, tris, name="mesh"):
    with open(path, "w") as f:
        f.write(f"solid {name}\n")
        for a, b, c in tris:
            f.write("  facet normal 0 0 0\n    outer loop\n")
            for v in (a, b, c):
                f.write("      vertex {:.9g} {:.9g} {:.9g}\n".format(*verts[v]))
            f.write("    endloop\n  endfacet\n")
        f.write(f"endsolid {name}\n")
