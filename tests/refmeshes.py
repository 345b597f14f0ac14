"""Fixture geometry restated from the reference's own test utilities.

- uniform_grid_nodes / uniform_grid_cell_boxes: mint::UniformMesh node
  coordinates x0 + i*h with h = (hi-lo)/(n-1), i fastest
  (mint/mesh/UniformMesh.hpp:271-279, UniformMesh.cpp:100-112) and the cell
  AABBs of spin/tests/spin_bvh.cpp:80-110.
- sphere_uv_mesh: quest/tests/quest_test_utilities.hpp:51-151 (including its
  duplicated pole row at phi = 0).
- plane_soup: the 4-triangle STL of quest_signed_distance_interface.cpp:61-116
  read as an unwelded triangle soup.
"""
import math

import numpy as np


def uniform_grid_nodes(lo, hi, n):
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    h = (hi - lo) / (n - 1)
    idx = np.arange(n)
    k, j, i = np.meshgrid(idx, idx, idx, indexing="ij")
    x = lo[0] + i.ravel() * h[0]
    y = lo[1] + j.ravel() * h[1]
    z = lo[2] + k.ravel() * h[2]
    return np.stack([x, y, z], 1)


def uniform_grid_cell_boxes(lo, hi, nodes):
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    nc = nodes - 1
    h = (hi - lo) / nc
    boxes = []
    for k in range(nc):
        for j in range(nc):
            for i in range(nc):
                mn = [lo[0] + i * h[0], lo[1] + j * h[1], lo[2] + k * h[2]]
                mx = [lo[0] + (i + 1) * h[0], lo[1] + (j + 1) * h[1], lo[2] + (k + 1) * h[2]]
                boxes.append(mn + mx)
    return np.array(boxes, float)


def sphere_uv_mesh(center, radius, theta_res, phi_res):
    deg = math.pi / 180.0
    theta_end = 360 * deg
    phi_end = 180 * deg
    verts = [[center[0], center[1], center[2] + radius],
             [center[0], center[1], center[2] - radius]]
    dphi = phi_end / float(phi_res - 1)
    dtheta = theta_end / float(theta_res - 1)
    for i in range(theta_res):
        theta = i * dtheta
        for j in range(phi_res - 2):
            phi = j * dphi
            r = radius * math.sin(phi)
            verts.append([r * math.cos(theta) + center[0], r * math.sin(theta) + center[1],
                          radius * math.cos(phi) + center[2]])
    pr = phi_res - 2
    stride = pr * theta_res
    tris = []
    for i in range(theta_res):
        tris.append([0, (pr * (i + 1) % stride) + 2, pr * i + 2])
    off = phi_res - 1
    for i in range(theta_res):
        tris.append([1, (pr * (i + 1) % stride) + off, pr * i + off])
    for i in range(theta_res):
        for j in range(phi_res - 3):
            c0 = pr * i + j + 2
            c2 = ((pr * (i + 1) + j) % stride) + 3
            tris.append([c0, c0 + 1, c2])
            tris.append([c0, c2, c2 - 1])
    return np.array(verts, float), np.array(tris, np.int32)


def plane_soup(tris):
    t = np.asarray(tris, float).reshape(-1, 3)
    return t, np.arange(t.shape[0], dtype=np.int32).reshape(-1, 3)
