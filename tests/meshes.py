"""Closed triangle meshes for the mesh-surface tests."""
import math

import numpy as np


def cube_mesh(lo, hi):
    """the box [lo, hi]: 8 vertices (bit a of the index picks hi on axis a), 12 triangles"""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    v = np.array([[(hi if (k >> a) & 1 else lo)[a] for a in range(3)] for k in range(8)])
    f = [[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6], [0, 1, 4], [1, 5, 4], [2, 6, 3], [3, 6, 7],
         [0, 4, 2], [2, 4, 6], [1, 3, 5], [3, 7, 5]]
    return v, np.array(f)


def octahedron(radius):
    """|x| + |y| + |z| <= radius"""
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float) * radius
    f = [[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]]
    return v, np.array(f)


def star_prism(points, radius, radius_internal, half_height, direction=(0.0, 0.0, 1.0)):
    """the prism |y| <= half_height over the star of `points` outer vertices at
    `radius` (the first along `direction`, in x-z) and inner vertices at
    `radius_internal` between them: the shape of the analytic star surface.
    Caps are fans from the axis (the star is star-shaped about it)."""
    d = np.array([direction[0], 0.0, direction[2]], float)
    d /= np.linalg.norm(d)
    perp = np.array([d[2], 0.0, -d[0]])
    a0 = 2 * math.pi / points
    ring = []
    for k in range(points):
        for j, rr in ((0, radius), (1, radius_internal)):
            t = (k + 0.5 * j) * a0
            ring.append(rr * (math.cos(t) * d + math.sin(t) * perp))
    ring = np.array(ring)
    m = len(ring)
    v = np.vstack([ring + [0, -half_height, 0], ring + [0, half_height, 0],
                   [[0, -half_height, 0], [0, half_height, 0]]])
    cb, ct = 2 * m, 2 * m + 1
    f = []
    for i in range(m):
        j = (i + 1) % m
        f += [[i, j, m + j], [i, m + j, m + i], [cb, j, i], [ct, m + i, m + j]]
    return v, np.array(f)
