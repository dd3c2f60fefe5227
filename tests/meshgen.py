"""Triangle meshes for the voxelizer tests: consistently oriented closed meshes
(box, icosphere, torus), a single tilted triangle and a triangle soup."""
from __future__ import annotations

import numpy as np


def box(lo, size):
    """Axis-aligned box, 8 shared vertices, 12 outward triangles."""
    lo = np.asarray(lo, dtype=np.float64)
    size = np.asarray(size, dtype=np.float64)
    c = np.array([[(i >> 0) & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], dtype=np.float64)
    pos = (lo + c * size).astype(np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6),   # z = lo, z = hi
             (0, 1, 5, 4), (2, 6, 7, 3),   # y = lo, y = hi
             (0, 4, 6, 2), (1, 3, 7, 5)]   # x = lo, x = hi
    tris = []
    for a, b, cc, d in quads:
        tris += [(a, b, cc), (a, cc, d)]
    return pos, np.array(tris, dtype=np.int32)


def icosphere(subdiv: int = 2, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    t = (1.0 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    pos = np.array(verts) * radius + np.asarray(center)
    return pos.astype(np.float32), np.array(f, dtype=np.int32)


def torus(R: float = 2.0, r: float = 0.7, nu: int = 24, nv: int = 12):
    u = np.arange(nu) * 2 * np.pi / nu
    w = np.arange(nv) * 2 * np.pi / nv
    uu, ww = np.meshgrid(u, w, indexing="ij")
    pos = np.stack([(R + r * np.cos(ww)) * np.cos(uu), (R + r * np.cos(ww)) * np.sin(uu), r * np.sin(ww)], -1)
    pos = pos.reshape(-1, 3)
    tris = []
    for i in range(nu):
        for j in range(nv):
            a, b = i * nv + j, ((i + 1) % nu) * nv + j
            c, d = ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv
            tris += [(a, b, c), (a, c, d)]
    return pos.astype(np.float32), np.array(tris, dtype=np.int32)


def rotate(pos, seed: int, shift=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return (pos.astype(np.float64) @ q.T + np.asarray(shift)).astype(np.float32)


def soup(n: int, seed: int):
    """Unrelated random triangles (open, self-intersecting)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-3, 3, size=(3 * n, 3)).astype(np.float32)
    return pos, np.arange(3 * n, dtype=np.int32).reshape(n, 3)


def unshare(pos, tris):
    """Three fresh vertices per triangle, the layout an STL file carries."""
    p = np.asarray(pos)[np.asarray(tris)].reshape(-1, 3)
    return p.astype(np.float32), np.arange(len(p), dtype=np.int32).reshape(-1, 3)
