"""Seeded synthetic workloads for tests and the benchmark.

  boxes      64^3 tray, 20 solid boxes with sides in [3, 12], value 1
  blobs      100^3 tray, 60 unions of 2-5 ellipsoids (semi-axes [3, 10],
             centres in a 30^3 box), cropped to their bounding box
  readme     240x123x100 tray, 350 unions of 3-6 ellipsoids/capsules with
             bounding boxes of 20-50 voxels and fill fraction in [0.30, 0.40]
  uneven     one of 48 independent instances: tray sides log-uniform in
             [48, 256] (z <= 128), 30-200 blob items sized 8-35 % of the
             tray's smallest side

All items are int32 0/1 grids; generation uses numpy.random.default_rng(seed).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple

import numpy as np


@dataclass
class Workload:
    name: str
    tray_size: Tuple[int, int, int]
    items: List[np.ndarray]
    num_orientations: int
    seed: int


def _crop(g: np.ndarray) -> np.ndarray:
    idx = np.argwhere(g)
    if len(idx) == 0:
        return g.astype(np.int32)
    lo, hi = idx.min(axis=0), idx.max(axis=0) + 1
    return np.ascontiguousarray(g[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]].astype(np.int32))


def _ellipsoid(shape, centre, axes):
    i, j, k = np.ogrid[:shape[0], :shape[1], :shape[2]]
    return (((i - centre[0]) / axes[0]) ** 2 + ((j - centre[1]) / axes[1]) ** 2 +
            ((k - centre[2]) / axes[2]) ** 2) <= 1.0


def _capsule(shape, a, b, r):
    i, j, k = np.ogrid[:shape[0], :shape[1], :shape[2]]
    p = [i - a[0], j - a[1], k - a[2]]
    d = np.asarray(b, float) - np.asarray(a, float)
    L2 = float(d @ d) or 1.0
    t = np.clip((p[0] * d[0] + p[1] * d[1] + p[2] * d[2]) / L2, 0.0, 1.0)
    q = [p[n] - t * d[n] for n in range(3)]
    return (q[0] ** 2 + q[1] ** 2 + q[2] ** 2) <= r * r


def boxes(seed: int = 1, n: int = 20, tray=(64, 64, 64)) -> Workload:
    rng = np.random.default_rng(seed)
    items = [np.ones(tuple(int(s) for s in rng.integers(3, 13, 3)), dtype=np.int32)
             for _ in range(n)]
    return Workload("boxes", tuple(tray), items, 1, seed)


def blob(rng, n_lo=2, n_hi=5, ax_lo=3.0, ax_hi=10.0, box=30) -> np.ndarray:
    shape = (box + 2 * int(np.ceil(ax_hi)),) * 3
    g = np.zeros(shape, dtype=bool)
    off = int(np.ceil(ax_hi))
    for _ in range(int(rng.integers(n_lo, n_hi + 1))):
        c = off + rng.uniform(0, box, 3)
        a = rng.uniform(ax_lo, ax_hi, 3)
        g |= _ellipsoid(shape, c, a)
    return _crop(g)


def blobs(seed: int = 2, n: int = 60, tray=(100, 100, 100)) -> Workload:
    rng = np.random.default_rng(seed)
    return Workload("blobs", tuple(tray), [blob(rng) for _ in range(n)], 1, seed)


def readme_item(rng, lo=20, hi=50, fill=(0.30, 0.40)) -> np.ndarray:
    while True:
        dims = rng.integers(lo, hi + 1, 3)
        shape = tuple(int(d) for d in dims)
        g = np.zeros(shape, dtype=bool)
        for _ in range(int(rng.integers(3, 7))):
            if rng.random() < 0.5:
                c = rng.uniform(0.2, 0.8, 3) * dims
                a = rng.uniform(0.15, 0.45, 3) * dims
                g |= _ellipsoid(shape, c, a)
            else:
                a = rng.uniform(0.1, 0.9, 3) * dims
                b = rng.uniform(0.1, 0.9, 3) * dims
                r = rng.uniform(0.08, 0.2) * dims.min()
                g |= _capsule(shape, a, b, r)
        c = _crop(g)
        if min(c.shape) < lo or max(c.shape) > hi:
            continue
        f = c.sum() / c.size
        if fill[0] <= f <= fill[1]:
            return c


def readme(seed: int = 3, n: int = 350, tray=(240, 123, 100)) -> Workload:
    rng = np.random.default_rng(seed)
    return Workload("readme", tuple(tray), [readme_item(rng) for _ in range(n)], 24, seed)


def uneven_instance(instance: int, base_seed: int = 4) -> Workload:
    rng = np.random.default_rng(base_seed + instance)
    lo, hi = np.log(48), np.log(256)
    tray = [int(round(np.exp(rng.uniform(lo, hi)))) for _ in range(3)]
    tray[2] = min(tray[2], 128)
    n = int(rng.integers(30, 201))
    smin = min(tray)
    items = []
    for _ in range(n):
        size = rng.uniform(0.08, 0.35) * smin
        ax = max(1.5, size / 4)
        items.append(blob(rng, 2, 5, ax * 0.6, ax, box=max(2, int(size / 2))))
    return Workload(f"uneven{instance}", tuple(tray), items, 24, base_seed + instance)


def by_name(name: str) -> Workload:
    return {"boxes": boxes, "blobs": blobs, "readme": readme}[name]()
