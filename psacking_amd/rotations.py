"""Orientations of voxel grids (reference: spectral_packer/rotations.py).

The list order is the reference's: 6 "up" faces (identity, RX, RX^2, RX^3, RY,
RY^3) times 4 z-turns (rotations.py:159-219); the 4-set is the z-turns, the
6-set the face bases (rotations.py:222-268).  A 90-degree turn is
``np.rot90(g, 1, axes)`` with axes (1, 2) about x, (2, 0) about y and (0, 1)
about z (rotations.py:141-156).

For the device every orientation is also an index map (perm, flip): output
axis ``a`` runs along base axis ``perm[a]``, reversed when ``flip[a]``.  The
maps are derived once from the same rot90 compositions applied to coordinate
grids, so the device reads the base item with permuted indices instead of
receiving 24 rotated copies.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np


def rotate_90_x(grid: np.ndarray) -> np.ndarray:
    return np.rot90(grid, k=1, axes=(1, 2))


def rotate_90_y(grid: np.ndarray) -> np.ndarray:
    return np.rot90(grid, k=1, axes=(2, 0))


def rotate_90_z(grid: np.ndarray) -> np.ndarray:
    return np.rot90(grid, k=1, axes=(0, 1))


def _face_bases(grid: np.ndarray) -> List[np.ndarray]:
    x = rotate_90_x
    y = rotate_90_y
    return [grid, x(grid), x(x(grid)), x(x(x(grid))), y(grid), y(y(y(grid)))]


def get_24_orientations(grid: np.ndarray) -> List[np.ndarray]:
    out = []
    for base in _face_bases(grid):
        g = base.copy()
        for _ in range(4):
            out.append(g.copy())
            g = rotate_90_z(g)
    return out


def get_orientations(grid: np.ndarray, num_orientations: int = 1) -> List[np.ndarray]:
    """Orientation list of the reference (views/copies, same order)."""
    if num_orientations == 1:
        return [grid]
    if num_orientations == 4:
        out, g = [], grid.copy()
        for _ in range(4):
            out.append(g.copy())
            g = rotate_90_z(g)
        return out
    if num_orientations == 6:
        return _face_bases(grid)
    if num_orientations == 24:
        return get_24_orientations(grid)
    raise ValueError(f"num_orientations must be 1, 4, 6, or 24, got {num_orientations}")


def make_contiguous(grid: np.ndarray) -> np.ndarray:
    if not grid.flags["C_CONTIGUOUS"]:
        return np.ascontiguousarray(grid)
    return grid


def _index_map(rotated_coords: List[np.ndarray]) -> Tuple[Tuple[int, int, int], Tuple[int, int, int]]:
    """(perm, flip) from the base coordinates seen at each rotated position."""
    perm, flip = [0, 0, 0], [0, 0, 0]
    shape = rotated_coords[0].shape
    for a in range(3):
        idx0 = [0, 0, 0]
        idx1 = [0, 0, 0]
        idx1[a] = 1
        for d in range(3):
            c0 = rotated_coords[d][tuple(idx0)]
            c1 = rotated_coords[d][tuple(idx1)] if shape[a] > 1 else c0
            if c1 != c0:
                perm[a] = d
                flip[a] = 1 if c1 < c0 else 0
    return tuple(perm), tuple(flip)


def _build_maps(num_orientations: int):
    base_shape = (2, 3, 4)   # distinct extents so every axis is identifiable
    coords = np.indices(base_shape)
    per_axis = [get_orientations(coords[d], num_orientations) for d in range(3)]
    maps = []
    for o in range(num_orientations):
        maps.append(_index_map([per_axis[d][o] for d in range(3)]))
    return maps


_MAPS = {n: _build_maps(n) for n in (1, 4, 6, 24)}


def orientation_maps(num_orientations: int) -> Tuple[np.ndarray, np.ndarray]:
    """(perms, flips) int32 arrays of shape (n, 3) in the reference order."""
    if num_orientations not in _MAPS:
        raise ValueError(f"num_orientations must be 1, 4, 6, or 24, got {num_orientations}")
    m = _MAPS[num_orientations]
    perms = np.array([p for p, _ in m], dtype=np.int32)
    flips = np.array([f for _, f in m], dtype=np.int32)
    return perms, flips


def apply_map(grid: np.ndarray, perm, flip) -> np.ndarray:
    """Host restatement of the device index map (for tests and placement)."""
    g = np.transpose(np.asarray(grid), axes=tuple(int(p) for p in perm))
    sl = tuple(slice(None, None, -1) if f else slice(None) for f in flip)
    return np.ascontiguousarray(g[sl])
