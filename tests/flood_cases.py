"""Known-answer cases of the interlocking-free mask.

The geometry and the asserted cells follow the reference's own flood-fill
tests (tests/test_flood_fill.py:331-453): a corner of an empty tray is free
and a cell where the object would clip is not, a full tray has no free cell,
a closed box keeps a small object out of its interior, a box with a 3x3 exit
lets a 2^3 object out but not a 4^3 one, an L-shaped object stays out of a
closed box, and a 3^3 object reaches the inside of a narrow corridor.  The
grids are built here from the sizes those tests state.

Each case is (name, tray, obj, expected) with expected a list of
((i, j, k), mask value), or the string "none" when no cell may be free.
"""
from __future__ import annotations

import numpy as np


def hollow_box(n: int, wall: int = 1) -> np.ndarray:
    """n^3 box of ones with a hollow n-2*wall interior (closed on all faces)."""
    g = np.ones((n, n, n), np.int32)
    if n > 2 * wall:
        g[wall:n - wall, wall:n - wall, wall:n - wall] = 0
    return g


def hollow_box_with_exit(n: int, wall: int = 1, exit_size: int = 1) -> np.ndarray:
    """hollow_box with a square exit_size hole through the high face of axis 1,
    centred on the other two axes."""
    g = hollow_box(n, wall)
    a = n // 2 - exit_size // 2
    g[a:a + exit_size, n - wall:, a:a + exit_size] = 0
    return g


def l_shape(n: int = 3) -> np.ndarray:
    """Staircase in the (0, 1) plane extruded along axis 2: cell (i, j, *) is
    set when j <= i."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return np.repeat((j <= i)[:, :, None], n, axis=2).astype(np.int32)


def cases():
    out = []
    t = np.zeros((20, 20, 20), np.int32)
    out.append(("empty tray, 3^3 object", t, np.ones((3, 3, 3), np.int32),
                [((0, 0, 0), 1), ((10, 10, 10), 1), ((19, 19, 19), 0)]))
    out.append(("full tray", np.ones((20, 20, 20), np.int32), np.ones((3, 3, 3), np.int32),
                "none"))
    t = np.zeros((20, 20, 20), np.int32)
    t[4:16, 4:16, 4:16] = hollow_box(12)
    out.append(("closed box, 2^3 object", t, np.ones((2, 2, 2), np.int32),
                [((9, 9, 9), 0), ((0, 0, 0), 1)]))
    t = np.zeros((20, 20, 20), np.int32)
    t[4:16, 4:16, 4:16] = hollow_box_with_exit(12, exit_size=3)
    out.append(("box with exit, 2^3 object", t, np.ones((2, 2, 2), np.int32),
                [((9, 9, 9), 1)]))
    out.append(("box with exit, 4^3 object", t, np.ones((4, 4, 4), np.int32),
                [((8, 8, 8), 0)]))
    t = np.zeros((24, 24, 24), np.int32)
    t[7:17, 7:17, 7:17] = hollow_box(10)
    out.append(("closed box, L object", t, l_shape(3), [((0, 0, 0), 1), ((11, 11, 11), 0)]))
    t = np.zeros((30, 30, 30), np.int32)
    t[5:25, 10:12, 5:25] = 1     # walls on both sides of the corridor
    t[5:25, 18:20, 5:25] = 1
    # back wall: the reference test fills it whole (its comments describe an
    # opening at i in [12, 17), which its first assignment closes again)
    t[5:25, 12:18, 22:24] = 1
    out.append(("narrow corridor, 3^3 object", t, np.ones((3, 3, 3), np.int32),
                [((14, 14, 10), 1), ((14, 14, 0), 1)]))
    return out


def check(mask: np.ndarray, expected) -> None:
    if isinstance(expected, str):
        assert expected == "none" and int(mask.sum()) == 0
        return
    for cell, v in expected:
        assert int(mask[cell]) == v, (cell, int(mask[cell]), v)
