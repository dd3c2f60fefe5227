"""Independent pins of oracle rules that the first pins left open.

* collision_grid's out-of-bounds clamp max(c, 1) (packing.cpp:241-251),
  against a loop over offsets and item voxels;
* the interlocking-free seeds and exits (fft3.cu:300-343), against a brute
  force that moves the object one voxel at a time through offsets whose
  overlap it sums directly (nothing from the oracle's labels), and against
  the reference's own flood-fill cases (tests/test_flood_fill.py:331-453);
* the not-found bound of the 1-orientation path (bestVal = INF = 1e9,
  packing.cpp:626-637) and of the cached path (< 1e29, fft3.cu:1361);
* TrayState (cached spectra, circular correlation on P >= N) against the
  direct search and mask on random cases, including items with empty
  trailing planes and items as large as the tray.
"""
import itertools

import numpy as np
import pytest

from oracle import spk_oracle as O
from tests import flood_cases as F


def _collision_loops(tray, item):
    N = tray.shape
    vox = [tuple(int(v) for v in y) for y in np.argwhere(item != 0)]
    pos = [y for y in vox if item[y] > 0]
    hi = [max(y[d] for y in pos) for d in range(3)] if pos else [-10 ** 9] * 3
    out = np.zeros(N, np.int64)
    for q in itertools.product(*map(range, N)):
        c = 0
        for y in vox:
            z = (q[0] + y[0], q[1] + y[1], q[2] + y[2])
            if z[0] < N[0] and z[1] < N[1] and z[2] < N[2]:
                c += int(tray[z]) * int(item[y])
        if any(q[d] + hi[d] > N[d] - 1 for d in range(3)):
            c = max(c, 1)
        out[q] = c
    return out


def test_collision_grid_oob_clamp_brute_force():
    rng = np.random.default_rng(101)
    for _ in range(10):
        N = tuple(int(v) for v in rng.integers(3, 8, 3))
        tray = (rng.integers(1, 4, N) * (rng.random(N) < 0.4)).astype(np.int32)
        # items no larger than the tray (the reference pads the item to the
        # tray's size, voxelGrid.cpp:93-100, which is undefined past it)
        dims = tuple(int(rng.integers(1, min(5, n) + 1)) for n in N)
        item = (rng.integers(1, 3, dims) * (rng.random(dims) < 0.6)).astype(np.int32)
        item[-1] = 0     # an empty trailing plane: hi comes from the voxels, not the box
        np.testing.assert_array_equal(O.collision_grid(tray, item), _collision_loops(tray, item))


def test_collision_grid_oob_indicator_on_empty_tray():
    # an empty tray correlates to 0 everywhere: the grid is exactly the
    # out-of-bounds indicator of the item's occupied box (hi = (2, 1, 3))
    item = np.zeros((4, 3, 5), np.int32)
    item[2, 1, 3] = 1
    item[0, 0, 0] = 1
    g = O.collision_grid(np.zeros((6, 5, 7), np.int32), item)
    i, j, k = np.indices((6, 5, 7))
    exp = ((i + 2 > 5) | (j + 1 > 4) | (k + 3 > 6)).astype(np.int32)
    np.testing.assert_array_equal(g, exp)
    # and an empty item (hi = -INF) has no out-of-bounds offset
    assert not O.collision_grid(np.zeros((4, 4, 4), np.int32), np.zeros((2, 2, 2), np.int32)).any()


def test_collision_grid_reference_cases():
    # tests/test_core.py:158-180 and tests/test_bindings.py:206-214
    a = np.zeros((4, 1, 4), np.int32)
    for p in [(0, 0, 1), (1, 0, 0), (2, 0, 0), (2, 0, 1), (2, 0, 3)]:
        a[p] = 1
    b = np.zeros((4, 1, 4), np.int32)
    for p in [(0, 0, 0), (0, 0, 1), (0, 0, 2), (1, 0, 1)]:
        b[p] = 1
    assert O.collision_grid(a, b)[0, 0, 0] > 0
    assert (O.collision_grid(np.zeros((10, 10, 10), np.int32), np.ones((3, 3, 3), np.int32)) == 0).sum() > 0
    assert (O.collision_grid(np.ones((10, 10, 10), np.int32), np.ones((3, 3, 3), np.int32)) > 0).all()


def _if_slide(tray, obj):
    """Interlocking-free offsets by moving the object voxel by voxel.

    An offset is free when the object lies inside the tray there and the
    ID-weighted overlap sum, taken voxel by voxel, is zero.  The object can
    leave the tray from a free offset that touches a face: offset 0 on an
    axis, or the largest offset at which its occupied box still fits.  The
    mask holds every free offset from which a chain of free one-voxel moves
    reaches such an offset.
    """
    N = tray.shape
    vox = [tuple(int(v) for v in y) for y in np.argwhere(obj != 0)]
    if not vox:
        return np.ones(N, np.int32)
    top = [max(y[d] for y in vox) for d in range(3)]

    def free(q):
        if any(q[d] < 0 or q[d] + top[d] >= N[d] for d in range(3)):
            return False
        return sum(int(tray[q[0] + y[0], q[1] + y[1], q[2] + y[2]]) * int(obj[y]) for y in vox) == 0

    fr = {q for q in itertools.product(*map(range, N)) if free(q)}
    out = np.zeros(N, np.int32)
    stack = [q for q in fr if any(q[d] == 0 or q[d] == N[d] - 1 - top[d] for d in range(3))]
    for q in stack:
        out[q] = 1
    while stack:
        q = stack.pop()
        for d in range(3):
            for s in (-1, 1):
                r = list(q)
                r[d] += s
                r = tuple(r)
                if r in fr and not out[r]:
                    out[r] = 1
                    stack.append(r)
    return out


def test_interlocking_mask_matches_sliding_brute_force():
    rng = np.random.default_rng(103)
    for trial in range(12):
        N = tuple(int(v) for v in rng.integers(5, 10, 3))
        tray = (rng.integers(1, 3, N) * (rng.random(N) < rng.uniform(0.1, 0.35))).astype(np.int32)
        dims = tuple(int(v) for v in rng.integers(1, 4, 3))
        obj = (rng.random(dims) < 0.7).astype(np.int32)
        if trial % 3 == 0:
            obj[:] = 0
            obj[0, 0, 0] = 1
            obj[-1, -1, -1] = 1   # two corners: a sparse object larger than 1^3
        exp = _if_slide(tray, obj)
        np.testing.assert_array_equal(O.interlocking_free_positions(tray, obj), exp)
        np.testing.assert_array_equal(O.interlocking_free_positions_bfs(tray, obj), exp)
        st = O.TrayState(tray, O.distance_field(tray))
        np.testing.assert_array_equal(O.interlocking_mask_state(st, obj), exp)


def test_interlocking_cage_with_large_objects():
    # a 9^3 cage with a 5^3 cavity and a 1-voxel window in its +x wall: a 1^3
    # object escapes through the window, a 2^3 one is trapped, and a 2x1x1
    # one escapes only when oriented along the window's axis
    cage = np.ones((9, 9, 9), np.int32)
    cage[2:7, 2:7, 2:7] = 0
    cage[7:9, 4, 4] = 0
    tray = np.zeros((13, 13, 13), np.int32)
    tray[2:11, 2:11, 2:11] = cage
    for obj, inside in ((np.ones((1, 1, 1), np.int32), 1), (np.ones((2, 2, 2), np.int32), 0),
                        (np.ones((2, 1, 1), np.int32), 1), (np.ones((1, 2, 1), np.int32), 0)):
        exp = _if_slide(tray, obj)
        assert exp[6, 6, 6] == inside and exp[0, 0, 0] == 1
        np.testing.assert_array_equal(O.interlocking_free_positions(tray, obj), exp)


@pytest.mark.parametrize("case", F.cases(), ids=lambda c: c[0])
def test_interlocking_reference_cases(case):
    name, tray, obj, expected = case
    m = O.interlocking_free_positions(tray, obj)
    F.check(m, expected)
    np.testing.assert_array_equal(O.interlocking_free_positions_bfs(tray, obj), m)


def test_interlocking_reference_cases_brute_force_small():
    # the three 20^3 cases again through the sliding brute force
    for name, tray, obj, expected in F.cases()[:3]:
        exp = _if_slide(tray, obj)
        F.check(exp, expected)
        np.testing.assert_array_equal(O.interlocking_free_positions(tray, obj), exp)


def test_not_found_bounds():
    # proximity 1e9 / 4 per voxel: an item of 4 voxels scores >= 1e9, which
    # the 1-orientation path rejects (bestVal = INF) and the cached path takes
    tray = np.zeros((6, 5, 4), np.int32)
    item = np.ones((2, 2, 1), np.int32)
    phi = np.full(tray.shape, 250_000_000, np.int32)
    assert not O.search(item, tray, phi, penalty="pow").found
    r = O.search(item, tray, phi, penalty="cube")
    assert r.found and r.score == 1e9 and r.position == (0, 0, 0)
    phi[:] = 249_999_999
    assert O.search(item, tray, phi, penalty="pow").found
    assert O.accept_bound("pow") == 1e9 and O.accept_bound("cube") == 1e29


def test_tray_state_matches_direct_search():
    rng = np.random.default_rng(105)
    for trial in range(10):
        N = tuple(int(v) for v in rng.integers(4, 13, 3))
        tray = (rng.integers(1, 5, N) * (rng.random(N) < 0.3)).astype(np.int32)
        phi = O.distance_field(tray, ["l1", "edt"][trial % 2], ["eq1", "nonzero"][trial % 2])
        st = O.TrayState(tray, phi)
        dims = [tuple(int(v) for v in rng.integers(1, 5, 3)) for _ in range(3)] + [N]
        for d in dims:
            item = (rng.random(d) < 0.6).astype(np.int32)
            if trial % 3 == 1:
                item[..., -1] = 0          # an empty trailing plane
            for pen in ("pow", "cube"):
                assert O.search_state(st, item, pen) == O.search(item, tray, phi, pen)
            m = O.interlocking_free_positions(tray, item)
            assert O.search_state(st, item, "cube", m) == O.search(item, tray, phi, "cube", m)
        assert O.search_state(st, np.zeros((2, 2, 2), np.int32), "cube") == \
            O.search(np.zeros((2, 2, 2), np.int32), tray, phi, "cube")
