"""Pins for the CPU oracle: reference KATs, brute force and closed forms."""
import itertools

import numpy as np
import pytest

from oracle import spk_oracle as O
from tests.conftest import grid_from_rows, piece


def _arr(x):
    return np.array(x, dtype=np.int32)


def test_conv_corr_kats(kats):
    for case in kats["dft_conv3"]:
        np.testing.assert_array_equal(O.dft_conv3(_arr(case["a"]), _arr(case["b"])), _arr(case["expected"]))
    for case in kats["dft_corr3"]:
        np.testing.assert_array_equal(O.dft_corr3(_arr(case["a"]), _arr(case["b"])), _arr(case["expected"]))


def test_distance_kats(kats):
    for case in kats["calculate_distance"]:
        g = np.zeros(case["shape"], dtype=np.int32)
        for p in case["ones"]:
            g[tuple(p)] = 1
        np.testing.assert_array_equal(O.calculate_distance(g), _arr(case["expected"]))
    for case in kats["calculate_distance_partial"]:
        g = np.zeros(case["shape"], dtype=np.int32)
        for p in case["ones"]:
            g[tuple(p)] = 1
        d = O.calculate_distance(g)
        assert d[tuple(case["center"])] == case["center_value"]
        assert d[0, 0, 0] >= case["corner_min"] and d[-1, -1, -1] >= case["corner_min"]


def test_bounds_tight_kats(kats):
    for case in kats["get_bounds"]:
        g = np.zeros(case["grid_shape"], dtype=np.int32)
        if "box" in case:
            b = case["box"]
            g[b[0][0]:b[0][1], b[1][0]:b[1][1], b[2][0]:b[2][1]] = 1
        for p in case.get("ones", []):
            g[tuple(p)] = 1
        lo, hi = O.get_bounds(g)
        assert lo == tuple(case["lo"]) and hi == tuple(case["hi"])
    for case in kats["make_tight"]:
        g = np.zeros(case["grid_shape"], dtype=np.int32)
        b = case["box"]
        g[b[0][0]:b[0][1], b[1][0]:b[1][1], b[2][0]:b[2][1]] = 1
        t = O.make_tight(g)
        assert t.shape == tuple(case["shape"]) and t.sum() == g.sum()
    with pytest.raises(ValueError):
        O.make_tight(np.zeros((3, 3, 3), np.int32))
    assert O.get_bounds(np.zeros((2, 2, 2), np.int32)) == ((O.INF_BOUND,) * 3, (-O.INF_BOUND,) * 3)


def _corr_loops(a, b):
    """Quadruple-loop definition r[q] = sum_y a[q+y] b[y] (tiny inputs)."""
    N = a.shape
    r = np.zeros(N, dtype=np.int64)
    for q in itertools.product(*map(range, N)):
        s = 0
        for y in itertools.product(*map(range, b.shape)):
            z = tuple(q[d] + y[d] for d in range(3))
            if all(z[d] < N[d] for d in range(3)):
                s += int(a[z]) * int(b[y])
        r[q] = s
    return r


def test_corr_matches_definition_and_fft():
    rng = np.random.default_rng(7)
    for _ in range(6):
        N = tuple(rng.integers(1, 6, 3))
        a = rng.integers(-3, 9, N)
        b = rng.integers(0, 4, tuple(rng.integers(1, 4, 3)))
        bp = O.pad_to(b, N)
        ref = _corr_loops(a, bp)
        np.testing.assert_array_equal(O.corr_bruteforce(a, bp), ref)
        np.testing.assert_array_equal(O._fft_exact_linear(a, bp, a.shape, corr=True), ref)
    # larger: FFT path against brute force
    a = rng.integers(0, 500, (40, 33, 29))
    b = (rng.random((9, 7, 11)) < 0.4).astype(np.int64)
    bp = O.pad_to(b, a.shape)
    np.testing.assert_array_equal(O._fft_exact_linear(a, bp, a.shape, True), O.corr_bruteforce(a, bp))


def test_conv_is_flipped_corr():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 5, (4, 3, 5))
    b = rng.integers(0, 5, (4, 3, 5))
    # conv[q] = sum_{y<=q} a[q-y] b[y]
    N = a.shape
    ref = np.zeros(N, np.int64)
    for q in itertools.product(*map(range, N)):
        for y in itertools.product(*map(range, N)):
            if all(y[d] <= q[d] for d in range(3)):
                ref[q] += a[tuple(q[d] - y[d] for d in range(3))] * b[y]
    np.testing.assert_array_equal(O.dft_conv3(a, b), ref)


def test_distance_l1_matches_bfs():
    rng = np.random.default_rng(11)
    for _ in range(5):
        N = tuple(rng.integers(1, 9, 3))
        g = rng.integers(0, 4, N) * (rng.random(N) < 0.15)
        for src in ("eq1", "nonzero"):
            np.testing.assert_array_equal(O.distance_field(g, "l1", src), O.distance_bfs(g, src))
    g = np.zeros((5, 6, 7), np.int32)
    assert (O.calculate_distance(g) == 5 + 6 + 7 + 10).all()
    g[0, 0, 0] = 2   # value 2 is not a source (fft3.cu:948)
    assert (O.calculate_distance(g) == 28).all()


def test_distance_edt_brute_force():
    rng = np.random.default_rng(5)
    N = (7, 6, 9)
    g = (rng.random(N) < 0.05).astype(np.int32)
    g[3, 3, 3] = 1
    src = np.argwhere(g == 1)
    exp = np.zeros(N, np.int64)
    for q in itertools.product(*map(range, N)):
        d2 = ((src - np.array(q)) ** 2).sum(axis=1).min()
        exp[q] = int(np.floor(np.sqrt(d2) + 0.5))
    np.testing.assert_array_equal(O.distance_field(g, "edt", "eq1"), exp)


def test_penalty_forms():
    p = O.penalty_table(100, "pow")
    c = O.penalty_table(100, "cube")
    assert p[0] == 0 and c[0] == 0
    assert abs(p[50] - 1e8 * 0.125) < 1e-6 and abs(c[50] - 1e8 * 0.125) < 1e-6
    assert np.all(np.diff(p) > 0)


def test_search_brute_force():
    rng = np.random.default_rng(21)
    for trial in range(8):
        N = tuple(rng.integers(3, 8, 3))
        tray = rng.integers(0, 3, N) * (rng.random(N) < 0.3)
        item = (rng.random(tuple(rng.integers(1, 4, 3))) < 0.7).astype(np.int32)
        phi = O.calculate_distance(tray)
        r = O.search(item, tray, phi, penalty="cube")
        # brute force over q
        best = (np.inf, None)
        _, hi = O.get_bounds(item)
        pen = O.penalty_table(N[2], "cube")
        for q in itertools.product(*map(range, N)):
            if any(q[d] + hi[d] > N[d] - 1 for d in range(3)):
                continue
            c = p = 0
            for y in np.argwhere(item != 0):
                z = tuple(q[d] + y[d] for d in range(3))
                if all(z[d] < N[d] for d in range(3)):
                    c += tray[z] * item[tuple(y)]
                    p += phi[z] * item[tuple(y)]
            if c != 0:
                continue
            s = float(p) + pen[q[2]]
            if s < best[0]:
                best = (s, q)
        if best[1] is None:
            assert not r.found
        else:
            assert r.found and r.position == best[1] and r.score == best[0]


def test_place_in_tray():
    t = np.zeros((4, 4, 4), np.int32)
    it = np.ones((2, 2, 2), np.int32)
    t2 = O.place_in_tray(it, t, (1, 1, 1), 5)
    assert t2.sum() == 40 and t2[1, 1, 1] == 5 and t.sum() == 0
    with pytest.raises(RuntimeError):
        O.place_in_tray(it, t2, (2, 2, 2), 6)
    t3 = O.place_in_tray(it, t2, (3, 3, 3), 6)   # clipped at the tray edge
    assert t3[3, 3, 3] == 6 and (t3 == 6).sum() == 1


def test_flood_fill_label_matches_bfs():
    rng = np.random.default_rng(9)
    for _ in range(6):
        N = tuple(rng.integers(4, 10, 3))
        tray = (rng.random(N) < 0.35).astype(np.int32)
        obj = (rng.random(tuple(rng.integers(1, 3, 3))) < 0.8).astype(np.int32)
        np.testing.assert_array_equal(O.interlocking_free_positions(tray, obj),
                                      O.interlocking_free_positions_bfs(tray, obj))


def test_flood_fill_cage():
    # closed 7^3 box with a 5^3 cavity: a 1-voxel object inside is blocked
    tray = np.ones((7, 7, 7), np.int32)
    tray[1:6, 1:6, 1:6] = 0
    big = np.zeros((11, 11, 11), np.int32)
    big[2:9, 2:9, 2:9] = tray
    m = O.interlocking_free_positions(big, np.ones((1, 1, 1), np.int32))
    assert m[5, 5, 5] == 0 and m[0, 0, 0] == 1
    empty = O.interlocking_free_positions(big, np.zeros((2, 2, 2), np.int32))
    assert empty.shape == big.shape


def test_tetromino_fixture(kats):
    shapes = kats["tetrominoes"]
    for run in kats["tetromino_runs"]:
        items = [piece(shapes[c]) for c in run["order"]]
        r = O.pack_voxels(items, tuple(run["tray"]), num_orientations=run["num_orientations"],
                          sort_by_volume=False)
        assert r.num_placed == 7
        assert abs(r.density - run["density"]) < 1e-12
        assert [tuple(p.position) for p in r.placements] == [tuple(p) for p in run["positions"]]
        assert [p.orientation_index for p in r.placements] == run["orientations"]
        exp = grid_from_rows(run["grid_cols_0_8"], (8, 9, 1))
        np.testing.assert_array_equal(r.tray[:, :9, :], exp)
        assert (r.tray[:, 9:, :] == 0).all()


def test_tetromino_sort_order_is_reference_argsort():
    # pieces dict order I,O,T,S,Z,J,L; np.argsort(volumes)[::-1] gives L..I
    vols = [4] * 7
    order = list(np.argsort(vols)[::-1])
    assert "".join("IOTSZJL"[i] for i in order) == "LJZSTOI"


def test_orientation_counts(kats):
    for case in kats["orientation_counts"]:
        item = np.ones(case["item_shape"], np.int32)
        r1 = O.pack_voxels([item], tuple(case["tray"]), num_orientations=1)
        r24 = O.pack_voxels([item], tuple(case["tray"]), num_orientations=24)
        assert r1.num_placed == case["placed_1"] and r24.num_placed == case["placed_24"]


def test_orientations_are_distinct_rotations():
    g = np.arange(2 * 3 * 4).reshape(2, 3, 4)
    outs = O.get_orientations(g, 24)
    assert len(outs) == 24
    keys = {(o.shape, o.tobytes()) for o in outs}
    assert len(keys) == 24
    assert np.array_equal(outs[0], g)
    # 4-set equals the first four of the 24-set; 6-set are the face bases
    for a, b in zip(O.get_orientations(g, 4), outs[:4]):
        assert np.array_equal(a, b)
    for a, b in zip(O.get_orientations(g, 6), outs[::4]):
        assert np.array_equal(a, b)


def test_reference_cpu_path_as_written_matches_oracle():
    from oracle import ref_cpu_path as R
    rng = np.random.default_rng(31)
    for _ in range(4):
        N = tuple(int(v) for v in rng.integers(4, 14, 3))
        tray = (rng.integers(1, 3, N) * (rng.random(N) < 0.3)).astype(np.int32)
        item = (rng.random(tuple(int(v) for v in rng.integers(1, 4, 3))) < 0.7).astype(np.int32)
        exp = O.search(item, tray, penalty="pow")
        got = R.search_as_written(item, tray)
        assert got == (exp.position, exp.found, exp.score)
        np.testing.assert_array_equal(R.corr3_as_written(tray, O.pad_to(item, N)),
                                      O.corr_exact(tray, O.pad_to(item, N)))


@pytest.mark.parametrize("metric", ["edt", "l1"])
def test_free_space_certificate_is_sound(metric):
    """The device skips an item when the tray's free-ball bound, the maximum
    over voxels of min(distance to the nearest occupied voxel, distance to the
    nearest voxel outside the tray), is below the item's inscribed-ball bound
    (largest distance from an item voxel to a voxel outside it).  Check on
    random trays and items that the rule never rejects an item with a feasible
    orientation, and that it does reject some."""
    rng = np.random.default_rng(77)
    certified = 0
    for case in range(60):
        shape = tuple(int(v) for v in rng.integers(10, 16, 3))
        tray = (rng.random(shape) < rng.uniform(0.05, 0.4)).astype(np.int32)
        dims = tuple(int(v) for v in rng.integers(5, 10, 3))
        item = np.ones(dims, np.int32)
        # a few holes on the faces keep the items irregular
        item[0][rng.random(item[0].shape) < 0.3] = 0
        comp = np.ones(tuple(d + 2 for d in dims), np.int32)
        comp[1:-1, 1:-1, 1:-1] = (item == 0)
        theta = int(O.distance_field(comp, metric, "nonzero").max())
        phi = O.distance_field(tray, metric, "nonzero")
        wall = np.minimum.reduce(np.meshgrid(*[np.minimum(np.arange(n) + 1, n - np.arange(n))
                                               for n in shape], indexing="ij"))
        psi_max = int(np.minimum(phi, wall).max())
        if psi_max >= theta:
            continue
        certified += 1
        for g in O.get_orientations(item, 24):
            g = O.as_grid(g)
            if any(g.shape[a] > shape[a] for a in range(3)):
                continue
            assert not O.search(g, tray, O.distance_field(tray, metric, "nonzero"),
                                penalty="cube").found, (case, metric)
    assert certified >= 10


def test_column_infeasibility_is_sound_and_stronger():
    """The column test (2x2 column blocks at full z resolution, x-y eroded
    core) never rejects an orientation with a collision-free offset, rejects
    some infeasible ones, and rejects everything the 2x2x2-block test does."""
    rng = np.random.default_rng(93)
    proved = cube_proved = infeasible = 0
    for case in range(140):
        shape = tuple(int(v) for v in rng.integers(9, 15, 3))
        tray = (rng.random(shape) < rng.uniform(0.02, 0.25)).astype(np.int32)
        tray[:, :, :int(rng.integers(0, shape[2] - 3))] = 2
        dims = tuple(int(v) for v in rng.integers(3, 9, 3))
        item = np.ones(dims, np.int32)
        item[rng.random(dims) < 0.1] = 0
        for g in O.get_orientations(item, 24)[:6]:
            g = O.as_grid(g)
            if any(g.shape[a] > shape[a] for a in range(3)):
                continue
            c = O.column_infeasible(g, tray)
            b = O.coarse_infeasible(g, tray)
            found = O.search(g, tray, O.distance_field(tray, "edt", "nonzero"), penalty="cube").found
            assert not (c and found), case
            assert c or not b, case
            infeasible += not found
            proved += c
            cube_proved += b
    assert proved > cube_proved >= 10 and infeasible > proved


def test_coarse_infeasibility_is_sound():
    """The coarse test (2x2x2 blocks: occupied-any tray, eroded item core)
    never rejects an orientation that has a collision-free offset, and it
    does reject some infeasible ones."""
    rng = np.random.default_rng(91)
    proved = infeasible = 0
    for case in range(140):
        shape = tuple(int(v) for v in rng.integers(9, 15, 3))
        tray = (rng.random(shape) < rng.uniform(0.02, 0.25)).astype(np.int32)
        # a solid slab from the floor leaves a free layer of random depth
        tray[:, :, :int(rng.integers(0, shape[2] - 3))] = 2
        dims = tuple(int(v) for v in rng.integers(3, 9, 3))
        item = np.ones(dims, np.int32)
        item[rng.random(dims) < 0.1] = 0
        for g in O.get_orientations(item, 24)[:6]:
            g = O.as_grid(g)
            if any(g.shape[a] > shape[a] for a in range(3)):
                continue
            c = O.coarse_infeasible(g, tray)
            found = O.search(g, tray, O.distance_field(tray, "edt", "nonzero"), penalty="cube").found
            assert not (c and found), case
            infeasible += not found
            proved += c
    assert proved >= 10 and infeasible > proved
