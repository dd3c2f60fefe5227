"""GPU parity: every _core entry point and the packer against the CPU oracle.

Integer outputs must be bit-exact; scores are doubles computed from exact
integers and must be equal.  Each test prints the case it runs first, so a
failing run names it.
"""
import numpy as np
import pytest

from oracle import spk_oracle as O
from tests.conftest import grid_from_rows, piece

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def P():
    import psacking_amd
    return psacking_amd


def rand_tray(rng, N, p=0.3, ids=3):
    return (rng.integers(1, ids + 1, N) * (rng.random(N) < p)).astype(np.int32)


def test_distance_kats(P, kats):
    say("distance kats")
    for case in kats["calculate_distance"]:
        g = np.zeros(case["shape"], dtype=np.int32)
        for p in case["ones"]:
            g[tuple(p)] = 1
        np.testing.assert_array_equal(P.calculate_distance(g), np.array(case["expected"]))


@pytest.mark.parametrize("metric,sources", [("l1", "eq1"), ("l1", "nonzero"), ("edt", "eq1"),
                                            ("edt", "nonzero")])
def test_distance_random(P, metric, sources):
    rng = np.random.default_rng(1)
    for N in [(1, 1, 1), (5, 1, 9), (17, 13, 11), (64, 64, 64), (100, 37, 53), (240, 123, 100)]:
        say("distance", metric, sources, N)
        g = rand_tray(rng, N, p=0.02)
        np.testing.assert_array_equal(P.distance_field(g, metric, sources),
                                      O.distance_field(g, metric, sources))
    g = np.zeros((7, 8, 9), np.int32)
    np.testing.assert_array_equal(P.distance_field(g, metric, sources), O.distance_field(g, metric, sources))


def test_correlation_kats(P, kats):
    say("correlation kats")
    for case in kats["dft_conv3"]:
        np.testing.assert_array_equal(P.dft_conv3(np.array(case["a"]), np.array(case["b"])),
                                      np.array(case["expected"]))
    for case in kats["dft_corr3"]:
        np.testing.assert_array_equal(P.dft_corr3(np.array(case["a"]), np.array(case["b"])),
                                      np.array(case["expected"]))
    with pytest.raises(RuntimeError):
        P.dft_corr3(np.ones((2, 2, 2)), np.ones((2, 2, 3)))


def test_correlation_random(P):
    rng = np.random.default_rng(2)
    for N in [(3, 4, 5), (16, 9, 21)]:
        say("correlation", N)
        a = rng.integers(-5, 9, N).astype(np.int32)
        b = rng.integers(-3, 4, N).astype(np.int32) * (rng.random(N) < 0.3)
        np.testing.assert_array_equal(P.dft_corr3(a, b), O.dft_corr3(a, b))
        np.testing.assert_array_equal(P.dft_conv3(a, b), O.dft_conv3(a, b))


def test_collision_grid(P):
    rng = np.random.default_rng(3)
    for N, I in [((10, 10, 10), (3, 3, 3)), ((20, 15, 12), (4, 6, 2)), ((8, 8, 8), (2, 9, 1))]:
        say("collision_grid", N, I)
        tray = rand_tray(rng, N, 0.2, 5)
        item = (rng.random(I) < 0.7).astype(np.int32)
        np.testing.assert_array_equal(P.collision_grid(tray, item), O.collision_grid(tray, item))


def test_interlocking_mask(P):
    rng = np.random.default_rng(4)
    # bit-column fill (z <= 128), z past 128 on the byte fill, and a tray
    # wider than one 32 x 32-column tile of the bit fill
    for N, I in [((12, 11, 10), (2, 2, 2)), ((20, 20, 20), (3, 1, 2)), ((9, 9, 9), (1, 1, 1)),
                 ((7, 6, 140), (2, 2, 3)), ((70, 40, 9), (3, 2, 2))]:
        say("interlock", N, I)
        tray = (rng.random(N) < 0.3).astype(np.int32)
        obj = (rng.random(I) < 0.8).astype(np.int32)
        np.testing.assert_array_equal(P.interlocking_free_positions(tray, obj),
                                      O.interlocking_free_positions(tray, obj))
    cage = np.zeros((15, 15, 15), np.int32)
    cage[3:12, 3:12, 3:12] = 1
    cage[4:11, 4:11, 4:11] = 0
    say("interlock cage")
    m = P.interlocking_free_positions(cage, np.ones((1, 1, 1), np.int32))
    np.testing.assert_array_equal(m, O.interlocking_free_positions(cage, np.ones((1, 1, 1))))
    assert m[7, 7, 7] == 0 and m[0, 0, 0] == 1


def test_place_in_tray(P):
    say("place_in_tray")
    rng = np.random.default_rng(5)
    tray = rand_tray(rng, (10, 11, 12), 0.1)
    item = (rng.random((3, 4, 2)) < 0.6).astype(np.int32)
    for pos in [(0, 0, 0), (8, 9, 11), (-1, 2, 3), (5, 5, 5)]:
        try:
            exp = O.place_in_tray(item, tray, pos, 9)
        except RuntimeError:
            with pytest.raises(RuntimeError):
                P.place_in_tray(item, tray, pos, 9)
            continue
        np.testing.assert_array_equal(P.place_in_tray(item, tray, pos, 9), exp)
    np.testing.assert_array_equal(P.place_in_tray(item, np.zeros_like(tray), (1, 1, 1)),
                                  O.place_in_tray(item, np.zeros_like(tray), (1, 1, 1), 1))


def test_search_single_random(P):
    rng = np.random.default_rng(6)
    for t in range(12):
        N = tuple(int(v) for v in rng.integers(4, 30, 3))
        I = tuple(int(min(v, n)) for v, n in zip(rng.integers(1, 8, 3), N))
        tray = rand_tray(rng, N, 0.25, 3)
        item = (rng.random(I) < 0.7).astype(np.int32)
        say("fft_search_placement", t, N, I)
        r = O.search(item, tray, penalty="pow")
        got = P.fft_search_placement(item, tray)
        assert got == (r.position, r.found, r.score)


def test_search_empty_and_oversize(P):
    say("search edge cases")
    tray = np.zeros((6, 7, 8), np.int32)
    assert P.fft_search_placement(np.ones((7, 1, 1), np.int32), tray) == ((-1, -1, -1), False, 0.0)
    assert P.fft_search_placement(np.zeros((2, 2, 2), np.int32), tray) == ((0, 0, 0), True, 0.0)
    assert P.fft_search_placement(np.ones((1, 1, 1), np.int32), np.ones((3, 3, 3), np.int32)) == \
        ((-1, -1, -1), False, 0.0)
    r = O.search(np.ones((6, 7, 8), np.int32), tray, penalty="pow")
    assert P.fft_search_placement(np.ones((6, 7, 8), np.int32), tray) == (r.position, r.found, r.score)


def test_search_batch_matches_oracle(P):
    rng = np.random.default_rng(7)
    from psacking_amd.rotations import get_orientations
    for t, (N, I) in enumerate([((30, 25, 20), (6, 4, 3)), ((41, 37, 33), (9, 7, 11)),
                                ((64, 50, 40), (12, 10, 8))]):
        tray = rand_tray(rng, N, 0.15, 4)
        item = (rng.random(I) < 0.6).astype(np.int32)
        phi = O.calculate_distance(tray)
        orients = [np.ascontiguousarray(g) for g in get_orientations(item, 24)]
        say("fft_search_batch", N, I)
        exp = O.search_batch(orients, tray, phi, penalty="cube")
        got = P.fft_search_batch(orients, tray, phi, 1000 + t)
        assert got == (exp[0], exp[1], exp[2])
        for o in (0, 5, 17):
            r = O.search(orients[o], tray, phi, penalty="cube")
            assert P.fft_search_placement_with_cache(orients[o], tray, phi, 1000 + t) == \
                (r.position, r.found, r.score)


def test_search_batch_interlocking_free(P):
    rng = np.random.default_rng(8)
    from psacking_amd.rotations import get_orientations
    N = (20, 18, 16)
    tray = (rng.random(N) < 0.25).astype(np.int32)
    tray[5:15, 5:15, 4:12] = 1
    tray[7:13, 7:13, 6:10] = 0
    item = np.ones((2, 3, 2), np.int32)
    phi = O.calculate_distance(tray)
    orients = [np.ascontiguousarray(g) for g in get_orientations(item, 24)]
    say("fft_search_batch_interlocking_free", N)
    exp = O.search_batch(orients, tray, phi, penalty="cube", interlocking_free=True)
    got = P.fft_search_batch_interlocking_free(orients, tray, phi, 77)
    assert got == exp


def test_search_interlocking_free_enclosed_winner(P):
    """Without the interlocking mask the best offset lies in a closed cavity
    (unreachable): the masked search must rank the reachable offsets (the
    shortcut that keeps a reachable unmasked winner does not apply); with the
    cavity opened to the top the two winners coincide."""
    N = (16, 16, 16)
    tray = np.zeros(N, np.int32)
    tray[:, :, 0] = 1                    # a floor (offsets at z = 0 are exits)
    tray[4:12, 4:12, 1:9] = 1
    tray[6:10, 6:10, 1:5] = 0            # a 4x4x4 cavity closed on every side
    item = [np.ones((4, 4, 4), np.int32)]
    phi = O.calculate_distance(tray)
    free = O.search_batch(item, tray, phi, penalty="cube", interlocking_free=False)
    masked = O.search_batch(item, tray, phi, penalty="cube", interlocking_free=True)
    assert free[0] == (6, 6, 1) and masked[0] != free[0]
    say("interlocking-free search, enclosed winner")
    assert P.fft_search_batch_interlocking_free(item, tray, phi, 5) == masked
    tray[6:10, 6:10, 5:9] = 0            # a shaft from the cavity to the top
    phi = O.calculate_distance(tray)
    free = O.search_batch(item, tray, phi, penalty="cube", interlocking_free=False)
    masked = O.search_batch(item, tray, phi, penalty="cube", interlocking_free=True)
    assert free == masked
    say("interlocking-free search, reachable winner")
    assert P.fft_search_batch_interlocking_free(item, tray, phi, 6) == masked


def test_gpu_tray_context(P):
    say("gpu_tray context")
    rng = np.random.default_rng(9)
    tray = rand_tray(rng, (24, 20, 16), 0.2)
    phi = O.calculate_distance(tray)
    P.gpu_tray_cleanup()
    assert not P.gpu_tray_is_initialized()
    with pytest.raises(RuntimeError):
        P.gpu_tray_search(np.ones((2, 2, 2), np.int32))
    P.gpu_tray_init(tray, phi)
    assert P.gpu_tray_is_initialized()
    item = np.ones((3, 2, 4), np.int32)
    r = O.search(item, tray, phi, penalty="cube")
    assert P.gpu_tray_search(item) == (r.position, r.found, r.score)
    P.gpu_tray_cleanup()
    assert not P.gpu_tray_is_initialized()


@pytest.mark.parametrize("engine", ["fused", "percall"])
def test_tetromino_fixture(P, kats, engine):
    shapes = kats["tetrominoes"]
    for run in kats["tetromino_runs"]:
        say("tetromino", engine, run["num_orientations"])
        items = [piece(shapes[c]) for c in run["order"]]
        bp = P.BinPacker(tuple(run["tray"]), num_orientations=run["num_orientations"],
                         engine=engine)
        r = bp.pack_voxels(items, sort_by_volume=False)
        assert r.num_placed == 7
        assert abs(r.density - run["density"]) < 1e-12
        assert [tuple(p.position) for p in r.placements] == [tuple(p) for p in run["positions"]]
        assert [p.orientation_index for p in r.placements] == run["orientations"]
        np.testing.assert_array_equal(r.tray[:, :9, :], grid_from_rows(run["grid_cols_0_8"], (8, 9, 1)))


def test_tetromino_reference_sort(P, kats):
    # the reference test passes the dict order with sort_by_volume=True
    shapes = kats["tetrominoes"]
    items = [piece(shapes[c]) for c in "IOTSZJL"]
    say("tetromino sorted")
    r = P.BinPacker((8, 16, 1), num_orientations=4).pack_voxels(items, sort_by_volume=True)
    run = kats["tetromino_runs"][1]
    np.testing.assert_array_equal(r.tray[:, :9, :], grid_from_rows(run["grid_cols_0_8"], (8, 9, 1)))


def _compare_pack(P, w, n_orient, interlocking_free=False, metric="l1", sources="eq1", n=None):
    items = w.items if n is None else w.items[:n]
    vols = [int((it > 0).sum()) for it in items]
    order = [int(i) for i in np.argsort(vols)[::-1]]
    say("pack", w.name, len(items), n_orient, interlocking_free, metric, sources)
    exp = O.pack_voxels(items, w.tray_size, num_orientations=n_orient,
                        interlocking_free=interlocking_free, order=order, metric=metric,
                        sources=sources)
    got = P.BinPacker(w.tray_size, num_orientations=n_orient, interlocking_free=interlocking_free,
                      metric=metric, sources=sources).pack_voxels(items)
    assert got.num_placed == exp.num_placed
    for a, b in zip(got.placements, exp.placements):
        assert (a.item_index, a.success, a.position, a.orientation_index, a.score) == \
            (b.item_index, b.success, b.position, b.orientation_index, b.score)
    np.testing.assert_array_equal(got.tray, exp.tray)
    assert got.density == exp.density


@pytest.mark.parametrize("nx, px", [(96, 96), (123, 128), (140, 144), (160, 160), (180, 192),
                                     (200, 224), (250, 256)])
def test_register_pass_a_lengths(P, nx, px):
    """Pass A in registers for x lengths 16 R (R = 6 .. 16): the tray's x
    length is padded to px (at most an eighth past the smallest 2^a 3^b 5^c
    7^d length), and the pipelined loop places as the oracle does."""
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    w = W.blobs(seed=30 + nx % 17, n=5, tray=(nx, 22, 18))
    s = P.Session(0)
    s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    assert s.plan()[0] == px
    say("register pass A, x length", px)
    s, res, exp, order = _session_pack(P, w, 24, metric="edt", sources="nonzero")
    by_item = {p.item_index: p for p in exp.placements}
    for t, r in enumerate(res):
        e = by_item[order[t]]
        assert r["found"] == e.success
        if e.success:
            assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                (e.position, e.orientation_index, e.score)
    np.testing.assert_array_equal(s.tray(), exp.tray)


def test_pack_boxes_64(P):
    from psacking_amd import workloads as W
    _compare_pack(P, W.boxes(), 1)


def test_pack_blobs_100(P):
    from psacking_amd import workloads as W
    _compare_pack(P, W.blobs(), 1, n=25)


def test_pack_blobs_24_orient(P):
    from psacking_amd import workloads as W
    w = W.blobs(seed=12, n=6, tray=(60, 50, 40))
    _compare_pack(P, w, 24)
    _compare_pack(P, w, 24, metric="edt", sources="nonzero")


def test_pack_interlocking_free(P):
    from psacking_amd import workloads as W
    w = W.blobs(seed=13, n=5, tray=(40, 36, 30))
    _compare_pack(P, w, 4, interlocking_free=True)


def test_orientation_count_case(P, kats):
    for case in kats["orientation_counts"]:
        say("orientation counts")
        item = np.ones(case["item_shape"], np.int32)
        assert P.BinPacker(tuple(case["tray"]), num_orientations=1).pack_voxels([item]).num_placed == case["placed_1"]
        assert P.BinPacker(tuple(case["tray"]), num_orientations=24).pack_voxels([item]).num_placed == case["placed_24"]


def _session_pack(P, w, n_orient, resid_limit=None, metric="l1", sources="eq1", n=None,
                  interlocking_free=False, options=(), timing=False):
    from psacking_amd.rotations import orientation_maps
    items = w.items if n is None else w.items[:n]
    vols = [int((it > 0).sum()) for it in items]
    order = [int(i) for i in np.argsort(vols)[::-1]]
    perms, flips = orientation_maps(n_orient)
    s = P.Session(0)
    if resid_limit is not None:
        s.set_option("resid_limit", resid_limit)
    for name, value in options:
        s.set_option(name, value)
    s.set_tray(np.zeros(w.tray_size, np.int32), metric, sources, 0)
    if timing:
        s.timing(True)
    res = s.pack(items, np.array(order, np.int32), perms, flips,
                 "pow" if n_orient == 1 and not interlocking_free else "cube", interlocking_free, 0)
    exp = O.pack_voxels(items, w.tray_size, num_orientations=n_orient, order=order,
                        metric=metric, sources=sources, interlocking_free=interlocking_free)
    return s, res, exp, order


def test_pack_forced_exact_path(P):
    """resid_limit < 0 hands every search back to the host's exact path."""
    from psacking_amd import workloads as W
    w = W.blobs(seed=14, n=6, tray=(40, 36, 30))
    say("pack forced exact path")
    s, res, exp, order = _session_pack(P, w, 24, resid_limit=-1.0)
    by_item = {p.item_index: p for p in exp.placements}
    for t, r in enumerate(res):
        e = by_item[order[t]]
        assert r["found"] == e.success
        # items with no orientation inside the tray are never searched
        assert r["path"] == 1 or not r["found"]
        if e.success:
            assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                (e.position, e.orientation_index, e.score)
    np.testing.assert_array_equal(s.tray(), exp.tray)


@pytest.mark.parametrize("options", [(), (("corr_chunk", 1), ("col_chunks", 3), ("score_radix", 25)),
                                     (("graph", 1), ("inc_dist", 0)),
                                     (("prescreen", 0), ("inc_masks", 0))])
def test_pack_pipelined_matches_oracle(P, options):
    """Default options; pass A one orientation per block, column-test
    blocks over several offset chunks, radix-25 score pass; the loop captured
    as a hipGraph, with the full distance transform after every placement;
    one round with the tray masks rebuilt whole after every placement."""
    from psacking_amd import workloads as W
    w = W.blobs(seed=15, n=10, tray=(64, 48, 40))
    say("pack pipelined", options)
    s, res, exp, order = _session_pack(P, w, 24, metric="edt", sources="nonzero", options=options)
    by_item = {p.item_index: p for p in exp.placements}
    for t, r in enumerate(res):
        e = by_item[order[t]]
        assert r["path"] == 0 and r["found"] == e.success
        if e.success:
            assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                (e.position, e.orientation_index, e.score)
            assert r["collision"] == 0
    np.testing.assert_array_equal(s.tray(), exp.tray)


def test_search_forced_exact_interlocking(P):
    from psacking_amd import workloads as W
    w = W.blobs(seed=16, n=3, tray=(30, 28, 24))
    say("interlocking forced exact path")
    s, res, exp, order = _session_pack(P, w, 4, resid_limit=-1.0, interlocking_free=True)
    by_item = {p.item_index: p for p in exp.placements}
    for t, r in enumerate(res):
        e = by_item[order[t]]
        assert r["found"] == e.success
        if e.success:
            assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                (e.position, e.orientation_index, e.score)
    np.testing.assert_array_equal(s.tray(), exp.tray)


def _pack_with_plan(P, w, n_orient, plan, metric="l1", sources="eq1", interlocking_free=False,
                    tile=0):
    from psacking_amd.rotations import orientation_maps
    items = w.items
    vols = [int((it > 0).sum()) for it in items]
    order = [int(i) for i in np.argsort(vols)[::-1]]
    perms, flips = orientation_maps(n_orient)
    s = P.Session(0)
    s.set_option("plan", plan)
    s.set_option("tile", tile)
    s.set_tray(np.zeros(w.tray_size, np.int32), metric, sources, 0)
    res = s.pack(items, np.array(order, np.int32), perms, flips,
                 "pow" if n_orient == 1 and not interlocking_free else "cube", interlocking_free, 0)
    exp = O.pack_voxels(items, w.tray_size, num_orientations=n_orient, order=order,
                        metric=metric, sources=sources, interlocking_free=interlocking_free)
    by_item = {p.item_index: p for p in exp.placements}
    for t, r in enumerate(res):
        e = by_item[order[t]]
        assert r["found"] == e.success
        if e.success:
            assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                (e.position, e.orientation_index, e.score)
    np.testing.assert_array_equal(s.tray(), exp.tray)
    if tile:
        assert s.tile()[0] == tile


@pytest.mark.parametrize("plan", [1, 2])
def test_plans_agree_with_oracle(P, plan):
    from psacking_amd import workloads as W
    say("plan", plan, "24 orientations edt")
    _pack_with_plan(P, W.blobs(seed=17, n=8, tray=(50, 44, 36)), 24, plan, "edt", "nonzero")
    say("plan", plan, "1 orientation l1")
    _pack_with_plan(P, W.boxes(seed=18, n=10, tray=(33, 29, 27)), 1, plan)
    say("plan", plan, "interlocking free")
    _pack_with_plan(P, W.blobs(seed=19, n=3, tray=(30, 26, 22)), 4, plan, interlocking_free=True)


def test_large_plane_tray_uses_split_plan(P):
    # y*z = 150*120 exceeds one block's LDS for the slab plan
    from psacking_amd import workloads as W
    w = W.blobs(seed=20, n=3, tray=(60, 150, 120))
    say("large-plane tray (auto split)")
    _pack_with_plan(P, w, 4, 0, "edt", "nonzero")


@pytest.mark.parametrize("ff_bits", [1, 0])
def test_pack_many_uneven_trays(P, ff_bits):
    """Independent instances of different tray sizes packed concurrently,
    interlocking-free check after each placement (flood fill on bit columns
    or bytes), each equal to its own oracle run."""
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    ws = [W.blobs(seed=21, n=4, tray=(34, 30, 26)),
          W.blobs(seed=22, n=5, tray=(48, 28, 20)),
          W.blobs(seed=23, n=3, tray=(26, 40, 30))]
    perms, flips = orientation_maps(4)
    sessions, lists, orders = [], [], []
    for w in ws:
        s = P.Session(0)
        s.set_option("ff_bits", ff_bits)
        s.set_tray(np.zeros(w.tray_size, np.int32), "l1", "eq1", 0)
        vols = [int((it > 0).sum()) for it in w.items]
        sessions.append(s)
        lists.append(list(w.items))
        orders.append(np.array([int(i) for i in np.argsort(vols)[::-1]], np.int32))
    say("pack_many 3 uneven trays, 4 orientations, interlocking free")
    got = P.pack_many(sessions, lists, orders, perms, flips, "cube", True)
    for w, s, res, order in zip(ws, sessions, got, orders):
        exp = O.pack_voxels(w.items, w.tray_size, num_orientations=4, order=list(order),
                            interlocking_free=True)
        by_item = {p.item_index: p for p in exp.placements}
        for t, r in enumerate(res):
            e = by_item[int(order[t])]
            assert r["found"] == e.success
            if e.success:
                assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                    (e.position, e.orientation_index, e.score)
        np.testing.assert_array_equal(s.tray(), exp.tray)


@pytest.mark.parametrize("feas,coarse,prescreen", [(0, 1, 16), (1, 0, 16), (1, 1, 16), (2, 0, 16),
                                                   (2, 1, 16), (1, 1, 2), (1, 1, 0)])
def test_feasibility_pass_keeps_results(P, feas, coarse, prescreen):
    """A tray that fills up: later items have no feasible orientation.  The
    feasibility pass (off, after a non-placement, every item), with and
    without its coarse 2x2x2-block pre-pass, must leave every outcome and the
    final tray equal to the oracle's; so must rounds of 2 items with a
    batched column test at the start of each (items that fit nowhere then
    are decided there) and a single round with none."""
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    w = W.blobs(seed=51, n=14, tray=(48, 40, 32))
    vols = [int((it > 0).sum()) for it in w.items]
    order = [int(i) for i in np.argsort(vols)[::-1]]
    perms, flips = orientation_maps(24)
    s = P.Session(0)
    s.set_option("feas", feas)
    s.set_option("coarse", coarse)
    s.set_option("prescreen", prescreen)
    s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    say("feasibility pass mode", feas, "coarse", coarse, "prescreen", prescreen)
    res = s.pack(w.items, np.array(order, np.int32), perms, flips, "cube", False, 0)
    exp = O.pack_voxels(w.items, w.tray_size, num_orientations=24, order=order,
                        metric="edt", sources="nonzero")
    by_item = {p.item_index: p for p in exp.placements}
    assert sum(1 for p in exp.placements if not p.success) >= 3
    for t, r in enumerate(res):
        e = by_item[order[t]]
        assert r["found"] == e.success and r["path"] == 0
        if e.success:
            assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                (e.position, e.orientation_index, e.score)
    np.testing.assert_array_equal(s.tray(), exp.tray)


def test_free_space_certificate_keeps_results(P):
    """Certificate on: a tray with scattered occupied voxels where large
    solid items fit nowhere (they are skipped without a search) and small
    ones fit; every outcome and the final tray equal the oracle's."""
    from psacking_amd.rotations import orientation_maps
    rng = np.random.default_rng(60)
    shape = (30, 28, 24)
    tray = (rng.random(shape) < 0.12).astype(np.int32)
    items = [np.ones((8, 7, 9), np.int32), np.ones((2, 1, 2), np.int32),
             np.ones((9, 9, 8), np.int32), np.ones((1, 2, 1), np.int32),
             np.ones((2, 2, 1), np.int32)]
    order = list(range(len(items)))
    perms, flips = orientation_maps(24)
    for cert in (0, 1):
        s = P.Session(0)
        s.set_option("certificate", cert)
        s.set_tray(tray, "edt", "nonzero", 0)
        say("certificate", cert)
        res = s.pack(items, np.array(order, np.int32), perms, flips, "cube", False,
                     int(tray.max()))
        exp = O.pack_voxels(items, shape, num_orientations=24, order=order, metric="edt",
                            sources="nonzero", initial_tray=tray)
        by_item = {p.item_index: p for p in exp.placements}
        assert not by_item[0].success and by_item[1].success
        for t, r in enumerate(res):
            e = by_item[order[t]]
            assert r["found"] == e.success
            if e.success:
                assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                    (e.position, e.orientation_index, e.score)
        np.testing.assert_array_equal(s.tray(), exp.tray)


def test_autotuned_tile_is_cached_and_exact(P):
    from psacking_amd import workloads as W
    w = W.blobs(seed=61, n=5, tray=(72, 54, 44))
    s1 = P.Session(0)
    s1.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    t1, split = s1.tile()
    say("autotuned tile", t1, "split", split)
    assert t1 in (4, 8, 16) and split == 0
    s2 = P.Session(0)
    s2.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    assert s2.tile() == (t1, split)
    s3 = P.Session(0)
    s3.set_option("autotune", 0)
    s3.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    assert s3.tile()[0] == 16
    # every admissible tile gives the oracle's placements
    for tile in (4, 8, 16):
        say("forced tile", tile)
        _pack_with_plan(P, w, 24, 0, "edt", "nonzero", tile=tile)


def _full_feasible(g, tray):
    """Oracle: a valid offset with no occupied voxel under the item."""
    n = np.array(tray.shape)
    lo, hi = O.get_bounds(g)
    h = np.array(hi)
    if h[0] < 0 or np.any(h > n - 1):
        return False
    c = O.corr_exact((tray != 0).astype(np.int64), O.pad_to((g > 0).astype(np.int64), tray.shape))
    return bool((c[:n[0] - h[0], :n[1] - h[1], :n[2] - h[2]] == 0).any())


def test_masks_after_placements_match_oracle(P):
    """The column test's tray masks are updated under each placed box only
    (option inc_masks, default): after a pack run the column flags of
    spk_feasible equal exhaustive feasibility on the packed tray, and the run
    with whole-tray mask rebuilds gives the same outcomes and tray."""
    from psacking_amd import proposer as PR
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    perms, flips = orientation_maps(24)
    w = W.blobs(seed=52, n=12, tray=(44, 38, 30))
    vols = [int((it > 0).sum()) for it in w.items]
    order = np.array([int(i) for i in np.argsort(vols)[::-1]], np.int32)
    out = []
    for inc in (1, 0):
        s = P.Session(0)
        s.set_option("inc_masks", inc)
        s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
        say("pack with inc_masks", inc)
        res = s.pack(w.items, order, perms, flips, "cube", False, 0)
        out.append(([(r["found"], tuple(r["position"]), r["orientation"], r["score"]) for r in res],
                    s.tray(), s))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
    tray, s = out[0][1], out[0][2]
    assert sum(1 for r in out[0][0] if r[0]) >= 5
    rng = np.random.default_rng(53)
    for case in range(5):
        dims = tuple(int(v) for v in rng.integers(4, 14, 3))
        item = np.ones(dims, np.int32)
        item[rng.random(dims) < 0.15] = 0
        say("column flags on the packed tray, case", case, dims)
        column, full, res = PR.feasible(s, item, perms, flips)
        for o, g in enumerate(O.get_orientations(item, 24)):
            g = O.as_grid(g)
            if any(g.shape[a] > tray.shape[a] for a in range(3)):
                continue
            exp = _full_feasible(g, tray)
            assert column[o] == exp and full[o] == exp, (case, o)


@pytest.mark.parametrize("fine_group", [64, 4])
def test_feasibility_flags_match_oracle(P, fine_group):
    """spk_feasible against the oracle on trays filled to several levels: the
    full-pass flags and the column test's flags (its block stage, then its
    full-resolution stage on the offsets the blocks leave open) both equal the
    exact feasibility of every orientation; the orientations the oracle's block
    stage alone excludes are among the infeasible ones.  Both lane-group sizes
    of the full-resolution stage (the default 64 and 4); two of the trays are
    empty above a height, where the tile-height test proves feasibility."""
    from psacking_amd import proposer as PR
    from psacking_amd.rotations import orientation_maps
    perms, flips = orientation_maps(24)
    rng = np.random.default_rng(71)
    shape = (36, 30, 26)
    checked = excluded = feasible = 0
    for level in (0, 8, 14, 18, 21):
        tray = (rng.random(shape) < 0.04).astype(np.int32)
        tray[:, :, :level] = 3
        if level in (8, 18):   # free space above: the tile-height test proves most orientations
            tray[:, :, level + 2:] = 0
        s = P.Session(0)
        s.set_option("fine_group", fine_group)
        s.set_tray(tray, "edt", "nonzero", 0)
        for case in range(6):
            dims = tuple(int(v) for v in rng.integers(4, 12, 3))
            item = np.ones(dims, np.int32)
            item[rng.random(dims) < 0.15] = 0
            say("feasibility flags level", level, "case", case, dims)
            column, full, res = PR.feasible(s, item, perms, flips)
            assert res <= 0.25
            for o, g in enumerate(O.get_orientations(item, 24)):
                g = O.as_grid(g)
                if any(g.shape[a] > shape[a] for a in range(3)):
                    assert column[o] == 0 and full[o] == 0
                    continue
                exp_full = _full_feasible(g, tray)
                assert full[o] == exp_full, (level, case, o)
                assert column[o] == exp_full, (level, case, o)
                blocks_exclude = O.column_infeasible(g, tray)
                assert not (blocks_exclude and exp_full), (level, case, o)
                checked += 1
                excluded += blocks_exclude
                feasible += bool(full[o])
    assert checked > 500 and excluded > 50 and feasible > 50


@pytest.mark.parametrize("dist_y,dist_x", [(0, 0), (1, 0), (2, 0), (2, 2)])
def test_distance_modes_match_oracle(P, dist_y, dist_x):
    """Every y-pass (outward search, Felzenszwalb envelope, full scan) and
    x-pass (outward search bounded by the line minimum, full scan) variant of
    the distance field gives the oracle's field, both metrics, on an empty, a
    scattered and a filled-from-below tray."""
    rng = np.random.default_rng(83)
    shape = (37, 29, 23)
    trays = [np.zeros(shape, np.int32), (rng.random(shape) < 0.02).astype(np.int32)]
    t = (rng.random(shape) < 0.01).astype(np.int32)
    t[:, :, :9] = 2
    trays.append(t)
    for metric in ("edt", "l1"):
        for k, tray in enumerate(trays):
            say("distance modes", dist_y, dist_x, metric, k)
            s = P.Session(0)
            s.set_option("dist_y", dist_y)
            s.set_option("dist_x", dist_x)
            s.set_tray(tray, metric, "nonzero", 0)
            np.testing.assert_array_equal(s.distance(), O.distance_field(tray, metric, "nonzero"))


def test_pack_flat_and_gather_agree(P):
    """spk_pack (items concatenated in one buffer) and spk_pack_gather (one
    pointer per item array, what Session.pack passes) give the same results
    and the same tray."""
    import ctypes as C
    from psacking_amd import _lib
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    w = W.blobs(seed=16, n=12, tray=(64, 48, 40))
    perms, flips = orientation_maps(24)
    order = np.arange(len(w.items), dtype=np.int32)
    say("pack flat vs gather, 12 items")
    s1 = P.Session(0)
    s1.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    res = s1.pack(list(w.items), order, perms, flips, "cube", False, 0)
    s2 = P.Session(0)
    s2.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    grids = [np.ascontiguousarray(it, dtype=np.int32) for it in w.items]
    flat = np.concatenate([g.reshape(-1) for g in grids])
    dims = np.array([g.shape for g in grids], np.int32).reshape(-1)
    orients = (_lib.SpkOrient * len(perms))()
    for o in range(len(perms)):
        orients[o].perm[:] = [int(v) for v in perms[o]]
        orients[o].flip[:] = [int(v) for v in flips[o]]
    out = (_lib.SpkResult * len(grids))()
    ip = C.POINTER(C.c_int32)
    rc = _lib.load().spk_pack(C.c_void_p(s2.handle()), flat.ctypes.data_as(ip), dims.ctypes.data_as(ip),
                              len(grids), order.ctypes.data_as(ip), orients, len(perms), 1, 0, 0, out)
    assert rc == 0, _lib.last_error()
    placed = 0
    for r, g in zip(res, out):
        assert r["found"] == bool(g.found)
        if r["found"]:
            placed += 1
            assert (tuple(r["position"]), r["orientation"], r["score"]) == (tuple(g.pos), g.orient, g.score)
    assert placed > 0
    np.testing.assert_array_equal(s1.tray(), s2.tray())


def test_feasibility_flags_large_footprint(P):
    """Items whose footprint exceeds what the item-mask kernel holds in LDS
    (ex * ey > 4096 columns) take the three-kernel path (core masks, item
    masks, run lists); the column test's flags still equal exact
    feasibility, on trays with sparse obstacles and one empty above a height."""
    from psacking_amd import proposer as PR
    from psacking_amd.rotations import orientation_maps
    perms, flips = orientation_maps(24)
    rng = np.random.default_rng(91)
    shape = (92, 84, 30)
    item = np.ones((70, 62, 5), np.int32)
    item[rng.random(item.shape) < 0.1] = 0
    item[10:22, 5:17, :] = 0
    item[:, :, 3:] *= (rng.random((70, 62, 1)) < 0.5).astype(np.int32)
    checked = big = feasible = 0
    for case, p in enumerate((0.002, 0.004, 0.03)):
        tray = (rng.random(shape) < p).astype(np.int32)
        tray[:, :, :6] = 2
        if case == 0:
            tray[:, :, 20:] = 0
        say("large-footprint feasibility, case", case)
        s = P.Session(0)
        s.set_tray(tray, "edt", "nonzero", 0)
        column, full, res = PR.feasible(s, item, perms, flips)
        assert res <= 0.25
        for o, g in enumerate(O.get_orientations(item, 24)):
            g = O.as_grid(g)
            if any(g.shape[a] > shape[a] for a in range(3)):
                assert column[o] == 0 and full[o] == 0
                continue
            exp = _full_feasible(g, tray)
            assert full[o] == exp and column[o] == exp, (case, o)
            checked += 1
            big += g.shape[0] * g.shape[1] > 4096
            feasible += bool(exp)
    assert big > 0 and feasible > 0 and feasible < checked


@pytest.mark.parametrize("metric,sources", [("edt", "nonzero"), ("l1", "nonzero"), ("l1", "eq1"),
                                            ("edt", "eq1")])
def test_incremental_distance_matches_oracle(P, metric, sources):
    """After each placement the context's field (updated from the placed
    voxels alone, option inc_dist on by default) equals the oracle's full
    transform of the tray: items placed one pack call at a time by the
    device loop, then by explicit placements on the host path."""
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    w = W.blobs(seed=24, n=8, tray=(50, 44, 36))
    vols = [int((it > 0).sum()) for it in w.items]
    order = [int(i) for i in np.argsort(vols)[::-1]]
    perms, flips = orientation_maps(24)
    s = P.Session(0)
    s.set_tray(np.zeros(w.tray_size, np.int32), metric, sources, 0)
    placed = 0
    say("incremental field, device loop", metric, sources)
    for i in order[:5]:
        r = s.pack([w.items[i]], np.zeros(1, np.int32), perms, flips, "cube", False, placed)[0]
        placed += int(r["found"])
        np.testing.assert_array_equal(s.distance(), O.distance_field(s.tray(), metric, sources))
        assert s.field_stats() == _field_stats(s.distance())
    assert placed >= 3
    say("incremental field, explicit placements", metric, sources)
    t = s.tray()
    for i in order[5:]:
        item = w.items[i]
        res = O.search_state(O.TrayState(t, O.distance_field(t, metric, sources)), item, "cube")
        if not res.found:
            continue
        placed += 1
        s.place(item, perms[0], flips[0], res.position, placed)
        t = O.place_in_tray(item, t, res.position, placed)
        np.testing.assert_array_equal(s.tray(), t)
        np.testing.assert_array_equal(s.distance(), O.distance_field(t, metric, sources))
        assert s.field_stats() == _field_stats(s.distance())


def _field_stats(phi):
    """sum, min, max of a field and the largest min(field, distance to the
    outside of the grid), the statistics the digit split is sized from."""
    w = None
    for a, n in enumerate(phi.shape):
        i = np.arange(n)
        d = np.minimum(i + 1, n - i).reshape([-1 if b == a else 1 for b in range(3)])
        w = d if w is None else np.minimum(w, d)
    return (int(phi.astype(np.int64).sum()), int(phi.min()), int(phi.max()),
            int(np.minimum(phi, w).max()))


@pytest.mark.parametrize("spec_inc", [0, 1, 16, 1000])
def test_incremental_spectrum_matches_oracle(P, spec_inc):
    """Sources "eq1": a placement with id != 1 adds no distance source, so the
    device loop updates the tray spectrum by the placed voxels alone (option
    spec_inc: at most that many updates in a row, then a full transform).
    Whole pack runs against the oracle with 1 and 24 orientations; every
    search stays on the FFT path and the update ran when it is on."""
    from psacking_amd import workloads as W
    runs = [(W.blobs(), 1, 25), (W.blobs(seed=15, n=12, tray=(64, 48, 40)), 24, None)]
    for w, n_orient, n in runs:
        say("incremental spectrum", spec_inc, w.name, n_orient)
        s, res, exp, order = _session_pack(P, w, n_orient, n=n, options=(("spec_inc", spec_inc),),
                                           timing=True)
        by_item = {p.item_index: p for p in exp.placements}
        for t, r in enumerate(res):
            e = by_item[order[t]]
            assert r["path"] == 0 and r["found"] == e.success, (t, r)
            if e.success:
                assert (tuple(r["position"]), r["orientation"], r["score"]) == \
                    (e.position, e.orientation_index, e.score)
        np.testing.assert_array_equal(s.tray(), exp.tray)
        placed = sum(1 for r in res if r["found"])
        updates = s.timing_units("spec_add")
        say("  placed", placed, "incremental updates", updates,
            "max residual", max(float(r["residual"]) for r in res))
        if spec_inc == 0:
            assert updates == 0
        else:
            # every placement after the first (id 1, which adds the sources)
            # updates incrementally except the full transform after each
            # spec_inc updates in a row
            assert updates == (placed - 1) - (placed - 1) // (spec_inc + 1)
