"""GPU parity at the large workloads, against stored oracle outcomes.

* readme: the 240x123x100 tray with the first 8 items of the seeded 350-item
  workload, 24 orientations, proximity edt/nonzero and l1/eq1, through the
  drop-in BinPacker.pack_voxels;
* uneven: three uneven instances (200x211x128 and two small trays), the first
  6 items of each, 24 orientations, interlocking-free, packed concurrently
  through pack_many;
* learned: the top-4 learned loop with the committed weights on the readme
  prefix, every placement replayed on the oracle (the proposals the device
  searched, then every orientation when none of them is feasible);
* the reference's flood-fill cases through interlocking_free_positions.

The readme and uneven outcomes come from tests/golden/large_fixtures.json,
written in advance by tests/golden/make_large_fixtures.py with the CPU
oracle; positions, orientation indices and scores must be equal, and the
final tray must have the stored SHA-256.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import spk_oracle as O
from tests import flood_cases as F
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int32).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def P():
    import psacking_amd
    return psacking_amd


@pytest.fixture(scope="module")
def fx():
    with open(os.path.join(GOLDEN, "large_fixtures.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="module")
def readme():
    from psacking_amd import workloads as W
    return W.readme()


def _check_outcomes(f, found, positions, orients, scores):
    assert len(found) == len(f["outcomes"])
    for t, (ok, pos, o, sc) in enumerate(f["outcomes"]):
        assert found[t] == ok, t
        if ok:
            assert (tuple(positions[t]), orients[t], scores[t]) == (tuple(pos), o, sc), t


@pytest.mark.parametrize("mode", ["edt/nonzero", "l1/eq1"])
def test_readme_prefix_matches_oracle(P, fx, readme, mode):
    f = fx["readme"][mode]
    items = [readme.items[i] for i in f["order"]]
    assert [sha(it) for it in items] == f["item_sha256"]
    metric, sources = mode.split("/")
    say("readme prefix", len(items), "items, 24 orientations", mode)
    r = P.BinPacker(tuple(f["tray_size"]), num_orientations=24, metric=metric,
                    sources=sources).pack_voxels(items, sort_by_volume=False)
    pl = r.placements
    _check_outcomes(f, [p.success for p in pl], [p.position for p in pl],
                    [p.orientation_index for p in pl], [p.score for p in pl])
    assert r.num_placed == f["num_placed"]
    assert sha(r.tray) == f["tray_sha256"] and int((r.tray > 0).sum()) == f["tray_occupied"]


@pytest.mark.parametrize("reg_x", [0, 1, 2])
def test_readme_prefix_pass_a_variants(P, fx, readme, reg_x):
    """Every pass-A kernel for the 240-point x length (0: LDS FFT, 1/2:
    registers with three or two LDS buffers) on the readme prefix."""
    from psacking_amd.rotations import orientation_maps
    f = fx["readme"]["edt/nonzero"]
    items = [readme.items[i] for i in f["order"]]
    perms, flips = orientation_maps(24)
    s = P.Session(0)
    s.set_option("reg_x", reg_x)
    s.set_tray(np.zeros(readme.tray_size, np.int32), "edt", "nonzero", 0)
    say("readme prefix, reg_x", reg_x)
    res = s.pack(items, np.arange(len(items), dtype=np.int32), perms, flips, "cube", False, 0)
    _check_outcomes(f, [bool(r["found"]) for r in res], [r["position"] for r in res],
                    [r["orientation"] for r in res], [r["score"] for r in res])
    assert sha(s.tray()) == f["tray_sha256"]


def test_uneven_instances_interlocking_free(P, fx):
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    perms, flips = orientation_maps(24)
    keys = sorted(fx["uneven"], key=int)
    sessions, lists = [], []
    for k in keys:
        f = fx["uneven"][k]
        w = W.uneven_instance(int(k))
        items = [w.items[i] for i in f["order"]]
        assert [sha(it) for it in items] == f["item_sha256"]
        assert list(w.tray_size) == f["tray_size"]
        s = P.Session(0)
        s.set_tray(np.zeros(w.tray_size, np.int32), f["metric"], f["sources"], 0)
        sessions.append(s)
        lists.append(items)
    say("pack_many", [fx["uneven"][k]["tray_size"] for k in keys], "24 orientations, IF")
    got = P.pack_many(sessions, lists, [np.arange(len(l), dtype=np.int32) for l in lists],
                      perms, flips, "cube", True)
    for k, s, res in zip(keys, sessions, got):
        f = fx["uneven"][k]
        _check_outcomes(f, [bool(r["found"]) for r in res], [r["position"] for r in res],
                        [r["orientation"] for r in res], [r["score"] for r in res])
        t = s.tray()
        assert sha(t) == f["tray_sha256"], k


def _replay_state(items, n_orient, tray_size, proposals, metric, sources):
    """Oracle loop that searches item t over proposals[t] (when one of them is
    feasible) else over every orientation, on cached spectra."""
    tray = np.zeros(tray_size, np.int32)
    out, placed = [], 0
    for t, item in enumerate(items):
        orients = [O.as_grid(g) for g in O.get_orientations(item, n_orient)]
        st = O.TrayState(tray, O.distance_field(tray, metric, sources))
        res = {}

        def best_over(olist):
            best, bs = None, np.inf
            for o in sorted(olist):
                if o not in res:
                    res[o] = O.search_state(st, orients[o], "cube")
                r = res[o]
                if r.found and r.score < bs:
                    best, bs = (r, o), r.score
            return best

        prop = [int(o) for o in proposals[t] if o >= 0]
        best = best_over(prop) if prop else None
        if best is None:
            best = best_over(range(n_orient))
        if best is None:
            out.append(None)
            continue
        r, o = best
        placed += 1
        tray = O.place_in_tray(orients[o], tray, r.position, placed)
        out.append((r.position, o, r.score))
    return out, tray


def test_learned_readme_prefix_replays_on_oracle(P, fx, readme):
    from psacking_amd import proposer as PR
    from psacking_amd.rotations import orientation_maps
    f = fx["readme"]["edt/nonzero"]
    items = [readme.items[i] for i in f["order"]]
    perms, flips = orientation_maps(24)
    nn = PR.load_trained(0, max_batch=8)
    s = P.Session(0)
    s.set_tray(np.zeros(readme.tray_size, np.int32), "edt", "nonzero", 0)
    say("learned top-4, readme prefix", len(items), "items")
    res, props, counts = PR.pack_learned(s, nn, items, np.arange(len(items)), perms, flips,
                                         "cube", topk=4)
    say("counts", list(counts))
    assert counts[0] > 0     # items decided on proposals
    exp, tray = _replay_state(items, 24, readme.tray_size, props, "edt", "nonzero")
    for t, (r, e) in enumerate(zip(res, exp)):
        if e is None:
            assert not r["found"], t
        else:
            assert r["found"] and (tuple(r["position"]), r["orientation"], r["score"]) == e, t
    np.testing.assert_array_equal(s.tray(), tray)


@pytest.mark.parametrize("case", F.cases(), ids=lambda c: c[0])
def test_interlocking_reference_cases(P, case):
    name, tray, obj, expected = case
    say("interlocking reference case", name)
    m = P.interlocking_free_positions(tray, obj)
    F.check(m, expected)
    np.testing.assert_array_equal(m, O.interlocking_free_positions(tray, obj))
