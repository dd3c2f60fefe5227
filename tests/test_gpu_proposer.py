"""GPU tests of the learned orientation proposer against the CPU oracles.

* features: bit-exact against oracle/spk_nn_oracle.py,
* network: logits and every parameter gradient against the bf16-emulating
  oracle (tolerances written below; accumulation order differs),
* training: the device's own Adam steps fit a fixed batch,
* records and the learned loop: every placement replayed with the packing
  oracle (the proposals the device searched, then the full set when none of
  them is feasible) must match exactly, as must the final tray.
"""
import numpy as np
import pytest

from oracle import spk_nn_oracle as N
from oracle import spk_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def P():
    import psacking_amd
    return psacking_amd


@pytest.fixture(scope="module")
def PR():
    from psacking_amd import proposer
    return proposer


def to_dev(x):
    from psacking_amd.proposer import DeviceArray
    return DeviceArray.from_numpy(N.bf16_bits(x))


def to_host_bits(t):
    return t.numpy().reshape(t.shape[0], -1)


def test_item_features_bitexact(PR):
    from psacking_amd.rotations import orientation_maps
    rng = np.random.default_rng(30)
    nn = PR.Proposer(0, 8)
    perms, flips = orientation_maps(24)
    for shape in [(37, 20, 70), (5, 9, 3), (64, 64, 64)]:
        say("item features", shape)
        item = (rng.random(shape) < 0.4).astype(np.int32) * rng.integers(1, 4, shape)
        got = to_host_bits(nn.item_features(item))[0]
        np.testing.assert_array_equal(got, N.bf16_bits(N.item_features(item)).reshape(-1))
        out = PR.DeviceArray((24, PR.ITEM_FEATURES), np.uint16)
        nn.item_features_all(item, perms, flips, out, 0)
        got = to_host_bits(out)
        for o in range(24):
            np.testing.assert_array_equal(
                got[o], N.bf16_bits(N.item_features(item, perms[o], flips[o])).reshape(-1))


@pytest.mark.parametrize("shape", [(240, 123, 100), (50, 20, 20)])
def test_tray_features_bitexact(P, PR, shape):
    rng = np.random.default_rng(31)
    tray = ((rng.random(shape) < 0.25) * rng.integers(1, 5, shape)).astype(np.int32)
    s = P.Session(0)
    s.set_tray(tray, "l1", "eq1", 0)
    say("tray features", shape)
    got = to_host_bits(PR.tray_features(s))[0]
    np.testing.assert_array_equal(got, N.bf16_bits(N.tray_features(tray)).reshape(-1))


def _batch(seed, B):
    rng = np.random.default_rng(seed)
    item = ((rng.random((B, 24, 32, 32, 32)) < 0.3) * rng.integers(1, 9, (B, 24, 32, 32, 32)) / 8.0)
    tray = N.bf16(rng.random((B, 64, 32, 32, 2)).astype(np.float32))
    scores = rng.uniform(0, 1e6, (B, 24))
    scores[rng.random((B, 24)) < 0.3] = np.inf
    q, mask = N.soft_targets(scores)
    return item.astype(np.float32), tray, q, mask


def _dev_batch(item, tray, q, mask):
    from psacking_amd.proposer import DeviceArray
    B = item.shape[0]
    return (to_dev(item.reshape(B * 24, -1)), to_dev(tray.reshape(B, -1)),
            DeviceArray.from_numpy(q.astype(np.float32)), DeviceArray.from_numpy(mask.astype(np.int32)))


def test_forward_matches_oracle(PR):
    params = PR.init_params(40)
    params[N.param_slices()["F2"][1]] = np.linspace(-1, 1, 8, dtype=np.float32)
    nn = PR.Proposer(0, 16, params)
    item, tray, q, mask = _batch(41, 5)
    say("proposer forward, batch 5")
    got = nn.forward(*_dev_batch(item, tray, q, mask)[:2])
    ref, _ = N.forward(params, item, tray, emulate=True)
    scale = np.abs(ref).max()
    err = float(np.abs(got - ref).max()) / scale
    say("forward max abs error / max |logit|", f"{err:.2e}")
    # accumulation order differs, so a bf16 rounding of an activation may
    # flip (2^-8 relative); the logits stay within a few 1e-3 of the scale
    assert err < 4e-3, err


def test_gradients_match_oracle(PR):
    rng = np.random.default_rng(42)
    params = PR.init_params(42) + rng.normal(0, 0.01, PR.N_PARAMS).astype(np.float32)
    nn = PR.Proposer(0, 16, params)
    item, tray, q, mask = _batch(43, 6)
    say("proposer gradient, batch 6")
    loss, g = nn.gradient(*_dev_batch(item, tray, q, mask))
    ref_loss, ref, _ = N.loss_grad(params.astype(np.float64), item, tray, q, mask, emulate=True)
    assert abs(loss - ref_loss) <= 1e-3 * max(1.0, abs(ref_loss))
    for name, (w, b) in N.param_slices().items():
        for part, sl in (("W", w), ("b", b)):
            err = np.linalg.norm(g[sl] - ref[sl]) / max(np.linalg.norm(ref[sl]), 1e-12)
            say(name, part, "relative L2 error", f"{err:.2e}")
            # measured on MI355X: 5.9e-4 .. 1.1e-3 per tensor (bf16
            # activations whose rounding flips with the accumulation order);
            # a gradient 5x off that would fail
            assert err < 5e-3, (name, part, err)


def test_training_fits_a_batch(PR):
    nn = PR.Proposer(0, 32, PR.init_params(44))
    item, tray, q, mask = _batch(45, 32)
    dev = _dev_batch(item, tray, q, mask)
    say("proposer training, 80 Adam steps on 32 records")
    first = nn.train_step(*dev, lr=1e-3)
    losses = [nn.train_step(*dev, lr=1e-3) for _ in range(80)]
    assert losses[-1] < 0.5 * first, (first, losses[-5:])
    # the parameters moved and round-trip through the host
    p = nn.get_params()
    assert not np.array_equal(p, PR.init_params(44))
    nn2 = PR.Proposer(0, 32, p)
    np.testing.assert_array_equal(nn2.forward(*dev[:2]), nn.forward(*dev[:2]))


def _workload(seed, n, tray):
    from psacking_amd import workloads as W
    w = W.blobs(seed=seed, n=n, tray=tray)
    vols = [int((it > 0).sum()) for it in w.items]
    order = [int(i) for i in np.argsort(vols)[::-1]]
    return w, order


def _replay(items, order, n_orient, tray_size, proposals=None, metric="edt", sources="nonzero"):
    """Oracle greedy loop that searches item t over proposals[t] (if any is
    feasible) else over every orientation; returns (results, tray, per-item
    per-orientation scores, per-item tray before its search)."""
    tray = np.zeros(tray_size, np.int32)
    out, all_scores, trays = [], [], []
    placed = 0
    for t, orig in enumerate(order):
        orients = [O.as_grid(g) for g in O.get_orientations(items[orig], n_orient)]
        phi = O.distance_field(tray, metric, sources)
        trays.append(tray.copy())
        scores = np.full(n_orient, np.inf)
        res = {}

        def search(o):
            g = orients[o]
            if o not in res:
                if any(g.shape[d] > tray.shape[d] for d in range(3)):
                    res[o] = None
                else:
                    res[o] = O.search(g, tray, phi, penalty="cube")
            return res[o]

        def best_over(olist):
            best, bs = None, np.inf
            for o in sorted(olist):
                r = search(o)
                if r is not None and r.found and r.score < bs:
                    best, bs = (r, o), r.score
            return best

        prop = None if proposals is None else [int(o) for o in proposals[t] if o >= 0]
        best = best_over(prop) if prop else None
        if best is None:
            best = best_over(range(n_orient))
        if proposals is None:
            for o in range(n_orient):
                r = search(o)
                scores[o] = r.score if (r is not None and r.found) else np.inf
        all_scores.append(scores)
        if best is None:
            out.append(None)
            continue
        r, o = best
        placed += 1
        tray = O.place_in_tray(orients[o], tray, r.position, placed)
        out.append((r.position, o, r.score))
    return out, tray, np.array(all_scores), trays


def _check(res, exp):
    for r, e in zip(res, exp):
        if e is None:
            assert not r["found"]
        else:
            assert r["found"] and (tuple(r["position"]), r["orientation"], r["score"]) == e


def test_pack_records_matches_oracle(P, PR):
    from psacking_amd.rotations import orientation_maps
    w, order = _workload(46, 6, (60, 50, 40))
    perms, flips = orientation_maps(24)
    s = P.Session(0)
    s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    say("pack records, 6 items, 24 orientations")
    res, feats, scores = PR.pack_records(s, w.items, order, perms, flips, "cube")
    exp, tray, exp_scores, trays = _replay(w.items, order, 24, w.tray_size)
    _check(res, exp)
    np.testing.assert_array_equal(s.tray(), tray)
    np.testing.assert_array_equal(scores, exp_scores)
    bits = to_host_bits(feats)
    for t in range(len(order)):
        np.testing.assert_array_equal(bits[t], N.bf16_bits(N.tray_features(trays[t])).reshape(-1))


@pytest.mark.parametrize("topk,min_conf,resid,screen_all,extra",
                         [(4, 0.0, None, 1, ()), (2, 0.0, -1.0, 1, ()), (24, 0.0, None, 1, ()),
                          (4, 1.01, None, 1, ()), (4, 0.0, None, 0, ()),
                          (4, 0.0, None, 1, (("graph", 1),)), (4, 0.0, None, 1, (("learn_dev", 0),))])
def test_pack_learned_replays_on_oracle(P, PR, topk, min_conf, resid, screen_all, extra):
    """screen_all 1 (default): every item is screened by the column test
    before proposing; 0: only after an item that placed nothing, so proposals
    may be infeasible and the remaining candidates are searched after them.
    The default configuration runs device-driven (learn_dev), also captured
    as a hipGraph, and host-driven."""
    from psacking_amd.rotations import orientation_maps
    w, order = _workload(47, 7, (56, 48, 40))
    perms, flips = orientation_maps(24)
    nn = PR.Proposer(0, 8, PR.init_params(48))
    s = P.Session(0)
    s.set_option("screen_all", screen_all)
    for name, value in extra:
        s.set_option(name, value)
    if resid is not None:
        s.set_option("resid_limit", resid)
    s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    say("pack learned, topk", topk, "min_conf", min_conf, "resid_limit", resid, "screen_all", screen_all)
    res, props, counts = PR.pack_learned(s, nn, w.items, order, perms, flips, "cube",
                                         topk=topk, min_conf=min_conf)
    if topk >= 24 or min_conf > 1.0:
        assert np.all(props == -1) and counts[0] == 0
    else:
        # some items were searched over proposals (on the FFT or the exact path)
        assert (props >= 0).any()
    if resid is not None:
        assert counts[3] > 0 and all(r["path"] == 1 for r in res if r["found"])
    exp, tray, _, _ = _replay(w.items, order, 24, w.tray_size, proposals=props)
    _check(res, exp)
    np.testing.assert_array_equal(s.tray(), tray)


@pytest.mark.parametrize("prescreen", [16, 2, 0])
def test_pack_learned_on_a_filling_tray(P, PR, prescreen):
    """Most late items fit nowhere: the column test decides them, at their
    turn or in the batched test at the start of a round (rounds of 16 and
    of 2 items, and one round with no batched test); every placement still
    replays exactly on the oracle."""
    from psacking_amd.rotations import orientation_maps
    w, order = _workload(51, 14, (48, 40, 32))
    perms, flips = orientation_maps(24)
    nn = PR.Proposer(0, 8, PR.init_params(52))
    s = P.Session(0)
    s.set_option("prescreen", prescreen)
    s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    say("pack learned on a filling tray, prescreen", prescreen)
    res, props, counts = PR.pack_learned(s, nn, w.items, order, perms, flips, "cube", topk=4)
    say("counts", list(counts))
    assert counts[4] > 0
    exp, tray, _, _ = _replay(w.items, order, 24, w.tray_size, proposals=props)
    _check(res, exp)
    np.testing.assert_array_equal(s.tray(), tray)


def test_trained_proposer_beats_untrained(P, PR):
    """What training gains: on fresh plain-packer records (README-scale items
    from a seed never used in training, 240x123x100 tray), the committed
    weights put the best orientation in their top 4 far more often than the
    untrained network does (chance is 4/24)."""
    import os
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    assert os.path.exists(PR.WEIGHTS)
    w = W.readme(seed=99, n=60)
    vols = [int((it > 0).sum()) for it in w.items]
    order = [int(i) for i in np.argsort(vols)[::-1]]
    perms, flips = orientation_maps(24)
    s = P.Session(0)
    s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
    say("records for the proposer gain test")
    res, feats, scores = PR.pack_records(s, w.items, order, perms, flips, "cube")
    rows = [t for t in range(len(order)) if np.isfinite(scores[t]).any()]
    assert len(rows) >= 20
    nn = PR.Proposer(0, 8)
    item_rows = PR.DeviceArray((len(rows) * 24, PR.ITEM_FEATURES), np.uint16)
    for j, t in enumerate(rows):
        nn.item_features_all(w.items[order[t]], perms, flips, item_rows, j * 24)
    tray_rows = PR.DeviceArray((len(rows), PR.TRAY_FEATURES), np.uint16)
    tray_rows.gather_from(feats, rows)
    best = np.argmin(scores[rows], axis=1)
    recall = {}
    for name, params in (("trained", np.load(PR.WEIGHTS, allow_pickle=False)),
                         ("untrained", PR.init_params(7))):
        nn.set_params(params)
        z = np.concatenate([
            nn.forward(item_rows.view(b0 * 24, min(8, len(rows) - b0) * 24),
                       tray_rows.view(b0, min(8, len(rows) - b0)))
            for b0 in range(0, len(rows), 8)])
        top4 = np.argsort(-z, axis=1, kind="stable")[:, :4]
        recall[name] = float(np.mean([best[i] in top4[i] for i in range(len(rows))]))
    say("top-4 recall", recall)
    assert recall["trained"] >= recall["untrained"] + 0.1



def test_learned_device_and_host_loops_agree(P, PR):
    """The device-driven learned loop (tray features from the column-test
    masks, head rows built in place, the pick as a kernel) proposes and
    places exactly as the host-driven loop (tray features from the int32
    tray, the pick on the host), item for item."""
    from psacking_amd.rotations import orientation_maps
    w, order = _workload(53, 12, (64, 48, 40))
    perms, flips = orientation_maps(24)
    nn = PR.Proposer(0, 8, PR.init_params(54))
    out = []
    for dev in (1, 0):
        s = P.Session(0)
        s.set_option("learn_dev", dev)
        s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
        say("pack learned, learn_dev", dev)
        res, props, counts = PR.pack_learned(s, nn, w.items, order, perms, flips, "cube", topk=4)
        out.append((res, props, s.tray()))
    (r1, p1, t1), (r0, p0, t0) = out
    assert (p1 >= 0).any()
    np.testing.assert_array_equal(p1, p0)
    for a, b in zip(r1, r0):
        assert (a["found"], tuple(a["position"]), a["orientation"], a["score"]) == \
            (b["found"], tuple(b["position"]), b["orientation"], b["score"])
    np.testing.assert_array_equal(t1, t0)
