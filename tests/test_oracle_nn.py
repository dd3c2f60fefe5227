"""Pins the proposer oracle (oracle/spk_nn_oracle.py) on the CPU.

* bf16 rounding against torch's float -> bfloat16 conversion,
* im2col convolution against a nested-loop convolution, col2im and unpool as
  adjoints (inner-product identities),
* the float64 backward pass against central finite differences,
* the loss gradient against finite differences of the loss,
* features against per-voxel loop definitions.
"""
import numpy as np
import pytest

from oracle import spk_nn_oracle as N


def test_param_count():
    from psacking_amd import proposer
    assert N.N_PARAMS == 260696 == proposer.N_PARAMS
    assert proposer.init_params(3).size == N.N_PARAMS


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 5000),
                        [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -1.0 - 2 ** -8, 0.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(N.bf16(x), ref)
    np.testing.assert_array_equal(N.from_bits(N.bf16_bits(x)), ref)


@pytest.mark.parametrize("i,cin,cout", [((5, 4, 6), 1, 3), ((7, 6, 5), 2, 4), ((4, 4, 4), 3, 2)])
def test_im2col_conv_matches_direct(i, cin, cout):
    rng = np.random.default_rng(1)
    l = N.Layer("x", i, cin, cout, 3, 2, 1, False)
    X = rng.normal(size=(2,) + i + (cin,))
    W = rng.normal(size=(cout, l.K))
    b = rng.normal(size=cout)
    got = (N.im2col(X, l) @ W.T + b).reshape((2,) + l.o + (cout,))
    np.testing.assert_allclose(got, N.conv_direct(X, W, b, l), rtol=1e-12, atol=1e-12)


def test_col2im_and_unpool_are_adjoints():
    rng = np.random.default_rng(2)
    l = N.Layer("x", (6, 5, 7), 3, 4, 3, 2, 1, True)
    X = rng.normal(size=(2, 6, 5, 7, 3))
    cols = N.im2col(X, l)
    D = rng.normal(size=cols.shape)
    assert np.isclose((cols * D).sum(), (X * N.col2im(D, l, 2)).sum())
    A = rng.normal(size=(2, 4, 6, 2, 5))
    G = rng.normal(size=(2, 2, 3, 1, 5))
    assert np.isclose((N.pool2(A) * G).sum(), (A * N.unpool2(G, A.shape)).sum())


def _case(seed, B=2):
    rng = np.random.default_rng(seed)
    params = N.init_params(seed).astype(np.float64)
    params += rng.normal(0, 0.01, params.size)   # non-zero biases
    item = (rng.random((B, 24, 32, 32, 32)) < 0.3) * rng.integers(1, 9, (B, 24, 32, 32, 32)) / 8.0
    tray = rng.random((B, 64, 32, 32, 2))
    scores = rng.uniform(0, 1e6, (B, 24))
    scores[rng.random((B, 24)) < 0.3] = np.inf
    q, mask = N.soft_targets(scores)
    return params, item, tray, q.astype(np.float64), mask


def test_backward_matches_finite_differences():
    params, item, tray, q, mask = _case(4, B=2)
    loss, g, _ = N.loss_grad(params, item, tray, q, mask, emulate=False)
    sl = N.param_slices()
    rng = np.random.default_rng(5)
    idx = []
    for name, (w, b) in sl.items():
        # F2's channel 0 is the logit; channels 1-7 carry no gradient
        hi = b.start + 1 if name == "F2" else b.stop
        idx += list(rng.integers(w.start, w.stop, 3)) + [b.start + int(rng.integers(0, hi - b.start))]
    eps = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        l1 = N.loss_and_dlogits(N.forward(p1, item, tray, False)[0], q, mask)[0]
        l2 = N.loss_and_dlogits(N.forward(p2, item, tray, False)[0], q, mask)[0]
        fd = (l1 - l2) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_loss_gradient_matches_finite_differences():
    from scipy.special import log_softmax
    rng = np.random.default_rng(6)
    z = rng.normal(size=(3, 24))
    scores = rng.uniform(0, 100, (3, 24))
    scores[:, ::3] = np.inf
    q, mask = N.soft_targets(scores)
    loss, d = N.loss_and_dlogits(z, q, mask)
    valid = np.isfinite(scores)
    ref = np.mean([-(q[b][valid[b]] * log_softmax(z[b][valid[b]])).sum() for b in range(3)])
    assert np.isclose(loss, ref)
    eps = 1e-6
    for b, o in [(0, 1), (1, 2), (2, 4), (0, 3)]:
        z1, z2 = z.copy(), z.copy()
        z1[b, o] += eps
        z2[b, o] -= eps
        fd = (N.loss_and_dlogits(z1, q, mask)[0] - N.loss_and_dlogits(z2, q, mask)[0]) / (2 * eps)
        assert abs(fd - d[b, o]) < 1e-8
    assert np.all(d[:, ::3] == 0)


def test_soft_targets():
    s = np.array([[5.0, 3.0, np.inf, 3.0], [np.inf] * 4])
    q, mask = N.soft_targets(s)
    assert mask[0] == 0b1011 and mask[1] == 0
    assert np.isclose(q[0].sum(), 1.0) and q[0, 2] == 0 and q[0, 1] == q[0, 3] > q[0, 0]
    assert np.all(q[1] == 0)


def test_item_features_definition():
    rng = np.random.default_rng(7)
    item = (rng.random((37, 20, 70)) < 0.4).astype(np.int32) * rng.integers(1, 4, (37, 20, 70))
    f = N.item_features(item)
    for a, b, c in [(0, 0, 0), (5, 3, 10), (18, 9, 31), (31, 31, 31), (10, 2, 20)]:
        cnt = 0
        for x in (2 * a, 2 * a + 1):
            for y in (2 * b, 2 * b + 1):
                for z in (2 * c, 2 * c + 1):
                    if x < 37 and y < 20 and z < 70 and z < 64 and item[x, y, z] > 0:
                        cnt += 1
        assert f[a, b, c] == cnt / 8


@pytest.mark.parametrize("shape", [(70, 40, 35), (50, 20, 20)])
def test_tray_features_definition(shape):
    rng = np.random.default_rng(8)
    tray = (rng.random(shape) < 0.2).astype(np.int32) * 3
    f = N.tray_features(tray)
    h = np.zeros(shape[:2], int)
    for x in range(shape[0]):
        for y in range(shape[1]):
            zs = np.nonzero(tray[x, y])[0]
            h[x, y] = zs.max() + 1 if len(zs) else 0
    np.testing.assert_array_equal(N.heightmap(tray), h)
    rx, ry, rz = (N.cell_ranges(n, G) for n, G in zip(shape, N.TRAY_GRID))
    for a, b, c in [(0, 0, 0), (63, 31, 31), (20, 7, 13), (41, 30, 2)]:
        occ = below = cnt = 0
        for x in range(*rx[a]):
            for y in range(*ry[b]):
                for z in range(*rz[c]):
                    cnt += 1
                    occ += tray[x, y, z] > 0
                    below += z < h[x, y]
        assert f[a, b, c, 0] == N.bf16(np.float32(occ) / np.float32(cnt))
        assert f[a, b, c, 1] == N.bf16(np.float32(below) / np.float32(cnt))


def test_product_soft_targets_match_oracle():
    from psacking_amd import proposer
    rng = np.random.default_rng(9)
    s = rng.uniform(0, 5e6, (50, 24))
    s[rng.random((50, 24)) < 0.4] = np.inf
    s[3] = np.inf
    q1, m1 = proposer.soft_targets(s)
    q2, m2 = N.soft_targets(s)
    np.testing.assert_array_equal(q1, q2)
    np.testing.assert_array_equal(m1, m2)


def test_oriented_items_follow_the_orientation_order():
    from oracle import spk_oracle as O
    from psacking_amd.rotations import orientation_maps
    rng = np.random.default_rng(10)
    item = rng.integers(0, 3, (5, 7, 4)).astype(np.int32)
    perms, flips = orientation_maps(24)
    ref = O.get_orientations(item, 24)
    for o in range(24):
        np.testing.assert_array_equal(N.oriented(item, perms[o], flips[o]), np.asarray(ref[o]))
