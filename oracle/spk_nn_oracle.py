"""CPU oracle of the learned orientation proposer.  TEST INFRASTRUCTURE ONLY.

The reference has no learned component (it searches every orientation,
spectral_packer/packer.py:487-505); this is the restatement of the
orientation scorer the HIP library trains and runs (psacking_amd/csrc/
spk_nn.hpp, spk_nn_host.hpp), used by tests to pin its features, forward
pass, loss and gradients.

* A record is one tray and the 24 orientations of one item: the item branch
  runs on each oriented item, the tray branch once, the head on rows
  [item_o | tray], channel 0 of its output is orientation o's logit, and the
  24 logits form one softmax.
* ``emulate=True`` rounds to bfloat16 (round to nearest even) exactly where the
  device stores bf16: input features, weights, every activation, the logit
  gradient and every activation gradient; products accumulate in float64
  (the device accumulates in fp32, so results agree to accumulation-order
  rounding).  ``emulate=False`` is the float64 network, which the finite-
  difference checks in tests/test_oracle_nn.py pin.
* Layouts are channels-last, conv weights W[cout][tap*cin + c] with
  tap = (t0*3 + t1)*3 + t2, parameters layer by layer (W then b).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Tuple

import numpy as np

ITEM_CELLS = 32
TRAY_GRID = (64, 32, 32)
N_LOGITS = 24


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def bf16_bits(x) -> np.ndarray:
    """uint16 bit patterns of bf16(x) (the device storage format)."""
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(h) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


# ---------------------------------------------------------------------------
# features
# ---------------------------------------------------------------------------
def oriented(item: np.ndarray, perm, flip) -> np.ndarray:
    """The item seen through an index map: output axis a runs along base axis
    perm[a], reversed when flip[a]."""
    g = np.transpose(np.asarray(item), perm)
    for a in range(3):
        if flip[a]:
            g = np.flip(g, axis=a)
    return np.ascontiguousarray(g)


def item_features(item: np.ndarray, perm=(0, 1, 2), flip=(0, 0, 0)) -> np.ndarray:
    """[32,32,32] occupancy of the (oriented) item at pitch 2: cell (a,b,c) =
    (# voxels > 0 among x in {2a,2a+1}, y in {2b,2b+1}, z in {2c,2c+1}) / 8;
    voxels beyond 64 per axis are outside the grid."""
    occ = (oriented(item, perm, flip) > 0)
    g = np.zeros((2 * ITEM_CELLS,) * 3, dtype=np.int32)
    sx, sy, sz = (min(s, 2 * ITEM_CELLS) for s in occ.shape)
    g[:sx, :sy, :sz] = occ[:sx, :sy, :sz]
    c = g.reshape(ITEM_CELLS, 2, ITEM_CELLS, 2, ITEM_CELLS, 2).sum(axis=(1, 3, 5))
    return (c.astype(np.float32) * np.float32(0.125)).astype(np.float32)


def heightmap(tray: np.ndarray) -> np.ndarray:
    """h[x][y] = 1 + max z with tray > 0, 0 for an empty column."""
    occ = np.asarray(tray) > 0
    nz = occ.shape[2]
    any_ = occ.any(axis=2)
    top = nz - 1 - np.argmax(occ[:, :, ::-1], axis=2)
    return np.where(any_, top + 1, 0).astype(np.int64)


def cell_ranges(n: int, G: int) -> List[Tuple[int, int]]:
    out = []
    for a in range(G):
        lo, hi = a * n // G, (a + 1) * n // G
        out.append((lo, max(hi, lo + 1)))
    return out


def tray_features(tray: np.ndarray) -> np.ndarray:
    """[64,32,32,2]: occupied fraction of each cell and fraction of the cell
    below its columns' heights, each float32(count)/float32(cells) rounded to
    bf16; cells partition each axis as [a*n//G, (a+1)*n//G) (at least one
    voxel)."""
    t = np.asarray(tray)
    occ = (t > 0).astype(np.int64)
    h = heightmap(t)
    nx, ny, nz = t.shape
    z = np.arange(nz)
    below = (z[None, None, :] < h[:, :, None]).astype(np.int64)
    rx, ry, rz = (cell_ranges(n, G) for n, G in zip(t.shape, TRAY_GRID))
    out = np.zeros(TRAY_GRID + (2,), dtype=np.float32)
    # prefix sums make each cell an O(1) box sum
    def box(P, x0, x1, y0, y1, z0, z1):
        return (P[x1, y1, z1] - P[x0, y1, z1] - P[x1, y0, z1] - P[x1, y1, z0] + P[x0, y0, z1] +
                P[x0, y1, z0] + P[x1, y0, z0] - P[x0, y0, z0])
    Po = np.zeros((nx + 1, ny + 1, nz + 1), np.int64)
    Po[1:, 1:, 1:] = occ.cumsum(0).cumsum(1).cumsum(2)
    Pb = np.zeros_like(Po)
    Pb[1:, 1:, 1:] = below.cumsum(0).cumsum(1).cumsum(2)
    for a, (x0, x1) in enumerate(rx):
        for b, (y0, y1) in enumerate(ry):
            for c, (z0, z1) in enumerate(rz):
                cnt = np.float32((x1 - x0) * (y1 - y0) * (z1 - z0))
                out[a, b, c, 0] = np.float32(box(Po, x0, x1, y0, y1, z0, z1)) / cnt
                out[a, b, c, 1] = np.float32(box(Pb, x0, x1, y0, y1, z0, z1)) / cnt
    return bf16(out)


# ---------------------------------------------------------------------------
# layers
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Layer:
    name: str
    i: Tuple[int, int, int]
    cin: int
    cout: int
    ks: int
    st: int
    pad: int
    relu: bool

    @property
    def o(self) -> Tuple[int, int, int]:
        if self.ks == 1:
            return self.i
        return tuple((d + 2 * self.pad - self.ks) // self.st + 1 for d in self.i)

    @property
    def K(self) -> int:
        return self.ks ** 3 * self.cin


def conv(name, i, cin, cout):
    return Layer(name, i, cin, cout, 3, 2, 1, True)


def dense(name, cin, cout, relu=True):
    return Layer(name, (1, 1, 1), cin, cout, 1, 1, 0, relu)


LAYERS: List[Layer] = [
    conv("I1", (32, 32, 32), 1, 16), conv("I2", (16, 16, 16), 16, 32),
    conv("I3", (8, 8, 8), 32, 48),
    conv("T1", (64, 32, 32), 2, 16), conv("T2", (32, 16, 16), 16, 32),
    conv("T3", (16, 8, 8), 32, 48),
    dense("F1", 1152, 128), dense("F2", 128, 8, relu=False),
]


def param_slices() -> Dict[str, Tuple[slice, slice]]:
    out, p = {}, 0
    for l in LAYERS:
        w = slice(p, p + l.cout * l.K)
        p += l.cout * l.K
        b = slice(p, p + l.cout)
        p += l.cout
        out[l.name] = (w, b)
    return out


N_PARAMS = sum(l.cout * l.K + l.cout for l in LAYERS)


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for l in LAYERS:
        w, _ = param_slices()[l.name]
        lim = np.sqrt(6.0 / l.K)
        p[w] = rng.uniform(-lim, lim, l.cout * l.K).astype(np.float32)
    return p


def im2col(X: np.ndarray, l: Layer) -> np.ndarray:
    """rows = output positions (b, o0, o1, o2), columns tap*cin + c."""
    B = X.shape[0]
    if l.ks == 1:
        return X.reshape(B, -1)
    o0, o1, o2 = l.o
    pad, st, ks = l.pad, l.st, l.ks
    shp = (B,) + tuple(d + 2 * pad + st for d in l.i) + (l.cin,)
    Xp = np.zeros(shp, X.dtype)
    Xp[:, pad:pad + l.i[0], pad:pad + l.i[1], pad:pad + l.i[2], :] = X
    cols = np.empty((B, o0, o1, o2, ks ** 3, l.cin), X.dtype)
    for t0 in range(ks):
        for t1 in range(ks):
            for t2 in range(ks):
                tap = (t0 * ks + t1) * ks + t2
                cols[:, :, :, :, tap, :] = Xp[:, t0:t0 + st * o0:st, t1:t1 + st * o1:st,
                                              t2:t2 + st * o2:st, :]
    return cols.reshape(B * o0 * o1 * o2, ks ** 3 * l.cin)


def col2im(D: np.ndarray, l: Layer, B: int) -> np.ndarray:
    """adjoint of im2col: scatter-add column gradients back to the input."""
    if l.ks == 1:
        return D.reshape((B,) + l.i + (l.cin,)) if l.i != (1, 1, 1) else D.reshape(B, l.cin)
    o0, o1, o2 = l.o
    pad, st, ks = l.pad, l.st, l.ks
    shp = (B,) + tuple(d + 2 * pad + st for d in l.i) + (l.cin,)
    Xp = np.zeros(shp, D.dtype)
    cols = D.reshape(B, o0, o1, o2, ks ** 3, l.cin)
    for t0 in range(ks):
        for t1 in range(ks):
            for t2 in range(ks):
                tap = (t0 * ks + t1) * ks + t2
                Xp[:, t0:t0 + st * o0:st, t1:t1 + st * o1:st, t2:t2 + st * o2:st, :] += \
                    cols[:, :, :, :, tap, :]
    return Xp[:, pad:pad + l.i[0], pad:pad + l.i[1], pad:pad + l.i[2], :]


def conv_direct(X: np.ndarray, W: np.ndarray, b: np.ndarray, l: Layer) -> np.ndarray:
    """Plain nested-loop convolution (pins im2col on small inputs)."""
    B = X.shape[0]
    o = l.o
    Y = np.zeros((B,) + o + (l.cout,), np.float64)
    Wr = W.reshape(l.cout, l.ks, l.ks, l.ks, l.cin)
    for bb in range(B):
        for p0 in range(o[0]):
            for p1 in range(o[1]):
                for p2 in range(o[2]):
                    acc = b.astype(np.float64).copy()
                    for t0 in range(l.ks):
                        for t1 in range(l.ks):
                            for t2 in range(l.ks):
                                x = (p0 * l.st - l.pad + t0, p1 * l.st - l.pad + t1,
                                     p2 * l.st - l.pad + t2)
                                if all(0 <= x[a] < l.i[a] for a in range(3)):
                                    acc += Wr[:, t0, t1, t2, :] @ X[bb, x[0], x[1], x[2], :]
                    Y[bb, p0, p1, p2] = acc
    return Y


def pool2(X: np.ndarray) -> np.ndarray:
    B, d0, d1, d2, C = X.shape
    return X.reshape(B, d0 // 2, 2, d1 // 2, 2, d2 // 2, 2, C).mean(axis=(2, 4, 6))


def unpool2(G: np.ndarray, shape) -> np.ndarray:
    B, d0, d1, d2, C = shape
    g = G.reshape(B, d0 // 2, 1, d1 // 2, 1, d2 // 2, 1, C) / 8.0
    return np.broadcast_to(g, (B, d0 // 2, 2, d1 // 2, 2, d2 // 2, 2, C)).reshape(shape)


# ---------------------------------------------------------------------------
# network
# ---------------------------------------------------------------------------
def _layer_params(params, l, emulate):
    w, b = param_slices()[l.name]
    W = np.asarray(params[w], np.float64 if not emulate else np.float32).reshape(l.cout, l.K)
    bias = np.asarray(params[b], np.float64)
    if emulate:
        W = bf16(W)
    return W.astype(np.float64), bias


def _act(v, emulate):
    return bf16(v.astype(np.float32)).astype(np.float64) if emulate else v


def forward(params, item_feat, tray_feat, emulate=True):
    """item_feat [B,24,32,32,32] (oriented items), tray_feat [B,64,32,32,2]
    -> logits [B,24], cache."""
    L = {l.name: l for l in LAYERS}
    B = tray_feat.shape[0]
    R = B * N_LOGITS
    xi = np.asarray(item_feat, np.float64).reshape(R, 32, 32, 32, 1)
    xt = np.asarray(tray_feat, np.float64).reshape(B, 64, 32, 32, 2)
    if emulate:
        xi, xt = _act(xi, True), _act(xt, True)
    cache = {"I0": xi, "T0": xt}

    def run(name, X):
        l = L[name]
        W, b = _layer_params(params, l, emulate)
        cols = im2col(X, l)
        pre = cols @ W.T + b
        y = np.maximum(pre, 0.0) if l.relu else pre
        return y.reshape((X.shape[0],) + l.o + (l.cout,)) if l.ks == 3 else y

    a = xi
    for n in ("I1", "I2", "I3"):
        a = _act(run(n, a), emulate)
        cache[n] = a
    pi = _act(pool2(a), emulate).reshape(R, -1)
    a = xt
    for n in ("T1", "T2", "T3"):
        a = _act(run(n, a), emulate)
        cache[n] = a
    pt = _act(pool2(a), emulate).reshape(B, -1)
    concat = np.concatenate([pi, np.repeat(pt, N_LOGITS, axis=0)], axis=1)
    cache["C"] = concat
    h1 = _act(run("F1", concat), emulate)
    cache["F1"] = h1
    out = run("F2", h1)
    if emulate:
        out = out.astype(np.float32).astype(np.float64)
    cache["F2"] = out
    return out[:, 0].reshape(B, N_LOGITS), cache


def loss_and_dlogits(logits, target, mask):
    """Soft-target cross entropy over the orientations whose mask bit is set;
    mean over rows; returns (loss, dlogits)."""
    B, n = logits.shape
    valid = ((np.asarray(mask, np.int64)[:, None] >> np.arange(n)[None, :]) & 1).astype(bool)
    z = np.where(valid, logits, -np.inf)
    mx = np.max(z, axis=1, keepdims=True)
    mx = np.where(np.isfinite(mx), mx, 0.0)
    e = np.where(valid, np.exp(z - mx), 0.0)
    se = e.sum(axis=1, keepdims=True)
    lse = mx + np.log(np.where(se > 0, se, 1.0))
    logp = np.where(valid, logits - lse, 0.0)
    q = np.where(valid, target, 0.0)
    loss_rows = -(q * logp).sum(axis=1)
    p = np.where(valid, np.exp(logp), 0.0)
    d = np.where(valid, (p - q) / B, 0.0)
    return loss_rows.mean(), d


def backward(params, cache, dlogits, emulate=True):
    """Gradient of the loss w.r.t. every parameter (flat, layer by layer)."""
    L = {l.name: l for l in LAYERS}
    sl = param_slices()
    g = np.zeros(N_PARAMS, np.float64)
    B = dlogits.shape[0]
    R = B * N_LOGITS
    dz = np.zeros((R, 8))
    dz[:, 0] = _act(dlogits, emulate).reshape(R)

    def wgrad(name, X, dY):
        l = L[name]
        cols = im2col(X, l)
        dY2 = dY.reshape(-1, l.cout)
        w, b = sl[name]
        g[w] = (dY2.T @ cols).reshape(-1)
        g[b] = dY2.sum(axis=0)

    def dgrad(name, dY, mask_act, rows):
        l = L[name]
        W, _ = _layer_params(params, l, emulate)
        dcols = dY.reshape(-1, l.cout) @ W
        dX = col2im(dcols, l, rows)
        if mask_act is not None:
            dX = np.where(mask_act > 0, dX, 0.0)
        return _act(dX, emulate)

    wgrad("F2", cache["F1"], dz)
    dh1 = dgrad("F2", dz, cache["F1"], R)
    wgrad("F1", cache["C"], dh1)
    dc = dgrad("F1", dh1, None, R)
    d_item = dc[:, :384]
    d_tray = _act(dc[:, 384:].reshape(B, N_LOGITS, 768).sum(axis=1), emulate)
    for branch, dp, rows in (("I", d_item, R), ("T", d_tray, B)):
        a3 = cache[branch + "3"]
        d3 = unpool2(dp.reshape((rows,) + tuple(s // 2 for s in a3.shape[1:4]) + (a3.shape[4],)),
                     a3.shape)
        d3 = _act(np.where(a3 > 0, d3, 0.0), emulate)
        wgrad(branch + "3", cache[branch + "2"], d3)
        d2 = dgrad(branch + "3", d3, cache[branch + "2"], rows)
        wgrad(branch + "2", cache[branch + "1"], d2)
        d1 = dgrad(branch + "2", d2, cache[branch + "1"], rows)
        wgrad(branch + "1", cache[branch + "0"], d1)
    return g


def loss_grad(params, item_feat, tray_feat, target, mask, emulate=True):
    logits, cache = forward(params, item_feat, tray_feat, emulate)
    loss, d = loss_and_dlogits(logits, target, mask)
    return loss, backward(params, cache, d, emulate), logits


def soft_targets(scores: np.ndarray, tau: float = 0.02) -> Tuple[np.ndarray, np.ndarray]:
    """Training targets from per-orientation best scores (+inf = infeasible):
    q_o proportional to exp(-(s_o - s_min) / (tau * (|s_min| + 1000))) over
    the feasible orientations; mask bit o set when orientation o is feasible.
    Rows with no feasible orientation get mask 0."""
    s = np.asarray(scores, np.float64)
    feas = np.isfinite(s)
    smin = np.where(feas, s, np.inf).min(axis=1, keepdims=True)
    smin = np.where(np.isfinite(smin), smin, 0.0)
    scale = tau * (np.abs(smin) + 1000.0)
    w = np.where(feas, np.exp(-(np.where(feas, s, smin) - smin) / scale), 0.0)
    tot = w.sum(axis=1, keepdims=True)
    q = np.where(tot > 0, w / np.where(tot > 0, tot, 1.0), 0.0)
    mask = (feas.astype(np.int64) << np.arange(s.shape[1])[None, :]).sum(axis=1)
    return q.astype(np.float32), mask.astype(np.int32)
