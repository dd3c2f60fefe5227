"""Learned orientation proposer: Python side of the spk_nn_* C-ABI.

No reference counterpart: the reference searches all orientations of every
item (spectral_packer/packer.py:487-505).  Here a 3-D CNN trained on the plain
packer's own per-orientation scores proposes the top-k orientations and the
exact FFT search evaluates them (``pack_learned``); the plain search over every
fitting orientation is the fallback when confidence is low or no proposal is
feasible.

Device buffers (features, targets, masks) are ``DeviceArray`` objects whose
memory libspk.so allocates (spk_dev_*), so one HIP runtime serves the whole
path; the network, its training and inference are the HIP kernels of
libspk.so.  Parameters are float32 vectors in the layout of
``spk_nn_set_params`` (layer by layer, W [cout][27*cin | cin] then b [cout]).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

ITEM_FEATURES = 32768
TRAY_FEATURES = 131072
N_LOGITS = 24
N_PARAMS = 260696
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "proposer.npy")


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(f"libspk error {rc}: {_lib.last_error()}")


class DeviceArray:
    """A device buffer of libspk.so with a numpy shape and dtype (bf16 data is
    held as uint16 bit patterns)."""

    def __init__(self, shape, dtype, device: int = 0):
        self.shape = tuple(int(v) for v in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.device = int(device)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._lib = _lib.load()
        p = C.c_void_p()
        _check(self._lib.spk_dev_alloc(self.device, max(self.nbytes, 16), C.byref(p)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a: np.ndarray, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(self._lib.spk_dev_copy(self.device, self.ptr, a.ctypes.data_as(C.c_void_p),
                                      self.nbytes, 0))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(self._lib.spk_dev_copy(self.device, out.ctypes.data_as(C.c_void_p), self.ptr,
                                      self.nbytes, 1))
        return out

    def row(self, i: int) -> C.c_void_p:
        """pointer to row i (first axis)"""
        per = self.nbytes // self.shape[0]
        return C.c_void_p(self.ptr.value + int(i) * per)

    def gather_from(self, src: "DeviceArray", idx, row0: int = 0) -> None:
        """rows row0 .. row0+len(idx)-1 of self = rows idx of src"""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        per = src.nbytes // src.shape[0]
        if per != self.nbytes // self.shape[0] or row0 < 0 or row0 + len(idx) > self.shape[0]:
            raise ValueError("row size or count mismatch")
        if len(idx) == 0:
            return
        _check(self._lib.spk_dev_gather(self.device, self.row(row0), src.ptr, per, src.shape[0],
                                        idx.ctypes.data_as(C.POINTER(C.c_int64)), len(idx)))

    def view(self, r0: int, n: int) -> "DeviceArray":
        """rows r0 .. r0+n-1 as an array that does not own its memory (keep
        this array alive while the view is used)"""
        if r0 < 0 or n <= 0 or r0 + n > self.shape[0]:
            raise ValueError("rows out of range")
        v = object.__new__(DeviceArray)
        v.shape = (n,) + self.shape[1:]
        v.dtype = self.dtype
        v.device = self.device
        v.nbytes = n * (self.nbytes // self.shape[0])
        v._lib = self._lib
        v.ptr = self.row(r0)
        v._owner = self
        return v

    def close(self) -> None:
        if getattr(self, "_owner", None) is not None:
            return
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            self._lib.spk_dev_free(self.device, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(t) -> C.c_void_p:
    return t.ptr if isinstance(t, DeviceArray) else C.c_void_p(int(t))


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases (float32, N_PARAMS)."""
    fans = [27, 432, 864, 54, 432, 864, 1152, 128]
    couts = [16, 32, 48, 16, 32, 48, 128, 8]
    rng = np.random.default_rng(seed)
    out = []
    for k, n in zip(fans, couts):
        lim = np.sqrt(6.0 / k)
        out.append(rng.uniform(-lim, lim, n * k).astype(np.float32))
        out.append(np.zeros(n, np.float32))
    p = np.concatenate(out)
    assert p.size == N_PARAMS
    return p


class Proposer:
    """One network on one device; one thread at a time."""

    def __init__(self, device: int = 0, max_batch: int = 256, params: Optional[np.ndarray] = None):
        self._lib = _lib.load()
        h = C.c_void_p()
        _check(self._lib.spk_nn_create(int(device), int(max_batch), C.byref(h)))
        self._h = h
        self.device = int(device)
        self.max_batch = int(max_batch)
        n = C.c_int64()
        _check(self._lib.spk_nn_num_params(h, C.byref(n)))
        self.n_params = int(n.value)
        if params is not None:
            self.set_params(params)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.spk_nn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(self._lib.spk_nn_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                           p.size))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        _check(self._lib.spk_nn_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                           p.size))
        return p

    def save(self, path: str = WEIGHTS) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, self.get_params())

    def load(self, path: str = WEIGHTS) -> None:
        self.set_params(np.load(path, allow_pickle=False))

    # ---- batches of device features -------------------------------------
    def forward(self, item_feat: DeviceArray, tray_feat: DeviceArray, batch: Optional[int] = None) -> np.ndarray:
        """logits [B, 24]; item_feat holds B*24 oriented item rows, tray_feat B rows."""
        B = int(tray_feat.shape[0] if batch is None else batch)
        out = np.empty((B, N_LOGITS), np.float32)
        _check(self._lib.spk_nn_forward(self._h, _ptr(item_feat), _ptr(tray_feat), B,
                                        out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def gradient(self, item_feat, tray_feat, target, mask) -> Tuple[float, np.ndarray]:
        B = int(tray_feat.shape[0])
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        _check(self._lib.spk_nn_gradient(self._h, _ptr(item_feat), _ptr(tray_feat), _ptr(target),
                                         _ptr(mask), B, g.ctypes.data_as(C.POINTER(C.c_float)),
                                         C.byref(loss)))
        return float(loss.value), g

    def train_step(self, item_feat, tray_feat, target, mask, lr: float,
                   batch: Optional[int] = None) -> float:
        B = int(tray_feat.shape[0] if batch is None else batch)
        loss = C.c_float()
        _check(self._lib.spk_nn_train_step(self._h, _ptr(item_feat), _ptr(tray_feat),
                                           _ptr(target), _ptr(mask), B, float(lr),
                                           C.byref(loss)))
        return float(loss.value)

    def item_features(self, item: np.ndarray, out: Optional[DeviceArray] = None, row: int = 0,
                      perm=None, flip=None):
        """bf16 bits [32768] of the item in one orientation (index map perm,
        flip; default as given) on the device (into row `row` of `out`)."""
        it = np.ascontiguousarray(item, dtype=np.int32)
        if out is None:
            out = DeviceArray((1, ITEM_FEATURES), np.uint16, self.device)
        o = None
        if perm is not None:
            o = _lib.SpkOrient()
            o.perm[:] = [int(v) for v in perm]
            o.flip[:] = [int(v) for v in (flip if flip is not None else (0, 0, 0))]
        _check(self._lib.spk_nn_item_features(self._h, it.ctypes.data_as(C.POINTER(C.c_int32)),
                                              *it.shape, C.byref(o) if o is not None else None,
                                              out.row(row)))
        return out

    def item_features_all(self, item: np.ndarray, perms, flips, out: DeviceArray, row0: int) -> None:
        """rows row0 .. row0+23 of `out` = the item in orientations 0..23 (a
        missing orientation repeats the base item)."""
        for o in range(N_LOGITS):
            if o < len(perms):
                self.item_features(item, out, row0 + o, perms[o], flips[o])
            else:
                self.item_features(item, out, row0 + o)


def tray_features(session, out: Optional[DeviceArray] = None, row: int = 0):
    """bf16 bits [131072] of the session's cached tray on the device."""
    if out is None:
        out = DeviceArray((1, TRAY_FEATURES), np.uint16)
    _check(_lib.load().spk_tray_features(C.c_void_p(session.handle()), out.row(row)))
    return out


class _Args:
    """spk_pack_args over numpy buffers kept alive by this object."""

    def __init__(self, items: Sequence[np.ndarray], order, perms, flips, penalty: str,
                 id_offset: int = 0):
        grids = [np.ascontiguousarray(it, dtype=np.int32) for it in items]
        for g in grids:
            if g.ndim != 3:
                raise ValueError("items must be 3-D")
        # one pointer per item array (spk_pack_args.item_ptrs): no concatenated copy
        self.grids = grids
        self.ptrs = np.array([g.__array_interface__["data"][0] for g in grids] or [0], np.uintp)
        self.dims = np.array([g.shape for g in grids], np.int32).reshape(-1)
        self.order = np.ascontiguousarray(order, dtype=np.int32)
        P = np.asarray(perms, np.int32).reshape(-1, 3)
        F = np.asarray(flips, np.int32).reshape(-1, 3)
        self.orients = (_lib.SpkOrient * len(P))()
        for o in range(len(P)):
            self.orients[o].perm[:] = [int(v) for v in P[o]]
            self.orients[o].flip[:] = [int(v) for v in F[o]]
        self.n = len(grids)
        self.results = (_lib.SpkResult * max(1, self.n))()
        a = _lib.SpkPackArgs()
        a.items = None
        a.item_ptrs = self.ptrs.ctypes.data_as(C.POINTER(C.POINTER(C.c_int32)))
        a.dims = self.dims.ctypes.data_as(C.POINTER(C.c_int32))
        a.n_items = self.n
        a.order = self.order.ctypes.data_as(C.POINTER(C.c_int32))
        a.orients = self.orients
        a.n_orient = len(P)
        a.penalty = {"pow": 0, "cube": 1}[penalty]
        a.interlocking_free = 0
        a.id_offset = int(id_offset)
        a.results = self.results
        self.args = a

    def result_dicts(self) -> List[dict]:
        out = []
        for t in range(self.n):
            r = self.results[t]
            out.append({"position": tuple(r.pos), "found": bool(r.found), "score": r.score,
                        "orientation": r.orient, "collision": r.collision,
                        "proximity": r.proximity, "residual": r.residual, "path": r.path})
        return out


def pack_records(session, items, order, perms, flips, penalty: str = "cube", id_offset: int = 0,
                 feats: Optional[DeviceArray] = None, row0: int = 0):
    """Plain greedy pack that also returns, per processed item, the tray
    features it saw (bf16 bits on the device, rows row0.. of `feats`, a new
    [n, 131072] array by default) and every orientation's best score (numpy
    [n, n_orient], +inf when infeasible)."""
    a = _Args(items, order, perms, flips, penalty, id_offset)
    if feats is None:
        feats = DeviceArray((a.n, TRAY_FEATURES), np.uint16)
    if row0 + a.n > feats.shape[0]:
        raise ValueError("feature buffer too small")
    scores = np.empty((a.n, a.args.n_orient), np.float64)
    _check(_lib.load().spk_pack_records(C.c_void_p(session.handle()), C.byref(a.args),
                                        feats.row(row0), scores.ctypes.data_as(C.POINTER(C.c_double))))
    return a.result_dicts(), feats, scores


def pack_learned(session, proposer: Proposer, items, order, perms, flips, penalty: str = "cube",
                 topk: int = 4, min_conf: float = 0.0, id_offset: int = 0):
    """Greedy pack with learned top-k orientation proposals.  Returns
    (results, proposals [n, topk] (-1 = all candidates searched), counts[5]:
    items decided on proposals, searches of all candidates, searches after
    infeasible proposals, exact host-path decisions, items the feasibility
    pass found infeasible)."""
    a = _Args(items, order, perms, flips, penalty, id_offset)
    props = np.empty((a.n, topk), np.int32)
    counts = np.zeros(5, np.int32)
    _check(_lib.load().spk_pack_learned(C.c_void_p(session.handle()), proposer.handle,
                                        C.byref(a.args), int(topk), float(min_conf),
                                        props.ctypes.data_as(C.POINTER(C.c_int32)),
                                        counts.ctypes.data_as(C.POINTER(C.c_int32))))
    return a.result_dicts(), props, counts


def soft_targets(scores: np.ndarray, tau: float = 0.02) -> Tuple[np.ndarray, np.ndarray]:
    """Training targets from per-orientation best scores (+inf = infeasible):
    q_o proportional to exp(-(s_o - s_min) / (tau * (|s_min| + 1000))) over
    the feasible orientations (sum 1); mask bit o set when orientation o is
    feasible.  Rows with no feasible orientation get mask 0 and q = 0."""
    s = np.asarray(scores, np.float64)
    feas = np.isfinite(s)
    smin = np.min(np.where(feas, s, np.inf), axis=1, keepdims=True)
    smin = np.where(np.isfinite(smin), smin, 0.0)
    scale = tau * (np.abs(smin) + 1000.0)
    w = np.where(feas, np.exp(-(np.where(feas, s, smin) - smin) / scale), 0.0)
    tot = w.sum(axis=1, keepdims=True)
    q = np.where(tot > 0, w / np.where(tot > 0, tot, 1.0), 0.0)
    mask = (feas.astype(np.int64) << np.arange(s.shape[1])[None, :]).sum(axis=1)
    return q.astype(np.float32), mask.astype(np.int32)


def load_trained(device: int = 0, path: str = WEIGHTS, max_batch: int = 64) -> Proposer:
    """The proposer with the committed trained weights."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: train with `python -m psacking_amd.train_proposer`")
    return Proposer(device, max_batch, np.load(path, allow_pickle=False))


def feasible(session, item, perms, flips):
    """Which orientations of `item` fit the session's tray with no collision:
    (column, full, residual) where full[o] = 1 when orientation o has a
    collision-free valid offset, column[o] = 0 when the column test (2x2
    column blocks, eroded core) already excludes it, residual = the FFT
    rounding distance of the full pass."""
    g = np.ascontiguousarray(item, dtype=np.int32)
    n = len(perms)
    orients = (_lib.SpkOrient * n)()
    for o in range(n):
        for a in range(3):
            orients[o].perm[a] = int(perms[o][a])
            orients[o].flip[a] = int(flips[o][a])
    column = np.zeros(n, np.int32)
    full = np.zeros(n, np.int32)
    res = C.c_float(0.0)
    ip = C.POINTER(C.c_int32)
    _check(_lib.load().spk_feasible(C.c_void_p(session.handle()), g.ctypes.data_as(ip),
                                    g.shape[0], g.shape[1], g.shape[2], orients, n,
                                    column.ctypes.data_as(ip), full.ctypes.data_as(ip),
                                    C.byref(res)))
    return column, full, float(res.value)
