"""CPU oracle for the spectral-packing hot path.  TEST INFRASTRUCTURE ONLY.

This module is a plain numpy/scipy restatement of the reference algorithm
(Vrroom/psacking).  It is the checker for the HIP implementation in
``psacking_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path never
calls it and must fail loudly when the HIP library is missing.

Every function cites the reference file:line whose behaviour it restates.
Semantics pinned (see tests/test_oracle.py for the pins of each function):

* correlation ``r[q] = sum_y a[q+y] * b[y]`` truncated to q in [0, N)
  (spectral_packing/packing.cpp:225-239, flip -> conv -> flip), pinned by the
  reference KATs (tests/test_core.py:57-91) and brute force on small inputs;
* truncated convolution (spectral_packing/fft3.cu:643-715), KATs
  tests/test_core.py:17-51;
* L1 distance with sources ``v == 1`` and fill ``Nx+Ny+Nz+10``
  (spectral_packing/fft3.cu:887-1004), pinned by the KATs
  tests/test_core.py:120-152 and a BFS restated from
  tests/test_distance_1d.py:14-35;
* score ``double(p) + P * pow(k/Nz, 3)`` (packing.cpp:620-641) or
  ``((hp*qz)*qz)*qz`` (fft3.cu:147-175), argmin with the lowest linear index
  on ties (packing.cpp:627-638 strict ``<`` in (i,j,k) order);
* interlocking-free mask: BFS fixpoint restated from
  tests/test_flood_fill.py:67-189 (the reference GPU loop is capped at
  nx+ny+nz sweeps, fft3.cu:380-389; the fixpoint is what its own tests
  check);
* greedy driver (spectral_packer/packer.py:366-544) and orientation order
  (spectral_packer/rotations.py:159-268).

Functions with no reference counterpart (Euclidean distance mode, occupied
sources) are pinned against scipy.ndimage.distance_transform_edt and brute
force.
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np
import scipy.fft as sfft
from scipy import ndimage

INF_BOUND = 1000000000          # constants.h:21 (INF)
HEIGHT_PENALTY = 100000000.0    # constants.h:7 (P)
INVALID_SCORE = 1e30            # fft3.cu:163

# ---------------------------------------------------------------------------
# grid helpers
# ---------------------------------------------------------------------------


def as_grid(a) -> np.ndarray:
    """int32 C-order copy, like the forcecast converters (bindings.cpp:72-97)."""
    a = np.asarray(a)
    if a.ndim != 3:
        raise RuntimeError(f"Input array must be 3-dimensional, got {a.ndim}D")
    return np.ascontiguousarray(a.astype(np.int32))


def get_bounds(grid) -> Tuple[Tuple[int, int, int], Tuple[int, int, int]]:
    """Bounds of voxels > 0; empty -> (INF..),(-INF..) (voxelGrid.cpp:134-148)."""
    g = as_grid(grid)
    idx = np.argwhere(g > 0)
    if len(idx) == 0:
        return (INF_BOUND,) * 3, (-INF_BOUND,) * 3
    lo = tuple(int(v) for v in idx.min(axis=0))
    hi = tuple(int(v) for v in idx.max(axis=0))
    return lo, hi


def make_tight(grid) -> np.ndarray:
    """Crop to the bounds of voxels > 0 (voxelGrid.cpp:164-178); empty -> ValueError."""
    g = as_grid(grid)
    lo, hi = get_bounds(g)
    if hi[0] == -INF_BOUND:
        raise ValueError("(make_voxel_grid_tight) Voxel grid doesn't have any occupied cells")
    return g[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1].copy()


def pad_to(a: np.ndarray, shape) -> np.ndarray:
    """Zero pad at the high end (voxelGrid.cpp:263-277 padto3d_flat)."""
    out = np.zeros(shape, dtype=np.int64)
    s = tuple(min(a.shape[i], shape[i]) for i in range(3))
    out[:s[0], :s[1], :s[2]] = a[:s[0], :s[1], :s[2]]
    return out


# ---------------------------------------------------------------------------
# exact correlation / convolution
# ---------------------------------------------------------------------------


def corr_bruteforce(a, b) -> np.ndarray:
    """r[q] = sum_y a[q+y]*b[y] for q in a's box, y in b's box, q+y in a's box.

    Direct integer sum (int64); the closed form that flip->conv->flip
    (packing.cpp:225-239) computes.  Cost O(|a| * nnz(b)).
    """
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    out = np.zeros(a.shape, dtype=np.int64)
    N = a.shape
    for y in np.argwhere(b != 0):
        w = b[tuple(y)]
        ext = tuple(N[d] - y[d] for d in range(3))
        if min(ext) <= 0:
            continue
        out[:ext[0], :ext[1], :ext[2]] += w * a[y[0]:y[0] + ext[0], y[1]:y[1] + ext[1], y[2]:y[2] + ext[2]]
    return out


def _fft_exact_linear(a: np.ndarray, b: np.ndarray, shape, corr: bool) -> np.ndarray:
    """Exact integer linear correlation/convolution via float64 FFT + rounding.

    Sizes are padded so that no circular wrap reaches the kept region; the
    rounding residual is checked so an inexact FP64 result raises instead of
    returning a wrong integer.
    """
    P = tuple(int(sfft.next_fast_len(a.shape[d] + b.shape[d])) for d in range(3))
    A = sfft.rfftn(a.astype(np.float64), P)
    B = sfft.rfftn(b.astype(np.float64), P)
    if corr:
        r = sfft.irfftn(A * np.conj(B), P)
    else:
        r = sfft.irfftn(A * B, P)
    r = r[:shape[0], :shape[1], :shape[2]]
    ri = np.rint(r)
    resid = float(np.abs(r - ri).max()) if r.size else 0.0
    if resid > 0.05:
        raise ArithmeticError(f"oracle FP64 FFT not exact (residual {resid})")
    return ri.astype(np.int64)


def corr_exact(a, b) -> np.ndarray:
    """Exact r[q] = sum_y a[q+y]*b[y], q over a's box (int64).

    Brute force for small work, FP64 FFT (checked residual) otherwise.
    """
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    if a.size * max(1, int(np.count_nonzero(b))) <= 4_000_000:
        return corr_bruteforce(a, b)
    return _fft_exact_linear(a, b, a.shape, corr=True)


def dft_corr3(a, b) -> np.ndarray:
    """Reference dft_corr3 (packing.cpp:225-239): flip(conv(flip(a), b)).

    Equals r[q] = sum_y a[q+y] b[y], truncated to a's box.  Both grids must
    have the same shape (fft3.cu:647-648 raises otherwise).
    """
    a = as_grid(a)
    b = as_grid(b)
    if a.shape != b.shape:
        raise RuntimeError("Input grids must be of the same size for convolution")
    return corr_exact(a, b).astype(np.int32)


def dft_conv3(a, b) -> np.ndarray:
    """Reference dft_conv3 (fft3.cu:643-715): first N taps of the linear convolution."""
    a = as_grid(a)
    b = as_grid(b)
    if a.shape != b.shape:
        raise RuntimeError("Input grids must be of the same size for convolution")
    # conv(a,b)[q] = corr(flip(a), b) flipped
    fa = a[::-1, ::-1, ::-1]
    r = corr_exact(fa, b)[::-1, ::-1, ::-1]
    return np.ascontiguousarray(r).astype(np.int32)


def collision_grid(tray, item) -> np.ndarray:
    """Reference collision_grid (packing.cpp:241-251 via bindings.cpp:361-372).

    Item padded to tray size, ID-weighted correlation, out-of-bound offsets
    (q + hi > N-1 on any axis, hi from voxels > 0) become max(c, 1).
    """
    t = as_grid(tray)
    it = as_grid(item)
    it = pad_to(it, t.shape)
    c = corr_exact(t, it)
    _, hi = get_bounds(it)
    oob = _oob_mask(t.shape, hi)
    c[oob] = np.maximum(c[oob], 1)
    return c.astype(np.int32)


def _oob_mask(shape, hi) -> np.ndarray:
    N = shape
    i = np.arange(N[0])[:, None, None]
    j = np.arange(N[1])[None, :, None]
    k = np.arange(N[2])[None, None, :]
    ok = (i + hi[0] <= N[0] - 1) & (j + hi[1] <= N[1] - 1) & (k + hi[2] <= N[2] - 1)
    return ~np.broadcast_to(ok, N)


# ---------------------------------------------------------------------------
# distance fields
# ---------------------------------------------------------------------------


def distance_fill(shape) -> int:
    """Unreached value Nx+Ny+Nz+10 (fft3.cu:944-949)."""
    return int(shape[0] + shape[1] + shape[2] + 10)


def calculate_distance(grid) -> np.ndarray:
    """Reference calculate_distance (fft3.cu:944-1004).

    L1 (Manhattan) distance to the nearest voxel whose value is exactly 1,
    capped at Nx+Ny+Nz+10 (the init value; sweeps only ever lower it).
    """
    return distance_field(grid, metric="l1", sources="eq1")


def _source_mask(g: np.ndarray, sources: str) -> np.ndarray:
    if sources == "eq1":
        return g == 1
    if sources == "nonzero":
        return g != 0
    raise ValueError(f"unknown sources {sources!r}")


def distance_field(grid, metric: str = "l1", sources: str = "eq1") -> np.ndarray:
    """Distance field used as the proximity kernel.

    metric "l1": taxicab distance (reference, fft3.cu:887-1004; sweeps are the
    exact separable min-plus with |d|).  metric "edt": round-to-nearest of the
    exact Euclidean distance (no reference counterpart; integer so the
    proximity correlation stays an exact integer).  Both capped at
    Nx+Ny+Nz+10, the fill value for grids without sources.
    """
    g = as_grid(grid)
    fill = distance_fill(g.shape)
    src = _source_mask(g, sources)
    if not src.any():
        return np.full(g.shape, fill, dtype=np.int32)
    if metric == "l1":
        d = ndimage.distance_transform_cdt(~src, metric="taxicab").astype(np.int64)
    elif metric == "edt":
        d2 = _squared_edt(src)
        d = np.rint(np.sqrt(d2.astype(np.float64))).astype(np.int64)
    else:
        raise ValueError(f"unknown metric {metric!r}")
    return np.minimum(d, fill).astype(np.int32)


def _squared_edt(src: np.ndarray) -> np.ndarray:
    """Exact squared Euclidean distance to the nearest source (int64).

    scipy's EDT returns float distances; squaring its index-based feature
    transform gives exact integers.
    """
    _, ind = ndimage.distance_transform_edt(~src, return_indices=True)
    coords = np.indices(src.shape)
    return ((coords - ind).astype(np.int64) ** 2).sum(axis=0)


def distance_bfs(grid, sources: str = "eq1") -> np.ndarray:
    """6-connected BFS distance (tests/test_distance_1d.py:14-35 restated).

    Pure Python, small grids only; the pin for ``distance_field(metric='l1')``.
    """
    g = as_grid(grid)
    N = g.shape
    fill = distance_fill(N)
    d = np.full(N, fill, dtype=np.int64)
    q = deque()
    for s in np.argwhere(_source_mask(g, sources)):
        d[tuple(s)] = 0
        q.append(tuple(int(v) for v in s))
    while q:
        x = q.popleft()
        for ax in range(3):
            for st in (-1, 1):
                y = list(x)
                y[ax] += st
                if 0 <= y[ax] < N[ax]:
                    y = tuple(y)
                    if d[y] > d[x] + 1:
                        d[y] = d[x] + 1
                        q.append(y)
    return d.astype(np.int32)


# ---------------------------------------------------------------------------
# score / argmin / search
# ---------------------------------------------------------------------------


def penalty_table(nz: int, mode: str = "pow", hp: float = HEIGHT_PENALTY) -> np.ndarray:
    """Height penalty per k.

    mode "pow":    P * pow(k/Nz, 3.0)       (packing.cpp:630-631)
    mode "cube":   ((hp*qz)*qz)*qz          (fft3.cu:173, left-to-right)
    Computed in IEEE double exactly as the reference expressions evaluate.
    """
    import math
    out = np.empty(nz, dtype=np.float64)
    for k in range(nz):
        qz = (k + 0.0) / (nz + 0.0)
        if mode == "pow":
            out[k] = hp * math.pow(qz, 3.0)
        elif mode == "cube":
            out[k] = ((hp * qz) * qz) * qz
        else:
            raise ValueError(mode)
    return out


@dataclass
class SearchResult:
    position: Tuple[int, int, int]
    found: bool
    score: float
    collision: Optional[int] = None
    proximity: Optional[int] = None


def accept_bound(penalty: str) -> float:
    """Scores at or above this bound count as not found.

    "pow" (the 1-orientation CPU path): the loop starts at bestVal = INF = 1e9
    and takes only strictly smaller scores (packing.cpp:626-637,
    constants.h:21).  "cube" (the cached GPU path): found iff the best score
    is < 1e29 (fft3.cu:1361, 1447).
    """
    return float(INF_BOUND) if penalty == "pow" else 1e29


def _pick(c, p, hi, penalty: str, mask=None) -> SearchResult:
    """Score, validity and argmin over the tray box from exact c and p."""
    N = c.shape
    valid = (c == 0) & ~_oob_mask(N, hi)
    if mask is not None:
        valid &= (np.asarray(mask) == 1)
    pen = penalty_table(N[2], penalty)
    score = p.astype(np.float64) + pen[None, None, :]
    score = np.where(valid, score, np.inf)
    flat = int(np.argmin(score))          # first occurrence = lowest linear index
    if not score.flat[flat] < accept_bound(penalty):
        return SearchResult((-1, -1, -1), False, 0.0)
    pos = tuple(int(v) for v in np.unravel_index(flat, N))
    return SearchResult(pos, True, float(score.flat[flat]), int(c.flat[flat]), int(p.flat[flat]))


def search(item, tray, phi=None, penalty: str = "pow", mask=None) -> SearchResult:
    """One-orientation search (packing.cpp:590-642 / fft3.cu:1299-1373).

    Valid offsets: q + hi <= N-1 per axis, collision c[q] == 0 (ID-weighted
    correlation of the raw tray with the raw item) and, if given, mask[q]==1.
    Score = double(p[q]) + penalty[k]; the minimum with the lowest linear
    index wins, if it is below accept_bound(penalty).  Not found ->
    ((-1,-1,-1), False, 0.0).
    """
    t = as_grid(tray)
    it = as_grid(item)
    if any(it.shape[d] > t.shape[d] for d in range(3)):
        return SearchResult((-1, -1, -1), False, 0.0)
    if phi is None:
        phi = calculate_distance(t)
    phi = as_grid(phi)
    itp = pad_to(it, t.shape)
    _, hi = get_bounds(itp)
    c = corr_exact(t, itp)
    p = corr_exact(phi, itp)
    return _pick(c, p, hi, penalty, mask)


# ---------------------------------------------------------------------------
# many searches against one tray state (large trays)
# ---------------------------------------------------------------------------


class TrayState:
    """FP64 spectra of one tray state, for searches that read valid offsets.

    A valid offset q has q + hi <= N - 1 on every axis, so every item voxel
    lands at q + y <= N - 1 < P: the circular correlation on P >= N (the
    smallest fast size) equals the linear one at every valid offset, which is
    all that search() and the interlocking mask read.  The tray and field
    spectra are computed once per state instead of once per call, which is
    what makes the oracle usable at 240x123x100 with 24 orientations.
    search_state / interlocking_mask_state equal search /
    interlocking_free_positions (pinned in tests/test_oracle.py).
    """

    def __init__(self, tray, phi):
        self.tray = as_grid(tray)
        self.phi = as_grid(phi)
        self.N = self.tray.shape
        self.P = tuple(int(sfft.next_fast_len(int(n), real=True)) for n in self.N)
        self.T = sfft.rfftn(self.tray.astype(np.float64), self.P, workers=-1)
        self.F = sfft.rfftn(self.phi.astype(np.float64), self.P, workers=-1)

    def _round(self, spec) -> np.ndarray:
        N = self.N
        r = sfft.irfftn(spec, self.P, workers=-1)[:N[0], :N[1], :N[2]]
        ri = np.rint(r)
        resid = float(np.abs(r - ri).max()) if r.size else 0.0
        if resid > 0.05:
            raise ArithmeticError(f"oracle FP64 FFT not exact (residual {resid})")
        return ri.astype(np.int64)

    def correlate(self, item, want_p: bool = True):
        """(c, p, hi): tray and field correlations with the item (int64 over
        the tray box, exact at valid offsets) and the item's upper bounds."""
        it = as_grid(item)
        _, hi = get_bounds(it)
        I = np.conj(sfft.rfftn(it.astype(np.float64), self.P, workers=-1))
        c = self._round(self.T * I)
        p = self._round(self.F * I) if want_p else None
        return c, p, hi


def search_state(state: TrayState, item, penalty: str = "cube", mask=None) -> SearchResult:
    """search(item, state.tray, state.phi, penalty, mask) on cached spectra."""
    it = as_grid(item)
    if any(it.shape[d] > state.N[d] for d in range(3)):
        return SearchResult((-1, -1, -1), False, 0.0)
    c, p, hi = state.correlate(it)
    return _pick(c, p, hi, penalty, mask)


def interlocking_mask_state(state: TrayState, obj) -> np.ndarray:
    """interlocking_free_positions(state.tray, obj) on cached spectra."""
    c, _, _ = state.correlate(obj, want_p=False)
    free, seed = _if_labels_from(c, as_grid(obj))
    return _reach(free, seed)


def search_batch(orientations: Sequence[np.ndarray], tray, phi, penalty: str = "cube",
                 interlocking_free: bool = False):
    """fft_search_batch[_interlocking_free] (fft3.cu:1549-1632).

    Orientations larger than the tray are skipped; a later orientation
    replaces the best only on a strictly smaller score.  Returns
    (pos, found, score, orient_idx) with orient_idx -1 when nothing fits.
    """
    t = as_grid(tray)
    best = ((-1, -1, -1), False, 1e30, -1)
    for o, it in enumerate(orientations):
        it = as_grid(it)
        if any(it.shape[d] > t.shape[d] for d in range(3)):
            continue
        mask = interlocking_free_positions(t, it) if interlocking_free else None
        r = search(it, t, phi, penalty=penalty, mask=mask)
        if r.found and r.score < best[2]:
            best = (r.position, True, r.score, o)
    if not best[1]:
        return (-1, -1, -1), False, 1e30, -1
    return best


def place_in_tray(item, tray, position, item_id: int = 1) -> np.ndarray:
    """Reference place_in_tray (packing.cpp:773-788).

    Writes item_id where item > 0 and tray == 0 inside the tray; raises
    RuntimeError if any item > 0 voxel lands on a tray > 0 voxel.  Item parts
    beyond the tray are ignored.
    """
    t = as_grid(tray).copy()
    it = as_grid(item)
    a, b, c = (int(v) for v in position)
    N = t.shape
    x0, y0, z0 = max(a, 0), max(b, 0), max(c, 0)
    x1 = min(N[0], a + it.shape[0])
    y1 = min(N[1], b + it.shape[1])
    z1 = min(N[2], c + it.shape[2])
    if x1 <= x0 or y1 <= y0 or z1 <= z0:
        return t
    sub = t[x0:x1, y0:y1, z0:z1]
    isub = it[x0 - a:x1 - a, y0 - b:y1 - b, z0 - c:z1 - c]
    clash = (sub > 0) & (isub > 0)
    if clash.any():
        ci = np.argwhere(clash)[0] + np.array([x0, y0, z0])
        raise RuntimeError(
            "(place_in_tray): Item to be placed intersects with contents of tray at position "
            f"({ci[0]}, {ci[1]}, {ci[2]})")
    w = (sub == 0) & (isub > 0)
    sub[w] = item_id
    return t


# ---------------------------------------------------------------------------
# coarse infeasibility test (no reference counterpart; exact pruning)
# ---------------------------------------------------------------------------


def coarse_core(item, s: int = 2) -> np.ndarray:
    """Cells Y (on the s-grid of the item) whose fine block may sit anywhere
    in the (2s-1)^3 neighbourhood of s*Y without leaving the item: K[Y] = 1
    iff every voxel of [sY-(s-1), sY+(s-1)]^3 is an item voxel (> 0)."""
    occ = as_grid(item) > 0
    er = ndimage.binary_erosion(np.pad(occ, s), structure=np.ones((2 * s - 1,) * 3),
                                border_value=0)[s:-s, s:-s, s:-s]
    return er[::s, ::s, ::s].astype(np.int32)


def coarse_tray(tray, s: int = 2) -> np.ndarray:
    """1 for every s^3 block of the tray holding an occupied voxel (blocks
    past the tray's end hold only the voxels inside it)."""
    t = as_grid(tray) != 0
    n = np.array(t.shape)
    nc = -(-n // s)
    pad = np.zeros(nc * s, bool)
    pad[:n[0], :n[1], :n[2]] = t
    return pad.reshape(nc[0], s, nc[1], s, nc[2], s).any(axis=(1, 3, 5)).astype(np.int32)


def coarse_columns(tray) -> np.ndarray:
    """Occupancy of the tray's 2x2 column blocks at full z resolution:
    C[X, Y, z] = 1 if any voxel (2X+u, 2Y+v, z), u, v in {0, 1}, inside the
    tray is occupied."""
    t = as_grid(tray) != 0
    nx, ny, nz = t.shape
    pad = np.zeros((nx + nx % 2, ny + ny % 2, nz), bool)
    pad[:nx, :ny] = t
    return pad.reshape(pad.shape[0] // 2, 2, pad.shape[1] // 2, 2, nz).any(axis=(1, 3))


def column_core(item) -> np.ndarray:
    """K[X, Y, z] = 1 iff the item has a voxel (> 0) at (2X+dx, 2Y+dy, z) for
    every dx, dy in {-1, 0, 1} (an x-y erosion by a 3x3 block, sampled at even
    x and y)."""
    occ = as_grid(item) > 0
    er = ndimage.binary_erosion(np.pad(occ, ((1, 1), (1, 1), (0, 0))),
                                structure=np.ones((3, 3, 1)), border_value=0)[1:-1, 1:-1, :]
    return er[::2, ::2, :]


def column_infeasible(item, tray) -> bool:
    """True only if the (oriented) item has no collision-free in-bounds offset.

    A placement at (2*Qx + rx, 2*Qy + ry, qz), rx, ry in {0, 1}, covers for
    every cell (X, Y, z) of column_core(item) the whole column block
    (Qx + X, Qy + Y) at height qz + z, so a collision-free one needs
    coarse_columns(tray)[Qx + X, Qy + Y, qz + z] = 0 there.  If every
    (Qx, Qy, qz) of a valid offset meets an occupied block, the item fits
    nowhere.  Never rejects a feasible item; may miss infeasible ones."""
    g = as_grid(item)
    t = as_grid(tray)
    n = np.array(t.shape)
    lo, hi = get_bounds(g)
    if hi[0] < 0:
        return False
    h = np.array(hi)
    if np.any(h > n - 1):
        return True
    K = column_core(g)
    if not K.any():
        return False
    C = coarse_columns(t).astype(np.int64)
    qx, qy = (n[0] - 1 - h[0]) // 2, (n[1] - 1 - h[1]) // 2
    qz = n[2] - 1 - h[2]
    Kp = np.zeros(C.shape, np.int64)
    kx, ky, kz = (min(K.shape[a], C.shape[a]) for a in range(3))
    Kp[:kx, :ky, :kz] = K[:kx, :ky, :kz]
    c = corr_exact(C, Kp)
    return not (c[:qx + 1, :qy + 1, :qz + 1] == 0).any()


def coarse_infeasible(item, tray, s: int = 2) -> bool:
    """True only if the (oriented) item has no collision-free in-bounds offset.

    A placement at q = s*Q + r (0 <= r < s per axis) covers, for every cell Y of
    coarse_core(item), the whole tray block Q + Y, so a collision-free one
    needs coarse_tray(tray)[Q + Y] = 0 for all those Y: a coarse correlation
    that is positive at every coarse offset Q = floor(q / s) of a valid q
    proves the item infeasible.  The test can miss infeasible items, never
    the reverse."""
    g = as_grid(item)
    t = as_grid(tray)
    n = np.array(t.shape)
    lo, hi = get_bounds(g)
    if hi[0] < 0:
        return False
    h = np.array(hi)
    if np.any(h > n - 1):
        return True
    K = coarse_core(g, s)
    if not K.any():
        return False
    B = coarse_tray(t, s)
    qmax = (n - 1 - h) // s           # largest coarse offset of a valid fine offset
    # core cells past the coarse tray would need item voxels past the tray
    K = pad_to(K[:B.shape[0], :B.shape[1], :B.shape[2]], B.shape)
    c = corr_exact(B, K)
    return not (c[:qmax[0] + 1, :qmax[1] + 1, :qmax[2] + 1] == 0).any()


# ---------------------------------------------------------------------------
# interlocking-free mask
# ---------------------------------------------------------------------------


def _if_labels_from(coll: np.ndarray, o: np.ndarray):
    """Free offsets (collision 0, in bounds) and seeds from a collision grid.

    Object max from voxels != 0, 0 for an empty object (the reference's own
    test oracle treats an empty object as fitting everywhere,
    tests/test_flood_fill.py:80-83); seeds are free cells on i|j|k == 0 or
    == max(0, N-1-objmax) (fft3.cu:300-343).
    """
    nz_idx = np.argwhere(o != 0)
    omax = nz_idx.max(axis=0) if len(nz_idx) else np.zeros(3, dtype=np.int64)
    N = coll.shape
    i = np.arange(N[0])[:, None, None]
    j = np.arange(N[1])[None, :, None]
    k = np.arange(N[2])[None, None, :]
    oob = (i + omax[0] >= N[0]) | (j + omax[1] >= N[1]) | (k + omax[2] >= N[2])
    free = (coll == 0) & ~oob
    ex = [max(0, N[d] - 1 - int(omax[d])) for d in range(3)]
    seed = (i == 0) | (j == 0) | (k == 0) | (i == ex[0]) | (j == ex[1]) | (k == ex[2])
    seed = free & np.broadcast_to(seed, N)
    return free, seed


def _if_labels(tray, obj):
    t = as_grid(tray)
    o = as_grid(obj)
    coll = corr_exact(t, pad_to(o, t.shape))
    return _if_labels_from(coll, o)


def _reach(free: np.ndarray, seed: np.ndarray) -> np.ndarray:
    """Free cells 6-connected to a seed (connected-component labelling)."""
    lab, n = ndimage.label(free, structure=ndimage.generate_binary_structure(3, 1))
    if n == 0:
        return np.zeros(free.shape, dtype=np.int32)
    keep = np.zeros(n + 1, dtype=bool)
    keep[np.unique(lab[seed])] = True
    keep[0] = False
    return keep[lab].astype(np.int32)


def interlocking_free_positions(tray, obj) -> np.ndarray:
    """Interlocking-free mask (bindings.cpp:561-648, fft3.cu:234-402).

    Free = collision 0 and in bounds (obj max from voxels != 0, 0 for an empty
    object); seeds = free cells on i|j|k == 0 or == max(0, N-1-objmax); mask =
    free cells 6-connected to a seed (uncapped fixpoint, as the reference's
    own oracle tests/test_flood_fill.py:67-189 computes).
    """
    free, seed = _if_labels(tray, obj)
    return _reach(free, seed)


def interlocking_free_positions_bfs(tray, obj) -> np.ndarray:
    """Queue BFS restated from tests/test_flood_fill.py:67-189 (small grids)."""
    free, seed = _if_labels(tray, obj)
    N = free.shape
    reached = np.zeros(N, dtype=bool)
    q = deque()
    for s in np.argwhere(seed):
        s = tuple(int(v) for v in s)
        reached[s] = True
        q.append(s)
    while q:
        x = q.popleft()
        for ax in range(3):
            for st in (-1, 1):
                y = list(x)
                y[ax] += st
                if 0 <= y[ax] < N[ax]:
                    y = tuple(y)
                    if free[y] and not reached[y]:
                        reached[y] = True
                        q.append(y)
    return reached.astype(np.int32)


# ---------------------------------------------------------------------------
# orientations
# ---------------------------------------------------------------------------


def _rx(g):
    return np.rot90(g, k=1, axes=(1, 2))   # rotations.py:141-144


def _ry(g):
    return np.rot90(g, k=1, axes=(2, 0))   # rotations.py:147-150


def _rz(g):
    return np.rot90(g, k=1, axes=(0, 1))   # rotations.py:153-156


def get_orientations(grid, num_orientations: int = 1) -> List[np.ndarray]:
    """Orientation list in the reference order (rotations.py:159-268)."""
    g = np.asarray(grid)
    if num_orientations == 1:
        return [g]
    if num_orientations == 4:
        out, h = [], g.copy()
        for _ in range(4):
            out.append(h.copy())
            h = _rz(h)
        return out
    if num_orientations == 6:
        return [g, _rx(g), _rx(_rx(g)), _rx(_rx(_rx(g))), _ry(g), _ry(_ry(_ry(g)))]
    if num_orientations == 24:
        out = []
        for base in (g, _rx(g), _rx(_rx(g)), _rx(_rx(_rx(g))), _ry(g), _ry(_ry(_ry(g)))):
            h = base.copy()
            for _ in range(4):
                out.append(h.copy())
                h = _rz(h)
        return out
    raise ValueError(f"num_orientations must be 1, 4, 6, or 24, got {num_orientations}")


# ---------------------------------------------------------------------------
# greedy driver
# ---------------------------------------------------------------------------


@dataclass
class OraclePlacement:
    item_index: int
    position: Optional[Tuple[int, int, int]]
    score: Optional[float]
    success: bool
    volume: int = 0
    orientation_index: int = 0


@dataclass
class OracleResult:
    tray: np.ndarray
    placements: List[OraclePlacement] = field(default_factory=list)
    num_placed: int = 0
    num_failed: int = 0
    density: float = 0.0
    total_volume: int = 0
    bounding_box: Optional[tuple] = None


def calculate_stats(tray):
    """Bounding-box density (packer.py:582-601)."""
    occ = tray > 0
    n = int(occ.sum())
    if n == 0:
        return 0.0, 0, None
    idx = np.argwhere(occ)
    lo = tuple(int(v) for v in idx.min(axis=0))
    hi = tuple(int(v) for v in idx.max(axis=0))
    vol = int(np.prod([h - l + 1 for l, h in zip(lo, hi)]))
    return float(n / vol), n, (lo, hi)


def pack_voxels(items: Sequence[np.ndarray], tray_size, num_orientations: int = 1,
                interlocking_free: bool = False, sort_by_volume: bool = True,
                initial_tray=None, metric: str = "l1", sources: str = "eq1",
                order: Optional[Sequence[int]] = None) -> OracleResult:
    """Greedy loop of BinPacker.pack_voxels (packer.py:366-544).

    ``order`` pins the processing order (the reference's np.argsort tie
    order is implementation defined); when None and sort_by_volume, a stable
    descending-volume order is used.  1 orientation without IF uses the
    ``pow`` penalty and an internally computed L1 field (packing.cpp:590-642);
    otherwise the field is computed once per item and the ``cube`` penalty is
    used (fft3.cu:173).
    """
    proc = [as_grid(it) for it in items]
    vols = [int((it > 0).sum()) for it in proc]
    if order is None:
        order = stable_volume_order(vols) if sort_by_volume else list(range(len(proc)))
    if initial_tray is not None:
        tray = as_grid(initial_tray).copy()
        id_offset = int(tray.max())
    else:
        tray = np.zeros(tuple(tray_size), dtype=np.int32)
        id_offset = 0
    placements = []
    num_placed = 0
    for orig in order:
        item = proc[orig]
        orients = [as_grid(g) for g in get_orientations(item, num_orientations)]
        single = (num_orientations == 1 and not interlocking_free)
        phi = distance_field(tray, metric, sources)
        state = TrayState(tray, phi)
        # orientations larger than the tray are skipped (fft3.cu:1563-1568,
        # packer.py:491-494); a later orientation replaces the best only on a
        # strictly smaller score (fft3.cu:1579, packer.py:501)
        best, best_score = None, float("inf")
        for o, g in enumerate(orients):
            if any(g.shape[d] > tray.shape[d] for d in range(3)):
                continue
            mask = interlocking_mask_state(state, g) if interlocking_free else None
            r = search_state(state, g, penalty="pow" if single else "cube", mask=mask)
            if r.found and r.score < best_score:
                best_score = r.score
                best = (r.position, r.score, o, g)
        if best is not None:
            item_id = id_offset + num_placed + 1
            tray = place_in_tray(best[3], tray, best[0], item_id)
            num_placed += 1
            placements.append(OraclePlacement(orig, best[0], best[1], True, vols[orig], best[2]))
        else:
            placements.append(OraclePlacement(orig, None, None, False, vols[orig], 0))
    placements.sort(key=lambda p: p.item_index)
    density, total, bbox = calculate_stats(tray)
    return OracleResult(tray, placements, num_placed, len(proc) - num_placed, density, total, bbox)


def stable_volume_order(volumes: Sequence[int]) -> List[int]:
    """Descending volume, ties in input order (stable)."""
    return sorted(range(len(volumes)), key=lambda i: (-volumes[i], i))
