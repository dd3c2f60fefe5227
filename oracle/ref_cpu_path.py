"""The reference CPU search path restated as written.  Test and CPU-baseline use only.

Used only by bench.py's ``cpu_baseline`` leg and by tests.  It follows the
reference's FFTW branch step by step, with pocketfft (scipy.fft, complex128)
standing in for FFTW3 (absent on this image):

* correlation = flip -> conv -> flip (packing.cpp:225-239); conv pads both
  operands to (2N+1) per axis, runs two forward and one inverse C2C double
  FFT with a fresh plan per call, scales by 1/P, rounds the real part and
  truncates to N (packing.cpp:174-223);
* collision grid: out-of-bounds offsets become max(c, 1) (packing.cpp:254-270);
* distance: L1 to value-1 voxels, fill Nx+Ny+Nz+10 (voxelGrid.cpp:181-210 with
  the compiled fill of fft3.cu:948);
* score loop over the zero-collision list in (i,j,k) order with strict <,
  score = p + 1e8 * pow(k/Nz, 3) (packing.cpp:620-641).
"""
from __future__ import annotations

import math
import time

import numpy as np
import scipy.fft as sfft

from .spk_oracle import calculate_distance, get_bounds, pad_to


def _conv3_as_written(a: np.ndarray, b: np.ndarray, workers: int) -> np.ndarray:
    N = a.shape
    P = tuple(2 * n + 1 for n in N)
    A = np.zeros(P, dtype=np.complex128)
    B = np.zeros(P, dtype=np.complex128)
    A[:N[0], :N[1], :N[2]] = a
    B[:N[0], :N[1], :N[2]] = b
    A = sfft.fftn(A, workers=workers)
    B = sfft.fftn(B, workers=workers)
    A *= B
    A = sfft.ifftn(A, workers=workers, norm="forward") * (1.0 / (P[0] * P[1] * P[2]))
    r = np.rint(A.real).astype(np.int64)
    return r[:N[0], :N[1], :N[2]]


def corr3_as_written(a: np.ndarray, b: np.ndarray, workers: int = 1) -> np.ndarray:
    fa = np.ascontiguousarray(a[::-1, ::-1, ::-1])
    r = _conv3_as_written(fa, b, workers)
    return np.ascontiguousarray(r[::-1, ::-1, ::-1])


def search_as_written(item, tray, phi=None, workers: int = 1, penalty: str = "pow"):
    """fft_search_placement (packing.cpp:590-642) with the CPU FFT path.

    penalty "pow" is that function's P * pow(k/L, 3) with the distance field
    computed inside; "cube" with a field passed in is the per-orientation
    step of the 24-orientation loop (packer.py:496-498), whose cached search
    scores ((hp*qz)*qz)*qz (fft3.cu:173), run here on the CPU FFT path."""
    tray = np.asarray(tray, dtype=np.int64)
    N = tray.shape
    L = N[2]
    item_p = pad_to(np.asarray(item), N)
    _, hi = get_bounds(item_p)
    coll = corr3_as_written(tray, item_p, workers)
    i = np.arange(N[0])[:, None, None]
    j = np.arange(N[1])[None, :, None]
    k = np.arange(N[2])[None, None, :]
    oob = ~((i + hi[0] <= N[0] - 1) & (j + hi[1] <= N[1] - 1) & (k + hi[2] <= N[2] - 1))
    coll = np.where(oob, np.maximum(coll, 1), coll)
    if phi is None:
        phi = calculate_distance(tray.astype(np.int32))
    prox = corr3_as_written(np.asarray(phi, dtype=np.int64), item_p, workers)
    # the strict-< loop over where3d_flat's (i, j, k) order keeps the first
    # minimum; vectorised here so the baseline is not charged Python loop time
    if penalty == "pow":
        pen = np.array([100000000.0 * math.pow((kk + 0.0) / (L + 0.0), 3.0) for kk in range(L)])
        bound = 1e9           # bestVal = INF (packing.cpp:626, constants.h:21)
    else:
        pen = np.array([((100000000.0 * (kk / L)) * (kk / L)) * (kk / L) for kk in range(L)])
        bound = 1e29          # fft3.cu:1361
    score = np.where(coll == 0, prox.astype(np.float64) + pen[None, None, :], np.inf)
    flat = int(np.argmin(score))
    if not score.flat[flat] < bound:
        return (-1, -1, -1), False, 0.0
    best_id = tuple(int(v) for v in np.unravel_index(flat, N))
    return best_id, True, float(score.flat[flat])


def time_item_sample(item, tray, n_orient: int = 24, workers: int = 1, budget_s: float = 20.0):
    """Seconds per item of the reference CPU path, from a bounded sample.

    Times the distance field once and as many single-orientation searches as
    fit in the budget (at least one), then extrapolates to n_orient
    orientations per item.  Returns (seconds_per_item, searches_timed,
    seconds_spent).
    """
    t0 = time.perf_counter()
    phi = calculate_distance(np.asarray(tray, dtype=np.int32))
    t_dist = time.perf_counter() - t0
    times = []
    from .spk_oracle import get_orientations
    orients = get_orientations(np.asarray(item), n_orient)
    start = time.perf_counter()
    for o in orients:
        t1 = time.perf_counter()
        search_as_written(np.ascontiguousarray(o), tray, phi, workers)
        times.append(time.perf_counter() - t1)
        if time.perf_counter() - start > budget_s:
            break
    per_search = float(np.mean(times))
    return t_dist + per_search * n_orient, len(times), time.perf_counter() - t0


def pack_as_written(items, tray_size, n_orient: int = 24, workers: int = 1):
    """The reference greedy loop (packer.py:447-529) on the CPU FFT path.

    Items in the given order; per item the L1 field of the tray once
    (packer.py:460-461) when n_orient > 1, each orientation that fits the
    tray searched as written (search_as_written, "cube" penalty; with one
    orientation the "pow" search with its own field), the first strictly
    smaller score wins (packer.py:501), the winner is placed with id
    num_placed + 1 (packer.py:509-511).  Returns ([(found, pos, orient,
    score)], final tray).
    """
    from .spk_oracle import get_orientations, place_in_tray
    tray = np.zeros(tuple(tray_size), np.int32)
    out = []
    placed = 0
    for item in items:
        orients = [np.ascontiguousarray(g.astype(np.int32))
                   for g in get_orientations(np.asarray(item), n_orient)]
        phi = calculate_distance(tray) if n_orient > 1 else None
        best = None
        for o, g in enumerate(orients):
            if any(g.shape[d] > tray.shape[d] for d in range(3)):
                continue
            if n_orient > 1:
                pos, found, score = search_as_written(g, tray, phi, workers, "cube")
            else:
                pos, found, score = search_as_written(g, tray, None, workers, "pow")
            if found and (best is None or score < best[2]):
                best = (pos, o, score, g)
        if best is None:
            out.append((False, None, -1, None))
            continue
        placed += 1
        tray = place_in_tray(best[3], tray, best[0], placed)
        out.append((True, tuple(best[0]), best[1], best[2]))
    return out, tray
