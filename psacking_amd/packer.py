"""High-level packing API (reference: spectral_packer/packer.py).

``BinPacker`` keeps the reference constructor, ``pack_voxels`` /
``pack_single`` signatures, ``PackingResult`` / ``PlacementInfo`` fields and
the bounding-box density (packer.py:582-601).  ``pack_voxels`` runs the whole
greedy loop (packer.py:447-529) in one native call: the tray, its distance
field and spectra stay resident on the GPU, orientations are index maps
applied on the device, and only per-item results come back.

Extensions (keyword-only, defaults reproduce the reference):
  metric   "l1" (reference, fft3.cu:887-1004) or "edt" (rounded Euclidean)
  sources  "eq1" (value-1 voxels, fft3.cu:948) or "nonzero" (all occupied)
  engine   "fused" (native loop) or "percall" (the reference's per-orientation
           Python loop over the _core functions, kept for parity checks)
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from .rotations import get_orientations, make_contiguous, orientation_maps


@dataclass
class PlacementInfo:
    """One item's outcome (packer.py:29-56)."""

    item_index: int
    position: Optional[Tuple[int, int, int]]
    score: Optional[float]
    success: bool
    volume: int = 0
    orientation_index: int = 0


@dataclass
class PackingResult:
    """Packing outcome (packer.py:86-166)."""

    tray: np.ndarray
    placements: List[PlacementInfo] = field(default_factory=list)
    num_placed: int = 0
    num_failed: int = 0
    density: float = 0.0
    total_volume: int = 0
    bounding_box: Optional[Tuple[Tuple[int, int, int], Tuple[int, int, int]]] = None
    mesh_placements: Optional[list] = None

    def get_item_mask(self, item_id: int) -> np.ndarray:
        return self.tray == item_id

    def save_vox(self, path: Union[str, Path]) -> None:
        from . import save_vox
        save_vox(self.tray, str(path))

    def summary(self) -> str:
        lines = [
            "Packing Result Summary",
            "=" * 40,
            f"Items placed:    {self.num_placed}",
            f"Items failed:    {self.num_failed}",
            f"Success rate:    {self.num_placed / max(1, self.num_placed + self.num_failed):.1%}",
            f"Packing density: {self.density:.1%}",
            f"Total volume:    {self.total_volume} voxels",
            f"Tray size:       {self.tray.shape}",
        ]
        if self.bounding_box:
            bbox_min, bbox_max = self.bounding_box
            lines.append(f"Bounding box:    {bbox_min} to {bbox_max}")
        return "\n".join(lines)


class BinPacker:
    """GPU spectral bin packer (packer.py:168-251)."""

    def __init__(
        self,
        tray_size: Tuple[int, int, int],
        voxel_resolution: int = 128,
        height_penalty: float = 1e8,
        num_orientations: int = 1,
        interlocking_free: bool = False,
        *,
        metric: str = "l1",
        sources: str = "eq1",
        engine: str = "fused",
        device: int = 0,
    ):
        if len(tray_size) != 3:
            raise ValueError(f"tray_size must be a 3-tuple, got {len(tray_size)} elements")
        if any(s <= 0 for s in tray_size):
            raise ValueError(f"tray_size dimensions must be positive, got {tray_size}")
        if voxel_resolution <= 0:
            raise ValueError(f"voxel_resolution must be positive, got {voxel_resolution}")
        if num_orientations not in (1, 4, 6, 24):
            raise ValueError(f"num_orientations must be 1, 4, 6, or 24, got {num_orientations}")
        if metric not in ("l1", "edt"):
            raise ValueError(f"metric must be 'l1' or 'edt', got {metric!r}")
        if sources not in ("eq1", "nonzero"):
            raise ValueError(f"sources must be 'eq1' or 'nonzero', got {sources!r}")
        if engine not in ("fused", "percall"):
            raise ValueError(f"engine must be 'fused' or 'percall', got {engine!r}")
        if engine == "percall" and (metric, sources) != ("l1", "eq1"):
            raise ValueError("engine='percall' reproduces the reference loop (l1 / eq1 only)")
        self.tray_size = tuple(int(s) for s in tray_size)
        self.voxel_resolution = voxel_resolution
        # stored like the reference; the reference never passes it down either
        # (packer.py:248, constants.h:7): the penalty is the constant 1e8.
        self.height_penalty = height_penalty
        self.num_orientations = num_orientations
        self.interlocking_free = interlocking_free
        self.metric = metric
        self.sources = sources
        self.engine = engine
        self.device = device
        self._session = None

    # -- file based entry points: mesh loading is not part of this build -------
    def pack_files(self, paths, sort_by_volume: bool = True, validate_meshes: bool = True,
                   repair_meshes: bool = True) -> PackingResult:
        raise NotImplementedError(
            "mesh loading/voxelization is not part of psacking_amd; voxelize meshes elsewhere "
            "and call pack_voxels")

    def pack_files_for_export(self, paths, sort_by_volume: bool = True,
                              validate_meshes: bool = True,
                              repair_meshes: bool = True) -> PackingResult:
        return self.pack_files(paths, sort_by_volume, validate_meshes, repair_meshes)

    # -- core ----------------------------------------------------------------------
    def _prepare(self, items, sort_by_volume, initial_tray):
        if len(items) == 0:
            raise ValueError("items list cannot be empty")
        processed, volumes = [], []
        for i, item in enumerate(items):
            if not isinstance(item, np.ndarray):
                raise ValueError(f"Item {i} is not a numpy array")
            if item.ndim != 3:
                raise ValueError(f"Item {i} must be 3D, got {item.ndim}D")
            it = np.ascontiguousarray(item.astype(np.int32))
            processed.append(it)
            volumes.append(int(np.sum(it > 0)))
        if sort_by_volume:
            order = [int(i) for i in np.argsort(volumes)[::-1]]   # packer.py:421-425
        else:
            order = list(range(len(items)))
        if initial_tray is not None:
            if initial_tray.shape != self.tray_size:
                raise ValueError(
                    f"initial_tray shape {initial_tray.shape} does not match "
                    f"tray_size {self.tray_size}")
            tray = np.ascontiguousarray(initial_tray.astype(np.int32).copy())
            id_offset = int(np.max(tray))
        else:
            tray = np.zeros(self.tray_size, dtype=np.int32)
            id_offset = 0
        return processed, volumes, order, tray, id_offset

    def session(self):
        from . import _core
        if self._session is None:
            self._session = _core.Session(self.device)
        return self._session

    def pack_voxels(
        self,
        items: Sequence[np.ndarray],
        sort_by_volume: bool = True,
        initial_tray: Optional[np.ndarray] = None,
    ) -> PackingResult:
        processed, volumes, order, tray, id_offset = self._prepare(items, sort_by_volume,
                                                                   initial_tray)
        if self.engine == "percall":
            return self._pack_percall(processed, volumes, order, tray, id_offset)
        single = self.num_orientations == 1 and not self.interlocking_free
        perms, flips = orientation_maps(self.num_orientations)
        s = self.session()
        s.set_tray(tray, self.metric, self.sources, 0)
        res = s.pack(processed, np.asarray(order, dtype=np.int32), perms, flips,
                     "pow" if single else "cube", bool(self.interlocking_free), id_offset)
        final = s.tray()
        placements = []
        num_placed = 0
        for t, r in enumerate(res):
            orig = order[t]
            if r["found"]:
                num_placed += 1
                placements.append(PlacementInfo(orig, tuple(r["position"]), float(r["score"]), True,
                                                volumes[orig], int(r["orientation"])))
            else:
                placements.append(PlacementInfo(orig, None, None, False, volumes[orig], 0))
        placements.sort(key=lambda p: p.item_index)
        density, total, bbox = self._calculate_stats(final)
        return PackingResult(final, placements, num_placed, len(items) - num_placed, density,
                             total, bbox)

    def _pack_percall(self, processed, volumes, order, tray, id_offset) -> PackingResult:
        """The reference loop (packer.py:447-529) over the _core entry points."""
        from . import (calculate_distance, fft_search_batch_interlocking_free,
                       fft_search_placement, fft_search_placement_with_cache, place_in_tray)
        generation = 0
        placements, num_placed = [], 0
        for orig in order:
            item, volume = processed[orig], volumes[orig]
            best_position, best_score, best_oi, best_item, found_any = None, float("inf"), 0, item, False
            orientations = get_orientations(item, self.num_orientations)
            if self.num_orientations > 1 or self.interlocking_free:
                tray_distance = calculate_distance(tray)
            if self.interlocking_free:
                valid, valid_idx = [], []
                for oi, rot in enumerate(orientations):
                    rot = make_contiguous(rot.astype(np.int32))
                    if not any(rot.shape[i] > self.tray_size[i] for i in range(3)):
                        valid.append(rot)
                        valid_idx.append(oi)
                if valid:
                    pos, found, score, bo = fft_search_batch_interlocking_free(
                        valid, tray, tray_distance, generation)
                    if found and bo >= 0:
                        best_position, best_score = pos, score
                        best_oi, best_item, found_any = valid_idx[bo], valid[bo], True
            else:
                for oi, rot in enumerate(orientations):
                    rot = make_contiguous(rot.astype(np.int32))
                    if any(rot.shape[i] > self.tray_size[i] for i in range(3)):
                        continue
                    if self.num_orientations > 1:
                        pos, found, score = fft_search_placement_with_cache(
                            rot, tray, tray_distance, generation)
                    else:
                        pos, found, score = fft_search_placement(rot, tray)
                    if found and score < best_score:
                        best_position, best_score, best_oi, best_item = pos, score, oi, rot
                        found_any = True
            if found_any:
                tray = place_in_tray(best_item, tray, best_position, id_offset + num_placed + 1)
                generation += 1
                num_placed += 1
                placements.append(PlacementInfo(orig, tuple(best_position), best_score, True,
                                                volume, best_oi))
            else:
                placements.append(PlacementInfo(orig, None, None, False, volume, 0))
        placements.sort(key=lambda p: p.item_index)
        density, total, bbox = self._calculate_stats(tray)
        return PackingResult(tray, placements, num_placed, len(processed) - num_placed, density,
                             total, bbox)

    def pack_single(
        self,
        item: np.ndarray,
        tray: Optional[np.ndarray] = None,
    ) -> Tuple[Optional[Tuple[int, int, int]], bool, float]:
        """packer.py:546-580."""
        from . import fft_search_placement
        if tray is None:
            tray = np.zeros(self.tray_size, dtype=np.int32)
        position, found, score = fft_search_placement(item.astype(np.int32), tray)
        if found:
            return tuple(position), True, score
        return None, False, 0.0

    @staticmethod
    def _calculate_stats(tray: np.ndarray):
        occupied = int(np.sum(tray > 0))
        if occupied == 0:
            return 0.0, 0, None
        idx = np.argwhere(tray > 0)
        bbox_min = tuple(idx.min(axis=0).tolist())
        bbox_max = tuple(idx.max(axis=0).tolist())
        bbox_volume = int(np.prod([mx - mn + 1 for mn, mx in zip(bbox_min, bbox_max)]))
        density = float(occupied / bbox_volume) if bbox_volume > 0 else 0.0
        return density, occupied, (bbox_min, bbox_max)

    def __repr__(self) -> str:
        return (f"BinPacker(tray_size={self.tray_size}, voxel_resolution={self.voxel_resolution}, "
                f"num_orientations={self.num_orientations})")
