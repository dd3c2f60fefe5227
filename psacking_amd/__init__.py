"""psacking_amd -- MI355X-native spectral 3D bin packing.

Drop-in for the reference package ``spectral_packer``
(spectral_packer/__init__.py:52-74): the same ``_core`` functions and
attributes, ``BinPacker`` / ``PackingResult`` / ``PlacementInfo`` and the
orientation helpers.  Every search, distance field, correlation, flood fill
and placement runs in HIP kernels for gfx950 behind the C-ABI of libspk.so;
importing the package fails when the native modules are not built.
"""
from __future__ import annotations

try:
    from . import _core
except ImportError as e:  # pragma: no cover - exercised only when unbuilt
    raise ImportError(
        "psacking_amd native modules are not built (libspk.so / _core): run "
        "`python psacking_amd/build.py`") from e

from ._core import (  # noqa: F401
    HEIGHT_PENALTY,
    VOXEL_RESOLUTION,
    Session,
    calculate_distance,
    collision_grid,
    dft_conv3,
    dft_corr3,
    distance_field,
    fft_search_batch,
    fft_search_batch_interlocking_free,
    fft_search_placement,
    fft_search_placement_with_cache,
    get_bounds,
    gpu_tray_cleanup,
    gpu_tray_init,
    gpu_tray_is_initialized,
    gpu_tray_search,
    interlocking_free_positions,
    make_tight,
    pack_many,
    place_in_tray,
    save_vox,
    voxelize_stl,
)
from ._core import __version__  # noqa: F401
from .mesh import voxelize_mesh  # noqa: F401
from .packer import BinPacker, PackingResult, PlacementInfo  # noqa: F401
from .rotations import (  # noqa: F401
    get_24_orientations,
    get_orientations,
    make_contiguous,
    orientation_maps,
    rotate_90_x,
    rotate_90_y,
    rotate_90_z,
)

__all__ = [
    "BinPacker", "PackingResult", "PlacementInfo", "Session",
    "fft_search_placement", "fft_search_placement_with_cache", "fft_search_batch",
    "fft_search_batch_interlocking_free", "place_in_tray", "voxelize_stl", "dft_conv3",
    "dft_corr3", "calculate_distance", "distance_field", "collision_grid", "make_tight",
    "get_bounds", "pack_many", "save_vox", "gpu_tray_init", "gpu_tray_cleanup", "gpu_tray_is_initialized",
    "gpu_tray_search", "interlocking_free_positions", "get_orientations",
    "get_24_orientations", "orientation_maps", "make_contiguous", "VOXEL_RESOLUTION",
    "HEIGHT_PENALTY", "voxelize_mesh",
]
