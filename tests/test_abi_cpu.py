"""CPU-only checks of the boundary: the C-ABI library loads and exports every
declared symbol, the pybind module mirrors the reference export list, and
the orientation index maps reproduce the reference orientation order."""
import os
import re

import numpy as np
import pytest

from oracle import spk_oracle as O
from tests.conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "spk.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(spk_\w+)\s*\(", txt, re.M)))


def test_header_matches_loader_list():
    from psacking_amd import _lib
    assert sorted(_lib.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    from psacking_amd import _lib
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_version_and_context_without_gpu():
    import ctypes as C

    from psacking_amd import _lib
    lib = _lib.load()
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    assert lib.spk_version(C.byref(a), C.byref(b), C.byref(c)) == 0
    assert (a.value, b.value) == (0, 1)
    # invalid device index is rejected without touching a kernel
    ctx = C.c_void_p()
    rc = lib.spk_ctx_create(-1, C.byref(ctx))
    assert rc != 0 and not ctx.value


def test_core_exports_reference_names():
    import psacking_amd as P
    names = ["fft_search_placement", "fft_search_placement_with_cache", "place_in_tray",
             "voxelize_stl", "dft_conv3", "dft_corr3", "calculate_distance", "collision_grid",
             "make_tight", "get_bounds", "save_vox", "gpu_tray_init", "gpu_tray_cleanup",
             "gpu_tray_is_initialized", "gpu_tray_search", "fft_search_batch",
             "fft_search_batch_interlocking_free", "interlocking_free_positions"]
    for n in names:
        assert callable(getattr(P._core, n)), n
    assert P.VOXEL_RESOLUTION == 128 and P.HEIGHT_PENALTY == 1e8 and P.__version__ == "0.1.0"


def test_host_utilities_match_oracle(kats):
    import psacking_amd as P
    for case in kats["get_bounds"]:
        g = np.zeros(case["grid_shape"], dtype=np.int32)
        if "box" in case:
            b = case["box"]
            g[b[0][0]:b[0][1], b[1][0]:b[1][1], b[2][0]:b[2][1]] = 1
        for p in case.get("ones", []):
            g[tuple(p)] = 1
        assert P.get_bounds(g) == O.get_bounds(g)
    assert P.get_bounds(np.zeros((2, 3, 4), np.int32)) == O.get_bounds(np.zeros((2, 3, 4)))
    rng = np.random.default_rng(0)
    g = (rng.random((9, 7, 8)) < 0.05).astype(np.int32)
    g[2, 3, 4] = 1
    np.testing.assert_array_equal(P.make_tight(g), O.make_tight(g))
    with pytest.raises(ValueError):
        P.make_tight(np.zeros((3, 3, 3), np.int32))
    with pytest.raises(RuntimeError):
        P.get_bounds(np.zeros((3, 3), np.int32))


@pytest.mark.parametrize("n", [1, 4, 6, 24])
def test_orientation_maps_match_rot90(n):
    from psacking_amd.rotations import apply_map, get_orientations, orientation_maps
    g = np.arange(3 * 5 * 7).reshape(3, 5, 7)
    perms, flips = orientation_maps(n)
    ref = O.get_orientations(g, n)
    ours = get_orientations(g, n)
    assert len(ref) == len(ours) == n == len(perms)
    for o in range(n):
        np.testing.assert_array_equal(ours[o], ref[o])
        np.testing.assert_array_equal(apply_map(g, perms[o], flips[o]), ref[o])


def test_save_vox_layout(tmp_path):
    import psacking_amd as P
    g = np.zeros((2, 3, 4), np.int32)
    g[1, 2, 3] = 7
    g[0, 0, 0] = 8
    p = tmp_path / "t.vox"
    P.save_vox(g, str(p))
    data = p.read_bytes()
    assert len(data) == 12 + 24 + 768
    assert np.frombuffer(data[:12], np.int32).tolist() == [2, 3, 4]
    vox = np.frombuffer(data[12:36], np.uint8).reshape(2, 3, 4)
    # written with x and z descending (packing.cpp:287-297)
    exp = np.full((2, 3, 4), 255, np.uint8)
    exp[0, 2, 0] = 123 + 7 % 5      # voxel (1,2,3)
    exp[1, 0, 3] = 123              # voxel (0,0,0) value 8 = INSIDE flag
    np.testing.assert_array_equal(vox, exp)


def test_binpacker_validation():
    import psacking_amd as P
    with pytest.raises(ValueError):
        P.BinPacker((10, 10))
    with pytest.raises(ValueError):
        P.BinPacker((10, 10, 0))
    with pytest.raises(ValueError):
        P.BinPacker((10, 10, 10), num_orientations=3)
    with pytest.raises(ValueError):
        P.BinPacker((10, 10, 10)).pack_voxels([])
    with pytest.raises(ValueError):
        P.BinPacker((10, 10, 10)).pack_voxels([np.ones((2, 2))])
