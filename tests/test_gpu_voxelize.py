"""GPU parity of the voxelizer (spk_voxelize_mesh / spk_voxelize_stl) with the
oracle restatement of the reference's STL voxelization: bit-exact grids."""
import time

import numpy as np
import pytest

from oracle import spk_vox_oracle as V
from tests import meshgen as M

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def P():
    import psacking_amd
    from psacking_amd import mesh
    return psacking_amd, mesh


def _mesh(name):
    if name == "cube":
        return M.box((0.25, 0.25, 0.25), (2.0, 2.0, 2.0))
    if name == "box":
        return M.box((0.3, -1.0, 2.0), (3.0, 2.2, 1.7))
    if name == "sphere":
        p, t = M.icosphere(3)
        return M.rotate(p, 11, shift=(1, 2, 3)), t
    if name == "torus":
        p, t = M.torus(nu=40, nv=16)
        return M.rotate(p, 12), t
    if name == "soup":
        return M.soup(300, 13)
    if name == "triangle":
        return np.array([[0, 0, 0], [10, 1, 3], [2, 9, 7]], np.float32), np.array([[0, 1, 2]], np.int32)
    if name == "tiny_far":   # small extent far from the origin
        p, t = M.icosphere(2, radius=0.05, center=(1e4, -2e4, 3e3))
        return p, t
    raise KeyError(name)


@pytest.mark.parametrize("name,res", [("cube", 24), ("cube", 128), ("box", 24), ("box", 61),
                                      ("sphere", 32), ("sphere", 128), ("torus", 50), ("torus", 257),
                                      ("soup", 40), ("triangle", 32), ("triangle", 2), ("cube", 2),
                                      ("tiny_far", 64)])
def test_voxelize_mesh_matches_oracle(P, name, res):
    _, mesh = P
    say("voxelize_mesh", name, res)
    pos, tris = _mesh(name)
    for p, t in ((pos, tris), M.unshare(pos, tris)):
        got = mesh.voxelize_mesh(p, t, res)
        exp = V.voxelize_mesh(p, t, res)
        assert got.dtype == np.int32 and got.shape == exp.shape
        np.testing.assert_array_equal(got, exp)


def test_voxelize_stl_both_encodings(P, tmp_path):
    psa, _ = P
    pos, tris = _mesh("torus")
    for enc in ("binary", "ascii"):
        path = str(tmp_path / f"torus_{enc}.stl")
        (V.write_stl_binary if enc == "binary" else V.write_stl_ascii)(path, pos, tris)
        for res in (16, 128):
            say("voxelize_stl", enc, res)
            got = psa.voxelize_stl(path, res)
            np.testing.assert_array_equal(got, V.voxelize_stl(path, res))
    # default resolution is 128, as in the reference binding
    np.testing.assert_array_equal(psa.voxelize_stl(path), V.voxelize_stl(path, 128))


def test_large_mesh_and_timing(P):
    _, mesh = P
    pos, tris = M.icosphere(5)   # 20480 triangles
    pos = M.rotate(pos, 21)
    say("voxelize_mesh sphere5", len(tris))
    mesh.voxelize_mesh(pos, tris, 128)
    t0 = time.perf_counter()
    got = mesh.voxelize_mesh(pos, tris, 128)
    t1 = time.perf_counter()
    exp = V.voxelize_mesh(pos, tris, 128)
    t2 = time.perf_counter()
    np.testing.assert_array_equal(got, exp)
    print(f"[vox] 20480 triangles @128: gpu path {1e3 * (t1 - t0):.1f} ms (two calls), "
          f"numpy oracle {1e3 * (t2 - t1):.0f} ms", flush=True)


def test_voxelize_errors(P, tmp_path):
    psa, mesh = P
    say("voxelize errors")
    with pytest.raises(RuntimeError, match="Voxelization failed for"):
        psa.voxelize_stl(str(tmp_path / "missing.stl"))
    bad = tmp_path / "bad.stl"
    bad.write_text("solid x\n facet normal 0 0 0\n outer lop\n")
    with pytest.raises(RuntimeError, match="Voxelization failed for"):
        psa.voxelize_stl(str(bad))
    collinear = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], np.float32)
    with pytest.raises(RuntimeError, match="occupied"):
        mesh.voxelize_mesh(collinear, [[0, 1, 2]], 16)
    pos, tris = _mesh("cube")
    with pytest.raises(RuntimeError, match="occupied"):   # the 2.5% margins hold no pixel centre
        mesh.voxelize_mesh(pos, tris, 1)
    for res in (0, 1025):
        with pytest.raises(RuntimeError, match="resolution"):
            mesh.voxelize_mesh(pos, tris, res)
    with pytest.raises(RuntimeError, match="index"):
        mesh.voxelize_mesh(pos, [[0, 1, 99]], 16)
