"""Pins the voxelizer oracle (oracle/spk_vox_oracle.py) with checks that do not
restate it: closed-form box shells, the even-crossing invariant of closed
meshes, exact plane depths of a tilted triangle, a hand-worked vertex merge and
both STL encodings. The reference ships no voxelizer outputs to compare with."""
from __future__ import annotations

import math
import struct

import numpy as np
import pytest

from oracle import spk_vox_oracle as V
from tests import meshgen as M

BITS = ((2, V.ALONG_Z), (0, V.ALONG_X), (1, V.ALONG_Y))   # (depth axis, bit) per view


def expected_box_raw(size, res):
    """Shell of an axis-aligned box in closed form, in voxel units: the box spans
    [0.025, 0.975] * e_a * res on axis a (e_a = extent / longest extent); a face
    at depth t lands in layer floor(t) and covers the pixel centres strictly
    inside the other two spans; layers at or beyond floor(e_a * res) are lost."""
    size = np.asarray(size, dtype=np.float64)
    e = size / size.max()
    g = np.floor(e * res).astype(int)
    lo, hi = 0.025 * e * res, 0.975 * e * res
    for a in range(3):   # no sampling tie within the float32 noise of the transform
        for t in (lo[a], hi[a]):
            assert min(abs(t - math.floor(t) - 0.5), t - math.floor(t), math.ceil(t) - t) > 1e-3
        assert abs(e[a] * res - round(e[a] * res)) > 1e-3 or e[a] == 1.0
    span = [np.arange(g[a])[(np.arange(g[a]) + 0.5 > lo[a]) & (np.arange(g[a]) + 0.5 < hi[a])] for a in range(3)]
    raw = np.zeros(tuple(g), dtype=np.uint8)
    for depth_axis, bit in BITS:
        p, q = [a for a in range(3) if a != depth_axis]
        for t in (lo[depth_axis], hi[depth_axis]):
            k = int(math.floor(t))
            if k >= g[depth_axis]:
                continue
            idx = [None, None, None]
            idx[p], idx[q], idx[depth_axis] = span[p][:, None], span[q][None, :], k
            raw[tuple(idx)] ^= bit
    return raw


@pytest.mark.parametrize("lo,size,res", [
    ((0.25, 0.25, 0.25), (2.0, 2.0, 2.0), 24),      # cube: face diagonals run through pixel centres
    ((0.3, -1.0, 2.0), (3.0, 2.2, 1.7), 24),        # top y and z layers fall outside the grid
    ((-5.0, 1.0, 0.5), (1.3, 4.1, 2.9), 37),
])
def test_box_shell_closed_form(lo, size, res):
    pos, tris = M.box(lo, size)
    raw = V.voxelize_mesh_raw(pos, tris, res)
    exp = expected_box_raw(size, res)
    assert raw.shape == exp.shape
    np.testing.assert_array_equal(raw, exp)
    np.testing.assert_array_equal(V.voxelize_mesh(pos, tris, res), V.crop_tight((exp > 0).astype(np.int32)))


def test_cube_diagonal_pixels_counted_once():
    """On the cube every face diagonal passes exactly through pixel centres; the
    aligned-edge rule must give each of those pixels to exactly one of the two
    triangles (counting both or neither would clear the voxel)."""
    pos, tris = M.box((0.25, 0.25, 0.25), (2.0, 2.0, 2.0))
    mpos, mtris = V.merge_vertices_exact(pos, tris)
    pts, _ = V.fixed_point(mpos, 24)
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    assert (lo == lo[0]).all() and (hi == hi[0]).all()
    raw = V.voxelize_mesh_raw(pos, tris, 24)
    k = int(lo[0]) >> V.FP_POW
    diag = np.arange(1, 23)
    assert (raw[diag, diag, k] & V.ALONG_Z).all()
    # the pixel centres on the diagonal are exactly on the shared edge
    c = (diag * V.FP_SCALE + V.FP_SCALE // 2).astype(np.int64)
    assert ((c - lo[0]) * (hi[1] - lo[1]) - (c - lo[1]) * (hi[0] - lo[0]) == 0).all()


def _dims_clear(pos, res):
    ext = pos.max(axis=0).astype(np.float64) - pos.min(axis=0)
    er = ext / ext.max() * res
    return all(er[a] - math.floor(er[a]) < 0.025 * er[a] or er[a] == res for a in range(3))


@pytest.mark.parametrize("mesh,seed,res", [("sphere", 1, 20), ("sphere", 2, 33), ("torus", 3, 28),
                                           ("torus", 4, 41), ("box", 5, 30)])
def test_closed_mesh_even_crossings(mesh, seed, res):
    """Every line of pixel centres crosses a closed surface an even number of
    times, so each view's bits XOR to zero along its depth axis."""
    if mesh == "sphere":
        pos, tris = M.icosphere(2)
    elif mesh == "torus":
        pos, tris = M.torus()
    else:
        pos, tris = M.box((0, 0, 0), (1.0, 0.8, 0.6))
    pos = M.rotate(pos, seed, shift=(seed, -seed, 0.5))
    # a resolution at which no top layer falls outside the grid (that would drop
    # one crossing of the lines through it)
    res = next(r for r in range(res, res + 64) if _dims_clear(pos, r))
    raw = V.voxelize_mesh_raw(*M.unshare(pos, tris), res)
    assert raw.any()
    for depth_axis, bit in BITS:
        par = np.bitwise_xor.reduce(raw & bit, axis=depth_axis)
        assert not par.any(), (depth_axis, int(np.count_nonzero(par)))


def test_tilted_triangle_depths_and_coverage():
    """A lone triangle: each view covers exactly the pixel centres inside its
    projection, at the depth of the plane there (10-bit barycentrics: within one
    voxel of the exact depth)."""
    pos = np.array([[0, 0, 0], [10, 1, 3], [2, 9, 7]], dtype=np.float32)
    tris = np.array([[0, 1, 2]], dtype=np.int32)
    res = 32
    mpos, mtris = V.merge_vertices_exact(pos, tris)
    pts, dims = V.fixed_point(mpos, res)
    raw = V.rasterize(pts, mtris, dims)
    P = pts[mtris[0]].astype(np.float64) / V.FP_SCALE
    for depth_axis, bit in BITS:
        p, q = {2: (0, 1), 0: (1, 2), 1: (2, 0)}[depth_axis]   # pixel axes of the view
        a2 = P[:, [p, q]]
        hits = np.argwhere(raw & bit)
        ii, jj = np.meshgrid(np.arange(dims[p]) + 0.5, np.arange(dims[q]) + 0.5, indexing="ij")
        pt = np.stack([ii, jj], -1)

        def cross(o, a, b):
            return (a[0] - o[..., 0]) * (b[1] - o[..., 1]) - (a[1] - o[..., 1]) * (b[0] - o[..., 0])

        w0, w1, w2 = cross(pt, a2[1], a2[2]), cross(pt, a2[2], a2[0]), cross(pt, a2[0], a2[1])
        area = w0 + w1 + w2
        strict = (w0 / area > 1e-6) & (w1 / area > 1e-6) & (w2 / area > 1e-6)
        outside = (w0 / area < -1e-6) | (w1 / area < -1e-6) | (w2 / area < -1e-6)
        depth = (w0 * P[0, depth_axis] + w1 * P[1, depth_axis] + w2 * P[2, depth_axis]) / area
        covered = np.zeros(strict.shape, dtype=bool)
        for h in hits:
            i, j, d = h[p], h[q], h[depth_axis]
            assert not covered[i, j]
            covered[i, j] = True
            assert abs(d - math.floor(depth[i, j])) <= 1, (depth_axis, i, j, d, depth[i, j])
        inrange = np.floor(depth) < dims[depth_axis]
        assert (covered[strict & inrange]).all()
        assert not covered[outside].any()


def test_merge_is_rank_of_sorted_position():
    pos = np.array([[1, 0, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0], [-0.0, 0, 0], [0, 0, -1]], dtype=np.float32)
    tris = np.array([[0, 1, 3], [2, 4, 5]])
    merged, mt = V.merge_vertices_exact(pos, tris)
    # sorted distinct positions: (0,0,-1) < (0,0,0) = (-0,0,0) < (0,1,0) < (1,0,0)
    np.testing.assert_array_equal(mt, [[3, 1, 2], [3, 1, 0]])
    np.testing.assert_array_equal(merged, [[0, 0, -1], [0, 0, 0], [0, 1, 0], [1, 0, 0]])


def test_stl_encodings_load_identically(tmp_path):
    pos, tris = M.rotate(M.torus(nu=10, nv=6)[0], 7), M.torus(nu=10, nv=6)[1]
    b, a = str(tmp_path / "m_bin.stl"), str(tmp_path / "m_ascii.stl")
    V.write_stl_binary(b, pos, tris)
    V.write_stl_ascii(a, pos, tris)
    pb, tb = V.load_stl(b)
    pa, ta = V.load_stl(a)
    np.testing.assert_array_equal(pb, pos[tris].reshape(-1, 3))
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ta, tb)
    np.testing.assert_array_equal(V.voxelize_stl(a, 20), V.voxelize_stl(b, 20))
    # a binary header that starts with "solid" is still binary when the size fits
    with open(b, "rb") as f:
        data = bytearray(f.read())
    data[:5] = b"solid"
    c = str(tmp_path / "solid_header.stl")
    open(c, "wb").write(bytes(data))
    np.testing.assert_array_equal(V.load_stl(c)[0], pb)


def test_stl_reader_rejections(tmp_path):
    pos, tris = M.box((0, 0, 0), (1, 1, 1))
    good = tmp_path / "good.stl"
    V.write_stl_ascii(str(good), pos, tris, fmt="{:.6e}")
    text = good.read_text()
    cases = {
        "upper_exponent": text.replace("e+", "E+").replace("e-", "E-"),
        "unnamed_solid": text.replace("solid test mesh\n", "solid\n", 1),
        "no_endsolid": text.replace("endsolid test mesh\n", ""),
        "bad_keyword": text.replace("outer loop", "outer lop", 1),
    }
    for name, body in cases.items():
        p = tmp_path / f"{name}.stl"
        p.write_text(body)
        with pytest.raises(V.StlError):
            V.load_stl(str(p))
    # binary with a wrong facet count is read as ASCII and fails
    bad = tmp_path / "short.stl"
    V.write_stl_binary(str(bad), pos, tris)
    data = bytearray(bad.read_bytes())
    data[80:84] = struct.pack("<I", 11)
    bad.write_bytes(bytes(data))
    with pytest.raises(V.StlError):
        V.load_stl(str(bad))


def test_float_tokens_follow_atof(tmp_path):
    body = ("solid x\n facet normal 0 0 0\n outer loop\n"
            " vertex 1.5e 2 3\n vertex +0.25 -.5 7.\n vertex 1e1 0 -2e-1\n"
            " endloop\n endfacet\nendsolid x\n")
    p = tmp_path / "tok.stl"
    p.write_text(body)
    pos, _ = V.load_stl(str(p))
    np.testing.assert_array_equal(pos, np.array([[1.5, 2, 3], [0.25, -0.5, 7], [10, 0, -0.2]], dtype=np.float32))


def test_nothing_marked_raises():
    pos = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], dtype=np.float32)   # collinear
    with pytest.raises(V.StlError, match="occupied"):
        V.voxelize_mesh(pos, np.array([[0, 1, 2]]), 16)
