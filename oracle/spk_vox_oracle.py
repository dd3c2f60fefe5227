"""CPU restatement of the reference's STL voxelizer (test infrastructure only).

Only tests may import this module; the product voxelizer is the HIP path behind
``spk_voxelize_stl`` / ``spk_voxelize_mesh``. It restates, in numpy:

* the STL reader of the vendored LibSL (MeshFormat_stl.cpp:204-232 picks binary
  when the file size is exactly 84 + 50 * count, ASCII otherwise;
  loadBinary :150-190, loadASCII :86-148 with BasicParser.h:231-351 tokenising),
* ``TriangleMesh::mergeVerticesExact`` (Mesh.cpp:350-391): vertices renumbered by
  the rank of their position in lexicographic order, exact duplicates merged,
* the fixed-point transform of ``voxelize`` (packing.cpp:526-560), in float32 with
  the association of LibSL's Matrix4x4 products (Matrix4x4.h:245-257, 283-294),
* ``rasterize`` / ``isInTriangle`` (packing.cpp:309-420) in three swizzled views,
  each covered pixel centre flipping the view's axis bit of one voxel, with the
  aligned-edge rule,
* ``array3d_to_binary_voxel_grid`` (packing.cpp:500-507) and
  ``make_voxel_grid_tight`` (voxelGrid.cpp:151-178).

Parity pinning (tests/test_oracle_vox.py): closed-form shells of axis-aligned
boxes (transform, swizzles, margins, aligned diagonals), the even-crossing
invariant along every line of closed meshes, plane depths of tilted triangles
against the exact plane, hand-worked merge order and both STL encodings. The
reference ships no voxelizer outputs, so no golden vector exists for it.
"""
from __future__ import annotations

import os
import struct

import numpy as np

FP_POW = 16                      # constants.h:12
FP_SCALE = 1 << FP_POW
ALONG_X, ALONG_Y, ALONG_Z = 1, 2, 4  # constants.h:13-15

_F = np.float32


# ---------------------------------------------------------------- STL reading

class StlError(ValueError):
    pass


class _Parser:
    """Character-level tokeniser with BasicParser's rules (BasicParser.h:231-351):
    spaces are ' ', '\\t', '\\r' and '\\n'; a float token is the longest run of
    characters from "0123456789.-+e" and is read with atof."""

    def __init__(self, data: bytes):
        self.d = data
        self.p = 0
        self.eof = False

    def _skip(self) -> bool:
        if self.eof:
            return False
        while self.p < len(self.d) and self.d[self.p] in b" \t\r\n":
            self.p += 1
        if self.p >= len(self.d):
            self.eof = True
            return False
        return True

    def string(self, stop: bytes | None = None, accepted: bytes | None = None) -> bytes:
        if not self._skip():
            return b""
        start = self.p
        while True:
            if self.p >= len(self.d):
                self.eof = True
                return self._cstr(start)
            c = self.d[self.p]
            # strchr also finds the terminator: NUL ends a stop-set token and is
            # accepted inside an accepted-set token
            if stop is not None:
                if c in stop or c == 0:
                    return self._cstr(start)
            elif accepted is not None:
                if c not in accepted and c != 0:
                    return self._cstr(start)
            elif c in b" \t\r\n":
                return self._cstr(start)
            self.p += 1

    def _cstr(self, start: int) -> bytes:
        """The token as the C string the reference reads back (cut at a NUL)."""
        return self.d[start:self.p].split(b"\0", 1)[0]

    def float(self) -> np.float32:
        tok = self.string(accepted=b"0123456789.-+e")
        return _F(_atof(tok))


def _atof(tok: bytes) -> float:
    """atof: the longest valid decimal prefix, 0.0 when there is none."""
    s = tok.decode("ascii", "replace")
    for end in range(len(s), 0, -1):
        try:
            return float(s[:end])
        except ValueError:
            continue
    return 0.0


def _expect(p: _Parser, word: bytes) -> None:
    s = p.string()
    if s != word:
        raise StlError(f"invalid ASCII STL: '{word.decode()}' expected, got '{s.decode(errors='replace')}'")


def load_stl(path: str):
    """Return (positions float32 (3n, 3), triangles int64 (n, 3)) with three fresh
    vertices per facet, as LibSL loads them (MeshFormat_stl.cpp:204-232)."""
    with open(path, "rb") as f:
        data = f.read()
    n = struct.unpack_from("<I", data, 80)[0] if len(data) >= 84 else 0
    expected = (n * 50 + 84) & 0xFFFFFFFF  # uint arithmetic (MeshFormat_stl.cpp:216)
    if len(data) >= 84 and expected == len(data):
        rec = np.frombuffer(data, dtype=np.uint8, count=50 * n, offset=84).reshape(n, 50)
        xyz = rec[:, 12:48].copy().view("<f4").reshape(n, 3, 3)
        pos = xyz.reshape(-1, 3).astype(np.float32)
        return pos, np.arange(3 * n, dtype=np.int64).reshape(n, 3)
    p = _Parser(data)
    if p.string() != b"solid":
        raise StlError(f"{path} is not an ASCII STL")
    p.string(stop=b"\n")
    verts = []
    while not p.eof:
        s = p.string()
        if s == b"endsolid":
            break
        if s != b"facet":
            raise StlError(f"invalid ASCII STL: 'facet' expected, got '{s.decode(errors='replace')}'")
        _expect(p, b"normal")
        p.float(), p.float(), p.float()
        _expect(p, b"outer")
        _expect(p, b"loop")
        for _ in range(3):
            _expect(p, b"vertex")
            verts.append((p.float(), p.float(), p.float()))
        _expect(p, b"endloop")
        _expect(p, b"endfacet")
    pos = np.array(verts, dtype=np.float32).reshape(-1, 3)
    return pos, np.arange(len(pos), dtype=np.int64).reshape(-1, 3)


# ---------------------------------------------------------------- mesh preparation

def merge_vertices_exact(pos: np.ndarray, tris: np.ndarray):
    """Mesh.cpp:350-391: sort (position, index) pairs lexicographically, number the
    distinct positions in that order, remap the triangles. Returns (positions of
    the merged vertices, remapped triangles)."""
    pos = np.asarray(pos, dtype=np.float32)
    n = len(pos)
    if n == 0:
        return pos, np.asarray(tris, dtype=np.int64)
    order = np.lexsort((np.arange(n), pos[:, 2], pos[:, 1], pos[:, 0]))
    sp = pos[order]
    new_pos = np.ones(n, dtype=bool)
    new_pos[1:] = np.any(sp[1:] != sp[:-1], axis=1)
    rank = np.cumsum(new_pos) - 1
    newindex = np.empty(n, dtype=np.int64)
    newindex[order] = rank
    merged = np.empty((rank[-1] + 1, 3), dtype=np.float32)
    merged[newindex] = pos            # the last original vertex of each rank wins
    return merged, newindex[np.asarray(tris, dtype=np.int64)]


def bbox(pos: np.ndarray):
    """Mesh.cpp:177-199 (min/max seeded with +-1e20)."""
    lo = np.minimum(pos.min(axis=0), _F(1e20)) if len(pos) else np.full(3, _F(1e20))
    hi = np.maximum(pos.max(axis=0), _F(-1e20)) if len(pos) else np.full(3, _F(-1e20))
    return lo.astype(np.float32), hi.astype(np.float32)


def fixed_point(pos: np.ndarray, resolution: int):
    """packing.cpp:530-560 in float32. Returns (int64 points (n, 3), grid dims)."""
    lo, hi = bbox(pos)
    ext = (hi - lo).astype(np.float32)
    mx = _F(ext.max())
    if not mx > 0:
        raise StlError("mesh has an empty bounding box")
    factor = _F(0.95)
    s_res = _F(resolution * FP_SCALE)
    inv = (_F(1.0) / mx)
    t1 = ((_F(1.0) - factor) * _F(0.5)) * ext
    # S(res*S) * S(1/mx) * T(t1) * S(0.95) * T(-lo), products left to right
    d = s_res * inv                      # diagonal of S*S
    a2_t = d * t1                        # translation after T(t1)
    a3_d = d * factor                    # diagonal after S(0.95)
    m_t = (a3_d * (-lo)).astype(np.float32) + a2_t   # translation after T(-lo)
    m_d = np.full(3, a3_d, dtype=np.float32)
    bx = (pos * m_d).astype(np.float32) + m_t.astype(np.float32)
    # std::round: half away from zero, done in float64 where x +- 0.5 is exact
    b64 = bx.astype(np.float64)
    r = np.where(b64 >= 0, np.floor(b64 + 0.5), np.ceil(b64 - 0.5)).astype(np.float32)
    top = _F(resolution * FP_SCALE) - _F(1.0)
    r = np.minimum(top, np.maximum(_F(0.0), r))
    pts = r.astype(np.int64)
    dims = ((ext / mx) * _F(resolution)).astype(np.float32)
    return pts, tuple(int(v) for v in dims.astype(np.uint32))


# ---------------------------------------------------------------- rasterization

# view -> (forward permutation of (x, y, z), axis bit); backward maps (i, j, depth)
# to voxel (x, y, z) (packing.cpp:342-366)
_VIEWS = (
    ((0, 1, 2), ALONG_Z),   # xy view, depth z
    ((1, 2, 0), ALONG_X),   # yz view, depth x
    ((2, 0, 1), ALONG_Y),   # zx view, depth y
)


def _cdiv(a: np.ndarray, b: int) -> np.ndarray:
    """C integer division (truncation toward zero)."""
    q = np.abs(a) // abs(b)
    return np.where((a < 0) != (b < 0), -q, q)


def rasterize(pts: np.ndarray, tris: np.ndarray, dims) -> np.ndarray:
    """Raw per-voxel axis bits (uint8) of packing.cpp:370-420 over all triangles and
    the three views."""
    xs, ys, zs = dims
    raw = np.zeros((xs, ys, zs), dtype=np.uint8)
    half = 1 << (FP_POW - 1)
    for tri in np.asarray(tris, dtype=np.int64):
        for perm, bit in _VIEWS:
            tp = [tuple(int(pts[v][a]) for a in perm) for v in tri]
            d10 = (tp[1][0] - tp[0][0], tp[1][1] - tp[0][1])
            d21 = (tp[2][0] - tp[1][0], tp[2][1] - tp[1][1])
            d02 = (tp[0][0] - tp[2][0], tp[0][1] - tp[2][1])
            if d10 == (0, 0) or d21 == (0, 0) or d02 == (0, 0):
                continue
            # the validity cross product is evaluated in 32-bit int arithmetic
            if (d02[0] * d10[1] - d02[1] * d10[0]) % (1 << 32) == 0:
                continue
            bx0 = min(t[0] // FP_SCALE for t in tp)
            bx1 = max(t[0] // FP_SCALE for t in tp)
            by0 = min(t[1] // FP_SCALE for t in tp)
            by1 = max(t[1] // FP_SCALE for t in tp)
            jj, ii = np.meshgrid(np.arange(by0, by1 + 1, dtype=np.int64),
                                 np.arange(bx0, bx1 + 1, dtype=np.int64), indexing="ij")
            pi = (ii << FP_POW) + half
            pj = (jj << FP_POW) + half
            c = []
            for (px, py, _), (ex, ey) in ((tp[0], d10), (tp[1], d21), (tp[2], d02)):
                c.append((pi - px) * ey - (pj - py) * ex)
            c0, c1, c2 = c
            inside = ((c0 <= 0) & (c1 <= 0) & (c2 <= 0)) | ((c0 >= 0) & (c1 >= 0) & (c2 >= 0))
            aligned = np.where(c0 == 0, 0, np.where(c1 == 0, 1, np.where(c2 == 0, 2, -1)))
            keep_edge = np.ones_like(inside)
            for a in range(3):
                keep_edge &= ~((aligned == a) & ~(tri[a] < tri[(a + 1) % 3]))
            sel = inside & keep_edge
            if not sel.any():
                continue
            c0, c1, c2 = c0[sel], c1[sel], c2[sel]
            area = int((c0 + c1 + c2)[0])
            b0 = _cdiv(c1 << 10, area)
            b1 = _cdiv(c2 << 10, area)
            b2 = (1 << 10) - b0 - b1
            depth = (b0 * tp[0][2] + b1 * tp[1][2] + b2 * tp[2][2]) >> 10
            vi, vj, vd = ii[sel], jj[sel], depth >> FP_POW
            vox = [None, None, None]
            vox[perm[0]], vox[perm[1]], vox[perm[2]] = vi, vj, vd
            vx, vy, vz = vox
            ok = (vx >= 0) & (vx < xs) & (vy >= 0) & (vy < ys) & (vz >= 0) & (vz < zs)
            raw[vx[ok], vy[ok], vz[ok]] ^= bit
    return raw


def crop_tight(grid: np.ndarray) -> np.ndarray:
    """voxelGrid.cpp:151-178; raises on a grid with no occupied cell."""
    nz = np.argwhere(grid > 0)
    if len(nz) == 0:
        raise StlError("(make_voxel_grid_tight) Voxel grid doesn't have any occupied cells")
    lo, hi = nz.min(axis=0), nz.max(axis=0)
    return np.ascontiguousarray(grid[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1])


def voxelize_mesh_raw(pos: np.ndarray, tris: np.ndarray, resolution: int):
    """merge + transform + rasterize (packing.cpp:526-588); returns the raw bit grid."""
    mpos, mtris = merge_vertices_exact(pos, tris)
    pts, dims = fixed_point(mpos, resolution)
    return rasterize(pts, mtris, dims)


def voxelize_mesh(pos: np.ndarray, tris: np.ndarray, resolution: int = 128) -> np.ndarray:
    raw = voxelize_mesh_raw(pos, tris, resolution)
    return crop_tight((raw > 0).astype(np.int32))


def voxelize_stl(path: str, resolution: int = 128) -> np.ndarray:
    """bindings.cpp:285-302: load, voxelize, tight crop; int32 0/1 grid."""
    if not os.path.exists(path):
        raise StlError(f"cannot open {path}")
    pos, tris = load_stl(path)
    return voxelize_mesh(pos, tris, resolution)


# ---------------------------------------------------------------- STL writers for tests

def write_stl_binary(path: str, pos: np.ndarray, tris: np.ndarray) -> None:
    tris = np.asarray(tris, dtype=np.int64)
    rec = np.zeros((len(tris), 50), dtype=np.uint8)
    xyz = np.asarray(pos, dtype=np.float32)[tris].reshape(len(tris), 9)
    rec[:, 12:48] = xyz.astype("<f4").view(np.uint8).reshape(len(tris), 36)
    with open(path, "wb") as f:
        f.write(b"binary test mesh".ljust(80, b" "))
        f.write(struct.pack("<I", len(tris)))
        f.write(rec.tobytes())


def write_stl_ascii(path: str, pos: np.ndarray, tris: np.ndarray, fmt: str = "{:.9e}") -> None:
    pos = np.asarray(pos, dtype=np.float32)
    lines = ["solid test mesh"]
    for t in np.asarray(tris, dtype=np.int64):
        lines.append("  facet normal 0 0 0")
        lines.append("    outer loop")
        for v in t:
            lines.append("      vertex " + " ".join(fmt.format(float(c)) for c in pos[v]))
        lines.append("    endloop")
        lines.append("  endfacet")
    lines.append("endsolid test mesh")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
