// spk_vox.hpp -- STL reading and surface voxelization (included by spk_api.hip).
//
// Behaviour of the reference voxelizer behind spectral_packer._core.voxelize_stl
// (bindings.cpp:285-302 -> packing.cpp:526-588 -> voxelGrid.cpp:166-178), built
// from these parts:
//   * STL reading with the vendored LibSL rules (MeshFormat_stl.cpp:204-232:
//     binary exactly when the file size is 84 + 50 * count; ASCII tokenised as
//     BasicParser.h:231-351 does), three fresh vertices per facet;
//   * exact vertex merge (Mesh.cpp:350-391): vertices numbered by the rank of
//     their position in lexicographic order, duplicates share a number;
//   * the fixed-point box transform of packing.cpp:530-548 in float32, with the
//     association of LibSL's Matrix4x4 products (Matrix4x4.h:245-257, 283-294);
//   * rasterization (packing.cpp:309-420) as one HIP kernel: a wavefront per
//     (triangle, view) walks the pixel box of the triangle's projection, and
//     every covered pixel centre flips the view's axis bit of one voxel with an
//     atomic XOR. XOR commutes, so the grid does not depend on launch order and
//     equals the sequential loop bit for bit;
//   * binarisation and the tight crop on the host (the result is a host array).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

namespace spkvox {

constexpr int kFpPow = 16;                 // constants.h:12
constexpr int kFpScale = 1 << kFpPow;
constexpr int kVoxThreads = 64;            // one wavefront per (triangle, view)
constexpr int kMaxResolution = 1024;       // keeps (i << 16) and the raw grid in range

// a triangle after the transform: fixed-point corners and merged vertex numbers
struct Tri {
  int p[9];          // x, y, z of corner 0, 1, 2
  unsigned id[3];    // merged vertex numbers (aligned-edge rule)
};

// view 0: pixels (x, y), depth z, bit 4; view 1: (y, z), depth x, bit 1;
// view 2: (z, x), depth y, bit 2 (swizzle_xyz / swizzle_yzx / swizzle_zxy,
// packing.cpp:342-366)
__global__ __launch_bounds__(kVoxThreads) void k_vox_raster(const Tri* __restrict__ tris,
                                                            unsigned* __restrict__ grid,
                                                            int xs, int ys, int zs) {
  const Tri t = tris[blockIdx.x];
  const int view = (int)blockIdx.y;
  const int ax = view, ay = (view + 1) % 3, az = (view + 2) % 3;
  const unsigned bit = view == 0 ? 4u : (view == 1 ? 1u : 2u);
  int px[3], py[3], pz[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    px[k] = t.p[3 * k + ax];
    py[k] = t.p[3 * k + ay];
    pz[k] = t.p[3 * k + az];
  }
  const int e10x = px[1] - px[0], e10y = py[1] - py[0];
  const int e21x = px[2] - px[1], e21y = py[2] - py[1];
  const int e02x = px[0] - px[2], e02y = py[0] - py[2];
  if ((e10x == 0 && e10y == 0) || (e21x == 0 && e21y == 0) || (e02x == 0 && e02y == 0)) return;
  // the validity test of packing.cpp:386 multiplies in 32-bit int; its wrap-around
  // is kept (unsigned arithmetic gives the same residue without overflow)
  if ((unsigned)e02x * (unsigned)e10y - (unsigned)e02y * (unsigned)e10x == 0u) return;
  const int x0 = min(px[0], min(px[1], px[2])) / kFpScale;
  const int x1 = max(px[0], max(px[1], px[2])) / kFpScale;
  const int y0 = min(py[0], min(py[1], py[2])) / kFpScale;
  const int y1 = max(py[0], max(py[1], py[2])) / kFpScale;
  const int w = x1 - x0 + 1, npix = w * (y1 - y0 + 1);
  for (int q = (int)threadIdx.x; q < npix; q += kVoxThreads) {
    const int i = x0 + q % w, j = y0 + q / w;
    const int ci = i * kFpScale + kFpScale / 2, cj = j * kFpScale + kFpScale / 2;
    const long long c0 = (long long)(ci - px[0]) * e10y - (long long)(cj - py[0]) * e10x;
    const long long c1 = (long long)(ci - px[1]) * e21y - (long long)(cj - py[1]) * e21x;
    const long long c2 = (long long)(ci - px[2]) * e02y - (long long)(cj - py[2]) * e02x;
    const bool inside = (c0 <= 0 && c1 <= 0 && c2 <= 0) || (c0 >= 0 && c1 >= 0 && c2 >= 0);
    if (!inside) continue;
    // a centre exactly on an edge counts for the triangle that lists that edge
    // with the smaller vertex number first (packing.cpp:328, 407)
    const int aligned = c0 == 0 ? 0 : (c1 == 0 ? 1 : (c2 == 0 ? 2 : -1));
    if (aligned >= 0 && !(t.id[aligned] < t.id[(aligned + 1) % 3])) continue;
    const long long area = c0 + c1 + c2;
    const long long b0 = (c1 * 1024) / area, b1 = (c2 * 1024) / area, b2 = 1024 - b0 - b1;
    const int depth = (int)((b0 * pz[0] + b1 * pz[1] + b2 * pz[2]) >> 10);
    int v[3];
    v[ax] = i;
    v[ay] = j;
    v[az] = depth >> kFpPow;
    if (v[0] < 0 || v[0] >= xs || v[1] < 0 || v[1] >= ys || v[2] < 0 || v[2] >= zs) continue;
    const size_t idx = ((size_t)v[0] * ys + v[1]) * zs + v[2];
    atomicXor(grid + (idx >> 2), bit << ((idx & 3) * 8));
  }
}

struct VoxError {
  int code;          // SPK_* status
  std::string msg;
};

#define VOX_HIP(expr)                                                                \
  do {                                                                               \
    const hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                            \
      throw spkvox::VoxError{e_ == hipErrorOutOfMemory ? SPK_ENOMEM : SPK_EDEVICE,   \
                             std::string(#expr) + ": " + hipGetErrorString(e_)};     \
  } while (0)

// ---------------------------------------------------------------- STL reading

// BasicParser over an in-memory file (BasicParser.h:231-351)
class StlParser {
 public:
  explicit StlParser(const std::vector<char>& d) : d_(d) {}
  bool eof() const { return eof_; }

  // stop == nullptr and accepted == nullptr: a token ends at a space or EOL
  std::string token(const char* stop = nullptr, const char* accepted = nullptr) {
    if (!skip()) return std::string();
    const size_t start = p_;
    while (true) {
      if (p_ >= d_.size()) {
        eof_ = true;
        break;
      }
      const char c = d_[p_];
      // strchr finds the terminator: NUL stops a stop-set token and is accepted
      // by an accepted-set token, as in the reference
      if (stop) {
        if (std::strchr(stop, c)) break;
      } else if (accepted) {
        if (!std::strchr(accepted, c)) break;
      } else if (c == ' ' || c == '\t' || c == '\r' || c == '\n') {
        break;
      }
      ++p_;
    }
    // the token is read back as a C string
    return std::string(d_.data() + start, strnlen(d_.data() + start, p_ - start));
  }

  float number() {
    const std::string s = token(nullptr, "0123456789.-+e");
    return (float)std::atof(s.c_str());
  }

 private:
  bool skip() {
    if (eof_) return false;
    while (p_ < d_.size() && (d_[p_] == ' ' || d_[p_] == '\t' || d_[p_] == '\r' || d_[p_] == '\n'))
      ++p_;
    if (p_ >= d_.size()) {
      eof_ = true;
      return false;
    }
    return true;
  }
  const std::vector<char>& d_;
  size_t p_ = 0;
  bool eof_ = false;
};

inline void expect(StlParser& p, const char* word) {
  const std::string s = p.token();
  if (s != word)
    throw VoxError{SPK_EINVAL, std::string("invalid ASCII STL ('") + word + "' expected, got '" + s + "')"};
}

// positions of 3 * n vertices, facet t owning vertices 3t, 3t+1, 3t+2
inline std::vector<float> load_stl(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) throw VoxError{SPK_EINVAL, std::string("cannot open file '") + path + "'"};
  std::vector<char> d;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) d.insert(d.end(), buf, buf + got);
  std::fclose(f);
  uint32_t n = 0;
  if (d.size() >= 84) std::memcpy(&n, d.data() + 80, 4);
  std::vector<float> pos;
  // MeshFormat_stl.cpp:216: the expected size is computed in unsigned 32-bit
  if (d.size() >= 84 && (uint64_t)(uint32_t)(n * 50u + 84u) == (uint64_t)d.size()) {
    pos.resize((size_t)n * 9);
    for (uint32_t t = 0; t < n; ++t)
      std::memcpy(&pos[(size_t)t * 9], d.data() + 84 + (size_t)t * 50 + 12, 36);
    return pos;
  }
  StlParser p(d);
  if (p.token() != "solid") throw VoxError{SPK_EINVAL, std::string("file '") + path + "' is not an ascii STL"};
  p.token("\n");
  while (!p.eof()) {
    const std::string s = p.token();
    if (s == "endsolid") break;
    if (s != "facet") throw VoxError{SPK_EINVAL, "invalid ASCII STL ('facet' expected, got '" + s + "')"};
    expect(p, "normal");
    p.number();
    p.number();
    p.number();
    expect(p, "outer");
    expect(p, "loop");
    for (int v = 0; v < 3; ++v) {
      expect(p, "vertex");
      for (int c = 0; c < 3; ++c) pos.push_back(p.number());
    }
    expect(p, "endloop");
    expect(p, "endfacet");
  }
  return pos;
}

// ---------------------------------------------------------------- mesh preparation

// Mesh.cpp:350-391: new number of each vertex and the merged positions
inline std::vector<unsigned> merge_exact(const float* pos, int64_t nv, std::vector<float>& merged) {
  std::vector<int64_t> order((size_t)nv);
  for (int64_t v = 0; v < nv; ++v) order[(size_t)v] = v;
  std::sort(order.begin(), order.end(), [pos](int64_t a, int64_t b) {
    for (int c = 0; c < 3; ++c) {
      if (pos[3 * a + c] < pos[3 * b + c]) return true;
      if (pos[3 * a + c] > pos[3 * b + c]) return false;
    }
    return a < b;
  });
  std::vector<unsigned> newindex((size_t)nv);
  int64_t rank = -1;
  const float* prev = nullptr;
  for (int64_t s = 0; s < nv; ++s) {
    const float* cur = pos + 3 * order[(size_t)s];
    if (!prev || cur[0] != prev[0] || cur[1] != prev[1] || cur[2] != prev[2]) ++rank;
    newindex[(size_t)order[(size_t)s]] = (unsigned)rank;
    prev = cur;
  }
  merged.assign((size_t)(rank + 1) * 3, 0.f);
  for (int64_t v = 0; v < nv; ++v)   // the last vertex of each number wins
    for (int c = 0; c < 3; ++c) merged[(size_t)newindex[(size_t)v] * 3 + c] = pos[3 * v + c];
  return newindex;
}

// packing.cpp:530-548 and :563: fixed-point corners and the grid dimensions.
// Each float operation is one IEEE single operation, in the order LibSL's
// matrix products perform the non-zero terms.
inline void to_fixed(const std::vector<float>& mp, int res, std::vector<int>& pts, int dims[3]) {
#pragma clang fp contract(off)
  const size_t n = mp.size() / 3;
  float lo[3] = {1e20f, 1e20f, 1e20f}, hi[3] = {-1e20f, -1e20f, -1e20f};
  for (size_t v = 0; v < n; ++v)
    for (int c = 0; c < 3; ++c) {
      const float x = mp[v * 3 + c];
      if (x > hi[c]) hi[c] = x;
      if (x < lo[c]) lo[c] = x;
    }
  float ext[3];
  for (int c = 0; c < 3; ++c) ext[c] = hi[c] - lo[c];
  const float mx = std::max(ext[0], std::max(ext[1], ext[2]));
  if (!(mx > 0.f)) throw VoxError{SPK_EINVAL, "mesh has an empty bounding box"};
  const float factor = 0.95f;
  // S(res * 2^16) * S(1 / mx) * T(0.025 * ext) * S(0.95) * T(-lo), left to right
  const float d = (float)(res * kFpScale) * (1.f / mx);
  const float dd = d * factor;
  const float half_margin = (1.f - factor) * 0.5f;
  float mt[3];
  for (int c = 0; c < 3; ++c) mt[c] = dd * (-lo[c]) + d * (half_margin * ext[c]);
  const float top = (float)(res * kFpScale) - 1.f;
  pts.resize(n * 3);
  for (size_t v = 0; v < n; ++v)
    for (int c = 0; c < 3; ++c) {
      const float bx = dd * mp[v * 3 + c] + mt[c];
      float r = std::round(bx);
      r = std::min(top, std::max(0.f, r));
      pts[v * 3 + c] = (int)r;
    }
  for (int c = 0; c < 3; ++c) {
    dims[c] = (int)(unsigned)(ext[c] / mx * (float)res);
  }
}

// Device buffers of one voxelization, released on every exit path.
struct VoxBuffers {
  Tri* tris = nullptr;
  unsigned* grid = nullptr;
  ~VoxBuffers() {
    if (tris) (void)hipFree(tris);
    if (grid) (void)hipFree(grid);
  }
};

// packing.cpp:526-588 followed by voxelGrid.cpp:166-178. tris == nullptr means
// facet t owns vertices 3t, 3t+1, 3t+2 (the layout the STL reader produces).
// dims receives the cropped shape; out (x-major, 0/1) is written when non-null.
inline void voxelize(const float* pos, int64_t nv, const int32_t* tris, int64_t nt, int res,
                     int32_t dims[3], int32_t* out, int64_t capacity) {
  if (res < 1 || res > kMaxResolution)
    throw VoxError{SPK_EUNSUPPORTED, "resolution must lie in [1, " + std::to_string(kMaxResolution) + "]"};
  if (nv < 0 || nt < 0 || nt > (int64_t)INT32_MAX || nv > (int64_t)UINT32_MAX)
    throw VoxError{SPK_EINVAL, "bad vertex or triangle count"};
  if ((nv && !pos) || (nt && !tris && nv < 3 * nt)) throw VoxError{SPK_EINVAL, "missing mesh arrays"};
  for (int64_t k = 0; k < 3 * nv; ++k)
    if (!std::isfinite(pos[k])) throw VoxError{SPK_EINVAL, "non-finite vertex coordinate"};
  if (tris)
    for (int64_t k = 0; k < 3 * nt; ++k)
      if (tris[k] < 0 || tris[k] >= nv) throw VoxError{SPK_EINVAL, "triangle index out of range"};
  const char* empty = "(make_voxel_grid_tight) Voxel grid doesn't have any occupied cells";
  if (nv == 0) throw VoxError{SPK_EINVAL, empty};
  std::vector<float> merged;
  const std::vector<unsigned> newindex = merge_exact(pos, nv, merged);
  std::vector<int> pts;
  int g[3];
  to_fixed(merged, res, pts, g);
  const size_t cells = (size_t)g[0] * g[1] * g[2];
  std::vector<unsigned> raw((cells + 3) / 4, 0u);
  if (cells && nt) {
    std::vector<Tri> ht((size_t)nt);
    for (int64_t t = 0; t < nt; ++t)
      for (int k = 0; k < 3; ++k) {
        const int64_t v = tris ? tris[3 * t + k] : 3 * t + k;
        const unsigned id = newindex[(size_t)v];
        ht[(size_t)t].id[k] = id;
        for (int c = 0; c < 3; ++c) ht[(size_t)t].p[3 * k + c] = pts[(size_t)id * 3 + c];
      }
    VoxBuffers b;
    VOX_HIP(hipMalloc((void**)&b.tris, ht.size() * sizeof(Tri)));
    VOX_HIP(hipMalloc((void**)&b.grid, raw.size() * sizeof(unsigned)));
    VOX_HIP(hipMemcpy(b.tris, ht.data(), ht.size() * sizeof(Tri), hipMemcpyHostToDevice));
    VOX_HIP(hipMemset(b.grid, 0, raw.size() * sizeof(unsigned)));
    hipLaunchKernelGGL(k_vox_raster, dim3((unsigned)nt, 3), dim3(kVoxThreads), 0, 0,
                       (const Tri*)b.tris, b.grid, g[0], g[1], g[2]);
    VOX_HIP(hipGetLastError());
    VOX_HIP(hipMemcpy(raw.data(), b.grid, raw.size() * sizeof(unsigned), hipMemcpyDeviceToHost));
  }
  const unsigned char* bytes = reinterpret_cast<const unsigned char*>(raw.data());
  int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {-1, -1, -1};
  for (int x = 0; x < g[0]; ++x)
    for (int y = 0; y < g[1]; ++y) {
      const unsigned char* row = bytes + ((size_t)x * g[1] + y) * g[2];
      for (int z = 0; z < g[2]; ++z)
        if (row[z]) {
          lo[0] = std::min(lo[0], x), hi[0] = std::max(hi[0], x);
          lo[1] = std::min(lo[1], y), hi[1] = std::max(hi[1], y);
          lo[2] = std::min(lo[2], z), hi[2] = std::max(hi[2], z);
        }
    }
  if (hi[0] < 0) throw VoxError{SPK_EINVAL, empty};
  for (int c = 0; c < 3; ++c) dims[c] = hi[c] - lo[c] + 1;
  if (!out) return;
  const int64_t need = (int64_t)dims[0] * dims[1] * dims[2];
  if (capacity < need)
    throw VoxError{SPK_EINVAL, "output capacity " + std::to_string(capacity) + " < " + std::to_string(need)};
  for (int x = 0; x < dims[0]; ++x)
    for (int y = 0; y < dims[1]; ++y) {
      const unsigned char* row = bytes + ((size_t)(x + lo[0]) * g[1] + (y + lo[1])) * g[2] + lo[2];
      int32_t* dst = out + ((size_t)x * dims[1] + y) * dims[2];
      for (int z = 0; z < dims[2]; ++z) dst[z] = row[z] ? 1 : 0;
    }
}

}  // namespace spkvox
