// core_bindings.cpp -- pybind11 module psacking_amd._core.
//
// Same exported names, argument names, defaults and return shapes as the
// reference module spectral_packer._core (spectral_packing/bindings.cpp:654-929).
// All compute goes through the C-ABI of libspk.so (include/spk.h); this file
// only converts numpy arrays and maps status codes to Python exceptions.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spk.h"

namespace py = pybind11;
using IntArray = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

namespace {

void check(int rc) {
  if (rc == SPK_OK) return;
  const std::string msg = spk_last_error();
  if (rc == SPK_EINVAL) throw py::value_error(msg);
  throw std::runtime_error(msg);   // RuntimeError, as the reference's std::runtime_error
}

struct Grid {
  IntArray arr;
  int nx = 0, ny = 0, nz = 0;
  const int32_t* ptr() const { return arr.data(); }
  size_t size() const { return (size_t)nx * ny * nz; }
};

Grid to_grid(py::handle obj, const char* what = "Input array") {
  IntArray a = IntArray::ensure(obj);
  if (!a) throw std::runtime_error(std::string(what) + " must be convertible to an int32 array");
  if (a.ndim() != 3)
    throw std::runtime_error(std::string(what) + " must be 3-dimensional, got " +
                             std::to_string(a.ndim()) + "D");
  Grid g;
  g.arr = a;
  g.nx = (int)a.shape(0);
  g.ny = (int)a.shape(1);
  g.nz = (int)a.shape(2);
  return g;
}

py::array_t<int32_t> new_grid(int nx, int ny, int nz) {
  return py::array_t<int32_t>({nx, ny, nz});
}

// one context per role, created on first use on the current HIP device
struct Contexts {
  spk_ctx* cached = nullptr;    // the tray cache behind the *_with_cache / batch / gpu_tray_* API
  spk_ctx* oneshot = nullptr;   // standalone calls that must not disturb the cache
  uint64_t generation = 0;
  bool have_generation = false;
  uint64_t key = 0;             // hash of (tray, distance) backing the cache
};
Contexts g_ctx;

spk_ctx* get_ctx(spk_ctx** slot) {
  if (!*slot) check(spk_ctx_create(0, slot));
  return *slot;
}

uint64_t hash_grid(uint64_t h, const Grid& g) {
  const size_t n = g.size();
  const uint32_t* p = reinterpret_cast<const uint32_t*>(g.ptr());
  h ^= (uint64_t)g.nx * 0x9E3779B97F4A7C15ULL + ((uint64_t)g.ny << 20) + ((uint64_t)g.nz << 40);
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001B3ULL;
  }
  return h;
}

py::tuple result_tuple(const spk_result& r) {
  return py::make_tuple(py::make_tuple(r.pos[0], r.pos[1], r.pos[2]), (bool)r.found,
                        r.found ? r.score : 0.0);
}

bool larger_than(const Grid& item, const Grid& tray) {
  return item.nx > tray.nx || item.ny > tray.ny || item.nz > tray.nz;
}

// Cache keyed on (generation, tray, distance).  The reference keys on the
// generation alone (bindings.cpp:236-239); the content hash makes a reused
// generation with a different tray recompute instead of returning stale spectra.
void ensure_cached_tray(const Grid& tray, const Grid& dist, uint64_t generation) {
  if (dist.nx != tray.nx || dist.ny != tray.ny || dist.nz != tray.nz)
    throw std::runtime_error("tray_distance must have the tray's shape");
  spk_ctx* c = get_ctx(&g_ctx.cached);
  int valid = 0;
  check(spk_tray_info(c, nullptr, nullptr, nullptr, nullptr, &valid));
  const uint64_t key = hash_grid(hash_grid(0xcbf29ce484222325ULL, tray), dist);
  if (!valid || !g_ctx.have_generation || generation != g_ctx.generation || key != g_ctx.key) {
    check(spk_tray_set(c, tray.ptr(), tray.nx, tray.ny, tray.nz, dist.ptr(), SPK_METRIC_L1,
                       SPK_SOURCES_EQ1, generation));
    g_ctx.generation = generation;
    g_ctx.have_generation = true;
    g_ctx.key = key;
  }
}

std::vector<Grid> grids_of(py::list orientations) {
  std::vector<Grid> out;
  for (auto o : orientations) out.push_back(to_grid(o, "Orientation array"));
  return out;
}

spk_result search_list(const std::vector<Grid>& items, int penalty, int interlocking_free) {
  std::vector<const int32_t*> ptrs;
  std::vector<int32_t> dims;
  for (const auto& g : items) {
    ptrs.push_back(g.ptr());
    dims.push_back(g.nx);
    dims.push_back(g.ny);
    dims.push_back(g.nz);
  }
  spk_result best;
  check(spk_search(g_ctx.cached, ptrs.data(), dims.data(), (int)items.size(), penalty,
                   interlocking_free, &best, nullptr));
  return best;
}

// ---------------------------------------------------------------------------

py::tuple py_fft_search_placement(py::object item, py::object tray) {
  Grid it = to_grid(item), tr = to_grid(tray);
  if (larger_than(it, tr)) return py::make_tuple(py::make_tuple(-1, -1, -1), false, 0.0);
  spk_ctx* c = get_ctx(&g_ctx.oneshot);
  check(spk_tray_set(c, tr.ptr(), tr.nx, tr.ny, tr.nz, nullptr, SPK_METRIC_L1, SPK_SOURCES_EQ1, 0));
  const int32_t* ptrs[1] = {it.ptr()};
  int32_t dims[3] = {it.nx, it.ny, it.nz};
  spk_result best;
  check(spk_search(c, ptrs, dims, 1, SPK_PENALTY_POW, 0, &best, nullptr));
  return result_tuple(best);
}

py::tuple py_fft_search_placement_with_cache(py::object item, py::object tray,
                                             py::object tray_distance, uint64_t generation) {
  Grid it = to_grid(item), tr = to_grid(tray), d = to_grid(tray_distance);
  if (larger_than(it, tr)) return py::make_tuple(py::make_tuple(-1, -1, -1), false, 0.0);
  ensure_cached_tray(tr, d, generation);
  return result_tuple(search_list({it}, SPK_PENALTY_CUBE, 0));
}

py::tuple py_fft_search_batch(py::list orientations, py::object tray, py::object tray_distance,
                              uint64_t generation) {
  Grid tr = to_grid(tray), d = to_grid(tray_distance);
  ensure_cached_tray(tr, d, generation);
  std::vector<Grid> items = grids_of(orientations);
  // fft3.cu:1560: the batch starts from best_score = 1e30
  if (items.empty()) return py::make_tuple(py::make_tuple(-1, -1, -1), false, 1e30);
  spk_result best = search_list(items, SPK_PENALTY_CUBE, 0);
  if (!best.found) return py::make_tuple(py::make_tuple(-1, -1, -1), false, 1e30);
  return result_tuple(best);
}

py::tuple py_fft_search_batch_interlocking_free(py::list orientations, py::object tray,
                                                py::object tray_distance, uint64_t generation) {
  Grid tr = to_grid(tray), d = to_grid(tray_distance);
  ensure_cached_tray(tr, d, generation);
  std::vector<Grid> items = grids_of(orientations);
  if (items.empty()) return py::make_tuple(py::make_tuple(-1, -1, -1), false, 1e30, -1);
  spk_result best = search_list(items, SPK_PENALTY_CUBE, 1);
  if (!best.found) return py::make_tuple(py::make_tuple(-1, -1, -1), false, 1e30, -1);
  return py::make_tuple(py::make_tuple(best.pos[0], best.pos[1], best.pos[2]), true, best.score,
                        best.orient);
}

py::array_t<int32_t> py_place_in_tray(py::object item, py::object tray, py::tuple position,
                                      int item_id) {
  Grid it = to_grid(item), tr = to_grid(tray);
  const int x = position[0].cast<int>(), y = position[1].cast<int>(), z = position[2].cast<int>();
  auto out = new_grid(tr.nx, tr.ny, tr.nz);
  check(spk_place_grid(get_ctx(&g_ctx.oneshot), tr.ptr(), tr.nx, tr.ny, tr.nz, it.ptr(), it.nx,
                       it.ny, it.nz, x, y, z, item_id, out.mutable_data()));
  return out;
}

py::array_t<int32_t> py_voxelize_stl(const std::string& path, int resolution) {
  // bindings.cpp:285-302: every failure surfaces as RuntimeError with this prefix
  auto fail_vox = [&]() {
    throw std::runtime_error("Voxelization failed for " + path + ": " + spk_last_error());
  };
  int32_t dims[3] = {0, 0, 0};
  if (spk_voxelize_stl(0, path.c_str(), resolution, dims, nullptr, 0) != SPK_OK) fail_vox();
  auto out = new_grid(dims[0], dims[1], dims[2]);
  if (spk_voxelize_stl(0, path.c_str(), resolution, dims, out.mutable_data(), (int64_t)out.size()) !=
      SPK_OK)
    fail_vox();
  return out;
}

py::array_t<int32_t> correlate(py::object a, py::object b, int convolve) {
  Grid ga = to_grid(a), gb = to_grid(b);
  if (ga.nx != gb.nx || ga.ny != gb.ny || ga.nz != gb.nz)
    throw std::runtime_error("Input grids must be of the same size for convolution");
  auto out = new_grid(ga.nx, ga.ny, ga.nz);
  check(spk_correlate(get_ctx(&g_ctx.oneshot), ga.ptr(), gb.ptr(), ga.nx, ga.ny, ga.nz, convolve,
                      out.mutable_data()));
  return out;
}

py::array_t<int32_t> py_calculate_distance(py::object grid) {
  Grid g = to_grid(grid);
  auto out = new_grid(g.nx, g.ny, g.nz);
  check(spk_distance(get_ctx(&g_ctx.oneshot), g.ptr(), g.nx, g.ny, g.nz, SPK_METRIC_L1,
                     SPK_SOURCES_EQ1, out.mutable_data()));
  return out;
}

py::array_t<int32_t> py_distance_field(py::object grid, const std::string& metric,
                                       const std::string& sources) {
  Grid g = to_grid(grid);
  int m = metric == "edt" ? SPK_METRIC_EDT : metric == "l1" ? SPK_METRIC_L1 : -1;
  int s = sources == "nonzero" ? SPK_SOURCES_NONZERO : sources == "eq1" ? SPK_SOURCES_EQ1 : -1;
  if (m < 0 || s < 0) throw py::value_error("metric must be 'l1'|'edt', sources 'eq1'|'nonzero'");
  auto out = new_grid(g.nx, g.ny, g.nz);
  check(spk_distance(get_ctx(&g_ctx.oneshot), g.ptr(), g.nx, g.ny, g.nz, m, s,
                     out.mutable_data()));
  return out;
}

py::array_t<int32_t> py_collision_grid(py::object tray, py::object item) {
  Grid tr = to_grid(tray), it = to_grid(item);
  auto out = new_grid(tr.nx, tr.ny, tr.nz);
  check(spk_collision_grid(get_ctx(&g_ctx.oneshot), tr.ptr(), tr.nx, tr.ny, tr.nz, it.ptr(), it.nx,
                           it.ny, it.nz, out.mutable_data()));
  return out;
}

py::tuple bounds(const Grid& g) {
  int32_t lo[3], hi[3];
  check(spk_bounds(g.ptr(), g.nx, g.ny, g.nz, lo, hi));
  return py::make_tuple(py::make_tuple(lo[0], lo[1], lo[2]), py::make_tuple(hi[0], hi[1], hi[2]));
}

py::array_t<int32_t> py_make_tight(py::object grid) {
  Grid g = to_grid(grid);
  int32_t lo[3], hi[3];
  check(spk_bounds(g.ptr(), g.nx, g.ny, g.nz, lo, hi));
  if (hi[0] < 0)
    throw py::value_error("(make_voxel_grid_tight) Voxel grid doesn't have any occupied cells");
  const int dx = hi[0] - lo[0] + 1, dy = hi[1] - lo[1] + 1, dz = hi[2] - lo[2] + 1;
  auto out = new_grid(dx, dy, dz);
  int32_t* o = out.mutable_data();
  const int32_t* p = g.ptr();
  for (int i = 0; i < dx; ++i)
    for (int j = 0; j < dy; ++j)
      std::memcpy(o + ((size_t)i * dy + j) * dz,
                  p + ((size_t)(i + lo[0]) * g.ny + (j + lo[1])) * g.nz + lo[2], dz * sizeof(int32_t));
  return out;
}

py::tuple py_get_bounds(py::object grid) { return bounds(to_grid(grid)); }

void py_save_vox(py::object grid, const std::string& path) {
  Grid g = to_grid(grid);
  check(spk_save_vox(g.ptr(), g.nx, g.ny, g.nz, path.c_str()));
}

void py_gpu_tray_init(py::object tray, py::object tray_distance) {
  Grid tr = to_grid(tray), d = to_grid(tray_distance);
  if (d.nx != tr.nx || d.ny != tr.ny || d.nz != tr.nz)
    throw std::runtime_error("tray_distance must have the tray's shape");
  spk_ctx* c = get_ctx(&g_ctx.cached);
  check(spk_tray_set(c, tr.ptr(), tr.nx, tr.ny, tr.nz, d.ptr(), SPK_METRIC_L1, SPK_SOURCES_EQ1,
                     g_ctx.generation));
  // an explicit init does not correspond to any generation (fft3.cu:1475-1480)
  g_ctx.have_generation = false;
}

void py_gpu_tray_cleanup() {
  if (g_ctx.cached) check(spk_tray_clear(g_ctx.cached));
  g_ctx.have_generation = false;
}

bool py_gpu_tray_is_initialized() {
  if (!g_ctx.cached) return false;
  int valid = 0;
  check(spk_tray_info(g_ctx.cached, nullptr, nullptr, nullptr, nullptr, &valid));
  return valid != 0;
}

py::tuple py_gpu_tray_search(py::object item) {
  if (!py_gpu_tray_is_initialized())
    throw std::runtime_error("GPU tray context not initialized. Call gpu_tray_init() first.");
  Grid it = to_grid(item);
  int nx, ny, nz;
  check(spk_tray_info(g_ctx.cached, &nx, &ny, &nz, nullptr, nullptr));
  if (it.nx > nx || it.ny > ny || it.nz > nz)
    return py::make_tuple(py::make_tuple(-1, -1, -1), false, 0.0);
  return result_tuple(search_list({it}, SPK_PENALTY_CUBE, 0));
}

py::array_t<int32_t> py_interlocking_free_positions(py::object tray, py::object obj) {
  Grid tr = to_grid(tray, "Tray array"), o = to_grid(obj, "Object array");
  auto out = new_grid(tr.nx, tr.ny, tr.nz);
  check(spk_interlock_mask(get_ctx(&g_ctx.oneshot), tr.ptr(), tr.nx, tr.ny, tr.nz, o.ptr(), o.nx,
                           o.ny, o.nz, out.mutable_data()));
  return out;
}

// ---- extensions: the fused device-resident greedy loop and helpers ---------

py::dict result_dict(const spk_result& r) {
  py::dict d;
  d["position"] = py::make_tuple(r.pos[0], r.pos[1], r.pos[2]);
  d["found"] = (bool)r.found;
  d["score"] = r.score;
  d["orientation"] = r.orient;
  d["collision"] = r.collision;
  d["proximity"] = r.proximity;
  d["residual"] = r.residual;
  d["path"] = r.path;
  return d;
}

std::vector<spk_orient> orients_of(py::object perms, py::object flips) {
  auto P = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(perms);
  auto F = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(flips);
  if (!P || !F || P.ndim() != 2 || F.ndim() != 2 || P.shape(1) != 3 || F.shape(1) != 3 ||
      P.shape(0) != F.shape(0))
    throw py::value_error("perms and flips must be (n, 3) integer arrays");
  std::vector<spk_orient> out(P.shape(0));
  for (ssize_t o = 0; o < P.shape(0); ++o)
    for (int a = 0; a < 3; ++a) {
      out[o].perm[a] = P.at(o, a);
      out[o].flip[a] = F.at(o, a);
    }
  return out;
}

int metric_of(const std::string& m) {
  if (m == "l1") return SPK_METRIC_L1;
  if (m == "edt") return SPK_METRIC_EDT;
  throw py::value_error("metric must be 'l1' or 'edt'");
}
int sources_of(const std::string& s) {
  if (s == "eq1") return SPK_SOURCES_EQ1;
  if (s == "nonzero") return SPK_SOURCES_NONZERO;
  throw py::value_error("sources must be 'eq1' or 'nonzero'");
}
int penalty_of(const std::string& p) {
  if (p == "pow") return SPK_PENALTY_POW;
  if (p == "cube") return SPK_PENALTY_CUBE;
  throw py::value_error("penalty must be 'pow' or 'cube'");
}

// A packing session: its own context holding the tray on the device.
class Session {
 public:
  explicit Session(int device) { check(spk_ctx_create(device, &ctx_)); }
  ~Session() { spk_ctx_destroy(ctx_); }
  void set_tray(py::object tray, const std::string& metric, const std::string& sources,
                uint64_t generation) {
    Grid tr = to_grid(tray);
    check(spk_tray_set(ctx_, tr.ptr(), tr.nx, tr.ny, tr.nz, nullptr, metric_of(metric),
                       sources_of(sources), generation));
  }
  py::array_t<int32_t> tray() {
    int nx, ny, nz;
    check(spk_tray_info(ctx_, &nx, &ny, &nz, nullptr, nullptr));
    auto out = new_grid(nx, ny, nz);
    check(spk_tray_get(ctx_, out.mutable_data(), nullptr));
    return out;
  }
  py::array_t<int32_t> distance() {
    int nx, ny, nz;
    check(spk_tray_info(ctx_, &nx, &ny, &nz, nullptr, nullptr));
    auto out = new_grid(nx, ny, nz);
    check(spk_tray_get(ctx_, nullptr, out.mutable_data()));
    return out;
  }
  py::tuple field_stats() {
    int64_t sum = 0;
    int32_t mn = 0, mx = 0, mxw = 0;
    check(spk_tray_stats(ctx_, &sum, &mn, &mx, &mxw));
    return py::make_tuple(sum, mn, mx, mxw);
  }
  py::list search(py::object item, py::object perms, py::object flips, const std::string& penalty,
                  bool interlocking_free) {
    Grid it = to_grid(item);
    auto os = orients_of(perms, flips);
    spk_result best;
    std::vector<spk_result> per(os.size());
    {
      py::gil_scoped_release nogil;
      check(spk_search_oriented(ctx_, it.ptr(), it.nx, it.ny, it.nz, os.data(), (int)os.size(),
                                penalty_of(penalty), interlocking_free ? 1 : 0, &best, per.data()));
    }
    py::list l;
    l.append(result_dict(best));
    py::list po;
    for (auto& r : per) po.append(result_dict(r));
    l.append(po);
    return l;
  }
  void place(py::object item, py::object perm, py::object flip, py::tuple pos, int id) {
    Grid it = to_grid(item);
    spk_orient o;
    auto P = py::array_t<int32_t, py::array::forcecast>::ensure(perm);
    auto F = py::array_t<int32_t, py::array::forcecast>::ensure(flip);
    if (!P || !F || P.size() != 3 || F.size() != 3) throw py::value_error("perm/flip need 3 entries");
    for (int a = 0; a < 3; ++a) {
      o.perm[a] = P.data()[a];
      o.flip[a] = F.data()[a];
    }
    check(spk_place(ctx_, it.ptr(), it.nx, it.ny, it.nz, &o, pos[0].cast<int>(), pos[1].cast<int>(),
                    pos[2].cast<int>(), id));
  }
  py::list pack(py::list items, py::object order, py::object perms, py::object flips,
                const std::string& penalty, bool interlocking_free, int id_offset) {
    std::vector<Grid> gs = grids_of(items);
    std::vector<int32_t> dims;
    std::vector<const int32_t*> ptrs;   // the arrays themselves: no concatenated copy
    for (auto& g : gs) {
      dims.push_back(g.nx);
      dims.push_back(g.ny);
      dims.push_back(g.nz);
      ptrs.push_back(g.ptr());
    }
    auto ord = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(order);
    if (!ord || ord.ndim() != 1 || (size_t)ord.shape(0) != gs.size())
      throw py::value_error("order must list every item once");
    auto os = orients_of(perms, flips);
    std::vector<spk_result> res(gs.size());
    {
      py::gil_scoped_release nogil;
      check(spk_pack_gather(ctx_, ptrs.data(), dims.data(), (int)gs.size(), ord.data(), os.data(),
                     (int)os.size(), penalty_of(penalty), interlocking_free ? 1 : 0, id_offset,
                     res.data()));
    }
    py::list out;
    for (auto& r : res) out.append(result_dict(r));
    return out;
  }
  void set_option(const std::string& name, double value) {
    check(spk_set_option(ctx_, name.c_str(), value));
  }
  void timing(bool on) { check(spk_timing_enable(ctx_, on ? 1 : 0)); }
  py::tuple timing_get(const std::string& name) {
    double ms = 0;
    int64_t n = 0;
    check(spk_timing_get(ctx_, name.c_str(), &ms, &n));
    return py::make_tuple(ms, n);
  }
  int64_t timing_units(const std::string& name) {
    int64_t u = 0;
    check(spk_timing_units(ctx_, name.c_str(), &u));
    return u;
  }
  void timing_reset() { check(spk_timing_reset(ctx_)); }
  py::tuple plan() {
    int px, py_, pz, b;
    check(spk_plan_info(ctx_, &px, &py_, &pz, &b));
    return py::make_tuple(px, py_, pz);
  }
  py::tuple tile() {
    int t = 0, split = 0;
    check(spk_tile_info(ctx_, &t, &split));
    return py::make_tuple(t, split);
  }

  spk_ctx* ctx() const { return ctx_; }

 private:
  spk_ctx* ctx_ = nullptr;
};

// Items of one instance flattened for the C-ABI.
struct PackJob {
  std::vector<int32_t> flat, dims, order;
  std::vector<spk_orient> orients;
  std::vector<spk_result> res;
};

py::list py_pack_many(py::list sessions, py::list item_lists, py::list orders, py::object perms,
                      py::object flips, const std::string& penalty, bool interlocking_free) {
  const size_t n = py::len(sessions);
  if (py::len(item_lists) != n || py::len(orders) != n)
    throw py::value_error("sessions, item_lists and orders must have the same length");
  std::vector<spk_ctx*> ctxs(n);
  std::vector<PackJob> jobs(n);
  std::vector<spk_pack_args> args(n);
  auto os = orients_of(perms, flips);
  for (size_t i = 0; i < n; ++i) {
    ctxs[i] = sessions[i].cast<Session&>().ctx();
    std::vector<Grid> gs = grids_of(item_lists[i].cast<py::list>());
    PackJob& j = jobs[i];
    for (auto& g : gs) {
      j.dims.push_back(g.nx);
      j.dims.push_back(g.ny);
      j.dims.push_back(g.nz);
      j.flat.insert(j.flat.end(), g.ptr(), g.ptr() + g.size());
    }
    auto ord = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(orders[i]);
    if (!ord || ord.ndim() != 1 || (size_t)ord.shape(0) != gs.size())
      throw py::value_error("order must list every item once");
    j.order.assign(ord.data(), ord.data() + gs.size());
    j.orients = os;
    j.res.resize(gs.size());
    spk_pack_args& a = args[i];
    a.items = j.flat.data();
    a.dims = j.dims.data();
    a.n_items = (int)gs.size();
    a.order = j.order.data();
    a.orients = j.orients.data();
    a.n_orient = (int)j.orients.size();
    a.penalty = penalty_of(penalty);
    a.interlocking_free = interlocking_free ? 1 : 0;
    a.id_offset = 0;
    a.results = j.res.data();
  }
  {
    py::gil_scoped_release nogil;
    check(spk_pack_many(ctxs.data(), (int)n, args.data()));
  }
  py::list out;
  for (auto& j : jobs) {
    py::list l;
    for (auto& r : j.res) l.append(result_dict(r));
    out.append(l);
  }
  return out;
}

}  // namespace

PYBIND11_MODULE(_core, m) {
  m.doc() = "MI355X spectral 3D bin packing core (HIP kernels behind a C-ABI)";

  m.def("fft_search_placement", &py_fft_search_placement, py::arg("item"), py::arg("tray"));
  m.def("fft_search_placement_with_cache", &py_fft_search_placement_with_cache, py::arg("item"),
        py::arg("tray"), py::arg("tray_distance"), py::arg("generation"));
  m.def("place_in_tray", &py_place_in_tray, py::arg("item"), py::arg("tray"), py::arg("position"),
        py::arg("item_id") = 1);
  m.def("voxelize_stl", &py_voxelize_stl, py::arg("path"), py::arg("resolution") = 128);
  m.def("dft_conv3", [](py::object a, py::object b) { return correlate(a, b, 1); }, py::arg("a"),
        py::arg("b"));
  m.def("dft_corr3", [](py::object a, py::object b) { return correlate(a, b, 0); }, py::arg("a"),
        py::arg("b"));
  m.def("calculate_distance", &py_calculate_distance, py::arg("grid"));
  m.def("collision_grid", &py_collision_grid, py::arg("tray"), py::arg("item"));
  m.def("make_tight", &py_make_tight, py::arg("grid"));
  m.def("get_bounds", &py_get_bounds, py::arg("grid"));
  m.def("save_vox", &py_save_vox, py::arg("grid"), py::arg("path"));
  m.def("gpu_tray_init", &py_gpu_tray_init, py::arg("tray"), py::arg("tray_distance"));
  m.def("gpu_tray_cleanup", &py_gpu_tray_cleanup);
  m.def("gpu_tray_is_initialized", &py_gpu_tray_is_initialized);
  m.def("gpu_tray_search", &py_gpu_tray_search, py::arg("item"));
  m.def("fft_search_batch", &py_fft_search_batch, py::arg("orientations"), py::arg("tray"),
        py::arg("tray_distance"), py::arg("generation"));
  m.def("fft_search_batch_interlocking_free", &py_fft_search_batch_interlocking_free,
        py::arg("orientations"), py::arg("tray"), py::arg("tray_distance"), py::arg("generation"));
  m.def("interlocking_free_positions", &py_interlocking_free_positions, py::arg("tray"),
        py::arg("obj"));

  // extensions
  m.def("distance_field", &py_distance_field, py::arg("grid"), py::arg("metric") = "l1",
        py::arg("sources") = "eq1");
  py::class_<Session>(m, "Session")
      .def(py::init<int>(), py::arg("device") = 0)
      .def("set_tray", &Session::set_tray, py::arg("tray"), py::arg("metric") = "l1",
           py::arg("sources") = "eq1", py::arg("generation") = 0)
      .def("tray", &Session::tray)
      .def("distance", &Session::distance)
      .def("field_stats", &Session::field_stats)
      .def("search", &Session::search, py::arg("item"), py::arg("perms"), py::arg("flips"),
           py::arg("penalty") = "cube", py::arg("interlocking_free") = false)
      .def("place", &Session::place, py::arg("item"), py::arg("perm"), py::arg("flip"),
           py::arg("position"), py::arg("item_id"))
      .def("pack", &Session::pack, py::arg("items"), py::arg("order"), py::arg("perms"),
           py::arg("flips"), py::arg("penalty") = "cube", py::arg("interlocking_free") = false,
           py::arg("id_offset") = 0)
      .def("set_option", &Session::set_option, py::arg("name"), py::arg("value"))
      .def("timing", &Session::timing, py::arg("on"))
      .def("timing_get", &Session::timing_get, py::arg("name"))
      .def("timing_units", &Session::timing_units, py::arg("name"))
      .def("timing_reset", &Session::timing_reset)
      .def("plan", &Session::plan)
      .def("tile", &Session::tile)
      .def("handle", [](const Session& s) { return (uintptr_t)s.ctx(); },
           "the spk_ctx* of this session, for C-ABI calls (ctypes)");

  m.def("pack_many", &py_pack_many, py::arg("sessions"), py::arg("item_lists"), py::arg("orders"),
        py::arg("perms"), py::arg("flips"), py::arg("penalty") = "cube",
        py::arg("interlocking_free") = false);

  m.attr("VOXEL_RESOLUTION") = 128;
  m.attr("HEIGHT_PENALTY") = 100000000.0;
  m.attr("__version__") = "0.1.0";
}
