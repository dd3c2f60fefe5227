// spk_api.hip -- host side of libspk.so: context, FFT planning, launches,
// exact fallback, and the extern "C" entry points declared in include/spk.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <thread>
#include <vector>

#include "../../include/spk.h"
#include "spk_kernels.hpp"
#include "spk_nn.hpp"

using namespace spk;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct HipError {
  int code;
  std::string msg;
};

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw HipError{e_ == hipErrorOutOfMemory ? SPK_ENOMEM : SPK_EDEVICE,        \
                     std::string(#expr) + ": " + hipGetErrorString(e_)};          \
  } while (0)

struct SpkError {
  int code;
  std::string msg;
};

// smooth sizes whose factors are all supported radices
bool smooth(int n) {
  for (int r : {2, 3, 5, 7})
    while (n % r == 0) n /= r;
  return n == 1;
}

// Radices with a register butterfly (spk_fft.hpp).
const int kRadices[] = {25, 16, 15, 14, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2};

// Fewest stages first (each stage is one LDS round trip and two barriers),
// then the most even split; larger radices run first.
AxisPlan make_plan(int n, int rmax = kMaxRadix) {
  std::map<int, std::vector<int>> memo;
  std::function<std::vector<int>(int)> best = [&](int m) -> std::vector<int> {
    if (m == 1) return {};
    auto it = memo.find(m);
    if (it != memo.end()) return it->second;
    std::vector<int> b;
    bool have = false;
    for (int r : kRadices) {
      if (r > rmax || m % r) continue;
      std::vector<int> c = best(m / r);
      if (!c.empty() && c[0] == -1) continue;
      c.push_back(r);
      auto score = [](const std::vector<int>& v) {
        int mx = 0;
        for (int x : v) mx = std::max(mx, x);
        return std::make_pair((int)v.size(), mx);
      };
      if (!have || score(c) < score(b)) { b = c; have = true; }
    }
    if (!have) b = {-1};
    memo[m] = b;
    return b;
  };
  std::vector<int> rs = best(n);
  if (!rs.empty() && rs[0] == -1) throw SpkError{SPK_EUNSUPPORTED, "FFT length has no plan"};
  std::sort(rs.begin(), rs.end(), [](int x, int y) { return x > y; });
  AxisPlan p{};
  p.n = n;
  p.nst = 0;
  p.rad = 0;
  for (int v : rs) p.rad |= (unsigned long long)v << (kRadixBits * p.nst++);
  return p;
}

int next_smooth(int n) {
  int m = std::max(n, 1);
  while (!smooth(m)) ++m;
  return m;
}

// A device buffer that grows on demand and frees itself: owners (contexts,
// proposers) are destroyed with their device current, after their stream
// has drained, so every member is released without being listed by hand.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap) { o.p = nullptr; o.cap = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; cap = o.cap;
      o.p = nullptr; o.cap = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void ensure(size_t n) {
    if (n <= cap) return;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    HIP_OK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
    cap = n;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Timed {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace

struct spk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // tray
  bool valid = false;
  int nx = 0, ny = 0, nz = 0;
  uint64_t generation = 0;
  int metric = SPK_METRIC_L1, sources = SPK_SOURCES_EQ1;
  AxisPlan px{}, py{}, pz{};
  int T = 16;                  // x-pencil tile
  int Tc = 16;                 // x-pencil tile of pass A (tuned on its own)
  int Ty = 16;                 // y-pencil tile (split plan)
  int Rf = 1, Rs = 1;          // rows per block: forward rows, score rows (split plan)
  bool split = false;          // split plan in use
  int plan_mode = 0;           // 0 auto, 1 slab, 2 split
  bool lds_attrs = false;
  DevBuf<int> tray, phi, tmpi;
  DevBuf<FieldStats> stats;
  DevBuf<FieldStats> stats_part;         // per-block partial field statistics
  DevBuf<float2> tw, t1, t2, g1, g2;
  DevBuf<double> pen;          // [2][nz]: pow, cube
  // search workspace
  DevBuf<int> items;
  DevBuf<OrientDesc> od;
  DevBuf<float2> ispec, y1, y2;
  DevBuf<WgBest> rec, per, rec2;   // rec2: interlocking-free masked records
  DevBuf<SearchOut> sout;
  DevBuf<long long> ver;
  DevBuf<int> flag;
  DevBuf<Tap> taps;
  DevBuf<long long> big;       // standalone full-grid outputs
  DevBuf<unsigned char> ffree, freach;   // IF: feasibility / reached grids [m][V]
  DevBuf<long long> pgrid;               // IF: proximity grids [m][V]
  DevBuf<int> if_keep;                   // IF: 1 when the unmasked winner is the masked one
  DevBuf<int> exits;                     // IF: high seed faces [m][3]
  DevBuf<unsigned char> ff_act;          // IF: tiles changed in the last round (two buffers)
  DevBuf<U128> fb_free, fb_reach;        // IF, nz <= 128: free / reached bit columns [m][nx][ny]
  bool ff_bits = true;                   // option "ff_bits": bit-column flood fill
  DevBuf<int> ff_flags;                  // IF: go / changed flags of a batch of rounds
  // greedy loop
  DevBuf<OrientDesc> od_all;             // descriptors of every item's fitting orientations
  DevBuf<PackOut> pout;                  // per-item outcome decided on the device
  DevBuf<int> ctl;                       // CTL_* control words
  DevBuf<int> zero;                      // an int that stays 0 (overlap flag of pipelined places)
  float resid_limit = 0.25f;             // largest accepted FFT rounding residual
  // learned proposals and training records
  DevBuf<int> heights;                   // column heights [nx][ny]
  DevBuf<spknn::bf16_t> icat;            // learned loop: head rows [item | tray] per item slot
  bool tray_nonneg = true;               // no tray voxel is negative (the tray features can
                                         // then come from the column-test masks)
  DevBuf<int> map_all;                   // orientation index of each od_all entry
  DevBuf<unsigned short> tfeat;          // tray features [64][32][32][2] (bf16)
  DevBuf<unsigned short> ifeat, ipool;   // oriented item features, pooled rows [n][24][384]
  DevBuf<spknn::ItemMap> imaps;          // orientation maps of the items [n][24]
  DevBuf<double> rscores;                // per-orientation best scores [n][n_orient]
  DevBuf<OrientDesc> od_prop;            // descriptors of the proposed orientations
  DevBuf<OrientDesc> od_rest;            // descriptors of the other orientations searched
  DevBuf<int> learn_gate;                // device-driven learned loop: proposer gate word,
  DevBuf<int> learn_pick, learn_used;    // per-item pick (fitting indices) and pick kind
  bool learn_dev = true;                 // option "learn_dev": device-driven learned loop
  int prescreen = 16;                    // option "prescreen": items per round of the device
                                         // loops, screened together first (0: one round, none)
  DevBuf<int> pre_fs;                    // flags of a batched screen
  DevBuf<int> pre_aux;                   // its per-item counts and orientation items
  int many_workers = 8;                  // option "many_workers": spk_pack_many host workers
  bool graph = false;                    // option "graph": device loops captured as a hipGraph
  bool fixed_fft = true;                 // option "fixed_fft": compile-time-plan score kernels
  int spec_inc = 16;                     // option "spec_inc": incremental G1 updates in a row
  int inc_split = 8;                     // option "inc_split": k_inc_x blocks per line group
  bool score2 = true;                    // option "score2": score pass with both slabs in LDS
  int lds_max = 0;                       // the device's LDS bytes per block
  int reg_x = 2;                         // option "reg_x": register pass A for px = 240
                                         // (0 off, 1 three LDS buffers, 2 two)
  // feasibility pass
  DevBuf<int> fstate;                    // residual bits, unused, per-orientation flags
  DevBuf<int> ctl_on;                    // control words that enable the pass outside the loop
  int feas_mode = 1;                     // 0 off, 1 after an item that placed nothing, 2 always
  bool autotune = true;                  // time x-pencil tiles per tray shape (option "autotune")
  int tile_force = 0;                    // option "tile": 0 tuned/default, else 4, 8 or 16
  int yz_rmax = kMaxRadix;               // option "yz_radix": largest y/z butterfly
  int score_rmax = 10;                   // option "score_radix": largest butterfly of the score pass
  int corr_chunk = 0;                    // option "corr_chunk": orientations per pass-A block (0 auto)
  AxisPlan sy{}, sz{};                   // y and z plans of the score pass
  int pad_y = 0, pad_z = 0;              // options "pad_y", "pad_z": padded sizes (0 = smallest smooth)
  bool pad_forced = false;               // pad_y / pad_z set by an option: not autotuned
  bool score_radix_forced = false;       // score_radix set by an option: not autotuned
  int feas_group = 0;                    // option "feas_group": pairs per feasibility launch pair
  bool screen_all = true;                // option "screen_all": learned loop screens every item
                                         // with the column test before proposing
  int dist_y = 2;                        // option "dist_y": 1 envelope, 2 full scan
  int dist_x = 0;                        // option "dist_x": 0 outward search, 2 full scan
  bool inc_dist = true;                  // option "inc_dist": incremental field after a placement
  bool inc_masks = true;                 // option "inc_masks": column-test masks updated under
                                         // the placed box only
  // column test before the feasibility pass (option "coarse")
  struct Coarse {
    bool on = true;
    bool ready = false;                  // tray z <= 128 and the slab plan
    int cx = 0, cy = 0;                  // 2x2 column blocks of the tray
    DevBuf<M128> cols, cm;               // block masks, core masks of a search
    DevBuf<M128> tf, im;                 // full-resolution tray columns, item columns
    DevBuf<unsigned> rl, frl;            // sorted run lists of the core and item columns
    DevBuf<int> nrun;                    // their lengths
    DevBuf<int> ht;                      // tile heights of the tray (kSkyTile columns square)
    DevBuf<M128> lines;                  // line masks of the packed items (k_item_lines)
    DevBuf<ItemLines> il;
    int tx = 0, ty = 0;
    DevBuf<int> fs;                      // residual slot (0), per-orientation flags
    int fine_group = 64;                 // option "fine_group": lanes per full-resolution candidate
    int col_chunks = 1;                  // option "col_chunks": 64-offset chunks per column-test block
  } co;
  // free-space certificate
  bool phi_internal = false;             // phi is this library's distance field of the tray
  bool certificate = false;              // option "certificate"
  DevBuf<FieldStats> theta;              // per item: inscribed-ball bound (mx)
  DevBuf<int> cgrid, cdist, ctmp;        // padded item complement and its distance field
  DevBuf<int> gates;                     // control words gating speculative screens
  // timing
  bool timing = false;
  std::vector<Timed> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long long>> acc;
  std::map<std::string, long long> units;   // orientations (pairs) handed to each kernel
  DevBuf<unsigned long long> pairs;          // feasibility pairs transformed (full, coarse),
                                             // orientations the score pass searched
  Mailbox* mail_h = nullptr;                 // pinned host mailbox and its device address
  Mailbox* mail_d = nullptr;
  unsigned mail_seq = 0;                     // sequence number of the last post
  int* stage = nullptr;                      // pinned staging buffer of the items
  size_t stage_cap = 0;
  int* pre_h = nullptr;                      // pinned: batched screen's aux upload and flags
  size_t pre_h_cap = 0;
  hipEvent_t stage_ev = nullptr;             // recorded after its last copy
  // host-side copy of item buffer bookkeeping
  std::vector<int> h_items;
  void* scratch = nullptr;                   // device address of this device's static scratch

  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    return e;
  }
  void flush_timing() {
    for (auto& t : pending) {
      HIP_OK(hipEventSynchronize(t.b));
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, t.a, t.b));
      auto& a = acc[t.name];
      a.first += ms;
      a.second += 1;
      pool.push_back(t.a);
      pool.push_back(t.b);
    }
    pending.clear();
  }
};

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    hipGetDevice(&prev);
    if (prev != dev) hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    hipGetDevice(&cur);
    if (prev >= 0 && cur != prev) hipSetDevice(prev);
  }
};

template <typename F>
int guarded(spk_ctx* ctx, F&& f) {
  if (!ctx) return fail(SPK_EINVAL, "null context");
  try {
    DeviceGuard g(ctx->device);
    return f();
  } catch (const HipError& e) {
    return fail(e.code, e.msg);
  } catch (const SpkError& e) {
    return fail(e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(SPK_ENOMEM, "host allocation failed");
  }
}

// launch with optional event timing
template <typename K, typename... Args>
void launch(spk_ctx* c, const char* name, K kernel, dim3 grid, dim3 block, size_t shmem,
            Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  hipEvent_t a = nullptr, b = nullptr;
  if (c->timing) {
    a = c->ev();
    b = c->ev();
    HIP_OK(hipEventRecord(a, c->stream));
  }
  hipLaunchKernelGGL(kernel, grid, block, shmem, c->stream, args...);
  HIP_OK(hipGetLastError());
  if (c->timing) {
    HIP_OK(hipEventRecord(b, c->stream));
    c->pending.push_back({name, a, b});
  }
}

template <typename K>
void allow_shmem(K kernel, size_t bytes) {
  if (bytes > 65536)
    HIP_OK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
}

void sync(spk_ctx* c) { HIP_OK(hipStreamSynchronize(c->stream)); }

// Host sources are often temporaries (pageable memory), so an upload returns
// only once the copy has been consumed.
void upload(spk_ctx* c, void* dst, const void* src, size_t bytes) {
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
}
void download(spk_ctx* c, void* dst, const void* src, size_t bytes) {
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
}

void check_dims(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0)
    throw SpkError{SPK_EINVAL, "grid dimensions must be positive"};
}

}  // namespace

#include "spk_nn_host.hpp"

namespace {

constexpr int kFfBatchMax = 64;   // flood-fill rounds enqueued between host reads

size_t slab_shmem(const spk_ctx* c) {
  return (size_t)(c->py.n + c->pz.n + c->py.n * c->pz.n) * sizeof(float2);
}
// pass B: slab + tables + the [ny][nz] low-digit sums
size_t score_shmem(const spk_ctx* c) {
  return slab_shmem(c) + (size_t)c->ny * c->nz * sizeof(int);
}
size_t pencil_shmem(const spk_ctx* c) {
  return (size_t)(c->px.n + c->px.n * c->T) * sizeof(float2);
}

// ---- host mailbox ------------------------------------------------------------
// kMailSlots pinned mailboxes.  post_async gathers the listed device words
// into one of them (k_post, stream-ordered) and returns the post's sequence
// number; wait_post waits until that slot holds that post.  Two slots let the
// learned loop keep one post in flight while it reads the other.
constexpr int kMailSlots = 2;
unsigned post_async(spk_ctx* c, int slot, const int* ctl, const PackOut* po, const FieldStats* st,
                    const float* logits, const int* fs, int nfs, int* gate = nullptr) {
  if (!c->mail_h) {
    HIP_OK(hipHostMalloc((void**)&c->mail_h, kMailSlots * sizeof(Mailbox),
                         hipHostMallocMapped | hipHostMallocCoherent));
    HIP_OK(hipHostGetDevicePointer((void**)&c->mail_d, c->mail_h, 0));
    for (int i = 0; i < kMailSlots; ++i) c->mail_h[i].seq = 0;
    c->mail_seq = 0;
  }
  const unsigned seq = ++c->mail_seq;
  launch(c, "post", k_post, dim3(1), dim3(64), 0, ctl, po, st, logits, fs, nfs, seq,
         c->mail_d + slot, gate);
  return seq;
}

// The host spins on the sequence word (a stream synchronisation wakes the
// host tens of microseconds later); after a bounded spin it falls back to the
// synchronisation, which also reports a failed kernel.
const Mailbox& wait_post(spk_ctx* c, int slot, unsigned seq) {
  const Mailbox* m = c->mail_h + slot;
  const auto t0 = std::chrono::steady_clock::now();
  bool seen = false;
  for (int it = 0; !seen; ++it) {
    seen = __atomic_load_n(&m->seq, __ATOMIC_ACQUIRE) == seq;
    if (!seen && (it & 4095) == 4095 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(100))
      break;
  }
  if (!seen) {
    sync(c);
    if (__atomic_load_n(&m->seq, __ATOMIC_ACQUIRE) != seq)
      throw SpkError{SPK_EDEVICE, "host mailbox not written"};
  }
  return *m;
}

const Mailbox& post_and_sync(spk_ctx* c, const int* ctl, const PackOut* po, const FieldStats* st,
                             const float* logits, const int* fs, int nfs) {
  return wait_post(c, 0, post_async(c, 0, ctl, po, st, logits, fs, nfs));
}

// ---- column test: 2x2 column blocks of the tray ----------------------------
void plan_coarse(spk_ctx* c) {
  spk_ctx::Coarse& q = c->co;
  q.ready = false;
  if (!q.on || c->split || c->nz > 128) return;
  q.cx = (c->nx + 1) / 2;
  q.cy = (c->ny + 1) / 2;
  q.cols.ensure((size_t)kColLevels * q.cx * q.cy);
  q.tf.ensure((size_t)kColLevels * c->nx * c->ny);
  q.tx = (c->nx + kSkyTile - 1) / kSkyTile;
  q.ty = (c->ny + kSkyTile - 1) / kSkyTile;
  q.ht.ensure((size_t)q.tx * q.ty + 1);   // tile heights, then the tray top
  q.fs.ensure(kFeasHeader + 24);
  q.ready = true;
}

// block masks of the tray, after every change of the tray
void build_coarse_spectra(spk_ctx* c, const int* ctl) {
  spk_ctx::Coarse& q = c->co;
  if (!q.ready) return;
  const int n = q.cx * q.cy;
  launch(c, "tray_columns", k_tray_columns, dim3((n + 3) / 4), dim3(256), 0, (const int*)c->tray.p,
         c->nx, c->ny, c->nz, q.cols.p, q.cx, q.cy, q.ht.p + (size_t)q.tx * q.ty, ctl);
  launch(c, "tray_fine", k_tray_fine, dim3((c->nx * c->ny + 3) / 4), dim3(256), 0,
         (const int*)c->tray.p, c->nx, c->ny, c->nz, q.tf.p, ctl);
  launch(c, "tile_heights", k_tile_heights, dim3((q.tx * q.ty + 3) / 4), dim3(256), 0,
         (const M128*)q.tf.p, c->nx, c->ny, q.ht.p, q.tx, q.ty, ctl);
}

// ---- FFT planning for a tray --------------------------------------------
// true when the runtime plan p is the compile-time plan FPlan<N, Rs...>
template <int N, int... Rs>
bool plan_is(const AxisPlan& p, FPlan<N, Rs...>) {
  const int rs[] = {Rs...};
  if (p.n != N || p.nst != (int)sizeof...(Rs)) return false;
  for (int s = 0; s < p.nst; ++s)
    if ((int)((p.rad >> (kRadixBits * s)) & kRadixMask) != rs[s]) return false;
  return true;
}

using SlabFwdKernel = decltype(&k_tray_slab_fwd<DynPlan, DynPlan>);
using ItemFwdKernel = decltype(&k_item_slab_fwd<DynPlan, DynPlan>);
using XFwdKernel = decltype(&k_tray_x_fwd<DynPlan, 0>);

// the forward slab and x-pencil kernels for the context's plans: the
// compile-time instantiation for the common padded sizes when the plans (and
// the tile) match, else the runtime one
SlabFwdKernel tray_slab_kernel(const spk_ctx* c) {
  using Y = FPlan<125, 25, 5>;
  using Z = FPlan<100, 10, 10>;
  if (c->fixed_fft && plan_is(c->py, Y{}) && plan_is(c->pz, Z{})) return k_tray_slab_fwd<Y, Z>;
  return k_tray_slab_fwd<DynPlan, DynPlan>;
}
ItemFwdKernel item_slab_kernel(const spk_ctx* c) {
  using Y = FPlan<125, 25, 5>;
  using Z = FPlan<100, 10, 10>;
  if (c->fixed_fft && plan_is(c->py, Y{}) && plan_is(c->pz, Z{})) return k_item_slab_fwd<Y, Z>;
  return k_item_slab_fwd<DynPlan, DynPlan>;
}
XFwdKernel tray_x_kernel(const spk_ctx* c) {
  using X = FPlan<240, 16, 15>;
  if (c->fixed_fft && plan_is(c->px, X{})) {
    if (c->T == 16) return k_tray_x_fwd<X, 16>;
    if (c->T == 8) return k_tray_x_fwd<X, 8>;
    if (c->T == 4) return k_tray_x_fwd<X, 4>;
  }
  return k_tray_x_fwd<DynPlan, 0>;
}

// x lengths of the register pass A (k_corr_xr): 16 R with a compiled R
bool reg_len(const spk_ctx* c, int n) {
  if (!c->fixed_fft || c->reg_x <= 0 || n % 16) return false;
  const int R = n / 16;
  return R == 6 || R == 8 || R == 9 || R == 10 || R == 12 || R == 14 || R == 15 || R == 16;
}

void plan_tray(spk_ctx* c) {
  // the x length: the smallest 2^a 3^b 5^c 7^d >= nx, or the register pass A's
  // next length when it pads by at most an eighth more
  int px = next_smooth(c->nx);
  for (int n = (c->nx + 15) / 16 * 16; n <= 256 && 8 * n <= 9 * px; n += 16)
    if (reg_len(c, n)) {
      px = n;
      break;
    }
  const int pz = c->pad_z >= c->nz && smooth(c->pad_z) ? c->pad_z : next_smooth(c->nz);
  int py = c->pad_y >= c->ny && smooth(c->pad_y) ? c->pad_y : next_smooth(c->ny);
  if (px > kMaxAxis || py > kMaxAxis || pz > kMaxAxis)
    throw SpkError{SPK_EUNSUPPORTED, "tray axis longer than 1024 is not supported"};
  if ((long long)c->ny * c->nz > 32768)
    throw SpkError{SPK_EUNSUPPORTED, "tray y*z plane larger than 32768 voxels"};
  // the x axis is transformed by the x-pencil kernels, whose butterflies stop
  // at kPencilRadix (so do y pencils of the split plan, re-planned below)
  c->px = make_plan(px, kPencilRadix);
  c->py = make_plan(py, c->yz_rmax);
  c->pz = make_plan(pz, c->yz_rmax);
  c->sy = make_plan(py, std::min(c->yz_rmax, c->score_rmax));
  c->sz = make_plan(pz, std::min(c->yz_rmax, c->score_rmax));
  auto tile = [](int n) {
    int T = 16;
    while (T > 1 && n * T > kPencilMax) T >>= 1;
    return T;
  };
  c->T = tile(px);
  c->Tc = reg_len(c, px) ? 4 : c->T;
  c->Ty = tile(py);
  c->Rf = std::max(1, kRowMax / pz);        // forward rows per block
  // score rows per block: R complex rows and ceil(R / 2) paired low-digit rows
  c->Rs = std::max(1, 2 * (kRowMax / pz) / 3);
  // twiddle tables exp(-2 pi i q / n), double -> float
  std::vector<float2> tw(px + py + pz);
  auto fill = [&](int off, int n) {
    for (int q = 0; q < n; ++q) {
      const double a = -2.0 * M_PI * (double)q / (double)n;
      tw[off + q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
  };
  fill(0, px);
  fill(px, py);
  fill(px + py, pz);
  c->tw.ensure(tw.size());
  upload(c, c->tw.p, tw.data(), tw.size() * sizeof(float2));
  // penalty tables: packing.cpp:630-631 (pow) and fft3.cu:173 (cube)
  std::vector<double> pen(2 * (size_t)c->nz);
  for (int k = 0; k < c->nz; ++k) {
    const double qz = (k + 0.0) / (c->nz + 0.0);
    pen[k] = 100000000.0 * std::pow(qz, 3.0);
    pen[c->nz + k] = ((100000000.0 * qz) * qz) * qz;
  }
  c->pen.ensure(pen.size());
  upload(c, c->pen.p, pen.data(), pen.size() * sizeof(double));
  int lds_max = 0;
  HIP_OK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
  c->lds_max = lds_max;
  if ((size_t)c->ny * c->nz * sizeof(int) > (size_t)lds_max)
    throw SpkError{SPK_EUNSUPPORTED, "tray y*z plane does not fit the device's LDS"};
  // slab plan when a whole yz slab (and its low-digit sums) fits one block's
  // LDS, else (or when asked) the split plan of z rows and y pencils
  const bool slab_fits = (long long)py * pz <= kSlabMax && score_shmem(c) + 512 <= (size_t)lds_max;
  c->split = c->plan_mode == 2 || (c->plan_mode == 0 && !slab_fits);
  if (!c->split && !slab_fits)
    throw SpkError{SPK_EUNSUPPORTED, "the slab plan does not fit this tray's yz plane"};
  if (c->split) c->py = make_plan(py, kPencilRadix);
  if (!c->lds_attrs) {
    const size_t dyn = (size_t)lds_max - 1024;   // leave room for static shared arrays
    using Y = FPlan<125, 25, 5>;
    using Z = FPlan<100, 10, 10>;
    using X = FPlan<240, 16, 15>;
    allow_shmem(k_tray_slab_fwd<DynPlan, DynPlan>, dyn);
    allow_shmem(k_tray_slab_fwd<Y, Z>, dyn);
    allow_shmem(k_item_slab_fwd<DynPlan, DynPlan>, dyn);
    allow_shmem(k_item_slab_fwd<Y, Z>, dyn);
    allow_shmem(k_score_yz<10>, dyn);
    allow_shmem(k_score_yz<25>, dyn);
    allow_shmem(k_score_yz_f<FPlan<125, 5, 5, 5>, FPlan<100, 10, 10>>, dyn);
    allow_shmem(k_score_yz_f<FPlan<125, 25, 5>, FPlan<100, 10, 10>>, dyn);
    allow_shmem(k_score_yz_f<FPlan<100, 10, 10>, FPlan<100, 10, 10>>, dyn);
    allow_shmem(k_score_yz_f<FPlan<64, 8, 8>, FPlan<64, 8, 8>>, dyn);
    allow_shmem(k_score_yz2<FPlan<125, 5, 5, 5>, FPlan<100, 10, 10>>, dyn);
    allow_shmem(k_score_yz2<FPlan<125, 25, 5>, FPlan<100, 10, 10>>, dyn);
    allow_shmem(k_score_yz2<FPlan<100, 10, 10>, FPlan<100, 10, 10>>, dyn);
    allow_shmem(k_score_yz2<FPlan<64, 8, 8>, FPlan<64, 8, 8>>, dyn);
    allow_shmem(k_corr_x, dyn);
    allow_shmem(k_corr_x_f<FPlan<240, 16, 15>, 16>, dyn);
    allow_shmem(k_corr_x_f<FPlan<240, 16, 15>, 8>, dyn);
    allow_shmem(k_corr_x_f<FPlan<240, 16, 15>, 4>, dyn);
    allow_shmem(k_corr_xr<15, 4, 3>, corr_xr_shmem(15, 4, 3));
    allow_shmem(k_corr_xr<15, 8, 3>, corr_xr_shmem(15, 8, 3));
    allow_shmem(k_corr_xr<15, 16, 3>, corr_xr_shmem(15, 16, 3));
    allow_shmem(k_corr_xr<15, 4, 2>, corr_xr_shmem(15, 4, 2));
    allow_shmem(k_corr_xr<15, 8, 2>, corr_xr_shmem(15, 8, 2));
    allow_shmem(k_corr_xr<15, 16, 2>, corr_xr_shmem(15, 16, 2));
    allow_shmem(k_tray_x_fwd<DynPlan, 0>, dyn);
    allow_shmem(k_tray_x_fwd<X, 16>, dyn);
    allow_shmem(k_tray_x_fwd<X, 8>, dyn);
    allow_shmem(k_tray_x_fwd<X, 4>, dyn);
    allow_shmem(k_feas_x, dyn);
    allow_shmem(k_feas_yz, dyn);
    c->lds_attrs = true;
  }
  plan_coarse(c);
}

// ---- distance field --------------------------------------------------------
void run_distance(spk_ctx* c, const int* d_grid, int* d_out, int* d_tmp, FieldStats* d_st,
                  int nx, int ny, int nz, int metric, int sources, const int* ctl = nullptr) {
  if ((long long)ny * nz > 32768)
    throw SpkError{SPK_EUNSUPPORTED, "distance: y*z plane larger than 32768 voxels"};
  if (nx > 1024) throw SpkError{SPK_EUNSUPPORTED, "distance: x axis longer than 1024 voxels"};
  const int fill = nx + ny + nz + 10;
  const int edt = metric == SPK_METRIC_EDT ? 1 : 0;
  const int big = edt ? fill * fill : fill;
  // y pass: c->dist_y (1 envelope, 2 full scan); the envelope's scratch
  // (ny*nz int2 behind the slab) needs planes up to 12288 voxels, else the
  // outward search (0)
  int ymode = c->dist_y;
  if (ymode == 1 && (long long)ny * nz > 12288) ymode = 0;
  const size_t sh = (size_t)ny * nz * (ymode == 1 ? 3 : 1) * sizeof(int);
  auto zy = [&](auto kern) {
    allow_shmem(kern, sh);
    launch(c, "dist_zy", kern, dim3(nx), dim3(kDistThreads), sh, (const int*)d_grid, d_tmp, nx,
           ny, nz, sources, big, d_st, ctl);
  };
  if (edt) {
    if (ymode == 1) zy(k_dist_zy<1, 1>); else if (ymode == 2) zy(k_dist_zy<1, 2>); else zy(k_dist_zy<1, 0>);
  } else {
    if (ymode == 1) zy(k_dist_zy<0, 1>); else if (ymode == 2) zy(k_dist_zy<0, 2>); else zy(k_dist_zy<0, 0>);
  }
  constexpr int TK = 16;
  auto xk = c->dist_x == 2 ? (edt ? k_dist_x<TK, 1, 2> : k_dist_x<TK, 0, 2>)
                           : (edt ? k_dist_x<TK, 1, 0> : k_dist_x<TK, 0, 0>);
  launch(c, "dist_x", xk, dim3((nz + TK - 1) / TK, ny), dim3(kDistThreads),
         (size_t)nx * TK * sizeof(int), (const int*)d_tmp, d_out, d_st, nx, ny, nz, fill, ctl);
}

// The placement a refresh follows, for the incremental distance update: the
// device's search outcome (so, with od the searched orientations) or an
// explicit orientation od[orient] at (qx, qy, qz); id as k_place takes it
// (the loop adds the placed count on the device).
struct PlaceRef {
  const SearchOut* so = nullptr;
  const OrientDesc* od = nullptr;
  int orient = 0, qx = 0, qy = 0, qz = 0, id = 0;
  int max_ex = 0;   // largest x extent among the orientations od may name
  int max_dim = 0;  // largest extent along any axis among them (0: unknown)
};

void update_coarse(spk_ctx* c, const int* ctl, const PlaceRef& pr);
// Consecutive incremental updates of G1 a device loop may make (k_decide's
// spec_max): the field must be this library's own transform of the tray with
// sources "eq1" (a placement with id != 1 then adds no source and leaves it
// unchanged), on the slab plan; 0 when that does not hold.
int spec_max(const spk_ctx* c) {
  return (!c->split && c->sources == SPK_SOURCES_EQ1 && c->phi_internal) ? c->spec_inc : 0;
}
// the cached spectra of the tray, then the column test's tray masks: rebuilt
// (pr == nullptr) or updated for the placement pr
void build_spectra(spk_ctx* c, const int* ctl = nullptr, const PlaceRef* pr = nullptr) {
  const int py = c->py.n, pz = c->pz.n, hz = pz / 2 + 1;
  c->t1.ensure((size_t)c->nx * py * pz);
  c->t2.ensure((size_t)c->nx * py * hz);
  c->g1.ensure((size_t)c->px.n * py * pz);
  c->g2.ensure((size_t)c->px.n * py * hz);
  const float2* tw = c->tw.p;
  if (c->split) {
    const int jb = (c->ny + c->Rf - 1) / c->Rf;
    launch(c, "tray_rows_fwd", k_tray_rows_fwd, dim3(jb * c->nx, 2), dim3(kRowThreads),
           (size_t)(pz + kRowMax) * sizeof(float2), (const int*)c->tray.p, (const int*)c->phi.p,
           (const FieldStats*)c->stats.p, c->t1.p, c->t2.p, c->nx, c->ny, c->nz, py, c->pz, c->Rf,
           tw + c->px.n + py, ctl);
    const size_t sh = (size_t)(py + py * c->Ty) * sizeof(float2);
    launch(c, "pencil_y_fwd", k_pencil<true>, dim3((pz + c->Ty - 1) / c->Ty, c->nx),
           dim3(kPencilThreads), sh, c->t1.p, c->py, pz, c->Ty, c->ny, py, tw + c->px.n,
           (const OrientDesc*)nullptr, c->nx, (int)PENCIL_PLAIN, ctl, 1);
    launch(c, "pencil_y_fwd", k_pencil<true>, dim3((hz + c->Ty - 1) / c->Ty, c->nx),
           dim3(kPencilThreads), sh, c->t2.p, c->py, hz, c->Ty, c->ny, py, tw + c->px.n,
           (const OrientDesc*)nullptr, c->nx, (int)PENCIL_PLAIN, ctl, 1);
  } else {
    launch(c, "tray_slab_fwd", tray_slab_kernel(c), dim3(c->nx, 2), dim3(kSlabThreads),
           slab_shmem(c), (const int*)c->tray.p, (const int*)c->phi.p,
           (const FieldStats*)c->stats.p, c->t1.p, c->t2.p, c->nx, c->ny, c->nz, c->py, c->pz,
           tw + c->px.n, tw + c->px.n + py, ctl);
  }
  launch(c, "tray_x_fwd", tray_x_kernel(c), dim3((pz + c->T - 1) / c->T, py, 2),
         dim3(kPencilThreads), pencil_shmem(c), (const float2*)c->t1.p, (const float2*)c->t2.p,
         c->g1.p, c->g2.p, c->nx, c->px, py, pz, c->T, tw, ctl);
  // after a placement in a device loop: the incremental update instead, when
  // k_decide chose it (the two kernels above then exit at once)
  if (ctl && pr && pr->so && spec_max(c) > 0)
    launch(c, "spec_add", k_spec_add, dim3((pz + c->T - 1) / c->T, py), dim3(kPencilThreads),
           pencil_shmem(c), c->g1.p, (const float2*)c->ispec.p, pr->od, pr->so, pr->max_ex,
           c->ny, c->nz, c->px, py, pz, c->T, tw,
           c->timing && c->pairs.p ? c->pairs.p + 3 : (unsigned long long*)nullptr, ctl);
  if (pr && c->inc_masks) update_coarse(c, ctl, *pr);
  else build_coarse_spectra(c, ctl);
}


// The column test's tray masks after the placement pr: only the columns under
// the placed box are rebuilt (k_tray_masks_inc), when the box side is known.
void update_coarse(spk_ctx* c, const int* ctl, const PlaceRef& pr) {
  spk_ctx::Coarse& q = c->co;
  if (!q.ready) return;
  if (pr.max_dim <= 0) {
    build_coarse_spectra(c, ctl);
    return;
  }
  const int bound = std::min(pr.max_dim, std::max(c->nx, c->ny));
  const int nb1 = bound / 2 + 1;
  const int waves = bound * bound + nb1 * nb1;
  launch(c, "tray_masks_inc", k_tray_masks_inc, dim3((waves + 3) / 4), dim3(256), 0,
         (const int*)c->tray.p, c->nx, c->ny, c->nz, q.tf.p, q.cols.p, q.cx, q.cy, q.ht.p, q.tx,
         q.ty, pr.so, pr.od, pr.orient, pr.qx, pr.qy, bound, ctl);
}

// k_inc_zy's LDS: the item box's column flags (rows x z extent) and the
// z-pass costs (rows x nz), rows and z extent bounded by the item's extents
size_t inc_zy_shmem(const spk_ctx* c, const PlaceRef& pr) {
  const size_t ry = pr.max_dim > 0 ? std::min(pr.max_dim, c->ny) : c->ny;
  const size_t rz = pr.max_dim > 0 ? std::min(pr.max_dim, c->nz) : c->nz;
  return (ry * rz + ry * c->nz) * sizeof(int);
}

// The incremental update applies when the field is this library's own
// transform of the tray (not a caller's) and the item's x extent fits its
// candidate buffer.
bool inc_applies(const spk_ctx* c, const PlaceRef* pr) {
  return pr && c->inc_dist && c->phi_internal && pr->max_ex <= kIncMaxEx && c->nx <= 1024 &&
         inc_zy_shmem(c, *pr) <= 160 * 1024;
}

// Phi = min(Phi, transform of the placed item's voxels) and the statistics
// of the result (k_inc_zy, k_inc_x); skipped on the device when the placement
// adds no source.
void run_distance_inc(spk_ctx* c, const PlaceRef& pr, const int* ctl) {
  const int fill = c->nx + c->ny + c->nz + 10;
  const int edt = c->metric == SPK_METRIC_EDT ? 1 : 0;
  const int big = edt ? fill * fill : fill;
  const size_t sh = inc_zy_shmem(c, pr);
  auto zy = edt ? k_inc_zy<1> : k_inc_zy<0>;
  allow_shmem(zy, sh);
  launch(c, "inc_zy", zy, dim3(std::min(pr.max_ex, c->nx), kIncSplit), dim3(kDistThreads), sh,
         (const int*)c->tray.p, c->tmpi.p, pr.so, pr.od, pr.orient, pr.qx, pr.qy, pr.qz, pr.id,
         c->nx, c->ny, c->nz, c->sources, big, ctl);
  const dim3 grid((c->nz + kIncLines - 1) / kIncLines, c->ny, c->inc_split);
  const int nb = (int)(grid.x * grid.y * grid.z);
  c->stats_part.ensure(nb);
  auto xk = edt ? k_inc_x<1> : k_inc_x<0>;
  const size_t xsh = (size_t)pr.max_ex * kIncLines * sizeof(int);
  allow_shmem(xk, xsh);
  launch(c, "inc_x", xk, grid, dim3(256), xsh,
         (const int*)c->tmpi.p, c->phi.p, c->stats_part.p, pr.so, pr.od, pr.orient, pr.qx, pr.qy,
         pr.qz, pr.id, c->nx, c->ny, c->nz, c->sources, big, fill, ctl);
  launch(c, "inc_stats", k_inc_stats, dim3(1), dim3(1024), 0, (const FieldStats*)c->stats_part.p,
         nb, c->stats.p, pr.so, pr.od, pr.orient, pr.qx, pr.qy, pr.qz, pr.id, c->ny, c->nz,
         c->sources, ctl);
}

// Distance field (computed, or the caller's) + statistics + cached spectra.
// With ctl (greedy loop) every kernel is skipped on the device when the last
// item placed nothing or the loop was handed back to the host.  After a
// placement (pr), the field is updated incrementally when that applies.
void refresh_field_and_spectra(spk_ctx* c, const int* dist_host, const int* ctl = nullptr,
                               const PlaceRef* pr = nullptr) {
  const size_t V = (size_t)c->nx * c->ny * c->nz;
  const bool inc = !dist_host && inc_applies(c, pr);
  if (dist_host) c->phi_internal = false;
  else if (!ctl) c->phi_internal = true;   // (a gated refresh keeps the flag as it was)
  if (dist_host) {
    upload(c, c->phi.p, dist_host, V * sizeof(int));
    launch(c, "stats_init", k_stats_init, dim3(1), dim3(64), 0, c->stats.p, (const int*)nullptr);
    const int nb = (int)std::min<size_t>(1024, (V + 255) / 256);
    launch(c, "field_stats", k_field_stats, dim3(nb), dim3(256), 0, (const int*)c->phi.p,
           (long long)V, c->stats.p);
  } else if (inc) {
    run_distance_inc(c, *pr, ctl);
  } else {
    run_distance(c, c->tray.p, c->phi.p, c->tmpi.p, c->stats.p, c->nx, c->ny, c->nz, c->metric,
                 c->sources, ctl);
  }
  build_spectra(c, ctl, pr);
}

// ---- orientation descriptors ----------------------------------------------
struct HostItem {
  const int* data;
  int dims[3];
  int lo[3], hi[3];   // bounds of voxels > 0, hi = -1 when empty
  long long sum;
  int vmax;           // largest voxel value
};

// Bounds and sum of the voxels > 0 of an item.  Row by row: the minimum and
// the sum of a row vectorise; only rows with a positive voxel look for their
// first and last one.
// copy (not null): the item is also copied there, in the same pass.
HostItem scan_item(const int* data, int ix, int iy, int iz, int* copy = nullptr) {
  if (!data && (long long)ix * iy * iz > 0) throw SpkError{SPK_EINVAL, "null item"};
  check_dims(ix, iy, iz);
  HostItem h{data, {ix, iy, iz}, {INT_MAX, INT_MAX, INT_MAX}, {-1, -1, -1}, 0, 0};
  for (int i = 0; i < ix; ++i)
    for (int j = 0; j < iy; ++j) {
      const size_t r0 = ((size_t)i * iy + j) * iz;
      const int* row = data + r0;
      int mn = INT_MAX, mx = 0;
      long long s = 0;
      if (copy) {
        int* dst = copy + r0;
        for (int k = 0; k < iz; ++k) {
          const int v = row[k];
          dst[k] = v;
          mn = std::min(mn, v);
          mx = std::max(mx, v);
          s += v;
        }
      } else {
        for (int k = 0; k < iz; ++k) {
          mn = std::min(mn, row[k]);
          mx = std::max(mx, row[k]);
          s += row[k];
        }
      }
      h.vmax = std::max(h.vmax, mx);
      if (mn < 0) throw SpkError{SPK_EINVAL, "negative item voxel values are not supported"};
      if (s == 0) continue;
      int k0 = 0, k1 = iz - 1;
      while (row[k0] == 0) ++k0;
      while (row[k1] == 0) --k1;
      h.sum += s;
      h.lo[0] = std::min(h.lo[0], i); h.hi[0] = std::max(h.hi[0], i);
      h.lo[1] = std::min(h.lo[1], j); h.hi[1] = std::max(h.hi[1], j);
      h.lo[2] = std::min(h.lo[2], k0); h.hi[2] = std::max(h.hi[2], k1);
    }
  return h;
}

OrientDesc make_desc(const spk_ctx* c, const HostItem& it, int item_off, const spk_orient* o) {
  OrientDesc d{};
  d.item_off = item_off;
  d.line_off = -1;
  d.bx = it.dims[0]; d.by = it.dims[1]; d.bz = it.dims[2];
  int perm[3] = {0, 1, 2}, flip[3] = {0, 0, 0};
  if (o) {
    bool seen[3] = {false, false, false};
    for (int a = 0; a < 3; ++a) {
      perm[a] = o->perm[a];
      flip[a] = o->flip[a] ? 1 : 0;
      if (perm[a] < 0 || perm[a] > 2 || seen[perm[a]])
        throw SpkError{SPK_EINVAL, "orientation perm must be a permutation of 0,1,2"};
      seen[perm[a]] = true;
    }
  }
  const bool empty = it.hi[0] < 0;
  const int N[3] = {c->nx, c->ny, c->nz};
  int ext[3], hi[3], nv[3];
  for (int a = 0; a < 3; ++a) {
    d.perm[a] = perm[a];
    d.flip[a] = flip[a];
    const int b = perm[a];
    ext[a] = it.dims[b];
    if (empty) hi[a] = -1;
    else hi[a] = flip[a] ? (it.dims[b] - 1 - it.lo[b]) : it.hi[b];
    nv[a] = std::min(N[a], N[a] - hi[a]);
  }
  d.ex = ext[0]; d.ey = ext[1]; d.ez = ext[2];
  d.hx = hi[0]; d.hy = hi[1]; d.hz = hi[2];
  d.nvx = nv[0]; d.nvy = nv[1]; d.nvz = nv[2];
  d.isum = it.sum;
  d.binary = it.vmax <= 1 ? 1 : 0;
  return d;
}

bool fits(const spk_ctx* c, const OrientDesc& d) {
  return d.ex <= c->nx && d.ey <= c->ny && d.ez <= c->nz && d.nvx > 0 && d.nvy > 0 && d.nvz > 0;
}

void to_result(const spk_ctx* c, bool found, int idx, double score, int orient, long long coll,
               long long prox, float resid, int path, spk_result* r) {
  std::memset(r, 0, sizeof(*r));
  if (found) {
    r->pos[2] = idx % c->nz;
    r->pos[1] = (idx / c->nz) % c->ny;
    r->pos[0] = idx / (c->ny * c->nz);
    r->found = 1;
    r->orient = orient;
    r->score = score;
    r->collision = coll;
    r->proximity = prox;
  } else {
    r->pos[0] = r->pos[1] = r->pos[2] = -1;
    r->found = 0;
    r->orient = -1;
    r->score = 0.0;
  }
  r->residual = resid;
  r->path = path;
}

// non-zero voxels of an oriented item as (offset, value) taps
std::vector<Tap> oriented_taps(const HostItem& it, const OrientDesc& d) {
  std::vector<Tap> taps;
  for (int w0 = 0; w0 < d.ex; ++w0)
    for (int w1 = 0; w1 < d.ey; ++w1)
      for (int w2 = 0; w2 < d.ez; ++w2) {
        int w[3] = {w0, w1, w2}, u[3];
        const int bd[3] = {d.bx, d.by, d.bz};
        for (int a = 0; a < 3; ++a) {
          const int ax = d.perm[a];
          u[ax] = d.flip[a] ? (bd[ax] - 1 - w[a]) : w[a];
        }
        const int v = it.data[((size_t)u[0] * d.by + u[1]) * d.bz + u[2]];
        if (v > 0) taps.push_back(Tap{w0, w1, w2, v});
      }
  return taps;
}

// exact direct-sum search of one orientation (fallback path)
WgBest exact_search_one(spk_ctx* c, const HostItem& it, const OrientDesc& d, int penalty) {
  const std::vector<Tap> taps = oriented_taps(it, d);
  const int nb = 1024;
  c->taps.ensure(std::max<size_t>(taps.size(), 1));
  if (!taps.empty()) upload(c, c->taps.p, taps.data(), taps.size() * sizeof(Tap));
  c->rec.ensure(nb);
  launch(c, "direct_search", k_direct_search, dim3(nb), dim3(256), 0, (const int*)c->tray.p,
         (const int*)c->phi.p, (const Tap*)c->taps.p, (int)taps.size(), d.nvx, d.nvy, d.nvz,
         c->ny, c->nz, (const double*)(c->pen.p + (penalty == SPK_PENALTY_CUBE ? c->nz : 0)),
         c->rec.p);
  std::vector<WgBest> rec(nb);
  download(c, rec.data(), c->rec.p, nb * sizeof(WgBest));
  sync(c);
  WgBest best{1e300, INT_MAX, 0, 0, 0.f, 0};
  for (const auto& r : rec)
    if (r.score < best.score || (r.score == best.score && r.idx < best.idx)) best = r;
  return best;
}

// Flood fill m feasibility grids (c->ffree, m*V bytes) from their seed faces to
// the fixpoint into c->freach.  exits[3o..] are the high seed faces.
// Flood fill of the m feasibility grids to their fixpoint.  Rounds are
// enqueued in batches that run back to back on the device (a round after
// one that changed nothing exits at once, and within a round only tiles next
// to a change are worked); the host reads one flag per batch: the last
// round's change flag.  Every round that reports a change reaches at least
// one new cell, so the loop ends.
// The same fixpoint on bit columns (nz <= 128, option "ff_bits"): pack the
// free offsets and seeds into 128-bit column masks, rounds of k_ffb_tile over
// 32 x 32-column tiles, then the reached bits back to one byte per offset.
void flood_fill_bits(spk_ctx* c, int m, int nx, int ny, int nz, bool bytes,
                     const int* skip = nullptr) {
  const size_t ncol = (size_t)nx * ny, V = ncol * nz;
  if (bytes) c->freach.ensure((size_t)m * V);
  c->fb_free.ensure((size_t)m * ncol);
  c->fb_reach.ensure((size_t)m * ncol);
  const dim3 cg((unsigned)((ncol + 3) / 4), m);
  launch(c, "ff_seed", k_ffb_pack, cg, dim3(256), 0, (const unsigned char*)c->ffree.p,
         c->fb_free.p, c->fb_reach.p, (const int*)c->exits.p, nx, ny, nz, skip);
  const int xt = (nx + FBT - 1) / FBT, yt = (ny + FBT - 1) / FBT;
  const size_t tiles = (size_t)xt * yt * m;
  c->ff_act.ensure(2 * tiles);
  c->ff_flags.ensure(kFfBatchMax + 1);
  int batch = 8;
  size_t round = 0;
  for (size_t it = 0; it <= (size_t)m * V; it += batch) {
    launch(c, "ff_batch_init", k_ff_batch_init, dim3(round == 0 ? 256 : 1), dim3(256), 0,
           c->ff_flags.p, batch, round == 0 ? c->ff_act.p + (round & 1) * tiles : nullptr,
           (long long)tiles, skip);
    for (int r = 0; r < batch; ++r, ++round)
      launch(c, "ff_tile", k_ffb_tile, dim3(xt * yt, m), dim3(FBT * FBT), 0,
             (const U128*)c->fb_free.p, c->fb_reach.p, nx, ny, xt, yt,
             (const int*)c->ff_flags.p + r, c->ff_flags.p + r + 1,
             (const unsigned char*)c->ff_act.p + (round & 1) * tiles,
             c->ff_act.p + ((round + 1) & 1) * tiles);
    int ch = 0;
    download(c, &ch, c->ff_flags.p + batch, sizeof(int));
    sync(c);
    if (!ch) break;
    batch = std::min(2 * batch, kFfBatchMax);   // long fills: fewer host reads
  }
  if (bytes)
    launch(c, "ff_unpack", k_ffb_unpack, cg, dim3(256), 0, (const U128*)c->fb_reach.p,
           c->freach.p, nx, ny, nz);
}

// Returns true when the bit fill ran: then c->fb_reach holds the result, and
// c->freach too only when bytes is set.
// skip (nullable, device): when set, the fill's kernels exit at once
bool flood_fill(spk_ctx* c, int m, const std::vector<int>& exits, int nx, int ny, int nz,
                bool bytes = true, const int* skip = nullptr) {
  const size_t V = (size_t)nx * ny * nz;
  c->freach.ensure((size_t)m * V);
  c->exits.ensure((size_t)3 * m);
  upload(c, c->exits.p, exits.data(), exits.size() * sizeof(int));
  if (c->ff_bits && nz <= 128) {
    flood_fill_bits(c, m, nx, ny, nz, bytes, skip);
    return true;
  }
  const int nb = (int)std::max<size_t>(1, std::min<size_t>(2048, (V + 255) / 256));
  launch(c, "ff_seed", k_ff_seed, dim3(nb, m), dim3(256), 0, (const unsigned char*)c->ffree.p,
         c->freach.p, (const int*)c->exits.p, nx, ny, nz, skip);
  const int xt = (nx + FFX - 1) / FFX;
  const dim3 grid((nz + FFZ - 1) / FFZ, (ny + FFY - 1) / FFY, xt * m);
  const size_t tiles = (size_t)grid.x * grid.y * grid.z;
  c->ff_act.ensure(2 * tiles);
  c->ff_flags.ensure(kFfBatchMax + 1);
  int batch = 8;
  size_t round = 0;
  for (size_t it = 0; it <= (size_t)m * V; it += batch) {
    launch(c, "ff_batch_init", k_ff_batch_init, dim3(round == 0 ? 256 : 1), dim3(256), 0,
           c->ff_flags.p, batch, round == 0 ? c->ff_act.p + (round & 1) * tiles : nullptr,
           (long long)tiles, skip);
    for (int r = 0; r < batch; ++r, ++round)
      launch(c, "ff_tile", k_ff_tile, grid, dim3(FFX * FFY * FFZ), 0,
             (const unsigned char*)c->ffree.p, c->freach.p, nx, ny, nz, xt,
             (const int*)c->ff_flags.p + r, c->ff_flags.p + r + 1,
             (const unsigned char*)c->ff_act.p + (round & 1) * tiles,
             c->ff_act.p + ((round + 1) & 1) * tiles);
    int ch = 0;
    download(c, &ch, c->ff_flags.p + batch, sizeof(int));
    sync(c);
    if (!ch) break;
    batch = std::min(2 * batch, kFfBatchMax);   // long fills: fewer host reads
  }
  return false;
}

// IF mode: feasibility grids -> reachable set -> masked argmin records
// winner_only: only the winner is wanted (no per-orientation results), so
// the masked argmin and the selection after it keep the unmasked winner when
// the fill reached it (c->if_keep = 1)
void if_masked_select(spk_ctx* c, const std::vector<OrientDesc>& ods, const double* pen,
                      bool winner_only = false) {
  const int m = (int)ods.size();
  std::vector<int> exits(3 * (size_t)m);
  for (int o = 0; o < m; ++o) {
    exits[3 * o] = c->nx - 1 - ods[o].hx;
    exits[3 * o + 1] = c->ny - 1 - ods[o].hy;
    exits[3 * o + 2] = c->nz - 1 - ods[o].hz;
  }
  c->if_keep.ensure(1);
  if (winner_only)   // a winner reachable on its own needs no fill
    launch(c, "if_direct", k_if_direct, dim3(1), dim3(64), 0, (const SearchOut*)c->sout.p,
           (const OrientDesc*)c->od.p, (const unsigned char*)c->ffree.p, c->nx, c->ny, c->nz,
           c->if_keep.p);
  const bool bits = flood_fill(c, m, exits, c->nx, c->ny, c->nz, false,
                               winner_only ? (const int*)c->if_keep.p : (const int*)nullptr);
  // c->per holds each orientation's unmasked summary (residual, range flag)
  c->rec2.ensure((size_t)m * c->nx);
  launch(c, "masked_argmin", k_masked_argmin, dim3(c->nx, m), dim3(256), 0,
         (const unsigned char*)c->freach.p, (const long long*)c->pgrid.p,
         (const OrientDesc*)c->od.p, pen, (const WgBest*)c->per.p, c->rec2.p, c->nx, c->ny,
         c->nz, bits ? (const U128*)c->fb_reach.p : (const U128*)nullptr,
         winner_only ? (const SearchOut*)c->sout.p : (const SearchOut*)nullptr, c->if_keep.p);
}

// records per orientation of the score pass: slab plan one per x slice,
// split plan one per block of Rs valid rows
int rec_stride(const spk_ctx* c) {
  return c->split ? (c->nx * c->ny + c->Rs - 1) / c->Rs : c->nx;
}

// ---- the search --------------------------------------------------------------
void ensure_search_ws(spk_ctx* c, int m, int max_ex) {
  const int py = c->py.n, pz = c->pz.n, hz = pz / 2 + 1;
  c->ispec.ensure((size_t)m * max_ex * py * pz);
  c->y1.ensure((size_t)m * c->nx * py * pz);
  c->y2.ensure((size_t)m * c->nx * py * hz);
  c->rec.ensure((size_t)m * rec_stride(c));
  c->per.ensure(m);
  c->sout.ensure(1);
  c->ver.ensure(3);
  c->fstate.ensure(kFeasHeader + std::max(m, 24));
}

// Enqueue one batched search of m orientation descriptors (device array od):
// item spectra, pass A, pass B, selection and the exact re-check of the
// winner.  Results land in c->per, c->sout and c->ver.  With fgrid/pgrid
// (interlocking-free) the feasibility and proximity grids are written and the
// flood fill + masked argmin run before the selection.
// item spectra of m orientation descriptors into c->ispec
void enqueue_item_spectra(spk_ctx* c, const OrientDesc* od, int m, int max_ex, const int* ctl) {
  const int py = c->py.n, pz = c->pz.n;
  const float2* tw = c->tw.p;
  const size_t ysh = (size_t)(py + py * c->Ty) * sizeof(float2);
  if (c->split) {
    const int max_ey = std::min(c->ny, py);
    const int jb = (max_ey + c->Rf - 1) / c->Rf;
    launch(c, "item_rows_fwd", k_item_rows_fwd, dim3(jb * max_ex, m), dim3(kRowThreads),
           (size_t)(pz + kRowMax) * sizeof(float2), (const int*)c->items.p, od, c->ispec.p,
           max_ex, max_ey, py, c->pz, c->Rf, tw + c->px.n + py, ctl);
    launch(c, "pencil_y_fwd", k_pencil<true>, dim3((pz + c->Ty - 1) / c->Ty, m * max_ex),
           dim3(kPencilThreads), ysh, c->ispec.p, c->py, pz, c->Ty, py, py, tw + c->px.n, od,
           max_ex, (int)PENCIL_ITEM, ctl, 0);
  } else {
    launch(c, "item_slab_fwd", item_slab_kernel(c), dim3(max_ex, m), dim3(kSlabThreads),
           slab_shmem(c), (const int*)c->items.p, od, c->ispec.p, max_ex, c->py, c->pz,
           tw + c->px.n, tw + c->px.n + py,
           c->co.ready ? (const int*)c->co.fs.p + kFeasHeader : (const int*)nullptr, ctl);
  }
}

// x-pass chunking: enough blocks to fill the chip, few orientations per block
int corr_chunks(const spk_ctx* c, int m, int T, int* chunk) {
  const int py = c->py.n, pz = c->pz.n;
  const int tiles = (pz + T - 1) / T;
  int nchunk = std::max(1, std::min(m, (2048 + tiles * py - 1) / (tiles * py)));
  if (c->corr_chunk > 0) nchunk = (m + c->corr_chunk - 1) / c->corr_chunk;   // forced
  *chunk = (m + nchunk - 1) / nchunk;
  return (m + *chunk - 1) / *chunk;
}

// Feasibility pass over m descriptors whose item spectra are in c->ispec
// (slab plan only): per-orientation flags and the residual in c->fstate.
// Pairs go through in groups of c->feas_group (0: all in one launch pair):
// every group reuses the start of y1, so a group's intermediate can stay in
// the last-level cache between its two kernels.
void enqueue_feas(spk_ctx* c, const OrientDesc* od, int m, int max_ex, const int* ctl) {
  const int py = c->py.n, pz = c->pz.n;
  const float2* tw = c->tw.p;
  const int np = (m + 1) / 2;
  const int tiles = (pz + c->T - 1) / c->T;
  const float scale = (float)(1.0 / ((double)c->px.n * py * pz));
  unsigned long long* pairs = c->timing ? c->pairs.p : nullptr;
  const int group = c->feas_group > 0 ? c->feas_group : np;
  for (int g0 = 0; g0 < np; g0 += group) {
    const int gp = std::min(group, np - g0);
    const int ob = 2 * g0, gm = std::min(m - ob, 2 * gp);
    int chunk = 0;
    const int nchunk = corr_chunks(c, gp, c->T, &chunk);
    const int* cf = c->co.ready ? c->co.fs.p + kFeasHeader + ob : nullptr;
    launch(c, "feas_x", k_feas_x, dim3(tiles, py, nchunk), dim3(kPencilThreads), pencil_shmem(c),
           (const float2*)c->g1.p, (const float2*)c->ispec.p + (size_t)ob * max_ex * py * pz,
           od + ob, c->y1.p, gm, chunk, max_ex, c->nx, c->px, py, pz, c->T, scale, tw, cf, 0,
           ctl);
    launch(c, "feas_yz", k_feas_yz, dim3(c->nx, gp), dim3(kSlabThreads), slab_shmem(c),
           (const float2*)c->y1.p, od + ob, gm, c->nx, c->py, c->pz, tw + c->px.n,
           tw + c->px.n + py, c->fstate.p, ob, cf, 0, pairs, ctl);
  }
}

// control words that enable the feasibility pass outside the greedy loop
const int* ctl_feas_on(spk_ctx* c) {
  c->ctl_on.ensure(CTL_WORDS);
  int on[CTL_WORDS] = {0};
  on[CTL_HALT_ITEM] = -1;
  on[CTL_FEAS] = 1;
  upload(c, c->ctl_on.p, on, sizeof(on));
  return c->ctl_on.p;
}

// Column test over the m descriptors od (device), before the feasibility
// pass and the scored search: core masks of the oriented items, the test,
// then k_coarse_decide (prune orientations, or decide the item: see there).
// ctl must not be null.
// With clear_flags the column test's flags are cleared here (no k_feas_begin
// before it); callers only pass it for max_exy <= kMaskCols.  Without decide
// the flags are only left for the host (no k_coarse_decide).
void enqueue_coarse(spk_ctx* c, const OrientDesc* od, int m, int max_ex, int max_exy,
                    const int* ctl, PackOut* out = nullptr, int t = 0, int halt_not_found = 0,
                    bool clear_flags = false, bool decide = true, int* flags = nullptr,
                    const int* pre = nullptr, int* cnt = nullptr, const int* od_item = nullptr,
                    int cap = 0, bool begin = false, int topk = 0, int* gate = nullptr) {
  // begin: k_item_masks opens the item in place of k_feas_begin (control words
  // and flags cleared, feasibility pass off), when it runs (coarse_begins);
  // gate: k_coarse_decide also sets the learned loop's proposer gate
  spk_ctx::Coarse& q = c->co;
  if (!q.ready) return;
  int* fs = flags ? flags : q.fs.p;
  // a batched screen (flags given) is timed under names of its own
  const char* n_masks = flags ? "pre_masks" : "item_masks";
  const char* n_test = flags ? "pre_column_test" : "column_test";
  const int ccx = (max_ex + 1) / 2, ccy = q.cy;
  const int mx = max_ex, my = c->ny;
  q.cm.ensure((size_t)m * ccx * ccy);
  q.im.ensure((size_t)m * mx * my);
  q.rl.ensure((size_t)m * kColRuns);
  q.frl.ensure((size_t)m * kFineRuns);
  q.nrun.ensure(2 * (size_t)m);
  if (max_exy <= kMaskCols) {
    launch(c, n_masks, k_item_masks, dim3(m), dim3(kMaskThreads), 0, (const int*)c->items.p,
           (const M128*)q.lines.p, od, m, q.cm.p, ccx, ccy, q.im.p, mx, my, q.rl.p, q.frl.p,
           q.nrun.p, clear_flags || begin ? fs : (int*)nullptr, (int*)ctl, pre,
           begin ? c->fstate.p : (int*)nullptr, begin ? 1 : 0);
  } else {
    launch(c, "core_columns", k_core_columns, dim3((ccx * ccy + 3) / 4, m), dim3(256), 0,
           (const int*)c->items.p, od, m, q.cm.p, ccx, ccy, ctl);
    launch(c, "item_columns", k_item_columns, dim3((mx * my + 3) / 4, m), dim3(256), 0,
           (const int*)c->items.p, od, m, q.im.p, mx, my, ctl);
    launch(c, "run_lists", k_run_lists, dim3(m), dim3(kListThreads), 0, (const M128*)q.cm.p, ccx,
           ccy, (const M128*)q.im.p, mx, my, od, m, 1, q.rl.p, q.frl.p, q.nrun.p, ctl);
  }
  auto test = [&](auto kernel) {
    const int chunks = (q.cx * q.cy + 63) / 64;
    launch(c, n_test, kernel, dim3((chunks + q.col_chunks - 1) / q.col_chunks, m),
           dim3(256), 0,
           (const M128*)q.cols.p, q.cx, q.cy, (const M128*)q.cm.p, ccx, ccy, (const M128*)q.tf.p,
           c->nx, c->ny, (const M128*)q.im.p, mx, my, od, m, (const unsigned*)q.rl.p,
           (const unsigned*)q.frl.p, (const int*)q.nrun.p, (const int*)q.ht.p, q.ty, fs, ctl,
           q.tx, pre, cnt, od_item, cap);
  };
  switch (q.fine_group) {
    case 4: test(k_column_test<4>); break;
    case 8: test(k_column_test<8>); break;
    case 32: test(k_column_test<32>); break;
    case 16: test(k_column_test<16>); break;
    default: test(k_column_test<64>); break;
  }
  if (decide)
    launch(c, "coarse_decide", k_coarse_decide, dim3(1), dim3(64), 0, (const int*)fs, m,
         c->fstate.p, out, t, halt_not_found, (int*)ctl, topk, gate);
}

// whether enqueue_coarse's first kernel can open an item (begin): the
// one-kernel item masks, used when every orientation's footprint fits its LDS
bool coarse_begins(const spk_ctx* c, int max_exy) { return c->co.ready && max_exy <= kMaskCols; }


using ScoreKernel = decltype(&k_score_yz<10>);

// the score pass for the context's plans: a compile-time instantiation when
// one matches (padded sizes of the common trays), else the runtime one
// (k_score_yz2, both slabs in LDS, when option "score2" is on and they fit)
struct ScoreLaunch {
  ScoreKernel k;
  size_t shmem;
};
template <class PY_, class PZ_>
bool fixed_score(const spk_ctx* c, ScoreLaunch* k) {
  if (!plan_is(c->sy, PY_{}) || !plan_is(c->sz, PZ_{})) return false;
  const size_t sh2 = score2_shmem(PY_::N, PZ_::N);
  if (c->score2 && sh2 + 1024 <= (size_t)c->lds_max) *k = {k_score_yz2<PY_, PZ_>, sh2};
  else *k = {k_score_yz_f<PY_, PZ_>, score_shmem(c)};
  return true;
}

ScoreLaunch score_launch(const spk_ctx* c) {
  ScoreLaunch k{c->score_rmax == 25 ? k_score_yz<25> : k_score_yz<10>, score_shmem(c)};
  if (c->fixed_fft) {
    fixed_score<FPlan<125, 5, 5, 5>, FPlan<100, 10, 10>>(c, &k) ||
        fixed_score<FPlan<125, 25, 5>, FPlan<100, 10, 10>>(c, &k) ||
        fixed_score<FPlan<100, 10, 10>, FPlan<100, 10, 10>>(c, &k) ||
        fixed_score<FPlan<64, 8, 8>, FPlan<64, 8, 8>>(c, &k);
  }
  return k;
}

using CorrKernel = decltype(&k_corr_x);

template <class PX_, int T_>
bool fixed_corr(const spk_ctx* c, CorrKernel* k) {
  if (c->Tc != T_ || !plan_is(c->px, PX_{})) return false;
  *k = k_corr_x_f<PX_, T_>;
  return true;
}

// pass A for the context's x plan and tile: the register kernel for
// px = 240, else a compile-time instantiation when one matches, else the
// runtime one; with its block size and LDS bytes
struct CorrLaunch {
  CorrKernel k;
  int threads;
  size_t shmem;
};
CorrLaunch corr_launch(const spk_ctx* c) {
  const int T = c->Tc;
  if (c->fixed_fft && c->reg_x > 0 && c->px.n == 240 && (T == 4 || T == 8 || T == 16)) {
    const int nb = c->reg_x == 2 ? 2 : 3;
    CorrKernel k;
    if (nb == 3) k = T == 4 ? k_corr_xr<15, 4, 3> : T == 8 ? k_corr_xr<15, 8, 3> : k_corr_xr<15, 16, 3>;
    else k = T == 4 ? k_corr_xr<15, 4, 2> : T == 8 ? k_corr_xr<15, 8, 2> : k_corr_xr<15, 16, 2>;
    return {k, 16 * T, corr_xr_shmem(15, T, nb)};
  }
  // other x lengths 16 R: tile 4, two LDS buffers
  if (reg_len(c, c->px.n) && T == 4) {
    const int R = c->px.n / 16;
    CorrKernel k = R == 6 ? k_corr_xr<6, 4, 2> : R == 8 ? k_corr_xr<8, 4, 2> :
                   R == 9 ? k_corr_xr<9, 4, 2> : R == 10 ? k_corr_xr<10, 4, 2> :
                   R == 12 ? k_corr_xr<12, 4, 2> : R == 14 ? k_corr_xr<14, 4, 2> :
                   k_corr_xr<16, 4, 2>;
    return {k, 64, corr_xr_shmem(R, 4, 2)};
  }
  CorrKernel k = k_corr_x;
  if (c->fixed_fft) {
    fixed_corr<FPlan<240, 16, 15>, 16>(c, &k) || fixed_corr<FPlan<240, 16, 15>, 8>(c, &k) ||
        fixed_corr<FPlan<240, 16, 15>, 4>(c, &k);
  }
  return {k, kPencilThreads, (size_t)(c->px.n + 2 * c->px.n * T) * sizeof(float2)};
}

// Scored search (pass A, pass B, selection, exact re-check of the winner) of
// m descriptors whose item spectra are in c->ispec.
void enqueue_scored(spk_ctx* c, const OrientDesc* od, int m, int max_ex, int max_e, int penalty,
                    const int* ctl, unsigned char* fgrid, long long* pgrid,
                    const std::vector<OrientDesc>* host_ods, bool winner_only = false) {
  const int py = c->py.n, pz = c->pz.n, hz = pz / 2 + 1;
  const float2* tw = c->tw.p;
  const size_t ysh = (size_t)(py + py * c->Ty) * sizeof(float2);
  const int tiles = (pz + c->Tc - 1) / c->Tc;
  int chunk = 0;
  const int nchunk = corr_chunks(c, m, c->Tc, &chunk);
  const float scale = (float)(1.0 / ((double)c->px.n * py * pz));
  if (c->timing) {
    c->units["corr_x"] += m;
    c->units[c->split ? "rows_score" : "score_yz"] += m;
  }
  // column-test flags of this list (pruning needs CTL_PRUNE set by its decide)
  const int* cf = c->co.ready && !fgrid ? c->co.fs.p + kFeasHeader : nullptr;
  const CorrLaunch cl = corr_launch(c);
  launch(c, "corr_x", cl.k, dim3(tiles, py, nchunk), dim3(cl.threads), cl.shmem,
         (const float2*)c->g1.p, (const float2*)c->g2.p, (const float2*)c->ispec.p, od, c->y1.p,
         c->y2.p, m, chunk, max_ex, c->nx, c->px, py, pz, c->Tc, scale, tw, cf, ctl);
  const double* pen = c->pen.p + (penalty == SPK_PENALTY_CUBE ? c->nz : 0);
  const int rs = rec_stride(c);
  if (c->split) {
    launch(c, "pencil_y_inv", k_pencil<false>, dim3((pz + c->Ty - 1) / c->Ty, m * c->nx),
           dim3(kPencilThreads), ysh, c->y1.p, c->py, pz, c->Ty, py, py, tw + c->px.n, od, c->nx,
           (int)PENCIL_SCORE, ctl, 0);
    launch(c, "pencil_y_inv", k_pencil<false>, dim3((hz + c->Ty - 1) / c->Ty, m * c->nx),
           dim3(kPencilThreads), ysh, c->y2.p, c->py, hz, c->Ty, py, py, tw + c->px.n, od, c->nx,
           (int)PENCIL_SCORE, ctl, 0);
    launch(c, "rows_score", k_rows_score, dim3(rs, m), dim3(kRowThreads),
           (size_t)(pz + (c->Rs + (c->Rs + 1) / 2) * pz) * sizeof(float2), (const float2*)c->y1.p,
           (const float2*)c->y2.p, od, (const FieldStats*)c->stats.p, pen, c->rec.p, rs, c->nx,
           c->ny, c->nz, py, c->pz, c->Rs, tw + c->px.n + py, fgrid, pgrid, ctl);
  } else {
    // the score pass has its own factorisations of py and pz (radices up to
    // c->score_rmax, compiled as the kernel's largest butterfly), or a
    // compile-time plan when one matches them
    const ScoreLaunch sl = score_launch(c);
    launch(c, "score_yz", sl.k, dim3(c->nx, m),
           dim3(kSlabThreads), sl.shmem, (const float2*)c->y1.p, (const float2*)c->y2.p, od,
           (const FieldStats*)c->stats.p, pen, c->rec.p, c->nx, c->ny, c->nz, c->sy, c->sz,
           tw + c->px.n, tw + c->px.n + py, fgrid, pgrid, cf,
           c->timing && c->pairs.p ? c->pairs.p + 2 : (unsigned long long*)nullptr, ctl);
  }
  launch(c, "select", k_select, dim3(1), dim3(kSelectThreads), 0, (const WgBest*)c->rec.p, od, m,
         rs, c->split ? c->Rs : 0, c->per.p, c->sout.p, c->ver.p, ctl, (const int*)nullptr);
  if (fgrid) {
    // interlocking-free: the summary above feeds the masked argmin, whose
    // per-slice records are selected again
    if_masked_select(c, *host_ods, pen, winner_only);
    launch(c, "select", k_select, dim3(1), dim3(kSelectThreads), 0, (const WgBest*)c->rec2.p, od,
           m, c->nx, 0, c->per.p, c->sout.p, c->ver.p, ctl,
           winner_only ? (const int*)c->if_keep.p : (const int*)nullptr);
  }
  launch(c, "verify", k_verify, dim3(std::min(512, (max_e + 255) / 256)), dim3(256), 0,
         (const int*)c->tray.p, (const int*)c->phi.p, (const int*)c->items.p, od,
         (const SearchOut*)c->sout.p, c->ver.p, c->ny, c->nz, ctl);
}

void enqueue_search(spk_ctx* c, const OrientDesc* od, int m, int max_ex, int max_e, int penalty,
                    const int* ctl, unsigned char* fgrid, long long* pgrid,
                    const std::vector<OrientDesc>* host_ods, bool winner_only = false) {
  enqueue_item_spectra(c, od, m, max_ex, ctl);
  enqueue_scored(c, od, m, max_ex, max_e, penalty, ctl, fgrid, pgrid, host_ods, winner_only);
}

// items: per entry base item + orientation; all on host.
struct Entry {
  HostItem item;
  const spk_orient* orient;
  int item_index;   // which distinct base item
};

void do_search(spk_ctx* c, const std::vector<HostItem>& bases, const std::vector<int>& base_off,
               bool bases_uploaded, const std::vector<Entry>& entries, int penalty,
               int interlocking_free, spk_result* best, spk_result* per_orient) {
  if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
  const int n = (int)entries.size();
  if (n <= 0) throw SpkError{SPK_EINVAL, "no orientations"};
  // upload base items
  if (!bases_uploaded) {
    size_t tot = 0;
    for (const auto& b : bases) tot += (size_t)b.dims[0] * b.dims[1] * b.dims[2];
    c->items.ensure(std::max<size_t>(tot, 1));
    for (size_t i = 0; i < bases.size(); ++i) {
      const auto& b = bases[i];
      const size_t sz = (size_t)b.dims[0] * b.dims[1] * b.dims[2];
      if (sz) upload(c, c->items.p + base_off[i], b.data, sz * sizeof(int));
    }
  }
  // descriptors of the orientations that fit
  std::vector<OrientDesc> ods;
  std::vector<int> map;   // ods index -> entry index
  int max_ex = 1;
  for (int e = 0; e < n; ++e) {
    const auto& en = entries[e];
    OrientDesc d = make_desc(c, en.item, base_off[en.item_index], en.orient);
    if (!fits(c, d)) continue;
    ods.push_back(d);
    map.push_back(e);
    max_ex = std::max(max_ex, d.ex);
  }
  if (per_orient)
    for (int e = 0; e < n; ++e) to_result(c, false, 0, 0.0, -1, 0, 0, 0.f, 0, &per_orient[e]);
  if (ods.empty()) {
    to_result(c, false, 0, 0.0, -1, 0, 0, 0.f, 0, best);
    return;
  }
  const int m = (int)ods.size();
  int max_e = 1;
  for (const auto& d : ods) max_e = std::max(max_e, d.ex * d.ey * d.ez);
  c->od.ensure(m);
  upload(c, c->od.p, ods.data(), m * sizeof(OrientDesc));
  ensure_search_ws(c, m, max_ex);
  const double* pen = c->pen.p + (penalty == SPK_PENALTY_CUBE ? c->nz : 0);
  const size_t V = (size_t)c->nx * c->ny * c->nz;
  if (interlocking_free) {
    c->ffree.ensure((size_t)m * V);
    c->pgrid.ensure((size_t)m * V);
    HIP_OK(hipMemsetAsync(c->ffree.p, 0, (size_t)m * V, c->stream));
  }
  // (without per-orientation results only the winner is wanted)
  enqueue_search(c, c->od.p, m, max_ex, max_e, penalty, nullptr,
                 interlocking_free ? c->ffree.p : nullptr, interlocking_free ? c->pgrid.p : nullptr,
                 interlocking_free ? &ods : nullptr, per_orient == nullptr);
  SearchOut so;
  std::vector<WgBest> per(m);
  long long ver[3];
  download(c, &so, c->sout.p, sizeof(so));
  download(c, per.data(), c->per.p, m * sizeof(WgBest));
  download(c, ver, c->ver.p, sizeof(ver));
  sync(c);
  if (c->timing) c->flush_timing();
  const bool bad = so.unsupported || !(so.resid <= c->resid_limit) ||
                   (so.found && (ver[1] != 0 || ver[2] != so.prox));
  if (!bad) {
    for (int i = 0; i < m; ++i) {
      const WgBest& r = per[i];
      const bool f = r.idx != INT_MAX;
      const long long p = ((long long)r.p_hi << 32) | (unsigned int)r.p_lo;
      if (per_orient)
        to_result(c, f, r.idx, r.score, map[i], 0, p, r.resid, 0, &per_orient[map[i]]);
    }
    to_result(c, so.found, so.idx, so.score, so.found ? map[so.orient] : -1, ver[0], so.prox,
              so.resid, 0, best);
    return;
  }
  // exact fallback: direct sums for every orientation
  if (interlocking_free) {
    HIP_OK(hipMemsetAsync(c->ffree.p, 0, (size_t)m * V, c->stream));
    for (int i = 0; i < m; ++i) {
      const std::vector<Tap> taps = oriented_taps(entries[map[i]].item, ods[i]);
      c->taps.ensure(std::max<size_t>(taps.size(), 1));
      if (!taps.empty()) upload(c, c->taps.p, taps.data(), taps.size() * sizeof(Tap));
      const int nb = 1024;
      launch(c, "direct_grid", k_direct_grid, dim3(nb), dim3(256), 0, (const int*)c->tray.p,
             (const int*)c->phi.p, (const Tap*)c->taps.p, (int)taps.size(), ods[i].nvx,
             ods[i].nvy, ods[i].nvz, c->ny, c->nz, c->ffree.p + (size_t)i * V,
             c->pgrid.p + (size_t)i * V);
      sync(c);
    }
    if_masked_select(c, ods, pen);
    launch(c, "select", k_select, dim3(1), dim3(kSelectThreads), 0, (const WgBest*)c->rec2.p,
           (const OrientDesc*)c->od.p, m, c->nx, 0, c->per.p, c->sout.p, c->ver.p,
           (const int*)nullptr, (const int*)nullptr);
    SearchOut s2;
    download(c, &s2, c->sout.p, sizeof(s2));
    download(c, per.data(), c->per.p, m * sizeof(WgBest));
    sync(c);
    for (int i = 0; i < m && per_orient; ++i) {
      const WgBest& r = per[i];
      const long long p = ((long long)r.p_hi << 32) | (unsigned int)r.p_lo;
      to_result(c, r.idx != INT_MAX, r.idx, r.score, map[i], 0, p, so.resid, 1, &per_orient[map[i]]);
    }
    to_result(c, s2.found, s2.idx, s2.score, s2.found ? map[s2.orient] : -1, 0, s2.prox,
              so.resid, 1, best);
    return;
  }
  double bs = 1e30;
  int bi = -1, bo = -1;
  long long bp = 0;
  for (int i = 0; i < m; ++i) {
    const Entry& en = entries[map[i]];
    WgBest r = exact_search_one(c, en.item, ods[i], penalty);
    const bool f = r.idx != INT_MAX;
    const long long p = ((long long)r.p_hi << 32) | (unsigned int)r.p_lo;
    if (per_orient) to_result(c, f, r.idx, r.score, map[i], 0, p, so.resid, 1, &per_orient[map[i]]);
    if (f && r.score < bs) { bs = r.score; bi = r.idx; bo = i; bp = p; }
  }
  if (bo < 0) {
    to_result(c, false, 0, 0.0, -1, 0, 0, so.resid, 1, best);
    return;
  }
  to_result(c, true, bi, bs, map[bo], 0, bp, so.resid, 1, best);
}

void do_place(spk_ctx* c, const OrientDesc& d, int x, int y, int z, int id, bool refresh) {
  c->flag.ensure(1);
  HIP_OK(hipMemsetAsync(c->flag.p, 0, sizeof(int), c->stream));
  c->od.ensure(1);
  upload(c, c->od.p, &d, sizeof(d));
  const int E = d.ex * d.ey * d.ez;
  const int nb = std::max(1, std::min(1024, (E + 255) / 256));
  launch(c, "place", k_place, dim3(nb), dim3(256), 0, c->tray.p, (const int*)c->items.p,
         (const OrientDesc*)c->od.p, (const SearchOut*)nullptr, 0, x, y, z, id, c->nx, c->ny,
         c->nz, 0, c->flag.p, (const int*)nullptr);
  int flag = 0;
  download(c, &flag, c->flag.p, sizeof(int));
  sync(c);
  if (flag) throw SpkError{SPK_EOVERLAP, "(place_in_tray): Item to be placed intersects with contents of tray"};
  launch(c, "place", k_place, dim3(nb), dim3(256), 0, c->tray.p, (const int*)c->items.p,
         (const OrientDesc*)c->od.p, (const SearchOut*)nullptr, 0, x, y, z, id, c->nx, c->ny,
         c->nz, 1, c->flag.p, (const int*)nullptr);
  if (refresh) {
    PlaceRef pr;
    pr.od = c->od.p;
    pr.qx = x; pr.qy = y; pr.qz = z; pr.id = id;
    pr.max_ex = d.ex;
    pr.max_dim = std::max(d.ex, std::max(d.ey, d.ez));
    refresh_field_and_spectra(c, nullptr, nullptr, &pr);
  }
}

// Orientations of every item that fit the tray, in processing order.
struct ItemPlan {
  int od_off = 0, m = 0, max_ex = 1, max_exy = 1, max_e = 1;
  int max_dim = 1;        // largest extent along any axis of its fitting orientations
  std::vector<int> map;   // fitting index -> orientation index
};

struct PackPlan {
  std::vector<ItemPlan> items;
  std::vector<OrientDesc> all;    // descriptors, item after item
};

// Builds the plan, uploads the descriptors (c->od_all) and their orientation
// indices (c->map_all) and sizes the search workspace and loop state.
PackPlan plan_pack(spk_ctx* c, const std::vector<HostItem>& bases, const std::vector<int>& off,
                   const int32_t* order, const spk_orient* orients, int n_orient) {
  const int n_items = (int)bases.size();
  PackPlan P;
  P.items.resize(n_items);
  std::vector<int> map_all;
  int ws_m = 1, ws_ex = 1;
  // line masks of the base items, read by the column test's item masks
  std::vector<int> line_off(n_items, -1);
  if (c->co.ready && n_items <= 65535) {
    std::vector<ItemLines> il(n_items);
    int tot = 0, maxl = 1;
    for (int i = 0; i < n_items; ++i) {
      const HostItem& b = bases[i];
      il[i] = ItemLines{off[i], b.dims[0], b.dims[1], b.dims[2], tot};
      line_off[i] = tot;
      const int nl = b.dims[0] * b.dims[1] + b.dims[0] * b.dims[2] + b.dims[1] * b.dims[2];
      tot += nl;
      maxl = std::max(maxl, nl);
    }
    c->co.lines.ensure(std::max(tot, 1));
    c->co.il.ensure(n_items);
    upload(c, c->co.il.p, il.data(), il.size() * sizeof(ItemLines));
    launch(c, "item_lines", k_item_lines, dim3((maxl + 255) / 256, n_items), dim3(256), 0,
           (const int*)c->items.p, (const ItemLines*)c->co.il.p, c->co.lines.p);
  }
  for (int t = 0; t < n_items; ++t) {
    const int it = order ? order[t] : t;
    ItemPlan& pl = P.items[t];
    pl.od_off = (int)P.all.size();
    for (int o = 0; o < n_orient; ++o) {
      OrientDesc d = make_desc(c, bases[it], off[it], orients ? &orients[o] : nullptr);
      if (!fits(c, d)) continue;
      d.line_off = line_off[it];
      P.all.push_back(d);
      map_all.push_back(o);
      pl.map.push_back(o);
      pl.max_ex = std::max(pl.max_ex, d.ex);
      pl.max_exy = std::max(pl.max_exy, d.ex * d.ey);
      pl.max_e = std::max(pl.max_e, d.ex * d.ey * d.ez);
      pl.max_dim = std::max(pl.max_dim, std::max(d.ex, std::max(d.ey, d.ez)));
    }
    pl.m = (int)pl.map.size();
    ws_m = std::max(ws_m, pl.m);
    ws_ex = std::max(ws_ex, pl.max_ex);
  }
  c->od_all.ensure(std::max<size_t>(P.all.size(), 1));
  c->map_all.ensure(std::max<size_t>(map_all.size(), 1));
  if (!P.all.empty()) {
    upload(c, c->od_all.p, P.all.data(), P.all.size() * sizeof(OrientDesc));
    upload(c, c->map_all.p, map_all.data(), map_all.size() * sizeof(int));
  }
  ensure_search_ws(c, ws_m, ws_ex);
  c->pout.ensure(std::max(n_items, 1));
  c->ctl.ensure(CTL_WORDS);
  c->zero.ensure(1);
  HIP_OK(hipMemsetAsync(c->zero.p, 0, sizeof(int), c->stream));
  return P;
}

// Inscribed-ball bounds of the items, in processing order, for the
// free-space certificate (k_feas_begin): theta[t].mx = largest distance from
// a voxel of item t to a voxel outside it, in the tray field's metric.  Null
// when the certificate is off (option "certificate") or does not apply: the
// tray field must be this library's distance to every occupied voxel
// (sources "nonzero", not caller-supplied).
const FieldStats* item_bounds(spk_ctx* c, const std::vector<HostItem>& bases,
                              const std::vector<int>& off, const int32_t* order) {
  if (!c->certificate || !c->phi_internal || c->sources != SPK_SOURCES_NONZERO) return nullptr;
  const int n = (int)bases.size();
  c->theta.ensure(std::max(n, 1));
  size_t maxv = 1;
  for (const auto& b : bases)
    maxv = std::max(maxv, (size_t)(b.dims[0] + 2) * (b.dims[1] + 2) * (b.dims[2] + 2));
  c->cgrid.ensure(maxv);
  c->cdist.ensure(maxv);
  c->ctmp.ensure(maxv);
  for (int t = 0; t < n; ++t) {
    const int it = order ? order[t] : t;
    const HostItem& h = bases[it];
    const int X = h.dims[0] + 2, Y = h.dims[1] + 2, Z = h.dims[2] + 2;
    if ((long long)Y * Z > 32768 || X > 1024) {
      // too large for the distance kernels: no bound (mx = INT_MIN never certifies)
      launch(c, "stats_init", k_stats_init, dim3(1), dim3(64), 0, c->theta.p + t,
             (const int*)nullptr);
      continue;
    }
    const int nb = std::min(1024, (X * Y * Z + 255) / 256);
    launch(c, "complement", k_complement, dim3(nb), dim3(256), 0, (const int*)c->items.p + off[it],
           h.dims[0], h.dims[1], h.dims[2], c->cgrid.p);
    run_distance(c, c->cgrid.p, c->cdist.p, c->ctmp.p, c->theta.p + t, X, Y, Z, c->metric,
                 SPK_SOURCES_NONZERO, nullptr);
  }
  return c->theta.p;
}

// One item of the device-driven loop: batched search over m descriptors (od),
// k_decide, optional per-orientation score record, placement of the accepted
// winner and refresh of the distance field and spectra.
//
// With the feasibility pass on (slab plan, feas_mode 1 or 2), k_feas_begin
// decides on the device whether this item runs it (mode 1: when the previous
// item placed nothing); when it runs and finds no feasible orientation,
// k_feas_decide records the item as not found and the scored search,
// placement and refresh are skipped.
void enqueue_step(spk_ctx* c, const OrientDesc* od, int m, int max_ex, int max_exy, int max_e,
                  int penalty, int t, int32_t id_offset, int halt_not_found, double* rec_row,
                  const int* rec_map, int feas_mode = -1,
                  const FieldStats* item_bound = nullptr, bool known_feasible = false,
                  int max_dim = 0, const int* pre = nullptr) {
  if (feas_mode < 0) feas_mode = c->feas_mode;
  const bool feas = feas_mode && !c->split;
  if (feas || item_bound || c->co.ready) {
    // with the column test exact (no FFT feasibility pass after it) and no
    // certificate, its first kernel opens the item in place of k_feas_begin
    const bool begin = !known_feasible && !item_bound && coarse_begins(c, max_exy);
    if (!begin)
      launch(c, "feas_begin", k_feas_begin, dim3(1), dim3(64), 0, c->fstate.p,
             c->co.ready ? c->co.fs.p : (int*)nullptr, m, feas ? feas_mode : 0, c->ctl.p,
             (const FieldStats*)c->stats.p, item_bound, c->pout.p, t);
    // orientations a column test already found feasible need no second one
    if (!known_feasible)
      enqueue_coarse(c, od, m, max_ex, max_exy, c->ctl.p, c->pout.p, t, halt_not_found, false,
                     true, nullptr, pre, nullptr, nullptr, 0, begin);
    enqueue_item_spectra(c, od, m, max_ex, c->ctl.p);
    // the column test's flags are exact: no FFT feasibility pass after it
    if (feas && !c->co.ready) {
      enqueue_feas(c, od, m, max_ex, c->ctl.p);
      launch(c, "feas_decide", k_feas_decide, dim3(1), dim3(64), 0, (const int*)c->fstate.p, m,
             c->pout.p, t, c->resid_limit, c->co.ready ? c->co.fs.p : (int*)nullptr, c->ctl.p);
    }
    enqueue_scored(c, od, m, max_ex, max_e, penalty, c->ctl.p, nullptr, nullptr, nullptr);
  } else {
    enqueue_search(c, od, m, max_ex, max_e, penalty, c->ctl.p, nullptr, nullptr, nullptr);
  }
  launch(c, "decide", k_decide, dim3(1), dim3(64), 0, (const SearchOut*)c->sout.p,
         (const long long*)c->ver.p, c->pout.p, t, c->resid_limit, c->ctl.p, halt_not_found,
         spec_max(c), od, (int)id_offset);
  if (rec_row)
    launch(c, "record", k_record_scores, dim3(1), dim3(64), 0, (const WgBest*)c->per.p, m, rec_map,
           rec_row, (const int*)c->ctl.p);
  const int nb = std::max(1, std::min(1024, (max_e + 255) / 256));
  launch(c, "place", k_place, dim3(nb), dim3(256), 0, c->tray.p, (const int*)c->items.p, od,
         (const SearchOut*)c->sout.p, 0, 0, 0, 0, (int)id_offset, c->nx, c->ny, c->nz, 1,
         c->zero.p, (const int*)c->ctl.p);
  PlaceRef pr;
  pr.so = c->sout.p;
  pr.od = od;
  pr.id = id_offset;
  pr.max_ex = max_ex;
  pr.max_dim = max_dim;
  refresh_field_and_spectra(c, nullptr, c->ctl.p, &pr);
}

// Optional hipGraph capture of a device-driven loop's enqueue (option
// "graph"): every launch of the items from the first undecided one to the
// last goes into one graph, instantiated and launched once.  The loops'
// buffers are sized before the capture begins (presize_loop), no copy,
// memset or synchronisation is enqueued inside it, and per-kernel timing
// (events) turns the capture off.  On an exception the capture is ended and
// its graph discarded.
struct LoopGraph {
  spk_ctx* c;
  bool on;
  bool capturing = false;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  LoopGraph(spk_ctx* c_, bool want) : c(c_), on(want && !c_->timing) {
    if (on) {
      HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
      capturing = true;
    }
  }
  // end the capture and launch the graph
  void run() {
    if (!on) return;
    capturing = false;
    HIP_OK(hipStreamEndCapture(c->stream, &g));
    HIP_OK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    HIP_OK(hipGraphLaunch(ge, c->stream));
  }
  ~LoopGraph() {
    if (capturing) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(c->stream, &dead);
      if (dead) (void)hipGraphDestroy(dead);
    }
    if (ge) {
      (void)hipStreamSynchronize(c->stream);   // the launched graph has finished
      (void)hipGraphExecDestroy(ge);
    }
    if (g) (void)hipGraphDestroy(g);
  }
};

// Sizes every buffer an item step of the device-driven loops may grow, for
// items whose largest search has m orientations of x extent up to max_ex,
// so that a captured enqueue allocates nothing.
void presize_loop(spk_ctx* c, int m, int max_ex) {
  spk_ctx::Coarse& q = c->co;
  if (q.ready) {
    const int ccx = (max_ex + 1) / 2;
    q.cm.ensure((size_t)m * ccx * q.cy);
    q.im.ensure((size_t)m * max_ex * c->ny);
    q.rl.ensure((size_t)m * kColRuns);
    q.frl.ensure((size_t)m * kFineRuns);
    q.nrun.ensure(2 * (size_t)m);
  }
  c->stats_part.ensure((size_t)((c->nz + kIncLines - 1) / kIncLines) * c->ny * kIncXSplitMax);
}

struct ExactBest {
  int fit = -1;            // index into the item's fitting orientations, -1 = none
  int idx = 0;
  double score = 0.0;
  long long prox = 0;
};

// Exact direct-sum search of the listed fitting orientations (ods indexed by
// fitting index); the first strictly lower score wins.  per receives each
// listed orientation's best score (+inf if none), in list order.
ExactBest exact_over(spk_ctx* c, const HostItem& item, const OrientDesc* ods,
                     const std::vector<int>& list, int penalty, std::vector<double>* per) {
  ExactBest b;
  double bs = 1e30;
  if (per) per->assign(list.size(), (double)INFINITY);
  for (size_t j = 0; j < list.size(); ++j) {
    const int i = list[j];
    const WgBest r = exact_search_one(c, item, ods[i], penalty);
    if (r.idx == INT_MAX) continue;
    if (per) (*per)[j] = r.score;
    if (r.score < bs) {
      bs = r.score;
      b.fit = i;
      b.idx = r.idx;
      b.score = r.score;
      b.prox = ((long long)r.p_hi << 32) | (unsigned int)r.p_lo;
    }
  }
  return b;
}

// Result and placement of an item decided on the host's exact path; returns
// 1 if it was placed.
int place_exact(spk_ctx* c, const ExactBest& eb, const ItemPlan& pl,
                const std::vector<OrientDesc>& all, float resid, int32_t id, spk_result* res) {
  if (eb.fit < 0) {
    to_result(c, false, 0, 0.0, -1, 0, 0, resid, 1, res);
    return 0;
  }
  to_result(c, true, eb.idx, eb.score, pl.map[eb.fit], 0, eb.prox, resid, 1, res);
  do_place(c, all[pl.od_off + eb.fit], res->pos[0], res->pos[1], res->pos[2], id, true);
  return 1;
}

// Items that fit nowhere on the current tray fit nowhere on any later tray of
// the same run (a placement only adds occupied voxels), so the greedy loops
// decide them (not found) before their turn, with no per-item work: one
// batched column test over the fitting orientations of items [t_begin, t_end)
// on the current tray.  Returns per item of the range the number of its
// fitting orientations feasible now, an upper bound for the rest of the run:
// 0 = the item fits nowhere; -1 = not tested (no fitting orientation, or the
// column test does not apply: split plan, tray z > 128, option "coarse" off).
// The flags stay in c->pre_fs (orientation i of item t at
// kFeasHeader + plans[t].od_off - *od0_out) until the next call, for the
// per-item column tests of the round (pre_flags): 1 feasible, 0 infeasible,
// 2 untested.  An item stops being tested once cap of its orientations are
// proven feasible (the count returned is then at least cap).
// The batched column test of a round, in two halves so that it can be
// enqueued behind the previous round's items and read with that round's
// outcomes (one host synchronisation per round): screen_enqueue uploads the
// per-item bookkeeping and enqueues the test and the copy of its flags into
// pinned memory; screen_finish, after the caller has synchronised the
// stream, turns the flags into live[t - t_begin] = feasible orientations of
// item t proven by the test (-1 when it did not run).
struct Screen {
  int t_begin = 0, t_end = 0, od0 = 0, m = 0;
  bool ran = false;
};
Screen screen_enqueue(spk_ctx* c, const std::vector<ItemPlan>& plans, int t_begin, int t_end,
                      int cap) {
  Screen sc;
  sc.t_begin = t_begin;
  sc.t_end = t_end;
  spk_ctx::Coarse& q = c->co;
  if (!q.ready || !q.on || t_end <= t_begin) return sc;
  int od0 = -1, od1 = 0, max_ex = 1, max_exy = 1;
  for (int t = t_begin; t < t_end; ++t) {
    const ItemPlan& pl = plans[t];
    if (pl.m == 0) continue;
    if (od0 < 0) od0 = pl.od_off;
    od1 = pl.od_off + pl.m;
    max_ex = std::max(max_ex, pl.max_ex);
    max_exy = std::max(max_exy, pl.max_exy);
  }
  if (od0 < 0) return sc;
  sc.od0 = od0;
  const int m = od1 - od0;   // the fitting orientations of a range are contiguous
  sc.m = m;
  const int ni = t_end - t_begin;
  // pinned: per-item feasible counts (zero), each orientation's item, then
  // the flags coming back
  const size_t need = (size_t)ni + 2 * (size_t)m;
  if (c->pre_h_cap < need) {
    if (c->pre_h) HIP_OK(hipHostFree(c->pre_h));
    c->pre_h = nullptr;
    c->pre_h_cap = 0;
    HIP_OK(hipHostMalloc((void**)&c->pre_h, need * sizeof(int)));
    c->pre_h_cap = need;
  }
  int* aux = c->pre_h;
  std::fill(aux, aux + ni + m, 0);
  for (int t = t_begin; t < t_end; ++t)
    for (int i = 0; i < plans[t].m; ++i) aux[(size_t)ni + plans[t].od_off - od0 + i] = t - t_begin;
  c->pre_aux.ensure((size_t)ni + m);
  HIP_OK(hipMemcpyAsync(c->pre_aux.p, aux, ((size_t)ni + m) * sizeof(int), hipMemcpyHostToDevice,
                        c->stream));
  c->pre_fs.ensure((size_t)kFeasHeader + m);
  launch(c, "pre_clear", k_fill_int, dim3(1), dim3(256), 0, c->pre_fs.p, kFeasHeader + m, 0);
  enqueue_coarse(c, c->od_all.p + od0, m, max_ex, max_exy, nullptr, nullptr, 0, 0, false, false,
                 c->pre_fs.p, nullptr, c->pre_aux.p, c->pre_aux.p + ni, std::max(1, cap));
  HIP_OK(hipMemcpyAsync(c->pre_h + ni + m, c->pre_fs.p + kFeasHeader, (size_t)m * sizeof(int),
                        hipMemcpyDeviceToHost, c->stream));
  sc.ran = true;
  return sc;
}
std::vector<int> screen_finish(const spk_ctx* c, const std::vector<ItemPlan>& plans,
                               const Screen& sc) {
  std::vector<int> live((size_t)std::max(0, sc.t_end - sc.t_begin), -1);
  if (!sc.ran) return live;
  const int ni = sc.t_end - sc.t_begin;
  const int* f = c->pre_h + ni + sc.m;
  for (int t = sc.t_begin; t < sc.t_end; ++t) {
    const ItemPlan& pl = plans[t];
    if (pl.m == 0) continue;
    int n = 0;
    for (int i = 0; i < pl.m; ++i) n += f[(size_t)(pl.od_off - sc.od0 + i)] != 0 ? 1 : 0;
    live[t - sc.t_begin] = n;
  }
  return live;
}
// both halves with a synchronisation between them
std::vector<int> prescreen(spk_ctx* c, const std::vector<ItemPlan>& plans, int t_begin,
                           int t_end, int* od0_out, int cap) {
  const Screen sc = screen_enqueue(c, plans, t_begin, t_end, cap);
  *od0_out = sc.od0;
  if (sc.ran) sync(c);
  return screen_finish(c, plans, sc);
}
// Buffers of the batched screens of every round sized up front, so that a
// screen enqueued behind a round's items never reallocates one in use.
void presize_screens(spk_ctx* c, const std::vector<ItemPlan>& plans, int n_items, int round) {
  spk_ctx::Coarse& q = c->co;
  if (!q.ready || round <= 0) return;
  int mmax = 1, exmax = 1, nimax = 1;
  for (int t0 = 0; t0 < n_items; t0 += round) {
    const int t1 = std::min(n_items, t0 + round);
    int m = 0;
    for (int t = t0; t < t1; ++t) {
      m += plans[t].m;
      exmax = std::max(exmax, plans[t].max_ex);
    }
    mmax = std::max(mmax, m);
    nimax = std::max(nimax, t1 - t0);
  }
  const int ccx = (exmax + 1) / 2;
  q.cm.ensure((size_t)mmax * ccx * q.cy);
  q.im.ensure((size_t)mmax * exmax * c->ny);
  q.rl.ensure((size_t)mmax * kColRuns);
  q.frl.ensure((size_t)mmax * kFineRuns);
  q.nrun.ensure(2 * (size_t)mmax);
  c->pre_aux.ensure((size_t)nimax + mmax);
  c->pre_fs.ensure((size_t)kFeasHeader + mmax);
  const size_t need = (size_t)nimax + 2 * (size_t)mmax;
  if (c->pre_h_cap < need) {
    if (c->pre_h) HIP_OK(hipHostFree(c->pre_h));
    c->pre_h = nullptr;
    c->pre_h_cap = 0;
    HIP_OK(hipHostMalloc((void**)&c->pre_h, need * sizeof(int)));
    c->pre_h_cap = need;
  }
}

const int* pre_flags(const spk_ctx* c, const std::vector<ItemPlan>& plans, int t, int live,
                     int od0) {
  return live > 0 ? c->pre_fs.p + kFeasHeader + (plans[t].od_off - od0) : nullptr;
}


// The greedy loop with every item enqueued back to back: search, k_decide,
// placement and refresh of item t+1 follow item t on the stream with no host
// round trip.  k_decide hands a search back to the host (CTL_HALT) when its
// FFT result cannot be trusted; the kernels of the remaining items then exit
// at once, the host redoes that item on the exact direct-sum path, places it
// and resumes the pipeline after it.  Returns the number of items placed.
//
// Training records (rec_scores != nullptr): before item t's search the tray
// features it sees are written to rec_feat + t * kNnTrayFeat (device), and
// rec_scores[t * n_orient + o] receives the best score of orientation o
// (+inf when it does not fit or has no feasible offset).
int pack_pipelined(spk_ctx* c, const std::vector<HostItem>& bases, const std::vector<int>& off,
                   const int32_t* order, const spk_orient* orients, int n_orient, int penalty,
                   int32_t id_offset, spk_result* results, double* rec_scores = nullptr,
                   unsigned short* rec_feat = nullptr) {
  const int n_items = (int)bases.size();
  PackPlan P = plan_pack(c, bases, off, order, orients, n_orient);
  const std::vector<ItemPlan>& plans = P.items;
  const std::vector<OrientDesc>& all = P.all;
  const FieldStats* bounds = item_bounds(c, bases, off, order);
  std::vector<std::pair<int, std::vector<double>>> host_rows;   // records of host-path items
  if (rec_scores) {
    c->rscores.ensure((size_t)n_items * n_orient);
    std::vector<double> inf((size_t)n_items * n_orient, (double)INFINITY);
    upload(c, c->rscores.p, inf.data(), inf.size() * sizeof(double));
    c->heights.ensure((size_t)c->nx * c->ny);
  }

  int placed = 0;
  int t0 = 0;
  {
    int mm = 1, mx = 1;
    for (const ItemPlan& pl : plans) {
      mm = std::max(mm, pl.m);
      mx = std::max(mx, pl.max_ex);
    }
    presize_loop(c, mm, mx);
  }
  // rounds of c->prescreen items (one round when it is 0 or records are
  // written): items a batched column test finds infeasible on the tray at
  // the start of their round are decided there (skip), the others enqueued
  const bool rounds = c->prescreen > 0 && !rec_feat && !rec_scores;
  std::vector<char> skip(n_items, 0);
  std::vector<const int*> pre(n_items, nullptr);
  int spec_n = 0;   // incremental spectrum updates in a row, carried across rounds
  // the next round's screen, enqueued behind a round's items and read with
  // its outcomes (void when the loop halts: the host then changes the tray)
  Screen next;
  bool have_next = false;
  if (rounds) presize_screens(c, plans, n_items, c->prescreen);
  while (t0 < n_items) {
    const int t_end = rounds ? std::min(n_items, t0 + c->prescreen) : n_items;
    if (rounds) {
      int od0 = 0;
      std::vector<int> live;
      if (have_next && next.t_begin == t0 && next.t_end == t_end) {
        live = screen_finish(c, plans, next);
        od0 = next.od0;
      } else {
        live = prescreen(c, plans, t0, t_end, &od0, 1);
      }
      have_next = false;
      for (int t = t0; t < t_end; ++t) {
        skip[t] = live[t - t0] == 0 ? 1 : 0;
        pre[t] = pre_flags(c, plans, t, live[t - t0], od0);
      }
    }
    int ctl0[CTL_WORDS] = {0, -1, 0, placed};
    ctl0[CTL_SPEC_N] = spec_n;
    upload(c, c->ctl.p, ctl0, sizeof(ctl0));
    LoopGraph lg(c, c->graph && !rec_feat && !rec_scores);
    for (int t = t0; t < t_end; ++t) {
      const ItemPlan& pl = plans[t];
      if (pl.m == 0 || skip[t]) continue;   // nothing fits / fits nowhere: tray unchanged
      if (rec_feat)
        nn_tray_features(c->stream, c->tray.p, c->nx, c->ny, c->nz, c->heights.p,
                         rec_feat + (size_t)t * kNnTrayFeat);
      const OrientDesc* od = c->od_all.p + pl.od_off;
      enqueue_step(c, od, pl.m, pl.max_ex, pl.max_exy, pl.max_e, penalty, t, id_offset, 0,
                   rec_scores ? c->rscores.p + (size_t)t * n_orient : nullptr,
                   c->map_all.p + pl.od_off, -1, bounds ? bounds + t : nullptr, false,
                   pl.max_dim, pre[t]);
    }
    lg.run();
    if (rounds && t_end < n_items) {
      next = screen_enqueue(c, plans, t_end, std::min(n_items, t_end + c->prescreen), 1);
      have_next = next.ran;
    }
    int ctl[CTL_WORDS];
    std::vector<PackOut> po(t_end - t0);
    download(c, ctl, c->ctl.p, sizeof(ctl));
    download(c, po.data(), c->pout.p + t0, po.size() * sizeof(PackOut));
    sync(c);
    if (c->timing) c->flush_timing();
    const int stop = ctl[CTL_HALT] ? ctl[CTL_HALT_ITEM] : t_end;
    for (int t = t0; t < stop; ++t) {
      const ItemPlan& pl = plans[t];
      if (pl.m == 0 || skip[t]) {
        to_result(c, false, 0, 0.0, -1, 0, 0, 0.f, 0, &results[t]);
        continue;
      }
      const SearchOut& s = po[t - t0].so;
      to_result(c, s.found, s.idx, s.score, s.found ? pl.map[s.orient] : -1,
                po[t - t0].collision, s.prox, s.resid, 0, &results[t]);
      if (s.found) ++placed;
    }
    // (an item placed on the host's exact path is followed by a full transform)
    spec_n = ctl[CTL_HALT] ? 0 : ctl[CTL_SPEC_N];
    if (!ctl[CTL_HALT]) {
      t0 = t_end;
      continue;
    }
    have_next = false;
    // item `stop` goes through the exact path on the host's orders
    const int h = stop;
    const ItemPlan& pl = plans[h];
    const int it = order ? order[h] : h;
    std::vector<int> fit(pl.m);
    for (int i = 0; i < pl.m; ++i) fit[i] = i;
    std::vector<double> row;
    const ExactBest eb = exact_over(c, bases[it], &all[pl.od_off], fit, penalty,
                                    rec_scores ? &row : nullptr);
    if (rec_scores) {
      std::vector<double> full(n_orient, (double)INFINITY);
      for (int i = 0; i < pl.m; ++i) full[pl.map[i]] = row[i];
      host_rows.emplace_back(h, std::move(full));
    }
    placed += place_exact(c, eb, pl, all, po[h - t0].so.resid, id_offset + placed + 1,
                          &results[h]);
    t0 = h + 1;
  }
  if (rec_scores) {
    download(c, rec_scores, c->rscores.p, (size_t)n_items * n_orient * sizeof(double));
    sync(c);
    for (auto& hr : host_rows)
      std::copy(hr.second.begin(), hr.second.end(), rec_scores + (size_t)hr.first * n_orient);
  }
  if (placed > 0) c->phi_internal = true;   // a placement recomputed the field on the device
  return placed;
}

// Time a group of launches on the context stream as one named entry.
template <typename F>
void timed_group(spk_ctx* c, const char* name, F&& f) {
  if (!c->timing) {
    f();
    return;
  }
  hipEvent_t a = c->ev(), b = c->ev();
  HIP_OK(hipEventRecord(a, c->stream));
  f();
  HIP_OK(hipEventRecord(b, c->stream));
  c->pending.push_back({name, a, b});
}

// Greedy loop with learned orientation proposals.  The proposer scores the
// orientations of an item against the tray as it is at that item; the exact
// search evaluates its top-k.  Candidates are the item's fitting orientations,
// narrowed to the feasible ones when the feasibility pass ran first.  When
// more than topk candidates remain and the softmax mass of the topk
// best-scored ones (over the candidates) is at least min_conf, only those
// are searched, in orientation order so that ties resolve as in the full
// search, and the outcome is the best placement among them; all candidates
// are searched otherwise, and the remaining ones also when no proposal is
// feasible.  The feasibility pass runs first after an item that placed
// nothing (most such items cannot be placed at all), else only to screen the
// remaining orientations after infeasible proposals.  One host
// synchronisation per item (two when the pass runs first) reads the outcome
// and the logits of the next item, computed right after this item's
// placement.  counts: [0] items decided on their proposals, [1] searches of
// all candidates, [2] searches after infeasible proposals, [3] items decided
// on the host's exact path, [4] items the feasibility pass found infeasible.
// Pooled item-branch features of every item in each of its orientations, on
// the device (c->ipool): 24 rows per item in processing order (rows past
// n_orient repeat the base item; their logits are unused).
// which (optional): only these items (positions in processing order), whose
// rows go to consecutive slots of c->ipool in the order listed.
// (cat: the pooled rows go to c->icat at the head's row pitch kNnConcat, whose
// tray parts nn_record_rows fills per evaluation; else to c->ipool)
void learned_item_pool(spk_ctx* c, spk_nn* nn, const std::vector<HostItem>& bases,
                       const std::vector<int>& off, const int32_t* order,
                       const spk_orient* orients, int n_orient,
                       const std::vector<int>* which = nullptr, bool cat = false) {
  const int n_items = which ? (int)which->size() : (int)bases.size();
  if (n_items == 0) return;
  hipStream_t s = c->stream;
  timed_group(c, "proposer_items", [&]() {
    nn_cast(nn, s);
    std::vector<spknn::ItemMap> maps((size_t)n_items * kNnLogits);
    for (int i = 0; i < n_items; ++i) {
      const int t = which ? (*which)[i] : i;
      const int it = order ? order[t] : t;
      const HostItem& h = bases[it];
      for (int o = 0; o < kNnLogits; ++o) {
        spknn::ItemMap& mp = maps[(size_t)i * kNnLogits + o];
        mp.bx = h.dims[0]; mp.by = h.dims[1]; mp.bz = h.dims[2];
        mp.off = off[it];
        const spk_orient* so = (orients && o < n_orient) ? &orients[o] : nullptr;
        for (int a = 0; a < 3; ++a) {
          mp.perm[a] = so ? so->perm[a] : a;
          mp.flip[a] = so ? (so->flip[a] ? 1 : 0) : 0;
        }
      }
    }
    c->imaps.ensure(maps.size());
    upload(c, c->imaps.p, maps.data(), maps.size() * sizeof(spknn::ItemMap));
    const int bmax = nn->max_batch;
    c->ifeat.ensure((size_t)std::min(n_items, bmax) * kNnLogits * kNnItemFeat);
    const int ld = cat ? kNnConcat : kNnItemPooled;
    DevBuf<spknn::bf16_t>& pool = cat ? c->icat : c->ipool;
    pool.ensure((size_t)n_items * kNnLogits * ld);
    for (int b0 = 0; b0 < n_items; b0 += bmax) {
      const int nb = std::min(bmax, n_items - b0);
      nn_item_features(s, c->items.p, c->imaps.p + (size_t)b0 * kNnLogits, nb * kNnLogits,
                       c->ifeat.p);
      nn_item_branch(nn, s, c->ifeat.p, nb * kNnLogits, pool.p + (size_t)b0 * kNnLogits * ld, ld);
    }
  });
}

int pack_learned(spk_ctx* c, spk_nn* nn, const std::vector<HostItem>& bases,
                 const std::vector<int>& off, const int32_t* order, const spk_orient* orients,
                 int n_orient, int penalty, int32_t id_offset, spk_result* results, int topk,
                 float min_conf, int32_t* proposals, int32_t* counts) {
  const int n_items = (int)bases.size();
  PackPlan P = plan_pack(c, bases, off, order, orients, n_orient);
  const std::vector<ItemPlan>& plans = P.items;
  const std::vector<OrientDesc>& all = P.all;
  hipStream_t s = c->stream;
  // free-space certificate, checked on the host with the tray field's
  // statistics read at every synchronisation
  const FieldStats* bounds = item_bounds(c, bases, off, order);
  std::vector<FieldStats> hb(bounds ? n_items : 0);
  FieldStats tst{};
  if (bounds) download(c, hb.data(), bounds, (size_t)n_items * sizeof(FieldStats));
  download(c, &tst, c->stats.p, sizeof(tst));
  sync(c);
  learned_item_pool(c, nn, bases, off, order, orients, n_orient);
  c->tfeat.ensure(kNnTrayFeat);
  c->heights.ensure((size_t)c->nx * c->ny);
  c->od_prop.ensure(kNnLogits);
  c->od_rest.ensure(kNnLogits);
  ctl_feas_on(c);
  const bool feas_plan = c->feas_mode != 0 && !c->split;
  float logits[kNnLogits];
  int fsv[kFeasHeader + kNnLogits];
  // nn->tpool holds the tray branch of the tray as the stream leaves it: the
  // head of an item then needs only its own rows (no step or exact placement
  // was enqueued since the tray branch was)
  bool tpool_valid = false;
  auto enqueue_head = [&](int t) {
    timed_group(c, "proposer", [&]() {
      if (!tpool_valid) {
        nn_tray_features(s, c->tray.p, c->nx, c->ny, c->nz, c->heights.p, c->tfeat.p);
        nn_tray_branch(nn, s, c->tfeat.p, 1);
        tpool_valid = true;
      }
      nn_head(nn, s, c->ipool.p + (size_t)t * kNnLogits * kNnItemPooled, 1);
    });
  };
  // descriptors of the listed fitting orientations of item t, on the device
  auto gather = [&](const ItemPlan& pl, const std::vector<int>& list, OrientDesc* dst,
                    int* max_ex, int* max_e) -> const OrientDesc* {
    *max_ex = 1;
    *max_e = 1;
    for (int i : list) {
      const OrientDesc& d = all[pl.od_off + i];
      *max_ex = std::max(*max_ex, d.ex);
      *max_e = std::max(*max_e, d.ex * d.ey * d.ez);
    }
    if ((int)list.size() == pl.m) return c->od_all.p + pl.od_off;   // the whole list, in order
    OrientPick op{};
    for (size_t j = 0; j < list.size(); ++j) op.idx[j] = list[j];
    launch(c, "gather_od", k_gather_od, dim3(1), dim3(32), 0,
           (const OrientDesc*)c->od_all.p + pl.od_off, op, (int)list.size(), dst);
    return dst;
  };
  // feasibility of the listed orientations: false when the pass' residual is
  // out of bounds (nothing learned), else `list` keeps the feasible ones
  auto screen = [&](const ItemPlan& pl, std::vector<int>& list) -> bool {
    int max_ex, max_e;
    const OrientDesc* od = gather(pl, list, c->od_rest.p, &max_ex, &max_e);
    const int k = (int)list.size();
    // the item masks clear the flags when they come from k_item_masks
    const bool fused = c->co.ready && pl.max_exy <= kMaskCols;
    if (!fused)
      launch(c, "feas_begin", k_feas_begin, dim3(1), dim3(64), 0, c->fstate.p,
             c->co.ready ? c->co.fs.p : (int*)nullptr, k, 2, (int*)nullptr,
             (const FieldStats*)nullptr, (const FieldStats*)nullptr, (PackOut*)nullptr, 0);
    if (c->co.ready) {
      // the column test alone: its flags are exact (residual slot 0)
      enqueue_coarse(c, od, k, max_ex, pl.max_exy, c->ctl_on.p, nullptr, 0, 0, fused, false);
      const Mailbox& mb = post_and_sync(c, nullptr, nullptr, nullptr, nullptr, c->co.fs.p,
                                        kFeasHeader + k);
      std::memcpy(fsv, mb.fs, (kFeasHeader + k) * sizeof(int));
      fsv[0] = 0;
    } else {
      enqueue_item_spectra(c, od, k, max_ex, nullptr);
      enqueue_feas(c, od, k, max_ex, c->ctl_on.p);
      const Mailbox& mb = post_and_sync(c, nullptr, nullptr, nullptr, nullptr, c->fstate.p,
                                        kFeasHeader + k);
      std::memcpy(fsv, mb.fs, (kFeasHeader + k) * sizeof(int));
    }
    float res;
    std::memcpy(&res, &fsv[0], sizeof(res));
    if (!(res <= c->resid_limit)) return false;
    std::vector<int> keep;
    for (int j = 0; j < k; ++j)
      if (fsv[kFeasHeader + j]) keep.push_back(list[j]);
    list.swap(keep);
    return true;
  };
  // Pipelined screens (column test, every item screened, no certificate):
  // behind an item's screen the next screened item's screen is enqueued
  // before the host has read the first one's flags.  The first one's post
  // (k_post) writes the gate that lets it run only when the first item fits
  // nowhere, i.e. when the tray stays as it is; otherwise it exits at once
  // and its post is ignored.  The two screens post into different mailbox
  // slots.
  const bool pipe = c->co.ready && c->screen_all && feas_plan && !bounds;
  c->gates.ensure(kMailSlots * CTL_WORDS);
  auto enqueue_screen = [&](const ItemPlan& pl, const int* gate, int slot) -> unsigned {
    const OrientDesc* od = c->od_all.p + pl.od_off;   // every fitting orientation
    const bool fused = pl.max_exy <= kMaskCols;
    if (!fused)
      launch(c, "feas_begin", k_feas_begin, dim3(1), dim3(64), 0, c->fstate.p, c->co.fs.p, pl.m, 2,
             (int*)nullptr, (const FieldStats*)nullptr, (const FieldStats*)nullptr,
             (PackOut*)nullptr, 0);
    enqueue_coarse(c, od, pl.m, pl.max_ex, pl.max_exy, gate, nullptr, 0, 0, fused, false);
    // the post also writes the gate of the screen that may follow
    return post_async(c, slot, nullptr, nullptr, nullptr, nullptr, c->co.fs.p, kFeasHeader + pl.m,
                      c->gates.p + slot * CTL_WORDS);
  };
  struct Spec {
    int t = -1, slot = 0;
    unsigned seq = 0;
  } spec;
  int placed = 0;
  auto reset_ctl = [&]() {
    int v[CTL_WORDS] = {0};
    v[CTL_HALT_ITEM] = -1;
    v[CTL_PLACED] = placed;
    upload(c, c->ctl.p, v, sizeof(v));
  };
  auto accept = [&](int t, const ItemPlan& pl, const std::vector<int>& list, const PackOut& po) {
    const SearchOut& so = po.so;
    to_result(c, so.found, so.idx, so.score, so.found ? pl.map[list[so.orient]] : -1,
              po.collision, so.prox, so.resid, 0, &results[t]);
    if (so.found) ++placed;
  };
  int cnt[5] = {0, 0, 0, 0, 0};
  reset_ctl();
  bool have = false;        // logits of item t are in `logits` and match the tray
  bool prev_placed = true;
  for (int t = 0; t < n_items; ++t) {
    const ItemPlan& pl = plans[t];
    if (proposals)
      for (int j = 0; j < topk; ++j) proposals[(size_t)t * topk + j] = -1;
    if (pl.m == 0) {
      to_result(c, false, 0, 0.0, -1, 0, 0, 0.f, 0, &results[t]);
      continue;
    }
    const bool have_t = have;   // logits of item t computed after the previous step
    have = false;
    if (bounds && tst.mxw < hb[t].mx) {
      ++cnt[4];
      to_result(c, false, 0, 0.0, -1, 0, 0, 0.f, 0, &results[t]);
      prev_placed = false;
      continue;
    }
    std::vector<int> cand(pl.m);
    for (int i = 0; i < pl.m; ++i) cand[i] = i;
    bool screened = false;
    if (feas_plan && (!prev_placed || (c->screen_all && c->co.ready))) {
      if (pipe) {
        int slot = 0;
        unsigned seq = 0;
        if (spec.t == t) {   // enqueued behind the previous screen, which found nothing
          slot = spec.slot;
          seq = spec.seq;
        } else {
          seq = enqueue_screen(pl, c->ctl_on.p, slot);
        }
        spec.t = -1;
        int tn = t + 1;
        while (tn < n_items && plans[tn].m == 0) ++tn;
        if (tn < n_items) {
          spec.t = tn;
          spec.slot = slot ^ 1;
          spec.seq = enqueue_screen(plans[tn], c->gates.p + slot * CTL_WORDS, spec.slot);
        }
        const Mailbox& mb = wait_post(c, slot, seq);
        std::memcpy(fsv, mb.fs, (kFeasHeader + pl.m) * sizeof(int));
        fsv[0] = 0;
        // as screen(): a residual (0 for the exact column test) above the
        // limit leaves the candidates as they are
        screened = 0.f <= c->resid_limit;
        if (screened) {
          std::vector<int> keep;
          for (int j = 0; j < pl.m; ++j)
            if (fsv[kFeasHeader + j]) keep.push_back(j);
          cand.swap(keep);
        }
        if (!cand.empty()) spec.t = -1;   // the item is searched: the next screen is redone
      } else {
        screened = screen(pl, cand);
      }
      if (screened && cand.empty()) {
        ++cnt[4];
        float res;
        std::memcpy(&res, &fsv[0], sizeof(res));
        to_result(c, false, 0, 0.0, -1, 0, 0, res, 0, &results[t]);
        continue;
      }
      if (!screened) {
        cand.resize(pl.m);
        for (int i = 0; i < pl.m; ++i) cand[i] = i;
      }
    }
    std::vector<int> pick;
    bool use_prop = false;
    if ((int)cand.size() > topk) {
      if (!have_t) {
        enqueue_head(t);
        const Mailbox& mb = post_and_sync(c, nullptr, nullptr, nullptr, nn->logits.p, nullptr, 0);
        std::memcpy(logits, mb.logits, sizeof(logits));
      }
      float mx = -INFINITY;
      for (int i : cand) mx = std::max(mx, logits[pl.map[i]]);
      std::vector<int> rank = cand;
      double se = 0.0;
      std::vector<double> e(pl.m, 0.0);
      for (int i : cand) {
        e[i] = std::exp((double)logits[pl.map[i]] - mx);
        se += e[i];
      }
      std::stable_sort(rank.begin(), rank.end(), [&](int a, int b) {
        return logits[pl.map[a]] > logits[pl.map[b]];
      });
      double top = 0.0;
      for (int j = 0; j < topk; ++j) top += e[rank[j]];
      if (top / se >= (double)min_conf) {
        use_prop = true;
        pick.assign(rank.begin(), rank.begin() + topk);
        std::sort(pick.begin(), pick.end());
      }
    }
    if (!use_prop) pick = cand;
    int max_ex, max_e;
    const OrientDesc* od = gather(pl, pick, c->od_prop.p, &max_ex, &max_e);
    if (use_prop && proposals)
      for (size_t j = 0; j < pick.size(); ++j) proposals[(size_t)t * topk + j] = pl.map[pick[j]];
    enqueue_step(c, od, (int)pick.size(), max_ex, pl.max_exy, max_e, penalty, t, id_offset,
                 use_prop ? 1 : 0, nullptr, nullptr, 0, nullptr, screened && c->co.ready,
                 pl.max_dim);
    tpool_valid = false;
    // the next searched item's logits, on the tray this step leaves
    int tn = t + 1;
    while (tn < n_items && plans[tn].m == 0) ++tn;
    const bool next_head = tn < n_items && plans[tn].m > topk;
    if (next_head) enqueue_head(tn);
    int ctl[CTL_WORDS];
    PackOut po;
    {
      const unsigned sq = post_async(c, 0, c->ctl.p, c->pout.p + t, c->stats.p,
                                     next_head ? nn->logits.p : nullptr, nullptr, 0);
      // the next screened item's screen, behind this step and its post: valid
      // whatever the step places (undone below if the host takes over)
      if (pipe && tn < n_items) {
        spec.t = tn;
        spec.slot = 1;
        spec.seq = enqueue_screen(plans[tn], c->ctl_on.p, 1);
      }
      const Mailbox& mb = wait_post(c, 0, sq);
      std::memcpy(ctl, mb.ctl, sizeof(ctl));
      po = mb.po;
      tst = mb.st;
      if (next_head) std::memcpy(logits, mb.logits, sizeof(logits));
    }
    if (c->timing) c->flush_timing();
    if (ctl[CTL_HALT] == 0) {
      accept(t, pl, pick, po);
      ++cnt[use_prop ? 0 : 1];
      prev_placed = po.so.found;
      have = next_head;
      continue;
    }
    spec.t = -1;   // the paths below may change the tray: the next screen is redone
    const int code = ctl[CTL_HALT];
    float resid = po.so.resid;
    reset_ctl();
    std::vector<int> rest;
    if (code == 2) {
      // no proposal is feasible: the other candidates, screened for
      // feasibility first unless they already were
      ++cnt[2];
      for (int i : cand)
        if (!std::binary_search(pick.begin(), pick.end(), i)) rest.push_back(i);
      if (feas_plan && !screened && screen(pl, rest) && rest.empty()) {
        ++cnt[4];
        to_result(c, false, 0, 0.0, -1, 0, 0, resid, 0, &results[t]);
        prev_placed = false;
        continue;
      }
      if (!rest.empty()) {
        const OrientDesc* odr = gather(pl, rest, c->od_rest.p, &max_ex, &max_e);
        enqueue_step(c, odr, (int)rest.size(), max_ex, pl.max_exy, max_e, penalty, t, id_offset, 0,
                     nullptr, nullptr, 0, nullptr, false, pl.max_dim);
        tpool_valid = false;
        {
          const Mailbox& mb = post_and_sync(c, c->ctl.p, c->pout.p + t, c->stats.p, nullptr,
                                            nullptr, 0);
          std::memcpy(ctl, mb.ctl, sizeof(ctl));
          po = mb.po;
          tst = mb.st;
        }
        if (c->timing) c->flush_timing();
        if (ctl[CTL_HALT] == 0) {
          accept(t, pl, rest, po);
          prev_placed = po.so.found;
          reset_ctl();
          continue;
        }
        resid = po.so.resid;
        reset_ctl();
      }
      pick.resize(pl.m);
      for (int i = 0; i < pl.m; ++i) pick[i] = i;
      use_prop = false;
    }
    // the FFT outcome was not trusted: exact direct sums over the same
    // orientations, then over all fitting ones if no proposal is feasible
    ++cnt[3];
    const int it = order ? order[t] : t;
    ExactBest eb = exact_over(c, bases[it], &all[pl.od_off], pick, penalty, nullptr);
    if (eb.fit < 0 && use_prop) {
      ++cnt[2];
      std::vector<int> every(pl.m);
      for (int i = 0; i < pl.m; ++i) every[i] = i;
      eb = exact_over(c, bases[it], &all[pl.od_off], every, penalty, nullptr);
    }
    placed += place_exact(c, eb, pl, all, resid, id_offset + placed + 1, &results[t]);
    tpool_valid = false;
    prev_placed = eb.fit >= 0;
    reset_ctl();
    tst = post_and_sync(c, nullptr, nullptr, c->stats.p, nullptr, nullptr, 0).st;
  }
  if (counts)
    for (int i = 0; i < 5; ++i) counts[i] = cnt[i];
  if (placed > 0) c->phi_internal = true;   // a placement recomputed the field on the device
  return placed;
}

// Device-driven learned loop.  Every item is enqueued back to back with no
// host round trip: its column test (exact feasibility flags; an item with no
// feasible orientation is decided there), the proposer on the current tray
// (its kernels exit at once unless more than topk orientations are
// feasible), the pick (k_learn_pick: the topk feasible orientations with the
// highest logits, or every feasible one), the search over the pick, the
// decision, the placement and the refresh.  The host reads the outcomes once
// at the end, or at a halt (an untrusted FFT outcome, or a search over
// proposals that found nothing), where it decides that item on the exact
// path and resumes.  Same results as pack_learned with the column test
// screening every item and no confidence threshold.
int pack_learned_dev(spk_ctx* c, spk_nn* nn, const std::vector<HostItem>& bases,
                     const std::vector<int>& off, const int32_t* order, const spk_orient* orients,
                     int n_orient, int penalty, int32_t id_offset, spk_result* results, int topk,
                     int32_t* proposals, int32_t* counts) {
  const int n_items = (int)bases.size();
  PackPlan P = plan_pack(c, bases, off, order, orients, n_orient);
  const std::vector<ItemPlan>& plans = P.items;
  const std::vector<OrientDesc>& all = P.all;
  hipStream_t s = c->stream;
  // every search runs over topk slots (padding skipped): size the workspace
  int ws_ex = 1;
  for (const ItemPlan& pl : plans) ws_ex = std::max(ws_ex, pl.max_ex);
  ensure_search_ws(c, topk, ws_ex);
  c->tfeat.ensure(kNnTrayFeat);
  c->heights.ensure((size_t)c->nx * c->ny);
  c->od_prop.ensure(topk);
  c->learn_gate.ensure(1);
  c->learn_pick.ensure((size_t)std::max(1, n_items) * topk);
  c->learn_used.ensure(std::max(1, n_items));
  int placed = 0, t0 = 0;
  int cnt[5] = {0, 0, 0, 0, 0};
  std::vector<int> lp, lu;
  {
    int mm = topk;
    for (const ItemPlan& pl : plans) mm = std::max(mm, pl.m);
    presize_loop(c, mm, ws_ex);
  }
  // rounds of c->prescreen items, as in pack_pipelined: items a batched
  // column test finds infeasible at the start of their round are decided
  // there (the column test would decide them at their turn)
  // An item with at most topk orientations feasible at the start of its round
  // has at most topk at its turn: the proposer would be gated off, so it is
  // not enqueued (k_learn_pick then takes every feasible orientation).
  // The item branch of the proposer (every orientation of an item) runs per
  // round for the round's items that may need proposals, into consecutive
  // slots of c->ipool (slot[t]).
  std::vector<char> skip(n_items, 0), propose(n_items, 1);
  std::vector<int> slot(n_items, 0);
  // tray features from the column test's masks (they mark voxels != 0, the
  // features count voxels > 0: equal while no voxel is negative)
  const bool feat_masks = c->tray_nonneg && id_offset >= 0 && c->co.ready;
  std::vector<const int*> pre(n_items, nullptr);
  int spec_n = 0;   // incremental spectrum updates in a row, carried across rounds
  // the next round's screen, enqueued behind a round's items (as in
  // pack_pipelined)
  Screen next;
  bool have_next = false;
  if (c->prescreen > 0) presize_screens(c, plans, n_items, c->prescreen);
  while (t0 < n_items) {
    const int t_end = c->prescreen > 0 ? std::min(n_items, t0 + c->prescreen) : n_items;
    if (c->prescreen > 0) {
      int od0 = 0;
      std::vector<int> live;
      if (have_next && next.t_begin == t0 && next.t_end == t_end) {
        live = screen_finish(c, plans, next);
        od0 = next.od0;
      } else {
        live = prescreen(c, plans, t0, t_end, &od0, topk + 1);
      }
      have_next = false;
      for (int t = t0; t < t_end; ++t) {
        skip[t] = live[t - t0] == 0 ? 1 : 0;
        propose[t] = live[t - t0] < 0 || live[t - t0] > topk ? 1 : 0;
        pre[t] = pre_flags(c, plans, t, live[t - t0], od0);
      }
    }
    {
      std::vector<int> need;
      for (int t = t0; t < t_end; ++t)
        if (plans[t].m > 0 && !skip[t] && propose[t]) {
          slot[t] = (int)need.size();
          need.push_back(t);
        }
      learned_item_pool(c, nn, bases, off, order, orients, n_orient, &need, true);
    }
    int ctl0[CTL_WORDS] = {0, -1, 0, placed};
    ctl0[CTL_SPEC_N] = spec_n;
    upload(c, c->ctl.p, ctl0, sizeof(ctl0));
    int* ctl = c->ctl.p;
    LoopGraph lg(c, c->graph);
    for (int t = t0; t < t_end; ++t) {
      const ItemPlan& pl = plans[t];
      if (pl.m == 0 || skip[t]) continue;   // nothing fits / fits nowhere: tray unchanged
      const OrientDesc* od_fit = c->od_all.p + pl.od_off;
      // the item opens in its first column-test kernel (else k_feas_begin),
      // and the proposer gate is set with the column test's decision
      const bool begin = coarse_begins(c, pl.max_exy);
      if (!begin)
        launch(c, "feas_begin", k_feas_begin, dim3(1), dim3(64), 0, c->fstate.p, c->co.fs.p, pl.m,
               0, ctl, (const FieldStats*)nullptr, (const FieldStats*)nullptr, c->pout.p, t);
      enqueue_coarse(c, od_fit, pl.m, pl.max_ex, pl.max_exy, ctl, c->pout.p, t, 0, false, true,
                     nullptr, pre[t], nullptr, nullptr, 0, begin, topk,
                     propose[t] ? c->learn_gate.p : (int*)nullptr);
      if (propose[t]) {
        nn->gate = c->learn_gate.p;
        timed_group(c, "proposer", [&]() {
          if (feat_masks)
            nn_tray_features_masks(s, c->co.tf.p, c->nx, c->ny, c->nz, c->tfeat.p,
                                   c->learn_gate.p);
          else
            nn_tray_features(s, c->tray.p, c->nx, c->ny, c->nz, c->heights.p, c->tfeat.p,
                             c->learn_gate.p);
          nn_record_rows(nn, s, c->tfeat.p, c->icat.p + (size_t)slot[t] * kNnLogits * kNnConcat);
        });
        nn->gate = nullptr;
      }
      launch(c, "learn_pick", k_learn_pick, dim3(1), dim3(64), 0, (const float*)nn->logits.p,
             (const int*)c->map_all.p + pl.od_off, od_fit, pl.m, topk, c->co.fs.p, c->od_prop.p,
             c->learn_pick.p + (size_t)t * topk, c->learn_used.p + t, (const int*)ctl);
      const OrientDesc* od = c->od_prop.p;
      enqueue_item_spectra(c, od, topk, pl.max_ex, ctl);
      enqueue_scored(c, od, topk, pl.max_ex, pl.max_e, penalty, ctl, nullptr, nullptr, nullptr);
      launch(c, "decide", k_decide, dim3(1), dim3(64), 0, (const SearchOut*)c->sout.p,
             (const long long*)c->ver.p, c->pout.p, t, c->resid_limit, ctl, 1, spec_max(c), od,
             (int)id_offset);
      const int nb = std::max(1, std::min(1024, (pl.max_e + 255) / 256));
      launch(c, "place", k_place, dim3(nb), dim3(256), 0, c->tray.p, (const int*)c->items.p, od,
             (const SearchOut*)c->sout.p, 0, 0, 0, 0, (int)id_offset, c->nx, c->ny, c->nz, 1,
             c->zero.p, (const int*)ctl);
      PlaceRef pr;
      pr.so = c->sout.p;
      pr.od = od;
      pr.id = id_offset;
      pr.max_ex = pl.max_ex;
      pr.max_dim = pl.max_dim;
      refresh_field_and_spectra(c, nullptr, ctl, &pr);
    }
    lg.run();
    if (c->prescreen > 0 && t_end < n_items) {
      next = screen_enqueue(c, plans, t_end, std::min(n_items, t_end + c->prescreen), topk + 1);
      have_next = next.ran;
    }
    int cw[CTL_WORDS];
    const int nrem = t_end - t0;
    std::vector<PackOut> po(nrem);
    lp.assign((size_t)nrem * topk, -1);
    lu.assign(nrem, -1);
    download(c, cw, c->ctl.p, sizeof(cw));
    download(c, po.data(), c->pout.p + t0, po.size() * sizeof(PackOut));
    download(c, lp.data(), c->learn_pick.p + (size_t)t0 * topk, lp.size() * sizeof(int));
    download(c, lu.data(), c->learn_used.p + t0, lu.size() * sizeof(int));
    sync(c);
    if (c->timing) c->flush_timing();
    const int stop = cw[CTL_HALT] ? cw[CTL_HALT_ITEM] : t_end;
    for (int t = t0; t < stop; ++t) {
      const ItemPlan& pl = plans[t];
      if (proposals)
        for (int j = 0; j < topk; ++j) proposals[(size_t)t * topk + j] = -1;
      const int r = t - t0;
      if (pl.m == 0) {
        to_result(c, false, 0, 0.0, -1, 0, 0, 0.f, 0, &results[t]);
        continue;
      }
      if (skip[t]) {   // decided by the batched screen, as the column test would
        ++cnt[4];
        to_result(c, false, 0, 0.0, -1, 0, 0, 0.f, 0, &results[t]);
        continue;
      }
      if (lu[r] < 0) {   // the column test found no feasible orientation
        ++cnt[4];
        to_result(c, false, 0, 0.0, -1, 0, 0, 0.f, 0, &results[t]);
        continue;
      }
      const SearchOut& so = po[r].so;
      const int* pk = lp.data() + (size_t)r * topk;
      to_result(c, so.found, so.idx, so.score, so.found ? pl.map[pk[so.orient]] : -1,
                po[r].collision, so.prox, so.resid, 0, &results[t]);
      if (so.found) ++placed;
      ++cnt[lu[r] ? 0 : 1];
      if (lu[r] && proposals)
        for (int j = 0; j < topk; ++j)
          proposals[(size_t)t * topk + j] = pk[j] >= 0 ? pl.map[pk[j]] : -1;
    }
    spec_n = cw[CTL_HALT] ? 0 : cw[CTL_SPEC_N];
    if (!cw[CTL_HALT]) {
      t0 = t_end;
      continue;
    }
    have_next = false;
    // item `stop` on the exact path: over its pick, then (no proposal
    // feasible) over every fitting orientation
    const int h = stop, r = h - t0;
    const ItemPlan& pl = plans[h];
    const int it = order ? order[h] : h;
    const int* pk = lp.data() + (size_t)r * topk;
    std::vector<int> pick, every(pl.m);
    for (int j = 0; j < topk; ++j)
      if (pk[j] >= 0) pick.push_back(pk[j]);
    for (int i = 0; i < pl.m; ++i) every[i] = i;
    if (proposals)
      for (int j = 0; j < topk; ++j)
        proposals[(size_t)h * topk + j] = (lu[r] > 0 && pk[j] >= 0) ? pl.map[pk[j]] : -1;
    ExactBest eb;
    if (cw[CTL_HALT] == 2) {
      ++cnt[2];
      eb = exact_over(c, bases[it], &all[pl.od_off], every, penalty, nullptr);
    } else {
      ++cnt[3];
      eb = exact_over(c, bases[it], &all[pl.od_off], pick.empty() ? every : pick, penalty,
                      nullptr);
      if (eb.fit < 0 && lu[r] > 0) {
        ++cnt[2];
        eb = exact_over(c, bases[it], &all[pl.od_off], every, penalty, nullptr);
      }
    }
    placed += place_exact(c, eb, pl, all, po[r].so.resid, id_offset + placed + 1, &results[h]);
    t0 = h + 1;
  }
  if (counts)
    for (int i = 0; i < 5; ++i) counts[i] = cnt[i];
  if (placed > 0) c->phi_internal = true;   // a placement recomputed the field on the device
  return placed;
}

// ---- autotuner -------------------------------------------------------------
// The x-pencil tile T (kz columns per block of the pencil kernels) trades
// blocks in flight against per-block work and write granularity, and the
// best value depends on the tray shape.  On the first spk_tray_set of a shape
// every admissible T in {4, 8, 16} is timed on a real 24-orientation search
// (and, on the slab plan, a feasibility pass) of a box item sized to the tray;
// the fastest is kept for that shape for the rest of the process.
struct TuneKey {
  int nx, ny, nz, split;
  bool operator<(const TuneKey& o) const {
    return std::tie(nx, ny, nz, split) < std::tie(o.nx, o.ny, o.nz, o.split);
  }
};
// the autotuner's choice for a tray shape: x-pencil tile, largest butterfly
// of the score pass, padded y and z FFT sizes
struct TuneChoice {
  int T, Tc, score_rmax, py, pz;
};
std::map<TuneKey, TuneChoice> g_tile;
std::mutex g_tile_mu;

// the 24 proper rotations as index maps
std::vector<spk_orient> rotation_maps() {
  std::vector<spk_orient> out;
  const int perms[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
  for (int p = 0; p < 6; ++p)
    for (int f = 0; f < 8; ++f) {
      const int flips = (f & 1) + ((f >> 1) & 1) + ((f >> 2) & 1);
      const bool even_perm = p < 3;
      if (((flips & 1) == 0) != even_perm) continue;   // determinant +1
      spk_orient o;
      for (int a = 0; a < 3; ++a) {
        o.perm[a] = perms[p][a];
        o.flip[a] = (f >> a) & 1;
      }
      out.push_back(o);
    }
  return out;
}

// Autotuner, per tray shape and process: x-pencil tile (4, 8, 16) of the
// pencil kernels, then pass A's own tile (4, 8, 16), the
// score pass's largest butterfly (radix 10 or 25: different stage counts and
// register pressure) and the padded y/z FFT sizes (the smallest
// 2^a 3^b 5^c 7^d size or the next one: more points, maybe fewer or cheaper
// stages), each timed on a batched 24-orientation search plus its
// feasibility pass with a box item about a fifth of the tray's smallest side.
// Coordinate search in that order; options that fix a choice ("tile",
// "score_radix", "pad_y", "pad_z") take it out of the search.  The choice
// changes no result (every plan gives exact integers).
void autotune_tile(spk_ctx* c) {
  const TuneKey key{c->nx, c->ny, c->nz, c->split ? 1 : 0};
  auto apply = [&](const TuneChoice& ch) {
    if (ch.py != c->py.n || ch.pz != c->pz.n) {
      c->pad_y = ch.py;
      c->pad_z = ch.pz;
      plan_tray(c);
      build_spectra(c);
    }
    c->score_rmax = ch.score_rmax;
    c->sy = make_plan(c->py.n, std::min(c->yz_rmax, c->score_rmax));
    c->sz = make_plan(c->pz.n, std::min(c->yz_rmax, c->score_rmax));
    c->T = ch.T;
    c->Tc = ch.Tc;
  };
  {
    std::lock_guard<std::mutex> g(g_tile_mu);
    auto it = g_tile.find(key);
    if (it != g_tile.end()) {
      apply(it->second);
      return;
    }
  }
  TuneChoice best{c->T, c->Tc, c->score_rmax, c->py.n, c->pz.n};
  // box item about a fifth of the tray's smallest side, slightly uneven
  const int a = std::max(2, std::min({c->nx, c->ny, c->nz}) / 5);
  const int d0 = std::min(c->nx, a), d1 = std::min(c->ny, std::max(2, a * 4 / 5)),
            d2 = std::min(c->nz, std::max(2, a * 6 / 5));
  std::vector<int> box((size_t)d0 * d1 * d2, 1);
  c->items.ensure(box.size());
  upload(c, c->items.p, box.data(), box.size() * sizeof(int));
  const HostItem h = scan_item(box.data(), d0, d1, d2);
  std::vector<OrientDesc> ods;
  int max_ex = 1, max_e = 1;
  for (const spk_orient& o : rotation_maps()) {
    OrientDesc d = make_desc(c, h, 0, &o);
    if (!fits(c, d)) continue;
    ods.push_back(d);
    max_ex = std::max(max_ex, d.ex);
    max_e = std::max(max_e, d.ex * d.ey * d.ez);
  }
  const int m = (int)ods.size();
  if (m > 0) {
    c->od.ensure(m);
    upload(c, c->od.p, ods.data(), m * sizeof(OrientDesc));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    const bool timing = c->timing;
    c->timing = false;
    auto measure = [&]() -> float {
      ensure_search_ws(c, m, max_ex);
      auto run = [&]() {
        enqueue_search(c, c->od.p, m, max_ex, max_e, SPK_PENALTY_CUBE, nullptr, nullptr,
                       nullptr, nullptr);
        if (!c->split) enqueue_feas(c, c->od.p, m, max_ex, nullptr);
      };
      run();   // warm-up
      HIP_OK(hipEventRecord(e0, c->stream));
      for (int r = 0; r < 2; ++r) run();
      HIP_OK(hipEventRecord(e1, c->stream));
      HIP_OK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      return ms;
    };
    float best_ms = measure();
    auto consider = [&](const TuneChoice& ch) {
      apply(ch);
      const float ms = measure();
      if (ms < best_ms) {
        best_ms = ms;
        best = ch;
      }
    };
    // 1. x-pencil tile
    if (!c->tile_force) {
      const TuneChoice base = best;
      for (int T : {4, 8, 16}) {
        if (T == base.T || c->px.n * T > kPencilMax || !(T <= 2 * c->pz.n || T == 4)) continue;
        TuneChoice ch = base;
        ch.T = T;
        consider(ch);
      }
      apply(best);
      // 1b. pass A's own tile (the register kernel keeps tile 4, measured
      // fastest for px = 240 at both 24 and 2-4 orientations per launch)
      const TuneChoice base2 = best;
      const bool reg = reg_len(c, c->px.n);
      for (int T : {4, 8, 16}) {
        if (reg) break;
        if (T == base2.Tc || c->px.n * T > kPencilMax || !(T <= 2 * c->pz.n || T == 4)) continue;
        TuneChoice ch = base2;
        ch.Tc = T;
        consider(ch);
      }
      apply(best);
    }
    // 2. largest butterfly of the score pass (slab plan)
    if (!c->split && !c->score_radix_forced) {
      TuneChoice ch = best;
      ch.score_rmax = best.score_rmax == 10 ? 25 : 10;
      const AxisPlan ay = make_plan(c->py.n, std::min(c->yz_rmax, ch.score_rmax));
      const AxisPlan az = make_plan(c->pz.n, std::min(c->yz_rmax, ch.score_rmax));
      if (ay.nst != c->sy.nst || ay.rad != c->sy.rad || az.nst != c->sz.nst || az.rad != c->sz.rad)
        consider(ch);
      apply(best);
    }
    // 3. padded y / z sizes (slab plan kept; the spectra are rebuilt)
    if (!c->split && !c->pad_forced) {
      const TuneChoice base = best;
      int lds_max = 0;
      HIP_OK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                   c->device));
      for (int axis = 0; axis < 2; ++axis) {
        TuneChoice ch = base;
        int& p = axis == 0 ? ch.py : ch.pz;
        p = next_smooth(p + 1);
        const size_t slab = (size_t)(ch.py + ch.pz + ch.py * ch.pz) * sizeof(float2) +
                            (size_t)c->ny * c->nz * sizeof(int);
        if ((long long)ch.py * ch.pz > kSlabMax || slab + 512 > (size_t)lds_max || p > kMaxAxis)
          continue;
        consider(ch);
        if (c->split) {   // the padded plan left the slab plan: not a candidate
          best_ms = 1e30f;
          apply(base);
          best_ms = measure();
          best = base;
        }
      }
      apply(best);
    }
    c->timing = timing;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  std::lock_guard<std::mutex> g(g_tile_mu);
  g_tile[key] = best;
}

// Parses and uploads the concatenated items of a pack call.
// Items of a pack call, concatenated in items or one pointer per item in
// ptrs: host threads scan them (bounds, sums) and copy them
// into the context's pinned staging buffer, which one asynchronous copy then
// moves to the device (a copy from pageable memory is staged and synchronous).
void load_items(spk_ctx* c, const int32_t* items, const int32_t* const* ptrs,
                const int32_t* dims, int n_items, const int32_t* order,
                std::vector<HostItem>& bases, std::vector<int>& off) {
  if ((!items && !ptrs) || !dims || n_items <= 0) throw SpkError{SPK_EINVAL, "bad pack arguments"};
  size_t tot = 0;
  off.resize(n_items);
  for (int t = 0; t < n_items; ++t) {
    check_dims(dims[3 * t], dims[3 * t + 1], dims[3 * t + 2]);
    off[t] = (int)tot;
    tot += (size_t)dims[3 * t] * dims[3 * t + 1] * dims[3 * t + 2];
  }
  bases.resize(n_items);
  c->items.ensure(std::max<size_t>(tot, 1));
  if (c->stage_ev) HIP_OK(hipEventSynchronize(c->stage_ev));   // the previous copy is done
  if (c->stage_cap < tot) {
    if (c->stage) HIP_OK(hipHostFree(c->stage));
    c->stage = nullptr;
    c->stage_cap = 0;
    HIP_OK(hipHostMalloc((void**)&c->stage, std::max<size_t>(tot, 1) * sizeof(int)));
    c->stage_cap = tot;
  }
  // helper threads: at most 15 per call and 15 across concurrent calls (the
  // multi-instance batch loads from one host thread per instance)
  static std::atomic<int> g_helpers{0};
  const int want = std::min({15, (int)std::thread::hardware_concurrency() - 1, (n_items - 1) / 8});
  int got = 0;
  while (got < want) {
    int cur = g_helpers.load();
    if (cur >= 15) break;
    if (g_helpers.compare_exchange_weak(cur, cur + 1)) ++got;
  }
  const int nth = got + 1;
  std::vector<SpkError> errs(nth, SpkError{SPK_OK, ""});
  auto work = [&](int w) {
    try {
      for (int t = w; t < n_items; t += nth) {
        const size_t n = (size_t)dims[3 * t] * dims[3 * t + 1] * dims[3 * t + 2];
        const int32_t* src = ptrs ? ptrs[t] : items + off[t];
        bases[t] = scan_item(src, dims[3 * t], dims[3 * t + 1], dims[3 * t + 2],
                             n ? c->stage + off[t] : nullptr);
      }
    } catch (const SpkError& e) {
      errs[w] = e;
    }
  };
  std::vector<std::thread> th;
  int w_next = 1;
  try {
    for (; w_next < nth; ++w_next) th.emplace_back(work, w_next);
  } catch (const std::exception&) {
    // no thread: the calling thread does the rest
  }
  work(0);
  for (int w = w_next; w < nth; ++w) work(w);
  for (auto& x : th) x.join();
  g_helpers -= got;
  for (const auto& e : errs)
    if (e.code != SPK_OK) throw e;
  if (tot)
    HIP_OK(hipMemcpyAsync(c->items.p, c->stage, tot * sizeof(int), hipMemcpyHostToDevice,
                          c->stream));
  if (!c->stage_ev) HIP_OK(hipEventCreateWithFlags(&c->stage_ev, hipEventDisableTiming));
  HIP_OK(hipEventRecord(c->stage_ev, c->stream));
  for (int t = 0; t < n_items; ++t) {
    const int it = order ? order[t] : t;
    if (it < 0 || it >= n_items) throw SpkError{SPK_EINVAL, "order index out of range"};
  }
}

}  // namespace

// ---- static scratch of the allocation-free entry point ---------------------
// spk_distance_noalloc keeps the grid (overwritten by the result) and the
// intermediate field of the distance transform in this file-scope array, so
// it calls no device allocator.  Each GPU the library runs on holds its own
// copy (32 MiB + 256 bytes) from module load to process end.  Host code never
// uses the array's name as a device address: spk_ctx_create looks up each
// device's address once, with that device current, and every launch and copy
// takes that address.
constexpr size_t kStaticVoxels = size_t(1) << 22;   // largest grid: 4,194,304 voxels
constexpr size_t kScratchBytes = 2 * kStaticVoxels * sizeof(int) + 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 256;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};
std::mutex g_scratch_call_mu[kMaxDevices];   // one allocation-free call per device at a time

// device address of `scratch` on `device` (which must be current): looked up
// once per device and kept
void* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) throw SpkError{SPK_EUNSUPPORTED, "device index too large"};
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw SpkError{SPK_EDEVICE, "hipGetSymbolAddress returned a null scratch address"};
    g_scratch_addr[device] = p;
  }
  return g_scratch_addr[device];
}
}  // namespace

namespace {
int dev_check(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(SPK_EDEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(SPK_EINVAL, "device index out of range");
  return SPK_OK;
}

template <typename F>
int dev_guarded(int device, F&& f) {
  const int rc = dev_check(device);
  if (rc != SPK_OK) return rc;
  try {
    DeviceGuard g(device);
    return f();
  } catch (const HipError& e) {
    return fail(e.code, e.msg);
  } catch (const SpkError& e) {
    return fail(e.code, e.msg);
  }
}
}  // namespace

#include "spk_vox.hpp"

// =============================================================================
// C-ABI
// =============================================================================
extern "C" {

const char* spk_last_error(void) { return g_err.c_str(); }

int spk_version(int* major, int* minor, int* patch) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  if (patch) *patch = 0;
  return SPK_OK;
}

int spk_ctx_create(int device, spk_ctx** out) {
  if (!out) return fail(SPK_EINVAL, "null out pointer");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(SPK_EDEVICE, "no HIP device available");
  if (device < 0 || device >= n) return fail(SPK_EINVAL, "device index out of range");
  spk_ctx* c = new (std::nothrow) spk_ctx();
  if (!c) return fail(SPK_ENOMEM, "host allocation failed");
  c->device = device;
  int rc = guarded(c, [&]() {
    HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->scratch = scratch_address(device);
    return SPK_OK;
  });
  if (rc != SPK_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return SPK_OK;
}

int spk_ctx_destroy(spk_ctx* c) {
  if (!c) return SPK_OK;
  {
    DeviceGuard g(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& t : c->pending) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
    for (auto e : c->pool) hipEventDestroy(e);
    if (c->mail_h) (void)hipHostFree(c->mail_h);
    if (c->stage_ev) (void)hipEventSynchronize(c->stage_ev);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->pre_h) (void)hipHostFree(c->pre_h);
    if (c->stage_ev) (void)hipEventDestroy(c->stage_ev);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;   // the device buffers free themselves, with the device current
  }
  return SPK_OK;
}

int spk_tray_set(spk_ctx* c, const int32_t* tray, int nx, int ny, int nz, const int32_t* dist,
                 int metric, int sources, uint64_t generation) {
  return guarded(c, [&]() {
    if (!tray) throw SpkError{SPK_EINVAL, "null tray"};
    check_dims(nx, ny, nz);
    if (metric != SPK_METRIC_L1 && metric != SPK_METRIC_EDT)
      throw SpkError{SPK_EINVAL, "metric must be SPK_METRIC_L1 or SPK_METRIC_EDT"};
    if (sources != SPK_SOURCES_EQ1 && sources != SPK_SOURCES_NONZERO)
      throw SpkError{SPK_EINVAL, "sources must be SPK_SOURCES_EQ1 or SPK_SOURCES_NONZERO"};
    c->valid = false;
    const bool reshape = !(c->nx == nx && c->ny == ny && c->nz == nz) || c->tw.p == nullptr;
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->metric = metric; c->sources = sources;
    if (reshape) {
      // sizes and radix the autotuner chose for another shape do not carry over
      if (!c->pad_forced) c->pad_y = c->pad_z = 0;
      if (!c->score_radix_forced) c->score_rmax = 10;
      plan_tray(c);
    }
    const size_t V = (size_t)nx * ny * nz;
    c->tray.ensure(V);
    c->phi.ensure(V);
    c->tmpi.ensure(V);
    c->stats.ensure(1);
    upload(c, c->tray.p, tray, V * sizeof(int));
    {
      int mn = 0;
      for (size_t i = 0; i < V; ++i) mn = std::min(mn, tray[i]);
      c->tray_nonneg = mn >= 0;
    }
    refresh_field_and_spectra(c, dist);
    if (c->tile_force) {
      if ((long long)c->px.n * c->tile_force > kPencilMax)
        throw SpkError{SPK_EUNSUPPORTED, "tile too large for this tray's x axis"};
      c->T = c->tile_force;
      c->Tc = c->tile_force;
    } else if (reshape && c->autotune) {
      c->valid = true;
      autotune_tile(c);
    }
    sync(c);
    if (c->timing) c->flush_timing();
    c->generation = generation;
    c->valid = true;
    return SPK_OK;
  });
}

int spk_tray_get(spk_ctx* c, int32_t* tray_out, int32_t* dist_out) {
  return guarded(c, [&]() {
    if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
    const size_t V = (size_t)c->nx * c->ny * c->nz;
    if (tray_out) download(c, tray_out, c->tray.p, V * sizeof(int));
    if (dist_out) download(c, dist_out, c->phi.p, V * sizeof(int));
    sync(c);
    return SPK_OK;
  });
}

int spk_tray_stats(spk_ctx* c, int64_t* sum, int32_t* mn, int32_t* mx, int32_t* mxw) {
  return guarded(c, [&]() {
    if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
    FieldStats f;
    download(c, &f, c->stats.p, sizeof(f));
    sync(c);
    if (sum) *sum = (int64_t)f.sum;
    if (mn) *mn = f.mn;
    if (mx) *mx = f.mx;
    if (mxw) *mxw = f.mxw;
    return SPK_OK;
  });
}

int spk_tray_info(spk_ctx* c, int* nx, int* ny, int* nz, uint64_t* generation, int* valid) {
  if (!c) return fail(SPK_EINVAL, "null context");
  if (nx) *nx = c->nx;
  if (ny) *ny = c->ny;
  if (nz) *nz = c->nz;
  if (generation) *generation = c->generation;
  if (valid) *valid = c->valid ? 1 : 0;
  return SPK_OK;
}

int spk_tray_clear(spk_ctx* c) {
  if (!c) return fail(SPK_EINVAL, "null context");
  c->valid = false;
  return SPK_OK;
}

int spk_search(spk_ctx* c, const int32_t* const* items, const int32_t* dims, int n, int penalty,
               int interlocking_free, spk_result* best, spk_result* per_orient) {
  return guarded(c, [&]() {
    if (!items || !dims || !best || n <= 0) throw SpkError{SPK_EINVAL, "bad search arguments"};
    std::vector<HostItem> bases;
    std::vector<int> off;
    std::vector<Entry> entries;
    int tot = 0;
    for (int o = 0; o < n; ++o) {
      bases.push_back(scan_item(items[o], dims[3 * o], dims[3 * o + 1], dims[3 * o + 2]));
      off.push_back(tot);
      tot += dims[3 * o] * dims[3 * o + 1] * dims[3 * o + 2];
    }
    for (int o = 0; o < n; ++o) entries.push_back(Entry{bases[o], nullptr, o});
    do_search(c, bases, off, false, entries, penalty, interlocking_free, best, per_orient);
    return SPK_OK;
  });
}

int spk_search_oriented(spk_ctx* c, const int32_t* item, int ix, int iy, int iz,
                        const spk_orient* orients, int n, int penalty, int interlocking_free,
                        spk_result* best, spk_result* per_orient) {
  return guarded(c, [&]() {
    if (!item || !best || n <= 0) throw SpkError{SPK_EINVAL, "bad search arguments"};
    std::vector<HostItem> bases{scan_item(item, ix, iy, iz)};
    std::vector<int> off{0};
    std::vector<Entry> entries;
    for (int o = 0; o < n; ++o) entries.push_back(Entry{bases[0], orients ? &orients[o] : nullptr, 0});
    do_search(c, bases, off, false, entries, penalty, interlocking_free, best, per_orient);
    return SPK_OK;
  });
}

int spk_feasible(spk_ctx* c, const int32_t* item, int ix, int iy, int iz,
                 const spk_orient* orients, int n, int32_t* column, int32_t* full,
                 float* residual) {
  return guarded(c, [&]() {
    if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
    if (!item || n <= 0 || n > 24 || !full) throw SpkError{SPK_EINVAL, "bad feasibility arguments"};
    if (c->split) throw SpkError{SPK_EUNSUPPORTED, "feasibility pass needs the slab plan"};
    const HostItem it = scan_item(item, ix, iy, iz);
    const size_t sz = (size_t)ix * iy * iz;
    c->items.ensure(std::max<size_t>(sz, 1));
    upload(c, c->items.p, item, sz * sizeof(int));
    std::vector<OrientDesc> ods;
    std::vector<int> fit;
    int max_ex = 1, max_exy = 1;
    for (int o = 0; o < n; ++o) {
      const OrientDesc d = make_desc(c, it, 0, orients ? &orients[o] : nullptr);
      const bool fits = it.hi[0] >= 0 && d.ex <= c->nx && d.ey <= c->ny && d.ez <= c->nz &&
                        d.nvx > 0 && d.nvy > 0 && d.nvz > 0;
      if (column) column[o] = 0;
      full[o] = 0;
      if (!fits) continue;
      ods.push_back(d);
      fit.push_back(o);
      max_ex = std::max(max_ex, d.ex);
      max_exy = std::max(max_exy, d.ex * d.ey);
    }
    if (residual) *residual = 0.f;
    const int k = (int)ods.size();
    if (k == 0) return SPK_OK;
    c->od_prop.ensure(24);
    upload(c, c->od_prop.p, ods.data(), ods.size() * sizeof(OrientDesc));
    ensure_search_ws(c, k, max_ex);
    launch(c, "feas_begin", k_feas_begin, dim3(1), dim3(64), 0, c->fstate.p,
           c->co.ready ? c->co.fs.p : (int*)nullptr, k, 2, (int*)nullptr,
           (const FieldStats*)nullptr, (const FieldStats*)nullptr, (PackOut*)nullptr, 0);
    enqueue_coarse(c, c->od_prop.p, k, max_ex, max_exy, ctl_feas_on(c));
    // the full-resolution pass over every fitting orientation (no pruning)
    enqueue_item_spectra(c, c->od_prop.p, k, max_ex, nullptr);
    enqueue_feas(c, c->od_prop.p, k, max_ex, nullptr);
    std::vector<int> cf(kFeasHeader + k, 1);
    if (c->co.ready) {
      const Mailbox& mc = post_and_sync(c, nullptr, nullptr, nullptr, nullptr, c->co.fs.p,
                                        kFeasHeader + k);
      std::memcpy(cf.data(), mc.fs, cf.size() * sizeof(int));
    }
    const Mailbox& mb = post_and_sync(c, nullptr, nullptr, nullptr, nullptr, c->fstate.p,
                                      kFeasHeader + k);
    for (int j = 0; j < k; ++j) {
      if (column) column[fit[j]] = cf[kFeasHeader + j] ? 1 : 0;
      full[fit[j]] = mb.fs[kFeasHeader + j] ? 1 : 0;
    }
    if (residual) std::memcpy(residual, &mb.fs[0], sizeof(float));
    if (c->timing) c->flush_timing();
    return SPK_OK;
  });
}

int spk_place(spk_ctx* c, const int32_t* item, int ix, int iy, int iz, const spk_orient* orient,
              int x, int y, int z, int32_t id) {
  return guarded(c, [&]() {
    if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
    HostItem it = scan_item(item, ix, iy, iz);
    const size_t sz = (size_t)ix * iy * iz;
    c->items.ensure(std::max<size_t>(sz, 1));
    upload(c, c->items.p, item, sz * sizeof(int));
    OrientDesc d = make_desc(c, it, 0, orient);
    do_place(c, d, x, y, z, id, true);
    if (id < 0) c->tray_nonneg = false;
    sync(c);
    if (c->timing) c->flush_timing();
    return SPK_OK;
  });
}

namespace {
int pack_impl(spk_ctx* c, const int32_t* items, const int32_t* const* ptrs, const int32_t* dims,
              int n_items, const int32_t* order, const spk_orient* orients, int n_orient,
              int penalty, int interlocking_free, int32_t id_offset, spk_result* results) {
  return guarded(c, [&]() {
    if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
    if (!results || n_orient <= 0) throw SpkError{SPK_EINVAL, "bad pack arguments"};
    std::vector<HostItem> bases;
    std::vector<int> off;
    load_items(c, items, ptrs, dims, n_items, order, bases, off);
    if (id_offset < 0) c->tray_nonneg = false;   // placed ids may be negative
    int placed = 0;
    if (interlocking_free) {
      // the flood fill iterates to its fixpoint under host control: one
      // synchronous search per item
      for (int t = 0; t < n_items; ++t) {
        const int it = order ? order[t] : t;
        std::vector<Entry> entries;
        for (int o = 0; o < n_orient; ++o)
          entries.push_back(Entry{bases[it], orients ? &orients[o] : nullptr, it});
        do_search(c, bases, off, true, entries, penalty, 1, &results[t], nullptr);
        if (results[t].found) {
          const int o = results[t].orient;
          OrientDesc d = make_desc(c, bases[it], off[it], orients ? &orients[o] : nullptr);
          do_place(c, d, results[t].pos[0], results[t].pos[1], results[t].pos[2],
                   id_offset + placed + 1, true);
          ++placed;
        }
      }
    } else {
      placed = pack_pipelined(c, bases, off, order, orients, n_orient, penalty, id_offset,
                              results);
    }
    sync(c);
    if (c->timing) c->flush_timing();
    c->generation += (uint64_t)placed;
    return SPK_OK;
  });
}
}  // namespace

int spk_pack(spk_ctx* c, const int32_t* items, const int32_t* dims, int n_items,
             const int32_t* order, const spk_orient* orients, int n_orient, int penalty,
             int interlocking_free, int32_t id_offset, spk_result* results) {
  return pack_impl(c, items, nullptr, dims, n_items, order, orients, n_orient, penalty,
                   interlocking_free, id_offset, results);
}

int spk_pack_gather(spk_ctx* c, const int32_t* const* items, const int32_t* dims, int n_items,
                    const int32_t* order, const spk_orient* orients, int n_orient, int penalty,
                    int interlocking_free, int32_t id_offset, spk_result* results) {
  if (!items) return fail(SPK_EINVAL, "bad pack arguments");
  return pack_impl(c, nullptr, items, dims, n_items, order, orients, n_orient, penalty,
                   interlocking_free, id_offset, results);
}

int spk_pack_many(spk_ctx* const* ctxs, int n, const spk_pack_args* args) {
  if (!ctxs || !args || n <= 0) return fail(SPK_EINVAL, "bad pack_many arguments");
  for (int i = 0; i < n; ++i)
    if (!ctxs[i]) return fail(SPK_EINVAL, "null context in pack_many");
  std::vector<int> rc(n, SPK_OK);
  std::vector<std::string> msg(n);
  // A pool of persistent host workers takes instances from a shared queue,
  // largest work first (tray voxels x items), so that the long instances
  // start at once and the short ones fill in behind them.  Each instance
  // runs on its own context and stream; the workers' kernels (and their
  // flood-fill batches) overlap on the device.  Pool size: option
  // "many_workers" of the first context.
  std::vector<int> queue(n);
  for (int i = 0; i < n; ++i) queue[i] = i;
  auto work = [&](int i) {
    const spk_ctx* c = ctxs[i];
    return (double)c->nx * c->ny * c->nz * std::max(1, args[i].n_items);
  };
  std::stable_sort(queue.begin(), queue.end(), [&](int a, int b) { return work(a) > work(b); });
  const int workers = std::max(1, std::min(n, ctxs[0]->many_workers));
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  th.reserve(workers);
  for (int w = 0; w < workers; ++w) {
    th.emplace_back([&]() {
      for (int q = next.fetch_add(1); q < n; q = next.fetch_add(1)) {
        const int i = queue[q];
        const spk_pack_args& a = args[i];
        rc[i] = pack_impl(ctxs[i], a.items, a.item_ptrs, a.dims, a.n_items, a.order, a.orients,
                          a.n_orient, a.penalty, a.interlocking_free, a.id_offset, a.results);
        if (rc[i] != SPK_OK) msg[i] = spk_last_error();
      }
    });
  }
  for (auto& t : th) t.join();
  for (int i = 0; i < n; ++i)
    if (rc[i] != SPK_OK) return fail(rc[i], "instance " + std::to_string(i) + ": " + msg[i]);
  return SPK_OK;
}

int spk_distance(spk_ctx* c, const int32_t* grid, int nx, int ny, int nz, int metric, int sources,
                 int32_t* out) {
  return guarded(c, [&]() {
    if (!grid || !out) throw SpkError{SPK_EINVAL, "null grid"};
    check_dims(nx, ny, nz);
    const size_t V = (size_t)nx * ny * nz;
    DevBuf<int> g, o, t;
    DevBuf<FieldStats> st;
    g.ensure(V); o.ensure(V); t.ensure(V); st.ensure(1);
    upload(c, g.p, grid, V * sizeof(int));
    run_distance(c, g.p, o.p, t.p, st.p, nx, ny, nz, metric, sources);
    download(c, out, o.p, V * sizeof(int));
    sync(c);
    if (c->timing) c->flush_timing();
    g.release(); o.release(); t.release(); st.release();
    return SPK_OK;
  });
}

int spk_distance_noalloc(spk_ctx* c, const int32_t* grid, int nx, int ny, int nz, int metric,
                         int sources, int32_t* out) {
  return guarded(c, [&]() {
    if (!grid || !out) throw SpkError{SPK_EINVAL, "null grid"};
    check_dims(nx, ny, nz);
    const size_t V = (size_t)nx * ny * nz;
    if (V > kStaticVoxels)
      throw SpkError{SPK_EUNSUPPORTED,
                     "spk_distance_noalloc: the grid has more voxels than the static scratch "
                     "holds (4194304)"};
    if (!c->scratch) throw SpkError{SPK_ESTATE, "static scratch address not set up"};
    std::lock_guard<std::mutex> lk(g_scratch_call_mu[c->device]);
    // per-kernel timing creates events on demand: the launches here are not timed
    struct TimingOff {
      spk_ctx* c;
      bool was;
      ~TimingOff() { c->timing = was; }
    } off{c, c->timing};
    c->timing = false;
    int* gbuf = static_cast<int*>(c->scratch);   // grid, then the result
    int* tbuf = gbuf + kStaticVoxels;            // field after the z and y passes
    FieldStats* st = reinterpret_cast<FieldStats*>(tbuf + kStaticVoxels);
    upload(c, gbuf, grid, V * sizeof(int));
    // every value of tbuf and of the result is written before it is read, and
    // the field statistics are reset by the first kernel: nothing to clear
    run_distance(c, gbuf, gbuf, tbuf, st, nx, ny, nz, metric, sources);
    download(c, out, gbuf, V * sizeof(int));
    sync(c);
    return SPK_OK;
  });
}

namespace {
std::vector<Tap> taps_of(const int* b, int nx, int ny, int nz) {
  std::vector<Tap> taps;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const int v = b[((size_t)i * ny + j) * nz + k];
        if (v != 0) taps.push_back(Tap{i, j, k, v});
      }
  return taps;
}

// direct correlation of grid a (device) with taps, int64 result downloaded
void direct_full(spk_ctx* c, const int* d_a, const std::vector<Tap>& taps, int nx, int ny, int nz,
                 int conv, std::vector<long long>& out) {
  const size_t V = (size_t)nx * ny * nz;
  c->taps.ensure(std::max<size_t>(taps.size(), 1));
  if (!taps.empty()) upload(c, c->taps.p, taps.data(), taps.size() * sizeof(Tap));
  c->big.ensure(V);
  const int nb = (int)std::max<size_t>(1, std::min<size_t>(4096, (V + 255) / 256));
  launch(c, "direct_corr", k_direct_corr, dim3(nb), dim3(256), 0, d_a, (const Tap*)c->taps.p,
         (int)taps.size(), c->big.p, nx, ny, nz, conv);
  out.resize(V);
  download(c, out.data(), c->big.p, V * sizeof(long long));
  sync(c);
}
}  // namespace

int spk_correlate(spk_ctx* c, const int32_t* a, const int32_t* b, int nx, int ny, int nz,
                  int convolve, int32_t* out) {
  return guarded(c, [&]() {
    if (!a || !b || !out) throw SpkError{SPK_EINVAL, "null grid"};
    check_dims(nx, ny, nz);
    const size_t V = (size_t)nx * ny * nz;
    DevBuf<int> g;
    g.ensure(V);
    upload(c, g.p, a, V * sizeof(int));
    std::vector<long long> r;
    direct_full(c, g.p, taps_of(b, nx, ny, nz), nx, ny, nz, convolve ? 1 : 0, r);
    for (size_t q = 0; q < V; ++q) out[q] = (int32_t)r[q];
    g.release();
    if (c->timing) c->flush_timing();
    return SPK_OK;
  });
}

int spk_collision_grid(spk_ctx* c, const int32_t* tray, int nx, int ny, int nz, const int32_t* item,
                       int ix, int iy, int iz, int32_t* out) {
  return guarded(c, [&]() {
    if (!tray || !item || !out) throw SpkError{SPK_EINVAL, "null grid"};
    check_dims(nx, ny, nz);
    check_dims(ix, iy, iz);
    const size_t V = (size_t)nx * ny * nz;
    // item padded/truncated to the tray box (bindings.cpp:366-368)
    std::vector<Tap> taps;
    const int kNegInf = -1000000000;   // constants.h:21 (-INF for an empty item)
    int hi[3] = {kNegInf, kNegInf, kNegInf};
    for (int i = 0; i < std::min(ix, nx); ++i)
      for (int j = 0; j < std::min(iy, ny); ++j)
        for (int k = 0; k < std::min(iz, nz); ++k) {
          const int v = item[((size_t)i * iy + j) * iz + k];
          if (v != 0) taps.push_back(Tap{i, j, k, v});
          if (v > 0) { hi[0] = std::max(hi[0], i); hi[1] = std::max(hi[1], j); hi[2] = std::max(hi[2], k); }
        }
    DevBuf<int> g;
    g.ensure(V);
    upload(c, g.p, tray, V * sizeof(int));
    std::vector<long long> r;
    direct_full(c, g.p, taps, nx, ny, nz, 0, r);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const size_t q = ((size_t)i * ny + j) * nz + k;
          int v = (int32_t)r[q];
          const bool ok = ((long long)i + hi[0] <= nx - 1) && ((long long)j + hi[1] <= ny - 1) &&
                          ((long long)k + hi[2] <= nz - 1);
          if (!ok) v = std::max(v, 1);
          out[q] = v;
        }
    g.release();
    if (c->timing) c->flush_timing();
    return SPK_OK;
  });
}

int spk_place_grid(spk_ctx* c, const int32_t* tray, int nx, int ny, int nz, const int32_t* item,
                   int ix, int iy, int iz, int x, int y, int z, int32_t id, int32_t* out) {
  return guarded(c, [&]() {
    if (!tray || !item || !out) throw SpkError{SPK_EINVAL, "null grid"};
    check_dims(nx, ny, nz);
    check_dims(ix, iy, iz);
    const size_t V = (size_t)nx * ny * nz;
    const size_t I = (size_t)ix * iy * iz;
    DevBuf<int> g, it;
    DevBuf<OrientDesc> od;
    DevBuf<int> flag;
    g.ensure(V); it.ensure(I); od.ensure(1); flag.ensure(1);
    upload(c, g.p, tray, V * sizeof(int));
    upload(c, it.p, item, I * sizeof(int));
    OrientDesc d{};
    d.line_off = -1;
    d.bx = ix; d.by = iy; d.bz = iz;
    d.perm[0] = 0; d.perm[1] = 1; d.perm[2] = 2;
    d.ex = ix; d.ey = iy; d.ez = iz;
    upload(c, od.p, &d, sizeof(d));
    HIP_OK(hipMemsetAsync(flag.p, 0, sizeof(int), c->stream));
    const int nb = (int)std::max<size_t>(1, std::min<size_t>(1024, (I + 255) / 256));
    for (int phase = 0; phase < 2; ++phase)
      launch(c, "place", k_place, dim3(nb), dim3(256), 0, g.p, (const int*)it.p,
             (const OrientDesc*)od.p, (const SearchOut*)nullptr, 0, x, y, z, (int)id, nx, ny, nz,
             phase, flag.p, (const int*)nullptr);
    int f = 0;
    download(c, &f, flag.p, sizeof(int));
    download(c, out, g.p, V * sizeof(int));
    sync(c);
    g.release(); it.release(); od.release(); flag.release();
    if (c->timing) c->flush_timing();
    if (f) {
      std::memcpy(out, tray, V * sizeof(int));
      throw SpkError{SPK_EOVERLAP,
                     "(place_in_tray): Item to be placed intersects with contents of tray"};
    }
    return SPK_OK;
  });
}

int spk_interlock_mask(spk_ctx* c, const int32_t* tray, int nx, int ny, int nz, const int32_t* obj,
                       int ox, int oy, int oz, int32_t* out) {
  return guarded(c, [&]() {
    if (!tray || !obj || !out) throw SpkError{SPK_EINVAL, "null grid"};
    check_dims(nx, ny, nz);
    check_dims(ox, oy, oz);
    const size_t V = (size_t)nx * ny * nz;
    int om[3] = {0, 0, 0};   // bindings.cpp:597-609: max over voxels != 0, from 0
    std::vector<Tap> taps;
    for (int i = 0; i < ox; ++i)
      for (int j = 0; j < oy; ++j)
        for (int k = 0; k < oz; ++k) {
          const int v = obj[((size_t)i * oy + j) * oz + k];
          if (v != 0) {
            om[0] = std::max(om[0], i); om[1] = std::max(om[1], j); om[2] = std::max(om[2], k);
            if (i < nx && j < ny && k < nz) taps.push_back(Tap{i, j, k, v});
          }
        }
    DevBuf<int> g;
    g.ensure(V);
    upload(c, g.p, tray, V * sizeof(int));
    std::vector<long long> r;
    direct_full(c, g.p, taps, nx, ny, nz, 0, r);
    std::vector<unsigned char> fr(V);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const size_t q = ((size_t)i * ny + j) * nz + k;
          const bool oob = i + om[0] >= nx || j + om[1] >= ny || k + om[2] >= nz;
          fr[q] = (!oob && r[q] == 0) ? 1 : 0;
        }
    c->ffree.ensure(V);
    upload(c, c->ffree.p, fr.data(), V);
    // exit faces max(0, N-1-objmax) (fft3.cu:276-278)
    const std::vector<int> exits{std::max(0, nx - 1 - om[0]), std::max(0, ny - 1 - om[1]),
                                 std::max(0, nz - 1 - om[2])};
    flood_fill(c, 1, exits, nx, ny, nz);
    std::vector<unsigned char> re(V);
    download(c, re.data(), c->freach.p, V);
    sync(c);
    for (size_t q = 0; q < V; ++q) out[q] = re[q] ? 1 : 0;
    g.release();
    if (c->timing) c->flush_timing();
    return SPK_OK;
  });
}

int spk_bounds(const int32_t* grid, int nx, int ny, int nz, int32_t lo[3], int32_t hi[3]) {
  if (!grid || !lo || !hi || nx < 0 || ny < 0 || nz < 0) return fail(SPK_EINVAL, "bad arguments");
  const int INF = 1000000000;   // constants.h:21
  lo[0] = lo[1] = lo[2] = INF;
  hi[0] = hi[1] = hi[2] = -INF;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k)
        if (grid[((size_t)i * ny + j) * nz + k] > 0) {
          lo[0] = std::min(lo[0], i); lo[1] = std::min(lo[1], j); lo[2] = std::min(lo[2], k);
          hi[0] = std::max(hi[0], i); hi[1] = std::max(hi[1], j); hi[2] = std::max(hi[2], k);
        }
  return SPK_OK;
}

int spk_save_vox(const int32_t* grid, int nx, int ny, int nz, const char* path) {
  if (!grid || !path || nx < 0 || ny < 0 || nz < 0) return fail(SPK_EINVAL, "bad arguments");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(SPK_EINVAL, std::string("cannot open ") + path);
  // slab .vox layout of saveAsVox (packing.cpp:275-305): three 4-byte sizes,
  // voxels with x and z descending, palette byte 123 + v%5 (255 empty, 123
  // for the INSIDE flag 8), then 256 RGB palette entries.
  const int32_t s[3] = {nx, ny, nz};
  std::fwrite(s, 4, 3, f);
  for (int i = nx - 1; i >= 0; --i)
    for (int j = 0; j < ny; ++j)
      for (int k = nz - 1; k >= 0; --k) {
        const unsigned char v = (unsigned char)grid[((size_t)i * ny + j) * nz + k];
        unsigned char pal = v != 0 ? (unsigned char)(123 + (v % 5)) : 255;
        if (v == 8) pal = 123;
        std::fwrite(&pal, 1, 1, f);
      }
  unsigned char palette[256][3];
  std::memset(palette, 0, sizeof(palette));
  const unsigned char colors[5][3] = {{127, 0, 127}, {255, 0, 0}, {0, 255, 0}, {0, 0, 255}, {255, 255, 255}};
  for (int q = 0; q < 5; ++q) std::memcpy(palette[123 + q], colors[q], 3);
  std::fwrite(palette, 3, 256, f);
  std::fclose(f);
  return SPK_OK;
}

int spk_timing_enable(spk_ctx* c, int on) {
  return guarded(c, [&]() {
    if (on && !c->pairs.p) {
      c->pairs.ensure(4);
      HIP_OK(hipMemsetAsync(c->pairs.p, 0, 4 * sizeof(unsigned long long), c->stream));
    }
    c->timing = on != 0;
    return SPK_OK;
  });
}

int spk_timing_get(spk_ctx* c, const char* kernel, double* total_ms, int64_t* launches) {
  return guarded(c, [&]() {
    c->flush_timing();
    auto it = c->acc.find(kernel ? kernel : "");
    if (total_ms) *total_ms = it == c->acc.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == c->acc.end() ? 0 : it->second.second;
    return SPK_OK;
  });
}

int spk_timing_units(spk_ctx* c, const char* kernel, int64_t* units) {
  return guarded(c, [&]() {
    const std::string k = kernel ? kernel : "";
    long long v = 0;
    if (k == "feas_x" || k == "feas_yz" ||
        ((k == "score_yz" || k == "corr_x" || k == "spec_add") && c->pairs.p)) {
      // counted on the device: pairs and orientations the column test
      // excluded never run; incremental spectrum updates that ran
      unsigned long long h[4] = {0, 0, 0, 0};
      if (c->pairs.p) {
        download(c, h, c->pairs.p, sizeof(h));
        sync(c);
      }
      v = (long long)(k == "spec_add" ? h[3] : (k == "score_yz" || k == "corr_x") ? h[2] : h[0]);
    } else {
      auto it = c->units.find(k);
      v = it == c->units.end() ? 0 : it->second;
    }
    if (units) *units = v;
    return SPK_OK;
  });
}

int spk_timing_reset(spk_ctx* c) {
  return guarded(c, [&]() {
    c->flush_timing();
    c->acc.clear();
    c->units.clear();
    if (c->pairs.p) HIP_OK(hipMemsetAsync(c->pairs.p, 0, 4 * sizeof(unsigned long long), c->stream));
    return SPK_OK;
  });
}

int spk_set_option(spk_ctx* c, const char* name, double value) {
  if (!c || !name) return fail(SPK_EINVAL, "null argument");
  if (std::string(name) == "resid_limit") {
    c->resid_limit = (float)value;
    return SPK_OK;
  }
  if (std::string(name) == "plan") {
    const int v = (int)value;
    if (v < 0 || v > 2) return fail(SPK_EINVAL, "plan must be 0 (auto), 1 (slab) or 2 (split)");
    c->plan_mode = v;
    c->tw.release();   // re-plan at the next spk_tray_set
    c->valid = false;
    return SPK_OK;
  }
  if (std::string(name) == "ff_bits") {
    c->ff_bits = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "spec_inc") {
    if (value < 0 || value > 1e6) return fail(SPK_EINVAL, "spec_inc must be in [0, 1e6]");
    c->spec_inc = (int)value;
    return SPK_OK;
  }
  if (std::string(name) == "score2") {
    c->score2 = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "inc_split") {
    const int v = (int)value;
    if (v < 1 || v > kIncXSplitMax) return fail(SPK_EINVAL, "inc_split must be 1 to 16");
    c->inc_split = v;
    return SPK_OK;
  }
  if (std::string(name) == "reg_x") {
    const int v = (int)value;
    if (v < 0 || v > 2) return fail(SPK_EINVAL, "reg_x must be 0, 1 or 2");
    c->reg_x = v;
    return SPK_OK;
  }
  if (std::string(name) == "autotune") {
    c->autotune = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "tile") {
    const int v = (int)value;
    if (v != 0 && v != 4 && v != 8 && v != 16)
      return fail(SPK_EINVAL, "tile must be 0 (tuned), 4, 8 or 16");
    c->tile_force = v;
    c->tw.release();   // re-plan at the next spk_tray_set
    c->valid = false;
    return SPK_OK;
  }
  if (std::string(name) == "score_radix") {
    const int v = (int)value;
    if (v != 10 && v != 25) return fail(SPK_EINVAL, "score_radix must be 10 or 25");
    c->score_rmax = v;
    c->score_radix_forced = true;
    c->tw.release();   // re-plan at the next spk_tray_set
    c->valid = false;
    return SPK_OK;
  }
  if (std::string(name) == "yz_radix" || std::string(name) == "pad_y" ||
      std::string(name) == "pad_z") {
    const int v = (int)value;
    if (std::string(name) == "yz_radix") {
      if (v < 2 || v > kMaxRadix) return fail(SPK_EINVAL, "yz_radix must lie in [2, 25]");
      c->yz_rmax = v;
    } else {
      if (v < 0 || v > kMaxAxis || (v && !smooth(v)))
        return fail(SPK_EINVAL, "padded size must be 0 or a 2^a 3^b 5^c 7^d length <= 1024");
      (std::string(name) == "pad_y" ? c->pad_y : c->pad_z) = v;
      c->pad_forced = c->pad_y != 0 || c->pad_z != 0;
    }
    c->tw.release();   // re-plan at the next spk_tray_set
    c->valid = false;
    return SPK_OK;
  }
  if (std::string(name) == "coarse") {
    c->co.on = value != 0.0;
    c->tw.release();   // re-plan at the next spk_tray_set
    c->valid = false;
    return SPK_OK;
  }
  if (std::string(name) == "corr_chunk") {
    if (value < 0 || value > 24) return fail(SPK_EINVAL, "corr_chunk must lie in [0, 24]");
    c->corr_chunk = (int)value;
    return SPK_OK;
  }
  if (std::string(name) == "screen_all") {
    c->screen_all = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "col_chunks") {
    const int v = (int)value;
    if (v < 1 || v > 64) return fail(SPK_EINVAL, "col_chunks must lie in [1, 64]");
    c->co.col_chunks = v;
    return SPK_OK;
  }
  if (std::string(name) == "fine_group") {
    const int v = (int)value;
    if (v != 4 && v != 8 && v != 16 && v != 32 && v != 64)
      return fail(SPK_EINVAL, "fine_group must be 4, 8, 16, 32 or 64");
    c->co.fine_group = v;
    return SPK_OK;
  }
  if (std::string(name) == "fixed_fft") {
    c->fixed_fft = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "graph") {
    c->graph = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "many_workers") {
    if (value < 1 || value > 256) return fail(SPK_EINVAL, "many_workers must lie in [1, 256]");
    c->many_workers = (int)value;
    return SPK_OK;
  }
  if (std::string(name) == "learn_dev") {
    c->learn_dev = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "prescreen") {
    if (value < 0 || value > 1e6) return fail(SPK_EINVAL, "prescreen must lie in [0, 1e6]");
    c->prescreen = (int)value;
    return SPK_OK;
  }
  if (std::string(name) == "inc_dist") {
    c->inc_dist = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "inc_masks") {
    c->inc_masks = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "dist_x") {
    if (value != 0 && value != 2) return fail(SPK_EINVAL, "dist_x must be 0 or 2");
    c->dist_x = (int)value;
    return SPK_OK;
  }
  if (std::string(name) == "dist_y") {
    if (value != 0 && value != 1 && value != 2)
      return fail(SPK_EINVAL, "dist_y must be 0 (outward), 1 (envelope) or 2 (full scan)");
    c->dist_y = (int)value;
    return SPK_OK;
  }
  if (std::string(name) == "feas_group") {
    if (value < 0 || value > 12) return fail(SPK_EINVAL, "feas_group must lie in [0, 12]");
    c->feas_group = (int)value;
    return SPK_OK;
  }
  if (std::string(name) == "certificate") {
    c->certificate = value != 0.0;
    return SPK_OK;
  }
  if (std::string(name) == "feas") {
    const int v = (int)value;
    if (v < 0 || v > 2)
      return fail(SPK_EINVAL, "feas must be 0 (off), 1 (after an item that placed nothing) or 2");
    c->feas_mode = v;
    return SPK_OK;
  }
  return fail(SPK_EINVAL, std::string("unknown option ") + name);
}

int spk_plan_info(spk_ctx* c, int* px, int* py, int* pz, int* digit_base) {
  if (!c) return fail(SPK_EINVAL, "null context");
  if (px) *px = c->px.n;
  if (py) *py = c->py.n;
  if (pz) *pz = c->pz.n;
  if (digit_base) *digit_base = 0;
  return SPK_OK;
}

int spk_tile_info(spk_ctx* c, int* tile, int* split) {
  if (!c) return fail(SPK_EINVAL, "null context");
  if (tile) *tile = c->T;
  if (split) *split = c->split ? 1 : 0;
  return SPK_OK;
}

// ---- device memory for callers without a HIP runtime of their own -----------


int spk_dev_alloc(int device, int64_t bytes, void** out) {
  if (!out || bytes <= 0) return fail(SPK_EINVAL, "bad allocation arguments");
  *out = nullptr;
  return dev_guarded(device, [&]() {
    HIP_OK(hipMalloc(out, (size_t)bytes));
    return SPK_OK;
  });
}

int spk_dev_free(int device, void* ptr) {
  if (!ptr) return SPK_OK;
  return dev_guarded(device, [&]() {
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipFree(ptr));
    return SPK_OK;
  });
}

int spk_dev_copy(int device, void* dst, const void* src, int64_t bytes, int kind) {
  if (!dst || !src || bytes < 0 || kind < 0 || kind > 2) return fail(SPK_EINVAL, "bad copy arguments");
  return dev_guarded(device, [&]() {
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIP_OK(hipDeviceSynchronize());
    if (bytes) HIP_OK(hipMemcpy(dst, src, (size_t)bytes, k));
    HIP_OK(hipDeviceSynchronize());
    return SPK_OK;
  });
}

int spk_dev_gather(int device, void* dst, const void* src, int64_t row_bytes, int64_t src_rows,
                   const int64_t* idx, int n) {
  if (!dst || !src || !idx || n <= 0 || row_bytes <= 0 || row_bytes % 4)
    return fail(SPK_EINVAL, "bad gather arguments (row size must be a multiple of 4 bytes)");
  for (int i = 0; i < n; ++i)
    if (idx[i] < 0 || idx[i] >= src_rows) return fail(SPK_EINVAL, "gather index out of range");
  return dev_guarded(device, [&]() {
    long long* d_idx = nullptr;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMalloc((void**)&d_idx, (size_t)n * sizeof(long long)));
    const hipError_t e1 = hipMemcpy(d_idx, idx, (size_t)n * sizeof(long long), hipMemcpyHostToDevice);
    hipError_t e2 = hipSuccess;
    if (e1 == hipSuccess) {
      const bool wide = row_bytes % 16 == 0 && (uintptr_t)dst % 16 == 0 &&
                        (uintptr_t)src % 16 == 0;
      const long long words = row_bytes / (wide ? 16 : 4);
      const dim3 grid((unsigned)std::min<long long>((words + 255) / 256, 1024), n);
      if (wide)
        hipLaunchKernelGGL(spknn::k_gather_rows<uint4>, grid, dim3(256), 0, nullptr, (uint4*)dst,
                           (const uint4*)src, words, (const long long*)d_idx, n);
      else
        hipLaunchKernelGGL(spknn::k_gather_rows<unsigned>, grid, dim3(256), 0, nullptr,
                           (unsigned*)dst, (const unsigned*)src, words, (const long long*)d_idx, n);
      e2 = hipGetLastError();
      if (e2 == hipSuccess) e2 = hipDeviceSynchronize();
    }
    hipFree(d_idx);
    HIP_OK(e1);
    HIP_OK(e2);
    return SPK_OK;
  });
}

// ---- learned orientation proposer -----------------------------------------

int spk_nn_create(int device, int max_batch, spk_nn** out) {
  if (!out) return fail(SPK_EINVAL, "null out pointer");
  *out = nullptr;
  if (max_batch <= 0 || max_batch > 4096) return fail(SPK_EINVAL, "max_batch must be in [1, 4096]");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(SPK_EDEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(SPK_EINVAL, "device index out of range");
  spk_nn* nn = new spk_nn;
  nn->device = device;
  const int rc = nn_guarded(nn, [&]() {
    HIP_OK(hipStreamCreate(&nn->stream));
    nn_build(nn, max_batch);
    return SPK_OK;
  });
  if (rc != SPK_OK) {
    const std::string msg = g_err;
    spk_nn_destroy(nn);
    return fail(rc, msg);
  }
  *out = nn;
  return SPK_OK;
}

int spk_nn_destroy(spk_nn* nn) {
  if (!nn) return SPK_OK;
  {
    DeviceGuard g(nn->device);
    if (nn->stream) hipStreamSynchronize(nn->stream);
    if (nn->stream) hipStreamDestroy(nn->stream);
    delete nn;   // the device buffers free themselves, with the device current
  }
  return SPK_OK;
}

int spk_nn_num_params(spk_nn* nn, int64_t* n) {
  if (!nn || !n) return fail(SPK_EINVAL, "null argument");
  *n = (int64_t)nn->n_params;
  return SPK_OK;
}

int spk_nn_set_params(spk_nn* nn, const float* params, int64_t n) {
  return nn_guarded(nn, [&]() {
    if (!params || n != (int64_t)nn->n_params)
      throw SpkError{SPK_EINVAL, "parameter count does not match the network"};
    HIP_OK(hipMemcpyAsync(nn->params.p, params, n * sizeof(float), hipMemcpyHostToDevice,
                          nn->stream));
    HIP_OK(hipMemsetAsync(nn->m1.p, 0, n * sizeof(float), nn->stream));
    HIP_OK(hipMemsetAsync(nn->m2.p, 0, n * sizeof(float), nn->stream));
    HIP_OK(hipStreamSynchronize(nn->stream));
    nn->step = 0;
    nn->dirty = true;
    return SPK_OK;
  });
}

int spk_nn_get_params(spk_nn* nn, float* params, int64_t n) {
  return nn_guarded(nn, [&]() {
    if (!params || n != (int64_t)nn->n_params)
      throw SpkError{SPK_EINVAL, "parameter count does not match the network"};
    HIP_OK(hipMemcpyAsync(params, nn->params.p, n * sizeof(float), hipMemcpyDeviceToHost,
                          nn->stream));
    HIP_OK(hipStreamSynchronize(nn->stream));
    return SPK_OK;
  });
}

int spk_nn_forward(spk_nn* nn, const uint16_t* item_feat, const uint16_t* tray_feat, int batch,
                   float* logits) {
  return nn_guarded(nn, [&]() {
    nn_check_batch(nn, batch);
    if (!item_feat || !tray_feat || !logits) throw SpkError{SPK_EINVAL, "null argument"};
    nn_forward(nn, nn->stream, item_feat, tray_feat, batch);
    HIP_OK(hipMemcpyAsync(logits, nn->logits.p, (size_t)batch * kNnLogits * sizeof(float),
                          hipMemcpyDeviceToHost, nn->stream));
    HIP_OK(hipStreamSynchronize(nn->stream));
    return SPK_OK;
  });
}

int spk_nn_gradient(spk_nn* nn, const uint16_t* item_feat, const uint16_t* tray_feat,
                    const float* target, const int32_t* mask, int batch, float* grads,
                    float* loss) {
  return nn_guarded(nn, [&]() {
    nn_check_batch(nn, batch);
    if (!item_feat || !tray_feat || !target || !mask || !grads)
      throw SpkError{SPK_EINVAL, "null argument"};
    nn_forward(nn, nn->stream, item_feat, tray_feat, batch);
    nn_backward(nn, nn->stream, item_feat, tray_feat, target, mask, batch);
    HIP_OK(hipMemcpyAsync(grads, nn->grads.p, nn->n_params * sizeof(float),
                          hipMemcpyDeviceToHost, nn->stream));
    const float l = nn_loss_sum(nn, nn->stream, batch);
    if (loss) *loss = l;
    return SPK_OK;
  });
}

int spk_nn_train_step(spk_nn* nn, const uint16_t* item_feat, const uint16_t* tray_feat,
                      const float* target, const int32_t* mask, int batch, float lr,
                      float* loss) {
  return nn_guarded(nn, [&]() {
    nn_check_batch(nn, batch);
    if (!item_feat || !tray_feat || !target || !mask)
      throw SpkError{SPK_EINVAL, "null argument"};
    if (!(lr >= 0.f)) throw SpkError{SPK_EINVAL, "learning rate must be >= 0"};
    hipStream_t s = nn->stream;
    nn_forward(nn, s, item_feat, tray_feat, batch);
    nn_backward(nn, s, item_feat, tray_feat, target, mask, batch);
    nn->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)nn->step), c2 = 1.f - std::pow(b2, (float)nn->step);
    const long long n = (long long)nn->n_params;
    nn_launch(s, spknn::k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nn->params.p,
              (const float*)nn->grads.p, nn->m1.p, nn->m2.p, n, lr, b1, b2, 1e-8f, c1, c2);
    nn->dirty = true;
    const float l = nn_loss_sum(nn, s, batch);
    if (loss) *loss = l;
    return SPK_OK;
  });
}

int spk_nn_item_features(spk_nn* nn, const int32_t* item, int ix, int iy, int iz,
                         const spk_orient* orient, uint16_t* out) {
  return nn_guarded(nn, [&]() {
    check_dims(ix, iy, iz);
    if (!item || !out) throw SpkError{SPK_EINVAL, "null argument"};
    spknn::ItemMap mp{};
    mp.bx = ix; mp.by = iy; mp.bz = iz;
    bool seen[3] = {false, false, false};
    for (int a = 0; a < 3; ++a) {
      mp.perm[a] = orient ? orient->perm[a] : a;
      mp.flip[a] = orient ? (orient->flip[a] ? 1 : 0) : 0;
      if (mp.perm[a] < 0 || mp.perm[a] > 2 || seen[mp.perm[a]])
        throw SpkError{SPK_EINVAL, "orientation perm must be a permutation of 0,1,2"};
      seen[mp.perm[a]] = true;
    }
    DevBuf<int> d;
    DevBuf<spknn::ItemMap> dm;
    const size_t n = (size_t)ix * iy * iz;
    d.ensure(n);
    dm.ensure(1);
    HIP_OK(hipMemcpyAsync(d.p, item, n * sizeof(int), hipMemcpyHostToDevice, nn->stream));
    HIP_OK(hipMemcpyAsync(dm.p, &mp, sizeof(mp), hipMemcpyHostToDevice, nn->stream));
    nn_item_features(nn->stream, d.p, dm.p, 1, out);
    HIP_OK(hipStreamSynchronize(nn->stream));
    d.release();
    dm.release();
    return SPK_OK;
  });
}

int spk_tray_features(spk_ctx* c, uint16_t* out) {
  return guarded(c, [&]() {
    if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
    if (!out) throw SpkError{SPK_EINVAL, "null argument"};
    c->heights.ensure((size_t)c->nx * c->ny);
    nn_tray_features(c->stream, c->tray.p, c->nx, c->ny, c->nz, c->heights.p, out);
    sync(c);
    return SPK_OK;
  });
}

int spk_pack_records(spk_ctx* c, const spk_pack_args* a, uint16_t* tray_feat, double* scores) {
  return guarded(c, [&]() {
    if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
    if (!a || !a->results || a->n_orient <= 0 || !tray_feat || !scores)
      throw SpkError{SPK_EINVAL, "bad pack arguments"};
    if (a->interlocking_free)
      throw SpkError{SPK_EINVAL, "training records come from the plain (not interlocking-free) loop"};
    std::vector<HostItem> bases;
    std::vector<int> off;
    load_items(c, a->items, a->item_ptrs, a->dims, a->n_items, a->order, bases, off);
    const int placed = pack_pipelined(c, bases, off, a->order, a->orients, a->n_orient, a->penalty,
                                      a->id_offset, a->results, scores, tray_feat);
    sync(c);
    if (c->timing) c->flush_timing();
    c->generation += (uint64_t)placed;
    return SPK_OK;
  });
}

int spk_pack_learned(spk_ctx* c, spk_nn* nn, const spk_pack_args* a, int topk, float min_conf,
                     int32_t* proposals, int32_t* counts) {
  return guarded(c, [&]() {
    if (!c->valid) throw SpkError{SPK_ESTATE, "no tray set in this context"};
    if (!nn) throw SpkError{SPK_EINVAL, "null proposer"};
    if (nn->device != c->device)
      throw SpkError{SPK_EINVAL, "proposer and context are on different devices"};
    if (!a || !a->results || a->n_orient <= 0 || a->n_orient > kNnLogits)
      throw SpkError{SPK_EINVAL, "bad pack arguments (1 to 24 orientations)"};
    if (topk < 1 || topk > kNnLogits) throw SpkError{SPK_EINVAL, "topk must be in [1, 24]"};
    if (a->interlocking_free)
      throw SpkError{SPK_EINVAL, "learned proposals run the plain (not interlocking-free) loop"};
    std::vector<HostItem> bases;
    std::vector<int> off;
    load_items(c, a->items, a->item_ptrs, a->dims, a->n_items, a->order, bases, off);
    if (a->id_offset < 0) c->tray_nonneg = false;   // placed ids may be negative
    // the device-driven loop covers the default configuration: the column
    // test screens every item, no confidence threshold, no certificate
    const bool dev = c->learn_dev && c->co.ready && c->screen_all && c->feas_mode != 0 &&
                     !c->split && !(min_conf > 0.f) && !c->certificate;
    const int placed =
        dev ? pack_learned_dev(c, nn, bases, off, a->order, a->orients, a->n_orient, a->penalty,
                               a->id_offset, a->results, topk, proposals, counts)
            : pack_learned(c, nn, bases, off, a->order, a->orients, a->n_orient, a->penalty,
                           a->id_offset, a->results, topk, min_conf, proposals, counts);
    sync(c);
    if (c->timing) c->flush_timing();
    c->generation += (uint64_t)placed;
    return SPK_OK;
  });
}

// ---------------------------------------------------------------- voxelizer

int spk_voxelize_mesh(int device, const float* positions, int64_t n_vertices,
                      const int32_t* triangles, int64_t n_triangles, int resolution,
                      int32_t* dims, int32_t* out, int64_t capacity) {
  if (!dims) return fail(SPK_EINVAL, "null dims");
  if (!triangles && n_triangles) return fail(SPK_EINVAL, "null triangles");
  return dev_guarded(device, [&]() -> int {
    try {
      spkvox::voxelize(positions, n_vertices, triangles, n_triangles, resolution, dims, out, capacity);
    } catch (const spkvox::VoxError& e) {
      return fail(e.code, e.msg);
    }
    return SPK_OK;
  });
}

int spk_voxelize_stl(int device, const char* path, int resolution, int32_t* dims, int32_t* out,
                     int64_t capacity) {
  if (!path || !dims) return fail(SPK_EINVAL, "null path or dims");
  return dev_guarded(device, [&]() -> int {
    try {
      const std::vector<float> pos = spkvox::load_stl(path);
      const int64_t nv = (int64_t)pos.size() / 3;
      spkvox::voxelize(pos.data(), nv, nullptr, nv / 3, resolution, dims, out, capacity);
    } catch (const spkvox::VoxError& e) {
      return fail(e.code, e.msg);
    }
    return SPK_OK;
  });
}

}  // extern "C"
