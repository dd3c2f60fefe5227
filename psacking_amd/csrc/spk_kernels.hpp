// spk_kernels.hpp -- HIP kernels of the spectral-packing hot path (gfx950).
//
// Correlation scheme (replaces the reference's padded C2C cuFFT pipeline,
// fft3.cu:1248-1294 and :1299-1373):
//   tray side, once per tray state, on an FFT grid P >= N (circular
//   correlation is exact at every in-bounds offset because q + y <= N-1):
//     B   = tray != 0                      (collision field)
//     Phi'= Phi - mu = b*h + l + b/2       (two-digit split, base b)
//     G1  = FFT3(B + i*h)                  (full complex spectrum)
//     G2  = FFT3(l)  kz in [0, Pz/2]       (real field -> half spectrum)
//   item side, per orientation: I~ = FFT_yz(item) for the item's x-slices.
//   pass A  (x pencils): I^ = FFT_x(I~); Y1 = G1*conj(I^)/P, Y2 = G2*conj(I^)/P;
//           inverse FFT_x; store only the x rows that are valid offsets.
//   pass B  (yz slabs):  inverse FFT_yz(Y2) -> l-sum, FFT_yz(Y1) -> (c, h-sum);
//           round each to an integer, p = b*h + l + (b/2+mu)*sum(I),
//           score = double(p) + penalty[k], workgroup argmin (score, index).
// Each rounded part has an error far below 0.5 (its field has a small
// range); the maximum distance to the nearest integer is reported so the
// host can route a search to the exact direct-sum path if it ever grows.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <climits>

#include "spk_fft.hpp"

namespace spk {

constexpr int kSlabThreads = 1024;   // yz-slab kernels
constexpr int kSlabMaxEl = 16;       // elements per thread -> slab <= 16384 complex
constexpr int kSlabMax = kSlabThreads * kSlabMaxEl;
constexpr int kPencilThreads = 512;  // x-pencil kernels
constexpr int kPencilMaxEl = 8;      // px * T <= 4096
constexpr int kPencilMax = kPencilThreads * kPencilMaxEl;
constexpr int kMaxAxis = 1024;       // largest FFT length along one axis
constexpr int kDistThreads = 1024;

// ---------------------------------------------------------------------------
// small shared structs (host and device)
// ---------------------------------------------------------------------------

struct FieldStats {          // reduction of the distance field
  unsigned long long sum;    // two's complement int64 sum
  int mn, mx;
  int mxw;                   // max over voxels of min(field, distance to the outside
                             // of the grid): the largest ball of free voxels inside
  int pad;
};

struct FieldParams {         // derived on device identically by every block
  int mu;                    // integer centre
  int lb;                    // log2 of the digit base
  int ok;                    // 0: range too large for the split
};

__device__ __forceinline__ FieldParams field_params(const FieldStats* st, long long vol) {
  FieldParams f;
  const long long s = (long long)st->sum;
  f.mu = (int)llrint((double)s / (double)vol);
  const long long r1 = (long long)st->mx - f.mu, r2 = (long long)f.mu - st->mn;
  const long long R = r1 > r2 ? r1 : r2;
  f.ok = 1;
  if (R <= 512) f.lb = 4;
  else if (R <= 1024) f.lb = 5;
  else if (R <= 4096) f.lb = 6;
  else { f.lb = 6; f.ok = 0; }
  return f;
}

struct OrientDesc {          // one searched orientation
  int item_off;              // offset of the base item in the item buffer
  int bx, by, bz;            // base item dims
  int perm[3], flip[3];      // output axis a reads base axis perm[a]
  int ex, ey, ez;            // oriented dims
  int hx, hy, hz;            // oriented max index of voxels > 0 (-1 if empty)
  int nvx, nvy, nvz;         // valid offsets per axis: N - 1 - h  (+1), clamped
  int line_off;              // the base item's line masks (k_item_lines), -1 if none
  int binary;                // every voxel value is 0 or 1
  long long isum;            // sum of item values
};

struct WgBest {              // per (orientation, x-slice) argmin record
  double score;
  int idx;                   // linear tray index, INT_MAX when nothing feasible
  int p_lo;                  // low 32 bits of the proximity at idx
  int p_hi;
  float resid;               // max |v - rint(v)| over the slice
  int unsupported;
};

struct SearchOut {           // written by k_select, read by verify/place
  int found;
  int orient;
  int idx;
  int pad;
  double score;
  long long prox;
  float resid;
  int unsupported;
};

// Control words of the device-driven greedy loop (spk_pack).  Kernels take
// a nullable pointer: outside that loop they always run.
// CTL_SKIP_SEARCH: the feasibility pass found no feasible orientation, so the
// item's scored search is skipped; CTL_FEAS: the feasibility pass runs for
// the current item; CTL_PRUNE: the coarse pass ran and the full-resolution
// pass may skip orientation pairs it proved infeasible; CTL_COARSE_DONE: the
// coarse pass proved every orientation infeasible (no full-resolution pass);
// CTL_SPEC: the placement left the distance field as it was, so the tray
// spectrum G1 is updated by the placed voxels alone (k_spec_add) and the
// tray's forward transforms are skipped; CTL_SPEC_N: such updates since the
// last full transform.
enum {
  CTL_HALT = 0, CTL_HALT_ITEM = 1, CTL_SKIP_REFRESH = 2, CTL_PLACED = 3, CTL_SKIP_SEARCH = 4,
  CTL_FEAS = 5, CTL_PRUNE = 6, CTL_COARSE_DONE = 7, CTL_SPEC = 8, CTL_SPEC_N = 9, CTL_WORDS = 10
};
// search kernels stop once a search has been handed back to the host, and
// for an item the feasibility pass has already decided
// A device-coherent read of a word that other workgroups of the running
// kernel update with atomics (no stale copy from a cache of another XCD).
__device__ __forceinline__ int load_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool search_off(const int* ctl) {
  return ctl && (ctl[CTL_HALT] | ctl[CTL_SKIP_SEARCH]);
}
// feasibility kernels run when the loop enabled them for the current item
__device__ __forceinline__ bool feas_off(const int* ctl) {
  return ctl && (ctl[CTL_HALT] || !ctl[CTL_FEAS]);
}
// refresh kernels also stop when the last item placed nothing
__device__ __forceinline__ bool refresh_off(const int* ctl) {
  return ctl && (ctl[CTL_HALT] | ctl[CTL_SKIP_REFRESH]);
}
// the tray's forward transforms also stop when k_spec_add updates G1
__device__ __forceinline__ bool spectra_off(const int* ctl) {
  return ctl && (ctl[CTL_HALT] | ctl[CTL_SKIP_REFRESH] | ctl[CTL_SPEC]);
}
// full-resolution work made unnecessary by the coarse pass
__device__ __forceinline__ bool coarse_done(const int* ctl) {
  return ctl && ctl[CTL_COARSE_DONE];
}

struct PackOut {             // one item of the greedy loop, as the device decided it
  SearchOut so;
  long long collision;
};

__device__ __forceinline__ int oriented_index(const OrientDesc& d, int w0, int w1, int w2) {
  int w[3] = {w0, w1, w2};
  int ext[3] = {d.ex, d.ey, d.ez};
  int u[3];
  int bd[3] = {d.bx, d.by, d.bz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int ax = d.perm[a];
    u[ax] = d.flip[a] ? (bd[ax] - 1 - w[a]) : w[a];
    (void)ext;
  }
  return (u[0] * d.by + u[1]) * d.bz + u[2];
}

// Workgroups of a launch go to the 8 XCDs in turn (workgroup b to XCD b mod 8)
// and each XCD has its own L2.  xcd_rank() renumbers them so that each XCD
// takes one contiguous run of ranks: the pencil kernels give neighbouring
// ranks neighbouring rows, whose 128-byte lines they share (row pitches such
// as 100 complex values = 800 bytes split lines between tiles), so a shared
// line is fetched into one L2 instead of two.
__device__ __forceinline__ int xcd_rank() {
  constexpr int kXcd = 8;
  const int nb = gridDim.x * gridDim.y * gridDim.z;
  const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int q = nb / kXcd, r = nb - q * kXcd;
  const int xcd = b % kXcd, slot = b / kXcd;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
}

__device__ __forceinline__ void load_table(float2* dst, const float2* src, int n, int tid, int nt) {
  for (int q = tid; q < n; q += nt) dst[q] = src[q];
}

// Global -> LDS copy of n contiguous complex values with every load issued
// before the first LDS store (one round trip of HBM latency per thread, not
// one per element).  Element e lands at lds[(e / w) * ld + e % w].
template <int NT, int MAXEL>
__device__ __forceinline__ void stage_rows(float2* lds, const float2* __restrict__ src, int n,
                                           int w, int ld, int tid) {
  float2 v[MAXEL];
  const float inv_w = 1.0f / (float)w;
#pragma unroll
  for (int s = 0; s < MAXEL; ++s) {
    const int e = tid + s * NT;
    if (e < n) v[s] = src[e];
  }
#pragma unroll
  for (int s = 0; s < MAXEL; ++s) {
    const int e = tid + s * NT;
    if (e < n) {
      const int r = fdiv(e, w, inv_w);
      lds[r * ld + (e - r * w)] = v[s];
    }
  }
}

// lexicographic (score, idx) minimum
__device__ __forceinline__ bool better(double s1, int i1, double s2, int i2) {
  return (s1 < s2) || (s1 == s2 && i1 < i2);
}

// ---------------------------------------------------------------------------
// distance field: separable exact min-plus transforms
//   L1 : g(k) = min_k' f(k') + |k-k'|      (reference sweeps, fft3.cu:887-942)
//   EDT: g(k) = min_k' f(k') + (k-k')^2    (squared distances, exact ints)
// The first (z) pass sees only sources (f = 0) and non-sources (f = big): two
// sweeps per row give the nearest source.  The y and x passes search outward
// from each voxel and stop at the first offset d whose cost alone reaches the
// best value found, which no farther k' can beat because f >= 0; the result is
// the same minimum as the full scan at a cost proportional to the distance.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int dcost(int d, int edt) { return edt ? d * d : (d < 0 ? -d : d); }

// min over k' of f(k') + cost(k - k') along a strided line of n values,
// f >= floor >= 0 (floor: a lower bound of the line, e.g. its minimum)
__device__ __forceinline__ int outward_min(const int* line, int stride, int n, int k, int edt,
                                           int floor = 0) {
  int best = line[k * stride];
  for (int d = 1;; ++d) {
    const int c = dcost(d, edt);
    const bool lo = k - d >= 0, hi = k + d < n;
    if (c + floor >= best || (!lo && !hi)) break;
    if (lo) {
      const int v = line[(k - d) * stride] + c;
      best = v < best ? v : best;
    }
    if (hi) {
      const int v = line[(k + d) * stride] + c;
      best = v < best ? v : best;
    }
  }
  return best;
}

// One line of the exact min-plus transform g(q) = min_r f(r) + cost(q - r),
// f >= 0 read at f[r * fs], r < n; put(q, g) receives every output in order.
// Squared Euclidean cost: the lower envelope of the parabolas f(r) + (q-r)^2
// (Felzenszwalb & Huttenlocher, "Distance transforms of sampled functions",
// 2012), breakpoints compared exactly in 64-bit integers; v (stride vs)
// holds the envelope's apexes.  L1 cost: a forward and a backward sweep
// (g = min(f, neighbour + 1)), both through v as scratch.  O(n) per line.
template <typename Put>
__device__ __forceinline__ void line_transform(const int* f, int fs, int n, int edt, int2* v,
                                               int vs, Put put) {
  if (!edt) {
    int g = INT_MAX - 1;
    for (int q = 0; q < n; ++q) {
      const int a = f[q * fs];
      g = a < g + 1 ? a : g + 1;
      v[q * vs].x = g;
    }
    g = INT_MAX - 1;
    for (int q = n - 1; q >= 0; --q) {
      const int a = v[q * vs].x;
      g = a < g + 1 ? a : g + 1;
      v[q * vs].x = g;
    }
    for (int q = 0; q < n; ++q) put(q, v[q * vs].x);
    return;
  }
  // apexes with their F = f(r) + r^2 (< 2^31 for the fields built here)
  int k = 0;
  v[0] = make_int2(0, f[0]);
  for (int q = 1; q < n; ++q) {
    const int Fq = f[q * fs] + q * q;
    int2 b = v[k * vs];
    while (k > 0) {
      // apex b leaves the envelope when the breakpoint of (b, q) is not past
      // the breakpoint of (a, b): (Fq-Fb)/(2(q-b)) <= (Fb-Fa)/(2(b-a))
      const int2 a = v[(k - 1) * vs];
      if ((long long)(Fq - b.y) * (b.x - a.x) <= (long long)(b.y - a.y) * (q - b.x)) {
        --k;
        b = a;
      } else {
        break;
      }
    }
    ++k;
    v[k * vs] = make_int2(q, Fq);
  }
  int j = 0;
  int2 a = v[0];
  int2 b = k > 0 ? v[vs] : a;
  for (int q = 0; q < n; ++q) {
    // advance while the breakpoint of (a, b) lies before q
    while (j < k && (long long)(b.y - a.y) < 2LL * q * (b.x - a.x)) {
      ++j;
      a = b;
      if (j < k) b = v[(j + 1) * vs];
    }
    put(q, a.y - a.x * a.x + (q - a.x) * (q - a.x));
  }
}

// one x-slice: init (sources) + z pass + y pass, in LDS; writes tmp int grid.
// Dynamic LDS: ny*nz ints of slab, then (envelope == 1) ny*nz int2 of
// envelope scratch for the y pass.
template <int EDT, int YMODE>
__global__ void __launch_bounds__(kDistThreads)
k_dist_zy(const int* __restrict__ tray, int* __restrict__ out, int nx, int ny, int nz,
          int sources, int big, FieldStats* __restrict__ st, const int* __restrict__ ctl) {
  constexpr int edt = EDT;
  if (refresh_off(ctl)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) {   // reduced into by the following k_dist_x
    st->sum = 0ULL;
    st->mn = INT_MAX;
    st->mx = INT_MIN;
    st->mxw = INT_MIN;
  }
  extern __shared__ int sh[];
  const int i = blockIdx.x;
  const int tid = threadIdx.x;
  const int S = ny * nz;
  const int* src = tray + (size_t)i * S;
  {
    int v[32];
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const int e = tid + t * kDistThreads;
      v[t] = e < S ? src[e] : 0;
    }
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const int e = tid + t * kDistThreads;
      if (e < S) {
        const bool s = sources ? (v[t] != 0) : (v[t] == 1);
        sh[e] = s ? 0 : 1;
      }
    }
  }
  __syncthreads();
  // z pass: each thread owns a run of consecutive elements; block-wide scans
  // give every element the index of the last source at or before it (max scan)
  // and of the first source at or after it (min scan).  A source in another z
  // row is farther than the row boundary and counts as none.
  {
    constexpr int kW = kDistThreads / 64;
    __shared__ int wlast[kW], wfirst[kW];
    const int per = (S + kDistThreads - 1) / kDistThreads;   // <= 32
    const int base = tid * per;
    int last = -1, first = INT_MAX;
    for (int t = 0; t < per; ++t) {
      const int e = base + t;
      if (e < S && sh[e] == 0) {
        last = e;
        first = first == INT_MAX ? e : first;
      }
    }
    // inclusive wave scans: max of last from the left, min of first from the right
    const int lane = tid & 63, wave = tid >> 6;
    int incl_last = last, incl_first = first;
    for (int d = 1; d < 64; d <<= 1) {
      const int a = __shfl_up(incl_last, d, 64);
      const int b = __shfl_down(incl_first, d, 64);
      if (lane >= d) incl_last = a > incl_last ? a : incl_last;
      if (lane + d < 64) incl_first = b < incl_first ? b : incl_first;
    }
    if (lane == 63) wlast[wave] = incl_last;
    if (lane == 0) wfirst[wave] = incl_first;
    __syncthreads();
    int before = __shfl_up(incl_last, 1, 64);     // exclusive within the wave
    if (lane == 0) before = -1;
    int after = __shfl_down(incl_first, 1, 64);
    if (lane == 63) after = INT_MAX;
    for (int w = 0; w < wave; ++w) before = wlast[w] > before ? wlast[w] : before;
    for (int w = wave + 1; w < kW; ++w) after = wfirst[w] < after ? wfirst[w] : after;
    int dz[32];
    int prev = before;
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const int e = base + t;
      if (t < per && e < S) {
        if (sh[e] == 0) prev = e;
        const int r0 = (e / nz) * nz;
        dz[t] = prev >= r0 ? e - prev : INT_MAX;
      }
    }
    int nxt = after;
#pragma unroll
    for (int t = 31; t >= 0; --t) {
      const int e = base + t;
      if (t < per && e < S) {
        const int src = sh[e] == 0;
        if (src) nxt = e;
        const int r1 = (e / nz) * nz + nz - 1;
        const int db = nxt <= r1 ? nxt - e : INT_MAX;
        const int d = dz[t] < db ? dz[t] : db;
        dz[t] = d == INT_MAX ? big : dcost(d, edt);
        if (dz[t] > big) dz[t] = big;
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const int e = base + t;
      if (t < per && e < S) sh[e] = dz[t];
    }
  }
  __syncthreads();
  int* dst = out + (size_t)i * S;
  if constexpr (YMODE == 1) {
    // y pass: one thread per z column, envelope scratch behind the slab
    int2* scratch = reinterpret_cast<int2*>(sh + S);
    for (int k = tid; k < nz; k += kDistThreads)
      line_transform(sh + k, nz, ny, edt, scratch + k, nz,
                     [&](int j, int g) { dst[j * nz + k] = g; });
  } else {
    // y pass per voxel: a full scan (YMODE 2) or an outward search (YMODE 0)
    // bounded by the column minimum (no offset can beat it)
    __shared__ int cmin[1024];
    if constexpr (YMODE == 0) {
      for (int k = tid; k < nz; k += kDistThreads) cmin[k] = INT_MAX;
      __syncthreads();
      for (int e = tid; e < S; e += kDistThreads) atomicMin(&cmin[e % nz], sh[e]);
      __syncthreads();
    }
    // one element per thread and step, stored at once (keeping the results
    // in registers until the end spills)
    for (int e = tid; e < S; e += kDistThreads) {
      const int j = e / nz, k = e - j * nz;
      int best;
      if constexpr (YMODE == 2) {
        best = sh[e];
        for (int jj = 0; jj < ny; ++jj) {
          const int c = sh[jj * nz + k] + dcost(j - jj, EDT);
          best = c < best ? c : best;
        }
      } else {
        best = outward_min(sh + k, nz, ny, j, EDT, cmin[k]);
      }
      dst[e] = best;
    }
  }
}

// x pass for a tile of TK consecutive k at fixed j; finalize and reduce stats.
// XMODE 0: outward search per voxel (cost ~ distance), 2: full scan (cost nx)
template <int TK, int EDT, int XMODE>
__global__ void __launch_bounds__(kDistThreads)
k_dist_x(const int* __restrict__ in, int* __restrict__ phi, FieldStats* __restrict__ st, int nx,
         int ny, int nz, int fill, const int* __restrict__ ctl) {
  constexpr int edt = EDT;
  if (refresh_off(ctl)) return;
  extern __shared__ int sh[];
  const int j = blockIdx.y;
  const int k0 = blockIdx.x * TK;
  const int tid = threadIdx.x;
  const int tk = min(TK, nz - k0);
  const int E = nx * TK;
  {
    int v[16];   // E = nx * TK <= 1024 * 16
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int e = tid + s * kDistThreads;
      const int x = e / TK, t = e - x * TK;
      v[s] = (e < E && t < tk) ? in[((size_t)x * ny + j) * nz + k0 + t] : 0;
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int e = tid + s * kDistThreads;
      if (e < E) sh[e] = v[s];
    }
  }
  __shared__ int lmin[TK];   // minimum of each line: no offset can beat it
  if (tid < TK) lmin[tid] = INT_MAX;
  __syncthreads();
  for (int e = tid; e < E; e += kDistThreads) {
    const int t = e % TK;
    if (t < tk) atomicMin(&lmin[t], sh[e]);
  }
  __syncthreads();
  long long lsum = 0;
  int lmn = INT_MAX, lmx = INT_MIN, lmw = INT_MIN;
  const int wy = min(j + 1, ny - j);   // distance to the outside across a y face
  for (int e = tid; e < E; e += kDistThreads) {
    const int x = e / TK, t = e - x * TK;
    if (t >= tk) continue;
    int best;
    if constexpr (XMODE == 2) {
      best = sh[e];
      for (int xx = 0; xx < nx; ++xx) {
        const int c = sh[xx * TK + t] + dcost(x - xx, EDT);
        best = c < best ? c : best;
      }
    } else {
      best = outward_min(sh + t, TK, nx, x, EDT, lmin[t]);
    }
    int v;
    if (edt) {
      v = (best >= fill * fill) ? fill : (int)rint(sqrt((double)best));
      v = v > fill ? fill : v;
    } else {
      v = best > fill ? fill : best;
    }
    phi[((size_t)x * ny + j) * nz + k0 + t] = v;
    lsum += v;
    lmn = v < lmn ? v : lmn;
    lmx = v > lmx ? v : lmx;
    // distance to the nearest voxel outside the grid, in any metric: the
    // nearest such voxel is straight across the closest face
    const int k = k0 + t;
    const int w = min(min(wy, min(k + 1, nz - k)), min(x + 1, nx - x));
    const int vw = v < w ? v : w;
    lmw = vw > lmw ? vw : lmw;
  }
  // block reduction of stats
  __shared__ long long rs[kDistThreads / 64];
  __shared__ int rmn[kDistThreads / 64], rmx[kDistThreads / 64], rmw[kDistThreads / 64];
  for (int o = 32; o > 0; o >>= 1) {
    lsum += __shfl_xor(lsum, o);
    const int a = __shfl_xor(lmn, o), b = __shfl_xor(lmx, o), c = __shfl_xor(lmw, o);
    lmn = a < lmn ? a : lmn;
    lmx = b > lmx ? b : lmx;
    lmw = c > lmw ? c : lmw;
  }
  __syncthreads();
  if ((tid & 63) == 0) {
    rs[tid >> 6] = lsum;
    rmn[tid >> 6] = lmn;
    rmx[tid >> 6] = lmx;
    rmw[tid >> 6] = lmw;
  }
  __syncthreads();
  if (tid == 0) {
    long long s = 0;
    int a = INT_MAX, b = INT_MIN, c = INT_MIN;
    for (int w = 0; w < kDistThreads / 64; ++w) {
      s += rs[w];
      a = rmn[w] < a ? rmn[w] : a;
      b = rmx[w] > b ? rmx[w] : b;
      c = rmw[w] > c ? rmw[w] : c;
    }
    atomicAdd(&st->sum, (unsigned long long)s);
    atomicMin(&st->mn, a);
    atomicMax(&st->mx, b);
    atomicMax(&st->mxw, c);
  }
}

// ---------------------------------------------------------------------------
// Incremental distance field after a placement.
//
// A placement only adds sources: with sources "nonzero" every voxel the item
// wrote, with sources "eq1" the voxels of id 1 (so only the placement of id 1
// adds any).  The field is a minimum over sources, and rounding
// (rint(sqrt()) of the squared Euclidean distance) and the fill cap are
// monotone, so  Phi_new = min(Phi_old, D)  with D the transform of the new
// sources alone.  The new sources lie in the item's occupied box
// [qx, qx+hx] x [qy, qy+hy] x [qz, qz+hz], which makes D separable with short
// candidate ranges:
//   k_inc_zy  one block per slab x in [qx, qx+hx]: z pass on the item's
//             columns, y pass over its hy+1 rows -> g[x-qx][y][z], all (y,z);
//   k_inc_x   per (y, z) line: the lower envelope of the hx+1 candidates
//             along x (L1: prefix/suffix minima), evaluated at every x,
//             minimum with Phi_old, and the field statistics of the result.
// Both take the placement as k_place does (the device's search outcome, or
// an explicit orientation and offset) and skip when it adds no source.
// ---------------------------------------------------------------------------
constexpr int kIncMaxEx = 128;       // largest x extent of an item updated this way
constexpr int kIncLines = 64;        // (y, z) lines per k_inc_x block (one per lane)
constexpr int kIncSplit = 4;         // k_inc_zy blocks per slab (shares of its y pass)
constexpr int kIncXSplitMax = 16;    // k_inc_x blocks per line group (shares of x), at most

struct PlacedBox {
  int qx, qy, qz;                    // offset of the oriented item
  int hx, hy, hz;                    // its occupied box, relative to the offset
  int id;                            // id written by the placement
};

// false when the placement adds no source (nothing placed, empty item, or a
// placement that is not id 1 under sources "eq1")
__device__ __forceinline__ bool placed_box(const SearchOut* so, const OrientDesc* od, int orient,
                                           int qx, int qy, int qz, int id, int ny, int nz,
                                           int sources, const int* ctl, PlacedBox& b) {
  if (ctl) id += ctl[CTL_PLACED];
  OrientDesc d;
  if (so) {
    const SearchOut s = *so;
    if (!s.found) return false;
    d = od[s.orient];
    qz = s.idx % nz; qy = (s.idx / nz) % ny; qx = s.idx / (ny * nz);
  } else {
    d = od[orient];
  }
  if (d.hx < 0 || d.hy < 0 || d.hz < 0) return false;
  if (!sources && id != 1) return false;
  b.qx = qx; b.qy = qy; b.qz = qz;
  b.hx = d.hx; b.hy = d.hy; b.hz = d.hz;
  b.id = id;
  return true;
}

// rint(sqrt(d2)) for 0 <= d2 < 2^31 without double arithmetic: r >= 1 is the
// rounded root exactly when r^2 - r + 1 <= d2 <= r^2 + r (sqrt(d2) lies in
// [r - 1/2, r + 1/2), ties are impossible for an integer d2), and r = 0 when
// d2 = 0; the float root is within one of r and the two tests correct it.
__device__ __forceinline__ int round_sqrt(int d2) {
  long long r = (long long)rintf(sqrtf((float)d2));
  if (r * r + r < d2) ++r;
  if (r > 0 && r * r - r + 1 > d2) --r;
  return (int)r;
}

// grid (kIncMaxEx, kIncSplit), kDistThreads threads: block (r, s) computes
// slab qx + r's z pass and share s of its y pass.  Dynamic LDS: the item
// box's source flags (hy+1)*(hz+1) ints, then g1 (hy+1)*nz ints.
// g: [hx+1][ny][nz] ints, >= big where a slab has no new source.
template <int EDT>
__global__ void __launch_bounds__(kDistThreads)
k_inc_zy(const int* __restrict__ tray, int* __restrict__ g, const SearchOut* __restrict__ so,
         const OrientDesc* __restrict__ od, int orient, int qx, int qy, int qz, int id, int nx,
         int ny, int nz, int sources, int big, const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  PlacedBox b;
  if (!placed_box(so, od, orient, qx, qy, qz, id, ny, nz, sources, ctl, b)) return;
  const int r = blockIdx.x;
  if (r > b.hx || b.qx + r >= nx) return;
  extern __shared__ int sm[];
  const int x = b.qx + r, tid = threadIdx.x;
  const int ry = min(b.hy + 1, ny - b.qy), rz = min(b.hz + 1, nz - b.qz);
  int* src = sm;                    // [ry][rz]: 1 where a new source lies
  int* g1 = sm + ry * rz;           // [ry][nz]: cost along z to the column's nearest source
  const int* slab = tray + (size_t)x * ny * nz;
  for (int e = tid; e < ry * rz; e += kDistThreads) {
    const int jy = e / rz, t = e - jy * rz;
    src[e] = slab[(size_t)(b.qy + jy) * nz + b.qz + t] == b.id ? 1 : 0;
  }
  __syncthreads();
  // z pass over the item's columns
  for (int e = tid; e < ry * nz; e += kDistThreads) {
    const int jy = e / nz, k = e - jy * nz;
    const int* cs = src + jy * rz;
    int best = big;
    for (int t = 0; t < rz; ++t) {
      const int c = dcost(k - (b.qz + t), EDT);
      best = (cs[t] && c < best) ? c : best;
    }
    g1[e] = best;
  }
  __syncthreads();
  // y pass over the item's rows, this block's share of the slab
  int* out = g + (size_t)r * ny * nz;
  const int share = (ny * nz + kIncSplit - 1) / kIncSplit;
  const int e1 = min(ny * nz, ((int)blockIdx.y + 1) * share);
  for (int e = blockIdx.y * share + tid; e < e1; e += kDistThreads) {
    const int j = e / nz, k = e - j * nz;
    int best = big;
    for (int jy = 0; jy < ry; ++jy) {
      const int v = g1[jy * nz + k];
      const int c = v + dcost(j - (b.qy + jy), EDT);
      best = (v < big && c < best) ? c : best;
    }
    out[e] = best;
  }
}

// grid (ceil(nz / kIncLines), ny, S), S <= kIncXSplitMax, 256 threads: lane l
// of every wave owns line (y = blockIdx.y, z = z0 + l); wave w evaluates
// x = x0 + w, x0 + w + 4, ... in the block's share of x (blockIdx.z of S).
// Every output is the minimum over the hx+1 candidate slabs (staged in LDS,
// dynamic: (hx+1) * kIncLines ints); an x whose distance to the candidate
// range alone reaches the old value cannot lower it and is skipped.
// part[block] receives the block's field statistics (k_inc_stats reduces).
template <int EDT>
__global__ void __launch_bounds__(256)
k_inc_x(const int* __restrict__ g, int* __restrict__ phi, FieldStats* __restrict__ part,
        const SearchOut* __restrict__ so, const OrientDesc* __restrict__ od, int orient, int qx,
        int qy, int qz, int id, int nx, int ny, int nz, int sources, int big, int fill,
        const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  PlacedBox b;
  if (!placed_box(so, od, orient, qx, qy, qz, id, ny, nz, sources, ctl, b)) return;
  extern __shared__ int cand[];     // [r][line]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int y = blockIdx.y, z0 = blockIdx.x * kIncLines, z = z0 + lane;
  const int m = min(b.hx + 1, nx - b.qx);      // candidate slabs
  for (int e = tid; e < m * kIncLines; e += 256) {
    const int rr = e / kIncLines, l = e - rr * kIncLines;
    cand[e] = (z0 + l < nz) ? g[((size_t)rr * ny + y) * nz + z0 + l] : big;
  }
  __syncthreads();
  const int xl = b.qx, xh = b.qx + m - 1;
  const int share = (nx + (int)gridDim.z - 1) / (int)gridDim.z;
  const int xa = blockIdx.z * share, xb = min(nx, xa + share);
  const int bid = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  // the smallest candidate of the line (big past nz): every candidate at
  // distance >= d from x costs at least fmin + dcost(d)
  int fmin = big;
  for (int rr = 0; rr < m; ++rr) fmin = min(fmin, cand[rr * kIncLines + lane]);
  long long lsum = 0;
  int lmn = INT_MAX, lmx = INT_MIN, lmw = INT_MIN;
  if (z < nz) {
    const int wy = min(y + 1, ny - y), wz = min(z + 1, nz - z);
#pragma unroll 4
    for (int x = xa + wave; x < xb; x += 4) {
      int* pv = phi + ((size_t)x * ny + y) * nz + z;
      const int old = *pv;
      int v = old;
      const int dx = x < xl ? xl - x : (x > xh ? x - xh : 0);
      // the new value can only be below old when the cheapest conceivable
      // candidate (fmin at the nearest slab) is: for EDT the rounded root of
      // s is >= old exactly when s >= old^2 - old + 1 (old >= 1)
      const long long lb = fmin >= big ? (long long)big : (long long)fmin + dcost(dx, EDT);
      const bool may = EDT ? (old > 0 && lb < (long long)old * old - old + 1) : lb < old;
      if (may) {
        int best = big;
        for (int rr = 0; rr < m; ++rr) {
          const int f = cand[rr * kIncLines + lane];
          const int c = f + dcost(x - (b.qx + rr), EDT);
          best = (f < big && c < best) ? c : best;
        }
        int d;
        if (EDT) d = best >= fill * fill ? fill : round_sqrt(best);
        else d = best;
        d = d > fill ? fill : d;
        if (d < old) {
          v = d;
          *pv = v;
        }
      }
      lsum += v;
      lmn = v < lmn ? v : lmn;
      lmx = v > lmx ? v : lmx;
      const int w = min(min(wy, wz), min(x + 1, nx - x));
      const int vw = v < w ? v : w;
      lmw = vw > lmw ? vw : lmw;
    }
  }
  __shared__ long long rs[4];
  __shared__ int rmn[4], rmx[4], rmw[4];
  for (int o = 32; o > 0; o >>= 1) {
    lsum += __shfl_xor(lsum, o);
    const int a = __shfl_xor(lmn, o), c = __shfl_xor(lmx, o), e = __shfl_xor(lmw, o);
    lmn = a < lmn ? a : lmn;
    lmx = c > lmx ? c : lmx;
    lmw = e > lmw ? e : lmw;
  }
  if (lane == 0) {
    rs[wave] = lsum; rmn[wave] = lmn; rmx[wave] = lmx; rmw[wave] = lmw;
  }
  __syncthreads();
  if (tid == 0) {
    FieldStats f;
    f.sum = 0ULL;
    f.mn = INT_MAX; f.mx = INT_MIN; f.mxw = INT_MIN; f.pad = 0;
    for (int w = 0; w < 4; ++w) {
      f.sum += (unsigned long long)rs[w];
      f.mn = rmn[w] < f.mn ? rmn[w] : f.mn;
      f.mx = rmx[w] > f.mx ? rmx[w] : f.mx;
      f.mxw = rmw[w] > f.mxw ? rmw[w] : f.mxw;
    }
    part[bid] = f;
  }
}

// the field statistics from k_inc_x's n block partials (one block of 1024)
__global__ void __launch_bounds__(1024)
k_inc_stats(const FieldStats* __restrict__ part, int n, FieldStats* __restrict__ st,
            const SearchOut* __restrict__ so, const OrientDesc* __restrict__ od, int orient,
            int qx, int qy, int qz, int id, int ny, int nz, int sources,
            const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  PlacedBox b;
  if (!placed_box(so, od, orient, qx, qy, qz, id, ny, nz, sources, ctl, b)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  unsigned long long sum = 0ULL;
  int mn = INT_MAX, mx = INT_MIN, mw = INT_MIN;
  for (int i = tid; i < n; i += 1024) {
    const FieldStats f = part[i];
    sum += f.sum;
    mn = f.mn < mn ? f.mn : mn;
    mx = f.mx > mx ? f.mx : mx;
    mw = f.mxw > mw ? f.mxw : mw;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o);
    const int a = __shfl_xor(mn, o), c = __shfl_xor(mx, o), e = __shfl_xor(mw, o);
    mn = a < mn ? a : mn;
    mx = c > mx ? c : mx;
    mw = e > mw ? e : mw;
  }
  __shared__ unsigned long long rs[16];
  __shared__ int rmn[16], rmx[16], rmw[16];
  if (lane == 0) {
    rs[wave] = sum; rmn[wave] = mn; rmx[wave] = mx; rmw[wave] = mw;
  }
  __syncthreads();
  if (tid == 0) {
    FieldStats f;
    f.sum = 0ULL;
    f.mn = INT_MAX; f.mx = INT_MIN; f.mxw = INT_MIN; f.pad = 0;
    for (int w = 0; w < 16; ++w) {
      f.sum += rs[w];
      f.mn = rmn[w] < f.mn ? rmn[w] : f.mn;
      f.mx = rmx[w] > f.mx ? rmx[w] : f.mx;
      f.mxw = rmw[w] > f.mxw ? rmw[w] : f.mxw;
    }
    *st = f;
  }
}

// stats of a caller-provided distance field
__global__ void __launch_bounds__(256)
k_field_stats(const int* __restrict__ phi, long long n, FieldStats* __restrict__ st) {
  long long lsum = 0;
  int lmn = INT_MAX, lmx = INT_MIN;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const int v = phi[e];
    lsum += v;
    lmn = v < lmn ? v : lmn;
    lmx = v > lmx ? v : lmx;
  }
  for (int o = 32; o > 0; o >>= 1) {
    lsum += __shfl_xor(lsum, o);
    const int a = __shfl_xor(lmn, o), b = __shfl_xor(lmx, o);
    lmn = a < lmn ? a : lmn;
    lmx = b > lmx ? b : lmx;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&st->sum, (unsigned long long)lsum);
    atomicMin(&st->mn, lmn);
    atomicMax(&st->mx, lmx);
    atomicMax(&st->mxw, INT_MAX);   // no bound for a field of unknown meaning
  }
}

// ---------------------------------------------------------------------------
// tray spectra
// ---------------------------------------------------------------------------

// grid (nx, 2): field 0 -> slab of B + i*h, full kz; field 1 -> slab of l,
// only kz < hz kept.  Forward 2D FFT (z rows < ny, then y columns).
// SY, SZ: compile-time plans of the padded y and z lengths, or DynPlan.
template <class SY, class SZ>
__global__ void __launch_bounds__(kSlabThreads)
k_tray_slab_fwd(const int* __restrict__ tray, const int* __restrict__ phi,
                const FieldStats* __restrict__ st, float2* __restrict__ t1,
                float2* __restrict__ t2, int nx, int ny, int nz, AxisPlan py_, AxisPlan pz_,
                const float2* __restrict__ tw_y, const float2* __restrict__ tw_z,
                const int* __restrict__ ctl) {
  constexpr int FY = SY::N, FZ = SZ::N, FH = FZ / 2 + 1;
  if (spectra_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int py = FY > 0 ? FY : py_.n, pz = FZ > 0 ? FZ : pz_.n, hz = pz / 2 + 1;
  float2* twy = smem;
  float2* twz = twy + py;
  float2* buf = twz + pz;
  const int x = blockIdx.x, field = blockIdx.y, tid = threadIdx.x;
  load_table(twy, tw_y, py, tid, kSlabThreads);
  load_table(twz, tw_z, pz, tid, kSlabThreads);
  const FieldParams fp = field_params(st, (long long)nx * ny * nz);
  const int b = 1 << fp.lb, half = b >> 1;
  const size_t so = (size_t)x * ny * nz;
  {
    // all global loads of the slab first, then the LDS stores
    int dv[kSlabMaxEl], tv[kSlabMaxEl];
    const float inv_pz = 1.0f / (float)pz;
#pragma unroll
    for (int s = 0; s < kSlabMaxEl; ++s) {
      const int e = tid + s * kSlabThreads;
      const int j = div_by<FZ>(e, pz, inv_pz), k = e - j * pz;
      dv[s] = 0;
      tv[s] = 0;
      if (e < py * pz && j < ny && k < nz) {
        dv[s] = phi[so + j * nz + k];
        if (field == 0) tv[s] = tray[so + j * nz + k];
      }
    }
#pragma unroll
    for (int s = 0; s < kSlabMaxEl; ++s) {
      const int e = tid + s * kSlabThreads;
      const int j = div_by<FZ>(e, pz, inv_pz), k = e - j * pz;
      if (e >= py * pz) continue;
      float2 v = make_float2(0.f, 0.f);
      if (j < ny && k < nz) {
        const int d = dv[s] - fp.mu;
        const int h = d >> fp.lb;            // floor division
        if (field == 0) {
          v.x = tv[s] != 0 ? 1.f : 0.f;
          v.y = (float)h;
        } else {
          v.x = (float)(d - (h << fp.lb) - half);
        }
      }
      buf[e] = v;
    }
  }
  __syncthreads();
  fft_axis<true, kSlabThreads, kSlabMaxEl, kMaxRadix, SZ, 1, FZ, true, 0, FY>(
      buf, pz_, 1, ny, pz, true, twz, tid);
  if (field == 0)
    fft_axis<true, kSlabThreads, kSlabMaxEl, kMaxRadix, SY, FZ, 1, false, FZ, 0>(
        buf, py_, pz, pz, 1, false, twy, tid);
  else
    fft_axis<true, kSlabThreads, kSlabMaxEl, kMaxRadix, SY, FZ, 1, false, FH, 0>(
        buf, py_, pz, hz, 1, false, twy, tid);
  if (field == 0) {
    float2* dst = t1 + (size_t)x * py * pz;
    for (int e = tid; e < py * pz; e += kSlabThreads) dst[e] = buf[e];
  } else {
    float2* dst = t2 + (size_t)x * py * hz;
    for (int e = tid; e < py * hz; e += kSlabThreads) {
      const int j = e / hz, k = e - j * hz;
      dst[e] = buf[j * pz + k];
    }
  }
}

// grid (ceil(ncols/T), py, 2): forward FFT along x of the slab spectra
// (rows x >= nx are zero), writes G[kx][ky][kz].  SX, TT: compile-time x plan
// and tile, or DynPlan and 0.
template <class SX, int TT>
__global__ void __launch_bounds__(kPencilThreads)
k_tray_x_fwd(const float2* __restrict__ t1, const float2* __restrict__ t2,
             float2* __restrict__ g1, float2* __restrict__ g2, int nx, AxisPlan px_, int py,
             int pz, int T_, const float2* __restrict__ tw_x, const int* __restrict__ ctl) {
  if (spectra_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int px = SX::N > 0 ? SX::N : px_.n, hz = pz / 2 + 1;
  const int T = TT > 0 ? TT : T_;
  float2* twx = smem;
  float2* buf = twx + px;
  // rank = (ky * tiles + tile) * 2 + field (xcd_rank): both fields of a tile,
  // the tiles of one row and neighbouring rows run on one XCD, and the
  // field-1 tiles past hz (which exit at once) are spread over all XCDs
  const int rk = xcd_rank(), ntile = gridDim.x;
  const int field = rk & 1, rest = rk >> 1;
  const int ky = rest / ntile, tid = threadIdx.x;
  const int ncol = field == 0 ? pz : hz;
  const int kz0 = (rest - ky * ntile) * T;
  if (kz0 >= ncol) return;
  load_table(twx, tw_x, px, tid, kPencilThreads);
  const float2* src = field == 0 ? t1 : t2;
  float2* dst = field == 0 ? g1 : g2;
  const size_t rs = (size_t)py * ncol;   // x stride
  {
    float2 v[kPencilMaxEl];
#pragma unroll
    for (int s = 0; s < kPencilMaxEl; ++s) {
      const int e = tid + s * kPencilThreads;
      const int x = e / T, t = e - x * T;
      const int kz = kz0 + t;
      v[s] = make_float2(0.f, 0.f);
      if (e < px * T && x < nx && kz < ncol) v[s] = src[x * rs + (size_t)ky * ncol + kz];
    }
#pragma unroll
    for (int s = 0; s < kPencilMaxEl; ++s) {
      const int e = tid + s * kPencilThreads;
      if (e < px * T) buf[e] = v[s];
    }
  }
  __syncthreads();
  fft_axis<true, kPencilThreads, kPencilMaxEl, kPencilRadix, SX, TT, 1, false, TT, 0>(
      buf, px_, T, T, 1, false, twx, tid);
  for (int e = tid; e < px * T; e += kPencilThreads) {
    const int x = e / T, t = e - x * T;
    const int kz = kz0 + t;
    if (kz < ncol) dst[x * rs + (size_t)ky * ncol + kz] = buf[e];
  }
}

// ---------------------------------------------------------------------------
// item spectra: grid (max_ex, n_orient).  I~[o][x][ky][kz] = FFT_yz(item_o[x])
// ---------------------------------------------------------------------------
// cf (nullable): column-test flags of the list; with CTL_PRUNE set, excluded
// orientations need no spectra (neither pass reads them)
// SY, SZ: compile-time plans of the padded y and z lengths, or DynPlan.
// Incremental tray spectrum.  After a placement that leaves the distance
// field (and so h, l and their statistics) as it was, B = (tray != 0) gains
// the placed voxels, the oriented item I (values 0 or 1) shifted by the
// offset q, and G1 = FFT(B + i h) gains
//     FFT(I shifted by q)[k] = e^{-2 pi i (kx qx / px + ky qy / py + kz qz / pz)} FFT(I)[k].
// FFT(I) comes from the winner orientation's yz spectra (ispec, written by
// the item's search) transformed along x here.  Grid (z tiles of T, py),
// kPencilThreads threads; runs only when k_decide set CTL_SPEC.
__global__ void __launch_bounds__(kPencilThreads)
k_spec_add(float2* __restrict__ g1, const float2* __restrict__ ispec,
           const OrientDesc* __restrict__ od, const SearchOut* __restrict__ so, int max_ex,
           int ny, int nz, AxisPlan px_, int py, int pz, int T, const float2* __restrict__ tw,
           unsigned long long* __restrict__ applied, const int* __restrict__ ctl) {
  if (!ctl || refresh_off(ctl) || !ctl[CTL_SPEC]) return;
  if (applied && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) atomicAdd(applied, 1ULL);
  extern __shared__ float2 smem[];
  const int px = px_.n, tid = threadIdx.x;
  float2* twx = smem;
  float2* buf = twx + px;
  const SearchOut s = *so;
  const OrientDesc d = od[s.orient];
  const int qz = s.idx % nz, qy = (s.idx / nz) % ny, qx = s.idx / (ny * nz);
  const int ky = blockIdx.y, kz0 = blockIdx.x * T;
  load_table(twx, tw, px, tid, kPencilThreads);
  const float2* is = ispec + (size_t)s.orient * max_ex * py * pz;
  for (int e = tid; e < px * T; e += kPencilThreads) {
    const int x = e / T, kz = kz0 + (e - x * T);
    buf[e] = (x < d.ex && kz < pz) ? is[((size_t)x * py + ky) * pz + kz] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<true, kPencilThreads, kPencilMaxEl, kPencilRadix>(buf, px_, T, T, 1, false, twx, tid);
  // phases from the tables exp(-2 pi i q / n) of the three axes (k and q
  // are below 1024 per axis: the products fit 32 bits)
  const float2 wy = tw[px + (ky * qy) % py];
  for (int e = tid; e < px * T; e += kPencilThreads) {
    const int kx = e / T, kz = kz0 + (e - kx * T);
    if (kz >= pz) continue;
    const float2 w = cmul(cmul(twx[(kx * qx) % px], wy), tw[px + py + (kz * qz) % pz]);
    float2* g = g1 + ((size_t)kx * py + ky) * pz + kz;
    *g = cadd(*g, cmul(buf[e], w));
  }
}

template <class SY, class SZ>
__global__ void __launch_bounds__(kSlabThreads)
k_item_slab_fwd(const int* __restrict__ items, const OrientDesc* __restrict__ od,
                float2* __restrict__ ispec, int max_ex, AxisPlan py_, AxisPlan pz_,
                const float2* __restrict__ tw_y, const float2* __restrict__ tw_z,
                const int* __restrict__ cf, const int* __restrict__ ctl) {
  constexpr int FY = SY::N, FZ = SZ::N;
  if (search_off(ctl) || coarse_done(ctl)) return;
  if (cf && ctl && ctl[CTL_PRUNE] && !cf[blockIdx.y]) return;
  extern __shared__ float2 smem[];
  const int py = FY > 0 ? FY : py_.n, pz = FZ > 0 ? FZ : pz_.n;
  const int x = blockIdx.x, o = blockIdx.y, tid = threadIdx.x;
  const OrientDesc d = od[o];
  if (x >= d.ex) return;
  float2* twy = smem;
  float2* twz = twy + py;
  float2* buf = twz + pz;
  load_table(twy, tw_y, py, tid, kSlabThreads);
  load_table(twz, tw_z, pz, tid, kSlabThreads);
  const int* base = items + d.item_off;
  {
    int iv[kSlabMaxEl];
    const float inv_pz = 1.0f / (float)pz;
#pragma unroll
    for (int s = 0; s < kSlabMaxEl; ++s) {
      const int e = tid + s * kSlabThreads;
      const int j = div_by<FZ>(e, pz, inv_pz), k = e - j * pz;
      iv[s] = (e < py * pz && j < d.ey && k < d.ez) ? base[oriented_index(d, x, j, k)] : 0;
    }
#pragma unroll
    for (int s = 0; s < kSlabMaxEl; ++s) {
      const int e = tid + s * kSlabThreads;
      if (e < py * pz) buf[e] = make_float2((float)iv[s], 0.f);
    }
  }
  __syncthreads();
  fft_axis<true, kSlabThreads, kSlabMaxEl, kMaxRadix, SZ, 1, FZ, true, 0, FY>(
      buf, pz_, 1, d.ey, pz, true, twz, tid);
  fft_axis<true, kSlabThreads, kSlabMaxEl, kMaxRadix, SY, FZ, 1, false, FZ, 0>(
      buf, py_, pz, pz, 1, false, twy, tid);
  float2* dst = ispec + ((size_t)o * max_ex + x) * py * pz;
  for (int e = tid; e < py * pz; e += kSlabThreads) dst[e] = buf[e];
}

// ---------------------------------------------------------------------------
// pass A: grid (ceil(pz/T), py, n_chunks).  x pencils of T consecutive kz.
// ---------------------------------------------------------------------------
// LDS: twiddles, then two pencil buffers (the G1 and G2 products of one
// orientation, both formed right after the item's forward FFT).
__global__ void __launch_bounds__(kPencilThreads)
k_corr_x(const float2* __restrict__ g1, const float2* __restrict__ g2,
         const float2* __restrict__ ispec, const OrientDesc* __restrict__ od,
         float2* __restrict__ y1, float2* __restrict__ y2, int n_orient, int chunk, int max_ex,
         int nx, AxisPlan px_, int py, int pz, int T, float scale, const float2* __restrict__ tw_x,
         const int* __restrict__ cf, const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  // orientations the column test excluded have no feasible offset: skipped
  // here, and k_score_yz records them as having none
  const bool prune = cf && ctl && ctl[CTL_PRUNE];
  extern __shared__ float2 smem[];
  const int px = px_.n, hz = pz / 2 + 1;
  float2* twx = smem;
  float2* buf = twx + px;
  float2* buf2 = buf + px * T;
  // rank = (ky * tiles + tile) * chunks + chunk: the chunks of one tile, the
  // tiles of one row and neighbouring rows run on one XCD (xcd_rank)
  const int rk = xcd_rank(), nch = gridDim.z, ntile = gridDim.x;
  const int rest = rk / nch;
  const int ky = rest / ntile, tid = threadIdx.x;
  const int kz0 = (rest - ky * ntile) * T;
  const int o0 = (rk - rest * nch) * chunk;
  const int o1 = min(n_orient, o0 + chunk);
  load_table(twx, tw_x, px, tid, kPencilThreads);
  const bool has2 = kz0 < hz;
  float2 G1[kPencilMaxEl], G2[kPencilMaxEl];
#pragma unroll
  for (int s = 0; s < kPencilMaxEl; ++s) {
    const int e = tid + s * kPencilThreads;
    const int x = e / T, t = e - x * T;
    const int kz = kz0 + t;
    G1[s] = make_float2(0.f, 0.f);
    G2[s] = make_float2(0.f, 0.f);
    if (e < px * T && kz < pz) {
      G1[s] = g1[((size_t)x * py + ky) * pz + kz];
      if (kz < hz) G2[s] = g2[((size_t)x * py + ky) * hz + kz];
    }
  }
  for (int o = o0; o < o1; ++o) {
    if (prune && !cf[o]) continue;
    const OrientDesc d = od[o];
    const float2* is = ispec + (size_t)o * max_ex * py * pz;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kPencilMaxEl; ++s) {
      const int e = tid + s * kPencilThreads;
      if (e < px * T) {
        const int x = e / T, t = e - x * T;
        const int kz = kz0 + t;
        float2 v = make_float2(0.f, 0.f);
        if (x < d.ex && kz < pz) v = is[((size_t)x * py + ky) * pz + kz];
        buf[e] = v;
      }
    }
    __syncthreads();
    lds_fft<true, kPencilThreads, kPencilMaxEl, kPencilRadix>(buf, px_, T, T, 1, false, twx, tid);
    // element e of both products belongs to the thread that read it
#pragma unroll
    for (int s = 0; s < kPencilMaxEl; ++s) {
      const int e = tid + s * kPencilThreads;
      if (e < px * T) {
        const float2 I = buf[e];
        buf[e] = cscale(cmulc(G1[s], I), scale);
        if (has2) buf2[e] = cscale(cmulc(G2[s], I), scale);
      }
    }
    __syncthreads();
    lds_fft<false, kPencilThreads, kPencilMaxEl, kPencilRadix>(buf, px_, T, T, 1, false, twx, tid);
    float2* d1 = y1 + (size_t)o * nx * py * pz;
    for (int e = tid; e < d.nvx * T; e += kPencilThreads) {
      const int x = e / T, t = e - x * T;
      const int kz = kz0 + t;
      if (kz < pz) d1[((size_t)x * py + ky) * pz + kz] = buf[e];
    }
    if (has2) {
      lds_fft<false, kPencilThreads, kPencilMaxEl, kPencilRadix>(buf2, px_, T, T, 1, false, twx,
                                                                 tid);
      float2* d2 = y2 + (size_t)o * nx * py * hz;
      for (int e = tid; e < d.nvx * T; e += kPencilThreads) {
        const int x = e / T, t = e - x * T;
        const int kz = kz0 + t;
        if (kz < hz) d2[((size_t)x * py + ky) * hz + kz] = buf2[e];
      }
    }
  }
}

// pass A with the x plan and the tile fixed at compile time (FPlan, TT): the
// same computation as k_corr_x, launched when the runtime plan and tile match.
template <class SX, int TT>
__global__ void __launch_bounds__(kPencilThreads)
k_corr_x_f(const float2* __restrict__ g1, const float2* __restrict__ g2,
           const float2* __restrict__ ispec, const OrientDesc* __restrict__ od,
           float2* __restrict__ y1, float2* __restrict__ y2, int n_orient, int chunk, int max_ex,
           int nx, AxisPlan px_, int py, int pz, int T_, float scale,
           const float2* __restrict__ tw_x, const int* __restrict__ cf,
           const int* __restrict__ ctl) {
  constexpr int PX = SX::N, T = TT, NT = kPencilThreads;
  constexpr int EL = CeilDiv<PX * T, NT>::v;
  if (search_off(ctl)) return;
  const bool prune = cf && ctl && ctl[CTL_PRUNE];
  extern __shared__ float2 smem[];
  const int hz = pz / 2 + 1;
  float2* twx = smem;
  float2* buf = twx + PX;
  float2* buf2 = buf + PX * T;
  // rank = (ky * tiles + tile) * chunks + chunk: the chunks of one tile, the
  // tiles of one row and neighbouring rows run on one XCD (xcd_rank)
  const int rk = xcd_rank(), nch = gridDim.z, ntile = gridDim.x;
  const int rest = rk / nch;
  const int ky = rest / ntile, tid = threadIdx.x;
  const int kz0 = (rest - ky * ntile) * T;
  const int o0 = (rk - rest * nch) * chunk;
  const int o1 = min(n_orient, o0 + chunk);
  load_table(twx, tw_x, PX, tid, NT);
  const bool has2 = kz0 < hz;
  float2 G1[EL], G2[EL];
#pragma unroll
  for (int s = 0; s < EL; ++s) {
    const int e = tid + s * NT;
    const int x = e / T, t = e - x * T;
    const int kz = kz0 + t;
    G1[s] = make_float2(0.f, 0.f);
    G2[s] = make_float2(0.f, 0.f);
    if (e < PX * T && kz < pz) {
      G1[s] = g1[((size_t)x * py + ky) * pz + kz];
      if (kz < hz) G2[s] = g2[((size_t)x * py + ky) * hz + kz];
    }
  }
  for (int o = o0; o < o1; ++o) {
    if (prune && !cf[o]) continue;
    const OrientDesc d = od[o];
    const float2* is = ispec + (size_t)o * max_ex * py * pz;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < EL; ++s) {
      const int e = tid + s * NT;
      if (e < PX * T) {
        const int x = e / T, t = e - x * T;
        const int kz = kz0 + t;
        float2 v = make_float2(0.f, 0.f);
        if (x < d.ex && kz < pz) v = is[((size_t)x * py + ky) * pz + kz];
        buf[e] = v;
      }
    }
    __syncthreads();
    FixedFft<true, NT, SX, T, 1, false, T, 0>::run(buf, 0, twx, tid);
#pragma unroll
    for (int s = 0; s < EL; ++s) {
      const int e = tid + s * NT;
      if (e < PX * T) {
        const float2 I = buf[e];
        buf[e] = cscale(cmulc(G1[s], I), scale);
        if (has2) buf2[e] = cscale(cmulc(G2[s], I), scale);
      }
    }
    __syncthreads();
    FixedFft<false, NT, SX, T, 1, false, T, 0>::run(buf, 0, twx, tid);
    float2* d1 = y1 + (size_t)o * nx * py * pz;
    for (int e = tid; e < d.nvx * T; e += NT) {
      const int x = e / T, t = e - x * T;
      const int kz = kz0 + t;
      if (kz < pz) d1[((size_t)x * py + ky) * pz + kz] = buf[e];
    }
    if (has2) {
      FixedFft<false, NT, SX, T, 1, false, T, 0>::run(buf2, 0, twx, tid);
      float2* d2 = y2 + (size_t)o * nx * py * hz;
      for (int e = tid; e < d.nvx * T; e += NT) {
        const int x = e / T, t = e - x * T;
        const int kz = kz0 + t;
        if (kz < hz) d2[((size_t)x * py + ky) * hz + kz] = buf2[e];
      }
    }
  }
}

// Pass A for px = N = 16 R, R <= 16 (240 = 16 * 15 for the common trays),
// with the x transforms in registers: the same products as k_corr_x.  Thread
// (b, c), b = tid / T < 16, c = tid % T, owns column kz0 + c of the tile.
//  forward, x = n + R r, k = m0 + 16 m1:
//    X[m0 + 16 m1] = sum_n wR^(n m1) [wN^(n m0) sum_r x[n + R r] w16^(r m0)]
//    step 1 (thread n = b < R): DFT16 over r, times wN^(n m0), to LDS;
//    step 2 (thread m0 = b): DFT-R over n, giving X[b + 16 m1], m1 < R;
//  the products P = G conj(X) scale at those R frequencies, then inverse,
//  x = x0 + R x1:
//    y[x0 + R x1] = sum_b w16^(-b x1) [wN^(-b x0) sum_m1 P[b + 16 m1] wR^(-m1 x0)]
//    step 3 (thread b): inverse DFT-R over m1, times wN^(-b x0), to LDS;
//    step 4 (thread x0 = b < R): inverse DFT16 over b, giving y[x0 + R x1].
// With NB = 3 LDS buffers (step 1's output and the two products) a block
// takes two barriers per orientation; with NB = 2 the low-digit product
// reuses step 1's buffer, for two more barriers and a third less LDS.  G1 and
// G2 at the thread's R frequencies stay in registers across the chunk's
// orientations.  LDS rows are padded to 17 T values so that the strided reads
// of steps 2 and 4 spread over the banks.
constexpr int kXrPitch = 17;
__host__ __device__ constexpr size_t corr_xr_shmem(int R, int T, int NB) {
  return (size_t)(16 * R + 16 * kXrPitch * T + (NB - 1) * R * kXrPitch * T) * sizeof(float2);
}

// The orientation loop of k_corr_xr; BOTH: the tile has columns with a
// low-digit product (kz < hz), else only the first product is formed.  The two
// cases are separate instantiations so that each keeps its whole loop body
// free of branches (the scheduler interleaves the two products' work).
template <int R, int T, int NB, bool BOTH>
__device__ __forceinline__ void xr_orients(
    const float2* __restrict__ ispec, const OrientDesc* __restrict__ od, float2* __restrict__ y1,
    float2* __restrict__ y2, int o0, int o1, bool prune, const int* __restrict__ cf, int max_ex,
    int nx, int rz, int rh, int ky, int pz, int hz, int kz, bool colok, bool has2,
    const float2 (&G1)[R], const float2 (&G2)[R], const float2* tw, float2* sa, float2* s1,
    float2* s2, int b, int c) {
  constexpr int PA = kXrPitch * T;
  for (int o = o0; o < o1; ++o) {
    if (prune && !cf[o]) continue;
    const OrientDesc d = od[o];
    // the thread's offsets, recomputed per orientation behind an opaque copy
    // (hoisted out of the loop, 48 of them would hold registers throughout)
    int fo = b * rz + ky * pz + kz, ho = b * rh + ky * hz + kz;
    asm volatile("" : "+v"(fo), "+v"(ho));
    {   // step 1 (threads b < R; the barriers stay outside the branch)
      const float2* is = ispec + (size_t)o * max_ex * rz;
      float2 v[16];
      if (b < R) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int x = b + R * r;
          v[r] = (x < d.ex && colok) ? is[fo + R * r * rz] : make_float2(0.f, 0.f);
        }
        Bfly<16, true>::run(v);
#pragma unroll
        for (int m0 = 1; m0 < 16; ++m0) v[m0] = cmul(v[m0], tw[b * m0]);
      }
      if (NB == 2 && BOTH) __syncthreads();   // the previous orientation's step 4 has read sa
      if (b < R) {
#pragma unroll
        for (int m0 = 0; m0 < 16; ++m0) sa[m0 * PA + b * T + c] = v[m0];
      }
    }
    __syncthreads();
    float2 X[R];   // step 2
#pragma unroll
    for (int n = 0; n < R; ++n) X[n] = sa[b * PA + n * T + c];
    Bfly<R, true>::run(X);
    // step 3
    {
      float2 P[R];
#pragma unroll
      for (int m1 = 0; m1 < R; ++m1) P[m1] = cmulc(G1[m1], X[m1]);
      Bfly<R, false>::run(P);
      if (BOTH && NB == 3) {
        float2 Q[R];
#pragma unroll
        for (int m1 = 0; m1 < R; ++m1) Q[m1] = cmulc(G2[m1], X[m1]);
        Bfly<R, false>::run(Q);
#pragma unroll
        for (int x0 = 0; x0 < R; ++x0) {
          float2 p = P[x0], q = Q[x0];
          if (x0 > 0) {
            float2 w = tw[b * x0];
            w.y = -w.y;
            p = cmul(p, w);
            q = cmul(q, w);
          }
          s1[x0 * PA + b * T + c] = p;
          s2[x0 * PA + b * T + c] = q;
        }
      } else {
#pragma unroll
        for (int x0 = 0; x0 < R; ++x0) {
          float2 p = P[x0];
          if (x0 > 0) {
            float2 w = tw[b * x0];
            w.y = -w.y;
            p = cmul(p, w);
          }
          s1[x0 * PA + b * T + c] = p;
        }
        if (BOTH) {   // NB == 2: the second product goes to sa once step 2 has read it
#pragma unroll
          for (int m1 = 0; m1 < R; ++m1) P[m1] = cmulc(G2[m1], X[m1]);
          Bfly<R, false>::run(P);
          __syncthreads();
#pragma unroll
          for (int x0 = 0; x0 < R; ++x0) {
            float2 p = P[x0];
            if (x0 > 0) {
              float2 w = tw[b * x0];
              w.y = -w.y;
              p = cmul(p, w);
            }
            s2[x0 * PA + b * T + c] = p;
          }
        }
      }
    }
    __syncthreads();
    if (b < R && colok) {   // step 4
      float2 v[16];
#pragma unroll
      for (int bb = 0; bb < 16; ++bb) v[bb] = s1[b * PA + bb * T + c];
      Bfly<16, false>::run(v);
      float2* d1 = y1 + (size_t)o * nx * rz;
#pragma unroll
      for (int x1 = 0; x1 < 16; ++x1) {
        const int x = b + R * x1;
        if (x < d.nvx) d1[fo + R * x1 * rz] = v[x1];
      }
      if (BOTH && has2) {
#pragma unroll
        for (int bb = 0; bb < 16; ++bb) v[bb] = s2[b * PA + bb * T + c];
        Bfly<16, false>::run(v);
        float2* d2 = y2 + (size_t)o * nx * rh;
#pragma unroll
        for (int x1 = 0; x1 < 16; ++x1) {
          const int x = b + R * x1;
          if (x < d.nvx) d2[ho + R * x1 * rh] = v[x1];
        }
      }
    }
  }
}

template <int R, int T, int NB>
__global__ void __launch_bounds__(16 * T)
k_corr_xr(const float2* __restrict__ g1, const float2* __restrict__ g2,
            const float2* __restrict__ ispec, const OrientDesc* __restrict__ od,
            float2* __restrict__ y1, float2* __restrict__ y2, int n_orient, int chunk,
            int max_ex, int nx, AxisPlan px_, int py, int pz, int T_, float scale,
            const float2* __restrict__ tw_x, const int* __restrict__ cf,
            const int* __restrict__ ctl) {
  constexpr int NT = 16 * T, PA = kXrPitch * T;
  if (search_off(ctl)) return;
  const bool prune = cf && ctl && ctl[CTL_PRUNE];
  // rank = (ky * tiles + tile) * chunks + chunk, as in k_corr_x_f
  const int rk = xcd_rank(), nch = gridDim.z, ntile = gridDim.x;
  const int rest = rk / nch;
  const int ky = rest / ntile, tid = threadIdx.x;
  const int kz0 = (rest - ky * ntile) * T;
  const int o0 = (rk - rest * nch) * chunk;
  const int o1 = min(n_orient, o0 + chunk);
  // a chunk with no orientation left to search loads nothing
  bool any = false;
  for (int o = o0; o < o1 && !any; ++o) any = !prune || cf[o] != 0;
  if (!any) return;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* sa = tw + 16 * R;
  float2* s1 = sa + 16 * PA;
  float2* s2 = NB == 3 ? s1 + R * PA : sa;
  for (int q = tid; q < 16 * R; q += NT) tw[q] = tw_x[q];
  const int b = tid / T, c = tid - b * T, kz = kz0 + c;
  const int hz = pz / 2 + 1;
  const bool colok = kz < pz, has2 = kz < hz;
  // 32-bit element offsets (a volume is below 2^31 elements); x rows of the
  // full (pz) and half (hz) spectra are py * pz and py * hz elements apart
  const int rz = py * pz, rh = py * hz;
  // the tray spectra with the inverse transform's 1 / (px py pz) folded in
  float2 G1[R], G2[R];
#pragma unroll
  for (int m1 = 0; m1 < R; ++m1) {
    const int k = b + 16 * m1;
    G1[m1] = colok ? cscale(g1[k * rz + ky * pz + kz], scale) : make_float2(0.f, 0.f);
    G2[m1] = has2 ? cscale(g2[k * rh + ky * hz + kz], scale) : make_float2(0.f, 0.f);
  }
  __syncthreads();   // twiddle table
  // every wave holds all T columns of the tile, so this branch is uniform
  if (kz0 < hz)
    xr_orients<R, T, NB, true>(ispec, od, y1, y2, o0, o1, prune, cf, max_ex, nx, rz, rh, ky, pz,
                              hz, kz, colok, has2, G1, G2, tw, sa, s1, s2, b, c);
  else
    xr_orients<R, T, NB, false>(ispec, od, y1, y2, o0, o1, prune, cf, max_ex, nx, rz, rh, ky, pz,
                               hz, kz, colok, has2, G1, G2, tw, sa, s1, s2, b, c);
}

// ---------------------------------------------------------------------------
// pass B: grid (nx, n_orient).  Inverse yz, round, score, argmin.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wg_argmin_write(double sc, int idx, long long p, float res,
                                                int unsup, WgBest* out, int tid, int nt) {
  for (int o = 32; o > 0; o >>= 1) {
    const double s2 = __shfl_xor(sc, o);
    const int i2 = __shfl_xor(idx, o);
    const long long p2 = __shfl_xor(p, o);
    const float r2 = __shfl_xor(res, o);
    if (better(s2, i2, sc, idx)) { sc = s2; idx = i2; p = p2; }
    res = r2 > res ? r2 : res;
  }
  __shared__ double ws[16];
  __shared__ int wi[16];
  __shared__ long long wp[16];
  __shared__ float wr[16];
  __syncthreads();
  if ((tid & 63) == 0) {
    ws[tid >> 6] = sc; wi[tid >> 6] = idx; wp[tid >> 6] = p; wr[tid >> 6] = res;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < nt / 64; ++w) {
      if (better(ws[w], wi[w], sc, idx)) { sc = ws[w]; idx = wi[w]; p = wp[w]; }
      res = wr[w] > res ? wr[w] : res;
    }
    WgBest b;
    b.score = sc;
    b.idx = idx;
    b.p_lo = (int)(p & 0xffffffffLL);
    b.p_hi = (int)(p >> 32);
    b.resid = res;
    b.unsupported = unsup;
    *out = b;
  }
}

// ---- score pass helpers ----------------------------------------------------
// Low digit, z pass over row pairs.  After the y pass, row j of the slab holds
// X_j, the half spectrum along z (kz < hz) of a real row; row j < h
// (h = ceil(nvy / 2)) takes Z = X_j + i X_{j+h} over every kz (X at pz - k
// is conj X at k), so the inverse z transform of row j carries the low-digit
// sums of row j in its real part and of row j + h in its imaginary part: one
// complex transform per two real rows.  In place: the thread of (j, k) alone
// reads X_j[k] and X_{j+h}[k] and writes Z_j[k] and Z_j[pz - k] (a column
// past the half spectrum).
// (row(j): the LDS row that holds logical row j, when the y pass leaves its
// rows permuted)
template <int NT, int PZ_, int HZ_, class RowMap = IdentityMap>
__device__ __forceinline__ void pair_low_rows(float2* buf, int pz_rt, int nvy, int tid,
                                              RowMap row = RowMap()) {
  const int pz = PZ_ > 0 ? PZ_ : pz_rt, hz = PZ_ > 0 ? HZ_ : pz_rt / 2 + 1;
  const int h = (nvy + 1) >> 1;
  const float inv_hz = 1.0f / (float)hz;
  for (int e = tid; e < h * hz; e += NT) {
    const int j = PZ_ > 0 ? e / HZ_ : fdiv(e, hz, inv_hz), k = e - j * hz;
    const int rj = row(j) * pz;
    const float2 a = buf[rj + k];
    const float2 b = j + h < nvy ? buf[row(j + h) * pz + k] : make_float2(0.f, 0.f);
    buf[rj + k] = make_float2(a.x - b.y, a.y + b.x);                      // a + i b
    if (k > 0 && pz - k >= hz) buf[rj + pz - k] = make_float2(a.x + b.y, b.x - a.y);
  }
}

// The valid offsets (j < nvy, k < nvz) of one slab, column by column: the
// thread walks columns k = c0, c0 + cols, ... and in each the rows
// j = g, g + G, ... in increasing order (cols = min(nvz, NT) threads per row
// group, G = NT / cols groups, threads past cols * G idle).
struct ColWalk {
  int cols, G, g, c0;
};
template <int NT>
__device__ __forceinline__ ColWalk col_walk(int nvz, int tid) {
  ColWalk w;
  w.cols = nvz < NT ? (nvz > 0 ? nvz : 1) : NT;
  w.G = NT / w.cols;
  w.g = tid / w.cols;
  w.c0 = w.g < w.G ? tid - w.g * w.cols : nvz;
  return w;
}

// rounded low-digit sums of the valid offsets into lo[j * nz + k]; returns the
// largest rounding residual (rows paired by pair_low_rows)
// (row, col: the LDS row and column of logical row j and column k, when the
// transforms leave them permuted)
template <class RowMap = IdentityMap, class ColMap = IdentityMap>
__device__ __forceinline__ float low_digit_sums(const float2* buf, int ldb, int* lo, int nz,
                                                int nvy, int nvz, const ColWalk& w,
                                                RowMap row = RowMap(), ColMap col = ColMap()) {
  const int h = (nvy + 1) >> 1;
  float res = 0.f;
  for (int k = w.c0; k < nvz; k += w.cols) {
    const int ck = col(k);
    for (int j = w.g; j < nvy; j += w.G) {
      const float v = j < h ? buf[row(j) * ldb + ck].x : buf[row(j - h) * ldb + ck].y;
      const float r = rintf(v);
      res = fmaxf(res, fabsf(v - r));
      lo[j * nz + k] = (int)r;
    }
  }
  return res;
}

// Collision and proximity of the valid offsets of slab x, and the slab's best
// (score, index).  Within a column the score double(p) + pen[k] is a strictly
// increasing function of the integer p while |p| + |pen| < 2^51 (doubles
// there are spaced at most 0.5 apart, so integers one apart never round to
// one value): the column's lowest score belongs to its smallest p, first in
// row order on ties, and only that offset's score is formed (score_order_ok
// checks the bound).  Ties across columns resolve by the lowest index.
template <class RowMap = IdentityMap, class ColMap = IdentityMap>
__device__ __forceinline__ void slab_scores(const float2* buf, int ldb, const int* lo, int x,
                                            int ny, int nz, int nvy, int nvz, int lb,
                                            long long addback, const double* __restrict__ pen,
                                            unsigned char* __restrict__ fgrid,
                                            long long* __restrict__ pgrid, size_t gbase,
                                            const ColWalk& w, float& res, double& best,
                                            int& bidx, long long& bp, RowMap row = RowMap(),
                                            ColMap col = ColMap()) {
  for (int k = w.c0; k < nvz; k += w.cols) {
    long long pm = 0;
    int jm = -1;
    const int ck = col(k);
    for (int j = w.g; j < nvy; j += w.G) {
      const float2 v = buf[row(j) * ldb + ck];
      const float c = rintf(v.x), h = rintf(v.y);
      res = fmaxf(res, fmaxf(fabsf(v.x - c), fabsf(v.y - h)));
      // b * h with b = 2^lb: a shift of the (two's complement) high digit
      const long long p =
          (long long)((unsigned long long)(long long)h << lb) + lo[j * nz + k] + addback;
      if (fgrid) {
        const size_t g = gbase + ((size_t)x * ny + j) * nz + k;
        fgrid[g] = c == 0.f ? 1 : 0;
        pgrid[g] = p;
      }
      if (c == 0.f && (jm < 0 || p < pm)) { pm = p; jm = j; }
    }
    if (jm >= 0) {
      const double sc = (double)pm + pen[k];
      const int idx = (x * ny + jm) * nz + k;
      if (better(sc, idx, best, bidx)) { best = sc; bidx = idx; bp = pm; }
    }
  }
}

// the bound under which slab_scores' column-wise minimum is exact: every
// proximity is a sum of item values times field values, and the penalty
// tables hold 1e8 (k / nz)^3 < 2^27
__device__ __forceinline__ bool score_order_ok(const OrientDesc& d, const FieldStats* st) {
  const double fm = fmax(fabs((double)st->mx), fabs((double)st->mn));
  return (double)d.isum * fm < 0x1p50;
}

template <int RMAX>
__global__ void __launch_bounds__(kSlabThreads)
k_score_yz(const float2* __restrict__ y1, const float2* __restrict__ y2,
           const OrientDesc* __restrict__ od, const FieldStats* __restrict__ st,
           const double* __restrict__ pen, WgBest* __restrict__ rec, int nx, int ny, int nz,
           AxisPlan py_, AxisPlan pz_, const float2* __restrict__ tw_y,
           const float2* __restrict__ tw_z, unsigned char* __restrict__ fgrid,
           long long* __restrict__ pgrid, const int* __restrict__ cf,
           unsigned long long* __restrict__ searched, const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int py = py_.n, pz = pz_.n, hz = pz / 2 + 1;
  const int x = blockIdx.x, o = blockIdx.y, tid = threadIdx.x;
  const OrientDesc d = od[o];
  WgBest* out = rec + (size_t)o * nx + x;
  if (x >= d.nvx) return;
  if (cf && ctl && ctl[CTL_PRUNE] && !cf[o]) {   // excluded by the column test
    if (tid == 0) {
      WgBest b;
      b.score = 1e300; b.idx = INT_MAX; b.p_lo = 0; b.p_hi = 0; b.resid = 0.f; b.unsupported = 0;
      *out = b;
    }
    return;
  }
  if (searched && x == 0 && tid == 0) atomicAdd(searched, 1ULL);   // an orientation searched
  float2* twy = smem;
  float2* twz = twy + py;
  float2* buf = twz + pz;
  load_table(twy, tw_y, py, tid, kSlabThreads);
  load_table(twz, tw_z, pz, tid, kSlabThreads);
  const FieldParams fp = field_params(st, (long long)nx * ny * nz);
  const long long addback = ((1LL << fp.lb) / 2 + fp.mu) * d.isum;
  const int nvy = d.nvy, nvz = d.nvz;
  const ColWalk w = col_walk<kSlabThreads>(nvz, tid);
  // --- low digit (real field, half spectrum): y pass, rows paired, z pass ---
  const float2* s2 = y2 + ((size_t)o * nx + x) * py * hz;
  stage_rows<kSlabThreads, kSlabMaxEl>(buf, s2, py * hz, hz, pz, tid);
  __syncthreads();
  lds_fft<false, kSlabThreads, kSlabMaxEl, RMAX>(buf, py_, pz, hz, 1, false, twy, tid);
  pair_low_rows<kSlabThreads, 0, 0>(buf, pz, nvy, tid);
  __syncthreads();
  lds_fft<false, kSlabThreads, kSlabMaxEl, RMAX>(buf, pz_, 1, (nvy + 1) >> 1, pz, true, twz, tid);
  // the rounded low-digit sums wait in LDS behind the slab ([ny][nz] ints),
  // not in registers: the second slab FFT needs the whole register budget
  int* lo = reinterpret_cast<int*>(buf + py * pz);
  float res = low_digit_sums(buf, pz, lo, nz, nvy, nvz, w);
  __syncthreads();
  // --- collision + high digit (complex field, full spectrum) ---
  const float2* s1 = y1 + ((size_t)o * nx + x) * py * pz;
  stage_rows<kSlabThreads, kSlabMaxEl>(buf, s1, py * pz, pz, pz, tid);
  __syncthreads();
  lds_fft<false, kSlabThreads, kSlabMaxEl, RMAX>(buf, py_, pz, pz, 1, false, twy, tid);
  lds_fft<false, kSlabThreads, kSlabMaxEl, RMAX>(buf, pz_, 1, nvy, pz, true, twz, tid);
  double best = 1e300;
  int bidx = INT_MAX;
  long long bp = 0;
  slab_scores(buf, pz, lo, x, ny, nz, nvy, nvz, fp.lb, addback, pen, fgrid, pgrid,
              (size_t)o * nx * ny * nz, w, res, best, bidx, bp);
  wg_argmin_write(best, bidx, bp, res, fp.ok && score_order_ok(d, st) ? 0 : 1, out, tid,
                  kSlabThreads);
}

// The score pass with the y and z plans fixed at compile time (FPlan): the
// same computation as k_score_yz, for the padded sizes of the common trays,
// with every index, stride and loop bound a constant.  The host launches it
// when the runtime plans equal SY and SZ.
template <class SY, class SZ>
__global__ void __launch_bounds__(kSlabThreads)
k_score_yz_f(const float2* __restrict__ y1, const float2* __restrict__ y2,
             const OrientDesc* __restrict__ od, const FieldStats* __restrict__ st,
             const double* __restrict__ pen, WgBest* __restrict__ rec, int nx, int ny, int nz,
             AxisPlan py_, AxisPlan pz_, const float2* __restrict__ tw_y,
             const float2* __restrict__ tw_z, unsigned char* __restrict__ fgrid,
             long long* __restrict__ pgrid, const int* __restrict__ cf,
             unsigned long long* __restrict__ searched, const int* __restrict__ ctl) {
  constexpr int PY = SY::N, PZ = SZ::N, HZ = PZ / 2 + 1, NT = kSlabThreads;
  constexpr int EL = CeilDiv<PY * PZ, NT>::v;       // slab elements per thread
  constexpr int ELH = CeilDiv<PY * HZ, NT>::v;      // half-spectrum elements per thread
  if (search_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int x = blockIdx.x, o = blockIdx.y, tid = threadIdx.x;
  const OrientDesc d = od[o];
  WgBest* out = rec + (size_t)o * nx + x;
  if (x >= d.nvx) return;
  if (cf && ctl && ctl[CTL_PRUNE] && !cf[o]) {   // excluded by the column test
    if (tid == 0) {
      WgBest b;
      b.score = 1e300; b.idx = INT_MAX; b.p_lo = 0; b.p_hi = 0; b.resid = 0.f; b.unsupported = 0;
      *out = b;
    }
    return;
  }
  if (searched && x == 0 && tid == 0) atomicAdd(searched, 1ULL);   // an orientation searched
  float2* twy = smem;
  float2* twz = twy + PY;
  float2* buf = twz + PZ;
  load_table(twy, tw_y, PY, tid, NT);
  load_table(twz, tw_z, PZ, tid, NT);
  const FieldParams fp = field_params(st, (long long)nx * ny * nz);
  const long long addback = ((1LL << fp.lb) / 2 + fp.mu) * d.isum;
  const int nvy = d.nvy, nvz = d.nvz;
  const ColWalk w = col_walk<NT>(nvz, tid);
  // --- low digit (real field, half spectrum): y pass, rows paired, z pass ---
  {
    const float2* s2 = y2 + ((size_t)o * nx + x) * PY * HZ;
    float2 v[ELH];
#pragma unroll
    for (int s = 0; s < ELH; ++s) {
      const int e = tid + s * NT;
      if (e < PY * HZ) v[s] = s2[e];
    }
#pragma unroll
    for (int s = 0; s < ELH; ++s) {
      const int e = tid + s * NT;
      if (e < PY * HZ) {
        const int j = e / HZ;
        buf[j * PZ + (e - j * HZ)] = v[s];
      }
    }
  }
  __syncthreads();
  // in-place transforms (one barrier per stage): the y pass leaves logical
  // row j in LDS row yrow(j), the z pass logical column k in LDS column zcol(k)
  const auto yrow = [](int j) { return PlanPos<SY>::at(j); };
  const auto zcol = [](int k) { return PlanPos<SZ>::at(k); };
  IpFft<false, NT, SY, PZ, 1, false, HZ, 0>::run(buf, 0, twy, tid, IdentityMap());
  pair_low_rows<NT, PZ, HZ>(buf, PZ, nvy, tid, yrow);
  __syncthreads();
  IpFft<false, NT, SZ, 1, PZ, true, 0, PY>::run(buf, (nvy + 1) >> 1, twz, tid, yrow);
  // the rounded low-digit sums wait in LDS behind the slab ([ny][nz] ints)
  int* lo = reinterpret_cast<int*>(buf + PY * PZ);
  float res = low_digit_sums(buf, PZ, lo, nz, nvy, nvz, w, yrow, zcol);
  __syncthreads();
  // --- collision + high digit (complex field, full spectrum) ---
  {
    const float2* s1 = y1 + ((size_t)o * nx + x) * PY * PZ;
    float2 v[EL];
#pragma unroll
    for (int s = 0; s < EL; ++s) {
      const int e = tid + s * NT;
      if (e < PY * PZ) v[s] = s1[e];
    }
#pragma unroll
    for (int s = 0; s < EL; ++s) {
      const int e = tid + s * NT;
      if (e < PY * PZ) buf[e] = v[s];
    }
  }
  __syncthreads();
  IpFft<false, NT, SY, PZ, 1, false, PZ, 0>::run(buf, 0, twy, tid, IdentityMap());
  IpFft<false, NT, SZ, 1, PZ, true, 0, PY>::run(buf, nvy, twz, tid, yrow);
  double best = 1e300;
  int bidx = INT_MAX;
  long long bp = 0;
  slab_scores(buf, PZ, lo, x, ny, nz, nvy, nvz, fp.lb, addback, pen, fgrid, pgrid,
              (size_t)o * nx * ny * nz, w, res, best, bidx, bp, yrow, zcol);
  wg_argmin_write(best, bidx, bp, res, fp.ok && score_order_ok(d, st) ? 0 : 1, out, tid, NT);
}

// ---- score pass with both digit slabs resident -----------------------------
// One in-place decimation-in-frequency butterfly (the element mapping of
// ipstage): R elements i0 + r S es, S = L / R, twiddled by w_N^(n r N / L).
template <int R, bool FWD, int N, int L>
__device__ __forceinline__ void ip_bfly(float2* buf, int i0, int es, int n, const float2* tw) {
  constexpr int S = L / R, TWS = N / L;
  float2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = buf[i0 + r * S * es];
  Bfly<R, FWD>::run(v);
  if (S > 1 && n != 0) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      float2 w = tw[n * r * TWS];
      if (!FWD) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) buf[i0 + r * S * es] = v[r];
}

// One stage over the columns of two slabs at once: B1 columns of a (element
// pitch B1) and B2 columns of b (pitch B2), every column of length N.
template <int R, bool FWD, int NT, int N, int L, int B1, int B2>
__device__ __forceinline__ void ipstage_cols2(float2* a, float2* b, const float2* tw, int tid) {
  constexpr int S = L / R, M = N / R;
  constexpr int T1 = B1 * M, TOT = (B1 + B2) * M;
  constexpr int MS = CeilDiv<TOT, NT>::v;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    const int e = tid + s * NT;
    if (e < T1) {
      const int j = e / B1, col = e - j * B1;
      const int blk = j / S, n = j - blk * S;
      ip_bfly<R, FWD, N, L>(a, col + (blk * L + n) * B1, B1, n, tw);
    } else if (e < TOT) {
      const int f = e - T1;
      const int j = f / B2, col = f - j * B2;
      const int blk = j / S, n = j - blk * S;
      ip_bfly<R, FWD, N, L>(b, col + (blk * L + n) * B2, B2, n, tw);
    }
  }
  __syncthreads();
}
template <bool FWD, int NT, int N, int L, int B1, int B2, int R, int... Rest>
__device__ __forceinline__ void ip_cols2(float2* a, float2* b, const float2* tw, int tid) {
  ipstage_cols2<R, FWD, NT, N, L, B1, B2>(a, b, tw, tid);
  if constexpr (sizeof...(Rest) > 0) ip_cols2<FWD, NT, N, L / R, B1, B2, Rest...>(a, b, tw, tid);
}

// One stage over the rows (length N, contiguous, pitch PITCH) of two slabs:
// n1 rows of a, logical row q at LDS row rowmap(q), and n2 rows of b at rows
// 0 .. n2-1 (both counts at most NMAX in total).
template <int R, bool FWD, int NT, int N, int L, int PITCH, int NMAX, class RowMap>
__device__ __forceinline__ void ipstage_rows2(float2* a, int n1, float2* b, int n2,
                                              const float2* tw, int tid, RowMap rowmap) {
  constexpr int S = L / R, M = N / R;
  constexpr int MS = CeilDiv<NMAX * M, NT>::v;
  const int t1 = n1 * M, tot = (n1 + n2) * M;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    const int e = tid + s * NT;
    if (e < tot) {
      const bool first = e < t1;
      const int f = first ? e : e - t1;
      const int q = f / M, j = f - q * M;
      const int blk = j / S, n = j - blk * S;
      float2* base = first ? a + rowmap(q) * PITCH : b + q * PITCH;
      ip_bfly<R, FWD, N, L>(base, blk * L + n, 1, n, tw);
    }
  }
  __syncthreads();
}
template <bool FWD, int NT, int N, int L, int PITCH, int NMAX, class RowMap, int R, int... Rest>
__device__ __forceinline__ void ip_rows2(float2* a, int n1, float2* b, int n2, const float2* tw,
                                         int tid, RowMap rm) {
  ipstage_rows2<R, FWD, NT, N, L, PITCH, NMAX, RowMap>(a, n1, b, n2, tw, tid, rm);
  if constexpr (sizeof...(Rest) > 0)
    ip_rows2<FWD, NT, N, L / R, PITCH, NMAX, RowMap, Rest...>(a, n1, b, n2, tw, tid, rm);
}
template <class P> struct IpCols2;
template <int N, int... Rs> struct IpCols2<FPlan<N, Rs...>> {
  template <bool FWD, int NT, int B1, int B2>
  __device__ __forceinline__ static void run(float2* a, float2* b, const float2* tw, int tid) {
    ip_cols2<FWD, NT, N, N, B1, B2, Rs...>(a, b, tw, tid);
  }
};
template <class P> struct IpRows2;
template <int N, int... Rs> struct IpRows2<FPlan<N, Rs...>> {
  template <bool FWD, int NT, int PITCH, int NMAX, class RowMap>
  __device__ __forceinline__ static void run(float2* a, int n1, float2* b, int n2,
                                             const float2* tw, int tid, RowMap rm) {
    ip_rows2<FWD, NT, N, N, PITCH, NMAX, RowMap, Rs...>(a, n1, b, n2, tw, tid, rm);
  }
};

// LDS of k_score_yz2: twiddles, the complex slab [PY][PZ] and the low digit's
// half-spectrum slab [PY][HZ] (after the row pairing: [ceil(nvy/2)][PZ])
__host__ __device__ constexpr size_t score2_shmem(int py, int pz) {
  return (size_t)(py + pz + py * pz + py * (pz / 2 + 1)) * sizeof(float2);
}

// The score pass (same results as k_score_yz_f) with both slabs in LDS at
// once, so that their transforms share stages: one load phase, the y pass of
// the PZ complex columns and the HZ low-digit columns in the same stages (a
// radix-25 y stage then has 5 (PZ + HZ) butterflies for the block's 1024
// threads instead of 5 PZ or 5 HZ), the low digit's rows paired as in
// pair_low_rows, and the z pass of the nvy complex rows and the paired rows
// in the same stages.  The scores then read both slabs directly.
template <class SY, class SZ>
__global__ void __launch_bounds__(kSlabThreads)
k_score_yz2(const float2* __restrict__ y1, const float2* __restrict__ y2,
            const OrientDesc* __restrict__ od, const FieldStats* __restrict__ st,
            const double* __restrict__ pen, WgBest* __restrict__ rec, int nx, int ny, int nz,
            AxisPlan py_, AxisPlan pz_, const float2* __restrict__ tw_y,
            const float2* __restrict__ tw_z, unsigned char* __restrict__ fgrid,
            long long* __restrict__ pgrid, const int* __restrict__ cf,
            unsigned long long* __restrict__ searched, const int* __restrict__ ctl) {
  constexpr int PY = SY::N, PZ = SZ::N, HZ = PZ / 2 + 1, NT = kSlabThreads;
  constexpr int EL = CeilDiv<PY * PZ, NT>::v;       // complex slab elements per thread
  constexpr int ELH = CeilDiv<PY * HZ, NT>::v;      // half-spectrum elements per thread
  constexpr int HMAX = (PY + 1) / 2;                // paired low-digit rows at most
  static_assert(HMAX * PZ <= PY * HZ, "paired rows must fit the half-spectrum slab");
  if (search_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int x = blockIdx.x, o = blockIdx.y, tid = threadIdx.x;
  const OrientDesc d = od[o];
  WgBest* out = rec + (size_t)o * nx + x;
  if (x >= d.nvx) return;
  if (cf && ctl && ctl[CTL_PRUNE] && !cf[o]) {   // excluded by the column test
    if (tid == 0) {
      WgBest b;
      b.score = 1e300; b.idx = INT_MAX; b.p_lo = 0; b.p_hi = 0; b.resid = 0.f; b.unsupported = 0;
      *out = b;
    }
    return;
  }
  if (searched && x == 0 && tid == 0) atomicAdd(searched, 1ULL);   // an orientation searched
  float2* twy = smem;
  float2* twz = twy + PY;
  float2* hb = twz + PZ;       // complex slab (collision + high digit)
  float2* lb = hb + PY * PZ;   // low digit
  load_table(twy, tw_y, PY, tid, NT);
  load_table(twz, tw_z, PZ, tid, NT);
  const FieldParams fp = field_params(st, (long long)nx * ny * nz);
  const long long addback = ((1LL << fp.lb) / 2 + fp.mu) * d.isum;
  const int nvy = d.nvy, nvz = d.nvz;
  const int h = (nvy + 1) >> 1;
  {   // both slabs: every load issued before the first LDS store
    const float2* s1 = y1 + ((size_t)o * nx + x) * PY * PZ;
    const float2* s2 = y2 + ((size_t)o * nx + x) * PY * HZ;
    float2 v[EL], u[ELH];
#pragma unroll
    for (int s = 0; s < EL; ++s) {
      const int e = tid + s * NT;
      if (e < PY * PZ) v[s] = s1[e];
    }
#pragma unroll
    for (int s = 0; s < ELH; ++s) {
      const int e = tid + s * NT;
      if (e < PY * HZ) u[s] = s2[e];
    }
#pragma unroll
    for (int s = 0; s < EL; ++s) {
      const int e = tid + s * NT;
      if (e < PY * PZ) hb[e] = v[s];
    }
#pragma unroll
    for (int s = 0; s < ELH; ++s) {
      const int e = tid + s * NT;
      if (e < PY * HZ) lb[e] = u[s];
    }
  }
  __syncthreads();
  // y pass of both slabs; logical row j then sits at LDS row yrow(j)
  const auto yrow = [](int j) { return PlanPos<SY>::at(j); };
  const auto zcol = [](int k) { return PlanPos<SZ>::at(k); };
  IpCols2<SY>::template run<false, NT, PZ, HZ>(hb, lb, twy, tid);
  {   // low digit: rows j and j + h paired into one complex row (pair_low_rows)
    constexpr int MP = CeilDiv<HMAX * HZ, NT>::v;
    float2 a[MP], bq[MP];
#pragma unroll
    for (int s = 0; s < MP; ++s) {
      const int e = tid + s * NT;
      if (e < h * HZ) {
        const int j = e / HZ, k = e - j * HZ;
        a[s] = lb[yrow(j) * HZ + k];
        bq[s] = j + h < nvy ? lb[yrow(j + h) * HZ + k] : make_float2(0.f, 0.f);
      }
    }
    __syncthreads();   // the paired rows overwrite the half-spectrum layout
#pragma unroll
    for (int s = 0; s < MP; ++s) {
      const int e = tid + s * NT;
      if (e < h * HZ) {
        const int j = e / HZ, k = e - j * HZ;
        lb[j * PZ + k] = make_float2(a[s].x - bq[s].y, a[s].y + bq[s].x);   // a + i b
        if (k > 0 && PZ - k >= HZ) lb[j * PZ + PZ - k] = make_float2(a[s].x + bq[s].y, bq[s].x - a[s].y);
      }
    }
  }
  __syncthreads();
  // z pass: the nvy complex rows (LDS rows yrow(j)) and the h paired rows
  IpRows2<SZ>::template run<false, NT, PZ, PY + HMAX>(hb, nvy, lb, h, twz, tid, yrow);
  // scores of the valid offsets (slab_scores, with the low digit read here)
  const ColWalk w = col_walk<NT>(nvz, tid);
  float res = 0.f;
  double best = 1e300;
  int bidx = INT_MAX;
  long long bp = 0;
  const size_t gbase = (size_t)o * nx * ny * nz;
  for (int k = w.c0; k < nvz; k += w.cols) {
    long long pm = 0;
    int jm = -1;
    const int ck = zcol(k);
    for (int j = w.g; j < nvy; j += w.G) {
      const float2 v = hb[yrow(j) * PZ + ck];
      const float lv = j < h ? lb[j * PZ + ck].x : lb[(j - h) * PZ + ck].y;
      const float c = rintf(v.x), hh = rintf(v.y), lr = rintf(lv);
      res = fmaxf(res, fmaxf(fabsf(lv - lr), fmaxf(fabsf(v.x - c), fabsf(v.y - hh))));
      // b * hh with b = 2^lb: a shift of the (two's complement) high digit
      const long long p =
          (long long)((unsigned long long)(long long)hh << fp.lb) + (long long)lr + addback;
      if (fgrid) {
        const size_t g = gbase + ((size_t)x * ny + j) * nz + k;
        fgrid[g] = c == 0.f ? 1 : 0;
        pgrid[g] = p;
      }
      if (c == 0.f && (jm < 0 || p < pm)) { pm = p; jm = j; }
    }
    if (jm >= 0) {
      const double sc = (double)pm + pen[k];
      const int idx = (x * ny + jm) * nz + k;
      if (better(sc, idx, best, bidx)) { best = sc; bidx = idx; bp = pm; }
    }
  }
  wg_argmin_write(best, bidx, bp, res, fp.ok && score_order_ok(d, st) ? 0 : 1, out, tid, NT);
}

// ---------------------------------------------------------------------------
// Feasibility pass: does any valid offset of an orientation have zero
// collision with the occupied voxels?  Collision-only correlations with the
// binary tray B, two orientations per complex transform (orientation 2p in
// the real part, 2p+1 in the imaginary part), about a third of the scored
// search's work per orientation.  FFT(B) comes from the cached spectrum
// G1 = FFT(B + i h) of two real fields: FFT(B)[k] = (G1[k] + conj(G1[-k])) / 2.
// State fs[0] = max rounding residual (float bits), fs[1] unused,
// fs[2 + o] = 1 when orientation o has a zero-collision valid offset.
// ---------------------------------------------------------------------------
constexpr int kFeasHeader = 2;

// start of an item in the greedy loop: the scored search runs unless the
// feasibility pass decides otherwise; mode 0 never runs the pass, 1 runs it
// after an item that placed nothing, 2 always.
//
// Free-space certificate (item_st != nullptr, the tray field being the
// distance to the nearest occupied voxel): item_st->mx = rint(sqrt(e)) for
// the largest e such that a discrete ball {|y - c|^2 <= e - 1} (L1 ball for
// the L1 metric) lies inside the item.  A collision-free placement puts that
// ball on free voxels inside the tray, so at its centre the distance to the
// nearest occupied or outside voxel, min(field, wall distance), is
// >= item_st->mx; when the maximum of that quantity (tray_st->mxw) is smaller,
// no orientation is feasible (the ball is invariant under the cube's
// symmetries) and the item is recorded as not found with no further work.
__global__ void k_feas_begin(int* __restrict__ fs, int* __restrict__ cfs, int m, int mode,
                             int* __restrict__ ctl, const FieldStats* __restrict__ tray_st,
                             const FieldStats* __restrict__ item_st, PackOut* __restrict__ out,
                             int t) {
  if (ctl && ctl[CTL_HALT]) return;
  const int i = threadIdx.x;
  const bool cert = ctl && item_st && tray_st->mxw < item_st->mx;
  if (i == 0 && ctl) {
    ctl[CTL_PRUNE] = 0;
    ctl[CTL_COARSE_DONE] = 0;
    if (cert) {
      PackOut r{};
      r.so.found = 0;
      r.so.idx = INT_MAX;
      out[t] = r;
      ctl[CTL_SKIP_SEARCH] = 1;
      ctl[CTL_SKIP_REFRESH] = 1;
      ctl[CTL_FEAS] = 0;
    } else {
      ctl[CTL_SKIP_SEARCH] = 0;
      ctl[CTL_FEAS] = mode == 2 ? 1 : (mode == 1 ? ctl[CTL_SKIP_REFRESH] : 0);
    }
  }
  for (int j = i; j < kFeasHeader + m; j += blockDim.x) {
    fs[j] = 0;
    if (cfs) cfs[j] = 0;
  }
}

// padded complement of an item: out (ix+2, iy+2, iz+2) = 1 where the item has
// no voxel > 0 (its one-voxel border included), 0 on the item's voxels
__global__ void k_complement(const int* __restrict__ item, int ix, int iy, int iz,
                             int* __restrict__ out) {
  const int X = ix + 2, Y = iy + 2, Z = iz + 2;
  const int n = X * Y * Z;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int z = e % Z, y = (e / Z) % Y, x = e / (Y * Z);
    int v = 1;
    if (x >= 1 && y >= 1 && z >= 1 && x <= ix && y <= iy && z <= iz)
      v = item[((size_t)(x - 1) * iy + (y - 1)) * iz + (z - 1)] > 0 ? 0 : 1;
    out[e] = v;
  }
}

// pass A: grid (ceil(pz/T), py, n_chunks); chunks of orientation pairs
__global__ void __launch_bounds__(kPencilThreads)
k_feas_x(const float2* __restrict__ g1, const float2* __restrict__ ispec,
         const OrientDesc* __restrict__ od, float2* __restrict__ y1, int m, int chunk, int max_ex,
         int nx, AxisPlan px_, int py, int pz, int T, float scale, const float2* __restrict__ tw_x,
         const int* __restrict__ cf, int coarse_pass, const int* __restrict__ ctl) {
  if (feas_off(ctl) || (!coarse_pass && coarse_done(ctl))) return;
  // pairs the coarse pass proved infeasible in both orientations are skipped
  // here and in k_feas_yz (cf: coarse flags of this launch's orientations)
  const bool prune = !coarse_pass && cf && ctl && ctl[CTL_PRUNE];
  extern __shared__ float2 smem[];
  const int px = px_.n;
  float2* twx = smem;
  float2* buf = twx + px;
  // rank = (ky * tiles + tile) * chunks + chunk, as in k_corr_x
  const int rk = xcd_rank(), nch = gridDim.z, ntile = gridDim.x;
  const int rest = rk / nch;
  const int ky = rest / ntile, tid = threadIdx.x;
  const int kz0 = (rest - ky * ntile) * T;
  const int np = (m + 1) / 2;
  const int p0 = (rk - rest * nch) * chunk;
  const int p1 = min(np, p0 + chunk);
  load_table(twx, tw_x, px, tid, kPencilThreads);
  const int kym = ky ? py - ky : 0;
  float2 G[kPencilMaxEl];
#pragma unroll
  for (int s = 0; s < kPencilMaxEl; ++s) {
    const int e = tid + s * kPencilThreads;
    const int x = e / T, t = e - x * T;
    const int kz = kz0 + t;
    G[s] = make_float2(0.f, 0.f);
    if (e < px * T && kz < pz) {
      const float2 a = g1[((size_t)x * py + ky) * pz + kz];
      const int xm = x ? px - x : 0, kzm = kz ? pz - kz : 0;
      const float2 b = g1[((size_t)xm * py + kym) * pz + kzm];
      G[s] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    }
  }
  for (int p = p0; p < p1; ++p) {
    const int oa = 2 * p, ob = 2 * p + 1;
    // an excluded orientation has no spectra (k_item_slab_fwd skipped it):
    // its half of the pair is zero
    const bool use_a = !prune || cf[oa];
    const bool hasb = ob < m && (!prune || cf[ob]);
    if (!use_a && !hasb) continue;
    const OrientDesc da = od[oa];
    const OrientDesc db = od[hasb ? ob : oa];
    const float2* sa = ispec + (size_t)oa * max_ex * py * pz;
    const float2* sb = ispec + (size_t)(hasb ? ob : oa) * max_ex * py * pz;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kPencilMaxEl; ++s) {
      const int e = tid + s * kPencilThreads;
      if (e < px * T) {
        const int x = e / T, t = e - x * T;
        const int kz = kz0 + t;
        float2 va = make_float2(0.f, 0.f), vb = make_float2(0.f, 0.f);
        if (kz < pz) {
          if (use_a && x < da.ex) va = sa[((size_t)x * py + ky) * pz + kz];
          if (hasb && x < db.ex) vb = sb[((size_t)x * py + ky) * pz + kz];
        }
        buf[e] = make_float2(va.x - vb.y, va.y + vb.x);   // va + i vb
      }
    }
    __syncthreads();
    lds_fft<true, kPencilThreads, kPencilMaxEl, kPencilRadix>(buf, px_, T, T, 1, false, twx, tid);
#pragma unroll
    for (int s = 0; s < kPencilMaxEl; ++s) {
      const int e = tid + s * kPencilThreads;
      if (e < px * T) buf[e] = cscale(cmulc(G[s], buf[e]), scale);
    }
    __syncthreads();
    lds_fft<false, kPencilThreads, kPencilMaxEl, kPencilRadix>(buf, px_, T, T, 1, false, twx, tid);
    const int nvx = max(use_a ? da.nvx : 0, hasb ? db.nvx : 0);
    float2* d1 = y1 + (size_t)p * nx * py * pz;
    for (int e = tid; e < nvx * T; e += kPencilThreads) {
      const int x = e / T, t = e - x * T;
      const int kz = kz0 + t;
      if (kz < pz) d1[((size_t)x * py + ky) * pz + kz] = buf[e];
    }
  }
}

// pass B: grid (nx, n_pairs).  Inverse yz of one x slice of a pair, then
// flags orientations with a zero-collision valid offset in it (od and the
// pairs start at orientation o_base of the item's list).
__global__ void __launch_bounds__(kSlabThreads)
k_feas_yz(const float2* __restrict__ y1, const OrientDesc* __restrict__ od, int m, int nx,
          AxisPlan py_, AxisPlan pz_, const float2* __restrict__ tw_y,
          const float2* __restrict__ tw_z, int* __restrict__ fs, int o_base,
          const int* __restrict__ cf, int coarse_pass, unsigned long long* __restrict__ pairs,
          const int* __restrict__ ctl) {
  if (feas_off(ctl) || (!coarse_pass && coarse_done(ctl))) return;
  extern __shared__ float2 smem[];
  const int py = py_.n, pz = pz_.n;
  const int x = blockIdx.x, p = blockIdx.y, tid = threadIdx.x;
  const int oa = 2 * p, ob = 2 * p + 1;
  const bool prune = !coarse_pass && cf && ctl && ctl[CTL_PRUNE];
  const bool use_a = !prune || cf[oa];
  const bool hasb = ob < m && (!prune || cf[ob]);
  if (!use_a && !hasb) return;
  // pairs transformed (both passes), counted when timing asks for it
  if (pairs && x == 0 && tid == 0) atomicAdd(pairs, 1ULL);
  const OrientDesc da = od[oa];
  const OrientDesc db = od[hasb ? ob : oa];
  const bool xa = use_a && x < da.nvx, xb = hasb && x < db.nvx;
  if (!xa && !xb) return;
  float2* twy = smem;
  float2* twz = twy + py;
  float2* buf = twz + pz;
  load_table(twy, tw_y, py, tid, kSlabThreads);
  load_table(twz, tw_z, pz, tid, kSlabThreads);
  const float2* s1 = y1 + ((size_t)p * nx + x) * py * pz;
  stage_rows<kSlabThreads, kSlabMaxEl>(buf, s1, py * pz, pz, pz, tid);
  __syncthreads();
  const int nvy = max(xa ? da.nvy : 0, xb ? db.nvy : 0);
  lds_fft<false, kSlabThreads, kSlabMaxEl>(buf, py_, pz, pz, 1, false, twy, tid);
  lds_fft<false, kSlabThreads, kSlabMaxEl>(buf, pz_, 1, nvy, pz, true, twz, tid);
  float res = 0.f;
  int fa = 0, fb = 0;
  const float inv_pz = 1.0f / (float)pz;
#pragma unroll
  for (int s = 0; s < kSlabMaxEl; ++s) {
    const int e = tid + s * kSlabThreads;
    const int j = fdiv(e, pz, inv_pz), k = e - j * pz;
    if (e < py * pz) {
      const float2 v = buf[e];
      if (xa && j < da.nvy && k < da.nvz) {
        const float r = rintf(v.x);
        res = fmaxf(res, fabsf(v.x - r));
        fa |= r == 0.f;
      }
      if (xb && j < db.nvy && k < db.nvz) {
        const float r = rintf(v.y);
        res = fmaxf(res, fabsf(v.y - r));
        fb |= r == 0.f;
      }
    }
  }
  fa = __syncthreads_or(fa);
  fb = __syncthreads_or(fb);
  for (int o = 32; o > 0; o >>= 1) res = fmaxf(res, __shfl_xor(res, o));
  __shared__ float wr[kSlabThreads / 64];
  if ((tid & 63) == 0) wr[tid >> 6] = res;
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < kSlabThreads / 64; ++w) res = fmaxf(res, wr[w]);
    atomicMax(&fs[0], __float_as_int(res));
    if (fa) fs[kFeasHeader + o_base + oa] = 1;
    if (fb) fs[kFeasHeader + o_base + ob] = 1;
  }
}

// What the host-driven loops read after each synchronisation, gathered by
// one kernel into pinned host memory (no device-to-host copies).
constexpr int kMailLogits = 24;
constexpr int kMailFlags = kFeasHeader + 24;
struct Mailbox {
  int ctl[CTL_WORDS];
  FieldStats st;
  PackOut po;
  float logits[kMailLogits];
  int fs[kMailFlags];
  unsigned seq;              // written last: the post this content belongs to
};

// any pointer may be null (that part is left as it was).  Every thread makes
// its stores visible to the host before thread 0 stores seq, so a host that
// sees seq can read the rest.  One wavefront.
//
// gate (with fs, the flags of a screen): the control words of the next
// item's speculative screen, gate[CTL_SKIP_SEARCH] = 1 when some orientation
// of this screen is feasible.  That screen, enqueued before the host read
// these flags, then exits at once; it runs only when this item fits nowhere,
// i.e. when the tray stays as it is.
__global__ void k_post(const int* __restrict__ ctl, const PackOut* __restrict__ po,
                       const FieldStats* __restrict__ st, const float* __restrict__ logits,
                       const int* __restrict__ fs, int nfs, unsigned seq,
                       Mailbox* __restrict__ mb, int* __restrict__ gate) {
  const int i = threadIdx.x;
  if (ctl && i < CTL_WORDS) mb->ctl[i] = ctl[i];
  if (logits && i < kMailLogits) mb->logits[i] = logits[i];
  if (fs)
    for (int j = i; j < nfs && j < kMailFlags; j += blockDim.x) mb->fs[j] = fs[j];
  if (gate && fs) {
    int v = 0;
    for (int j = kFeasHeader + i; j < nfs; j += blockDim.x) v |= fs[j];
    const bool any = __any(v != 0);
    if (i < CTL_WORDS) gate[i] = i == CTL_SKIP_SEARCH && any ? 1 : 0;
  }
  if (i == 0) {
    if (st) mb->st = *st;
    if (po) mb->po = *po;
  }
  __threadfence_system();
  __syncthreads();
  if (i == 0) {
    __threadfence_system();
    *reinterpret_cast<volatile unsigned*>(&mb->seq) = seq;
  }
}

// ---------------------------------------------------------------------------
// Column test: an exact, conservative pre-pass of the feasibility pass on bit
// masks.  The tray is reduced to 2x2 column blocks at full z resolution (bit z
// of block (X, Y) set when any of its voxels at height z is occupied), each
// oriented item to its column core (bit z of core column (X, Y) set when the
// item has a voxel at (2X+dx, 2Y+dy, z) for all dx, dy in {-1, 0, 1}).  A
// placement at (2Qx + rx, 2Qy + ry, qz) fills every block (Qx+X, Qy+Y) at the
// heights qz + z of the core completely, so a collision-free placement needs
// those block bits clear; an orientation for which every (Qx, Qy, qz) of a
// valid offset meets a set bit has no collision-free offset.  Heights are
// 128-bit masks (tray z <= 128); per core column the set bits form runs
// [a, a+L), and the blocked qz of a run are (window-OR of length L of the
// block's mask) >> a.
// ---------------------------------------------------------------------------
struct M128 {
  unsigned long long lo, hi;
};
__device__ __forceinline__ M128 m_or(M128 a, M128 b) { return {a.lo | b.lo, a.hi | b.hi}; }
__device__ __forceinline__ M128 m_shr(M128 a, int s) {   // bit i of the result = bit i+s of a
  if (s == 0) return a;
  if (s >= 64) return {a.hi >> (s - 64), 0ULL};
  return {(a.lo >> s) | (a.hi << (64 - s)), a.hi >> s};
}
// bit i = OR of bits i .. i+L-1 of t (L >= 1)
__device__ __forceinline__ M128 m_window(M128 t, int L) {
  M128 w = t;
  int len = 1;
  while (2 * len <= L) {
    w = m_or(w, m_shr(w, len));
    len *= 2;
  }
  if (len < L) w = m_or(w, m_shr(w, L - len));
  return w;
}
// bits [lo, hi) of a 64-bit word, 0 <= lo, hi <= 64
__device__ __forceinline__ unsigned long long bits64(int lo, int hi) {
  if (hi <= lo) return 0ULL;
  const unsigned long long upper = hi >= 64 ? ~0ULL : ((1ULL << hi) - 1ULL);
  const unsigned long long lower = lo >= 64 ? ~0ULL : ((1ULL << lo) - 1ULL);
  return upper & ~lower;
}
__device__ __forceinline__ int m_ctz(M128 a) {
  return a.lo ? __ffsll((long long)a.lo) - 1 : (a.hi ? 64 + __ffsll((long long)a.hi) - 1 : 128);
}

// one wavefront per 2x2 column block: lanes over z, ballots form the masks.
// cols holds kColLevels planes of cx*cy masks: plane k is the window-OR of
// length 2^k of the block masks (plane 0 the masks themselves).
constexpr int kColLevels = 8;
__global__ void __launch_bounds__(256)
k_tray_columns(const int* __restrict__ tray, int nx, int ny, int nz, M128* __restrict__ cols,
               int cx, int cy, int* __restrict__ top, const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) *top = 0;   // k_tile_heights raises it
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (col >= cx * cy) return;
  const int X = col / cy, Y = col - X * cy;
  M128 m{0ULL, 0ULL};
  for (int w = 0; w < 2; ++w) {
    const int z = w * 64 + lane;
    bool occ = false;
    for (int u = 0; u < 2; ++u)
      for (int v = 0; v < 2; ++v) {
        const int x = 2 * X + u, y = 2 * Y + v;
        if (x < nx && y < ny && z < nz && tray[((size_t)x * ny + y) * nz + z] != 0) occ = true;
      }
    const unsigned long long b = __ballot(occ);
    if (w == 0) m.lo = b; else m.hi = b;
  }
  if (lane < kColLevels) {
    M128 w = m;
    for (int k = 0; k < lane; ++k) w = m_or(w, m_shr(w, 1 << k));
    cols[(size_t)lane * cx * cy + col] = w;
  }
}

// full-resolution column masks of the tray (bit z of column (x, y) set when
// voxel (x, y, z) is occupied) with their window-OR levels, tf[level][x*ny+y];
// one wavefront per column
__global__ void __launch_bounds__(256)
k_tray_fine(const int* __restrict__ tray, int nx, int ny, int nz, M128* __restrict__ tf,
            const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (col >= nx * ny) return;
  M128 m{0ULL, 0ULL};
  for (int w = 0; w < 2; ++w) {
    const int z = w * 64 + lane;
    const bool occ = z < nz && tray[(size_t)col * nz + z] != 0;
    const unsigned long long b = __ballot(occ);
    if (w == 0) m.lo = b; else m.hi = b;
  }
  if (lane < kColLevels) {
    M128 w = m;
    for (int k = 0; k < lane; ++k) w = m_or(w, m_shr(w, 1 << k));
    tf[(size_t)lane * nx * ny + col] = w;
  }
}

// tile heights of the tray, after k_tray_fine: ht[X * ty + Y] = max over the
// columns (x, y) of tile X = x / kSkyTile, Y = y / kSkyTile of 1 + the highest
// occupied z (0 for empty columns); one wavefront per tile, a lane per column.
// ht[tx * ty] is the tray top, the largest of them (reset by k_tray_columns).
constexpr int kSkyTile = 8;
__global__ void __launch_bounds__(256)
k_tile_heights(const M128* __restrict__ tf, int nx, int ny, int* __restrict__ ht, int tx, int ty,
               const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tx * ty) return;
  const int X = t / ty, Y = t - X * ty;
  const int x = X * kSkyTile + (lane >> 3), y = Y * kSkyTile + (lane & 7);
  int top = 0;
  if (x < nx && y < ny) {
    const M128 m = tf[(size_t)x * ny + y];   // level 0: the column's own mask
    top = m.hi ? 128 - __clzll((long long)m.hi) : (m.lo ? 64 - __clzll((long long)m.lo) : 0);
  }
  for (int s = 1; s < 64; s <<= 1) top = max(top, __shfl_xor(top, s, 64));
  if (lane == 0) {
    ht[t] = top;
    if (top > load_agent(&ht[tx * ty])) atomicMax(&ht[tx * ty], top);   // (it only grows)
  }
}

// The tray masks after one placement.  A placement only adds occupied voxels
// inside the placed item's box, so only the columns under the box change:
// their full-resolution masks and the 2x2 block masks over them are rebuilt
// from the tray, and the tile heights, which can only grow, take the new
// column tops by atomic maximum.  Waves [0, bound^2) take the box's columns,
// waves [bound^2, bound^2 + nb1^2) its 2x2 blocks (nb1 = bound / 2 + 1);
// bound >= the box side along x and y (the grid is sized for it).  The
// placement is the device's search outcome (so != nullptr: orientation
// od[so->orient] at so->idx) or orientation od[orient] at (qx, qy).
__global__ void __launch_bounds__(256)
k_tray_masks_inc(const int* __restrict__ tray, int nx, int ny, int nz, M128* __restrict__ tf,
                 M128* __restrict__ cols, int cx, int cy, int* __restrict__ ht, int tx, int ty,
                 const SearchOut* __restrict__ so, const OrientDesc* __restrict__ od, int orient,
                 int qx, int qy, int bound, const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  OrientDesc d;
  if (so) {
    const SearchOut s = *so;
    if (!s.found) return;
    d = od[s.orient];
    qy = (s.idx / nz) % ny;
    qx = s.idx / (ny * nz);
  } else {
    d = od[orient];
  }
  if (d.hx < 0 || d.hy < 0 || d.hz < 0) return;
  const int ex = d.hx + 1, ey = d.hy + 1;   // columns [qx, qx + ex) x [qy, qy + ey)
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int ncol = bound * bound, nb1 = bound / 2 + 1;
  if (w < ncol) {
    const int u = w / bound, v = w - u * bound;
    const int x = qx + u, y = qy + v;
    if (u >= ex || v >= ey || x >= nx || y >= ny) return;
    const size_t col = (size_t)x * ny + y;
    M128 m{0ULL, 0ULL};
    for (int h = 0; h < 2; ++h) {
      const int z = h * 64 + lane;
      const bool occ = z < nz && tray[col * nz + z] != 0;
      const unsigned long long b = __ballot(occ);
      if (h == 0) m.lo = b; else m.hi = b;
    }
    if (lane < kColLevels) {
      M128 wv = m;
      for (int k = 0; k < lane; ++k) wv = m_or(wv, m_shr(wv, 1 << k));
      tf[(size_t)lane * nx * ny + col] = wv;
    }
    if (lane == 0) {
      // heights only grow, so a value read here (device-coherent) is at most
      // the current one: a top not above it needs no atomic (the tray top is
      // one word that every column of the box would otherwise update)
      const int top = m.hi ? 128 - __clzll((long long)m.hi) : (m.lo ? 64 - __clzll((long long)m.lo) : 0);
      int* tile = &ht[(x / kSkyTile) * ty + y / kSkyTile];
      if (top > load_agent(tile)) atomicMax(tile, top);
      if (top > load_agent(&ht[tx * ty])) atomicMax(&ht[tx * ty], top);
    }
    return;
  }
  const int b = w - ncol;
  if (b >= nb1 * nb1) return;
  const int X = qx / 2 + b / nb1, Y = qy / 2 + (b - (b / nb1) * nb1);
  if (X > (qx + ex - 1) / 2 || Y > (qy + ey - 1) / 2 || X >= cx || Y >= cy) return;
  M128 m{0ULL, 0ULL};
  for (int h = 0; h < 2; ++h) {
    const int z = h * 64 + lane;
    bool occ = false;
    for (int u = 0; u < 2; ++u)
      for (int v = 0; v < 2; ++v) {
        const int x = 2 * X + u, y = 2 * Y + v;
        if (x < nx && y < ny && z < nz && tray[((size_t)x * ny + y) * nz + z] != 0) occ = true;
      }
    const unsigned long long bb = __ballot(occ);
    if (h == 0) m.lo = bb; else m.hi = bb;
  }
  if (lane < kColLevels) {
    M128 wv = m;
    for (int k = 0; k < lane; ++k) wv = m_or(wv, m_shr(wv, 1 << k));
    cols[(size_t)lane * cx * cy + (size_t)X * cy + Y] = wv;
  }
}

// full-resolution column masks of m oriented items (bit z set when the item
// has a voxel > 0 at (x, y, z)), im[(o * mx + x) * my + y] for x < ex, y < ey
// (the only entries the column test reads); one wavefront per column, lanes
// over z
__global__ void __launch_bounds__(256)
k_item_columns(const int* __restrict__ items, const OrientDesc* __restrict__ od, int m,
               M128* __restrict__ im, int mx, int my, const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int o = blockIdx.y;
  if (o >= m || col >= mx * my) return;
  const OrientDesc d = od[o];
  const int x = col / my, y = col - x * my;
  if (x >= d.ex || y >= d.ey) return;
  const int* base = items + d.item_off;
  M128 r{0ULL, 0ULL};
  for (int w = 0; w < 2; ++w) {
    const int z = w * 64 + lane;
    const bool v = x < d.ex && y < d.ey && z < d.ez && base[oriented_index(d, x, y, z)] > 0;
    const unsigned long long b = __ballot(v);
    if (w == 0) r.lo = b; else r.hi = b;
  }
  if (lane == 0) im[((size_t)o * mx + x) * my + y] = r;
}

// core columns of m oriented items, cm[(o * ccx + X) * ccy + Y] for
// X < (ex + 1) / 2, Y < (ey + 1) / 2 (the entries the column test reads); one
// wavefront per core column, lanes over z
__global__ void __launch_bounds__(256)
k_core_columns(const int* __restrict__ items, const OrientDesc* __restrict__ od, int m,
               M128* __restrict__ cm, int ccx, int ccy, const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int o = blockIdx.y;
  if (o >= m || col >= ccx * ccy) return;
  const OrientDesc d = od[o];
  const int X = col / ccy, Y = col - X * ccy;
  if (X >= (d.ex + 1) / 2 || Y >= (d.ey + 1) / 2) return;
  const int* base = items + d.item_off;
  M128 r{0ULL, 0ULL};
  for (int w = 0; w < 2; ++w) {
    const int z = w * 64 + lane;
    bool full = z < d.ez && 2 * X + 1 < d.ex && 2 * Y + 1 < d.ey && X >= 0;
    for (int dx = -1; dx <= 1 && full; ++dx)
      for (int dy = -1; dy <= 1 && full; ++dy) {
        const int x = 2 * X + dx, y = 2 * Y + dy;
        if (x < 0 || y < 0 || base[oriented_index(d, x, y, z)] <= 0) full = false;
      }
    const unsigned long long b = __ballot(full);
    if (w == 0) r.lo = b; else r.hi = b;
  }
  if (lane == 0) cm[((size_t)o * ccx + X) * ccy + Y] = r;
}

// grid (chunks, m).  Block (b, o): orientation o, chunks of 64 offsets
// q = 64 c + thread / 4 over the valid (Qx, Qy) rectangle of the column
// blocks, c = b, b + chunks, ..., four lanes per offset.  The core's runs come longest first
// from k_run_lists; lane j of an offset ORs the heights that runs j,
// j+4, ... block (window of length L = level 2^k, plus a shifted copy when L
// is not a power of two), the four lanes combine after rounds of growing size
// and stop once every valid height is blocked.
// Without the full-resolution stage (tf == nullptr) an offset with a height
// left flags the orientation as possibly feasible.  With it, such an offset's
// four full-resolution offsets (2 Qx + rx, 2 Qy + ry) become candidates that
// start from the heights the blocks already exclude and OR in every run of
// the item's own columns (masks im); one with a height left proves the
// orientation feasible, and the flag then means exactly that: an orientation
// left unflagged has no collision-free offset.  A block starts a chunk only
// while the orientation is unflagged.
constexpr int kColRuns = 2048;
constexpr int kRunsLds = 256;   // run-list entries the column test keeps in LDS
constexpr int kFineRuns = 4096;

// every run [a, a+L) of set bits of a column mask, passed to f(a, L)
template <typename F>
__device__ __forceinline__ void for_runs(M128 k, F f) {
  while (k.lo | k.hi) {
    const int a = m_ctz(k);
    const M128 rest = m_shr(k, a);
    const int L = m_ctz(M128{~rest.lo, ~rest.hi});
    f(a, L);
    k.lo &= ~bits64(min(a, 64), min(a + L, 64));
    k.hi &= ~bits64(max(a - 64, 0), max(a + L - 64, 0));
  }
}

// a run in an LDS list: x (7 bits), y (8 bits), start a (7 bits), length - 1 (7 bits)
__device__ __forceinline__ unsigned pack_run(int x, int y, int a, int L) {
  return (unsigned)x | ((unsigned)y << 7) | ((unsigned)a << 15) | ((unsigned)(L - 1) << 22);
}

// heights blocked by a run [a, a+L) at column c of the plane-strided levels lv
__device__ __forceinline__ M128 run_blocks_at(const M128* __restrict__ lv, size_t plane, size_t c,
                                              int a, int L) {
  const int k = 31 - __clz(L);
  M128 w = lv[(size_t)k * plane + c];
  const int p = 1 << k;
  if (L > p) w = m_or(w, m_shr(w, L - p));
  return m_shr(w, a);
}

__device__ __forceinline__ bool m_any_free(M128 bad, unsigned long long vlo,
                                           unsigned long long vhi) {
  return ((~bad.lo & vlo) | (~bad.hi & vhi)) != 0ULL;
}

// Run lists of m oriented items for the column test, one block per
// orientation: the runs of the core masks (cm) into runs[o * kColRuns ...] and
// of the item's own column masks (im, when fine) into fruns[o * kFineRuns ...],
// each sorted longest first by a counting sort over the 128 lengths (a long
// run blocks the most heights, so most offsets are refuted by the first few).
// nrun[2 o] and nrun[2 o + 1] receive the list lengths; a length above the
// capacity (or kColRuns + 1 when the core's x or y extent exceeds the packed
// fields) means the list was not written and the test reads the masks.
constexpr int kListThreads = 1024;
constexpr int kListHold = 4;   // masks per thread kept in registers between the passes
__global__ void __launch_bounds__(kListThreads)
k_run_lists(const M128* __restrict__ cm, int ccx, int ccy, const M128* __restrict__ im, int mx,
            int my, const OrientDesc* __restrict__ od, int m, int fine_on,
            unsigned* __restrict__ runs, unsigned* __restrict__ fruns, int* __restrict__ nrun,
            const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  const int o = blockIdx.x, tid = threadIdx.x;
  if (o >= m) return;
  const OrientDesc d = od[o];
  __shared__ int cnt[2][128];   // per run length: count, then next slot
  __shared__ int ntot[2];
  if (tid < 256) cnt[tid >> 7][tid & 127] = 0;
  const int kx = (d.ex + 1) / 2, ky = (d.ey + 1) / 2;
  const M128* core = cm + (size_t)o * ccx * ccy;
  const M128* own = im + (size_t)o * mx * my;
  const bool clist = kx <= 128 && ky <= 256;   // the packed run's x and y fields
  const bool fine = fine_on && d.ex <= 128 && d.ey <= 256;
  const int nc = clist ? kx * ky : 0, nf = fine ? d.ex * d.ey : 0;
  const M128 zero{0ULL, 0ULL};
  M128 hc[kListHold], hf[kListHold];
#pragma unroll
  for (int i = 0; i < kListHold; ++i) {
    const int c = tid + i * kListThreads;
    hc[i] = c < nc ? core[(c / ky) * ccy + (c % ky)] : zero;
    hf[i] = c < nf ? own[(c / d.ey) * my + (c % d.ey)] : zero;
  }
  __syncthreads();
  auto count0 = [&](int, int L) { atomicAdd(&cnt[0][L - 1], 1); };
  auto count1 = [&](int, int L) { atomicAdd(&cnt[1][L - 1], 1); };
#pragma unroll
  for (int i = 0; i < kListHold; ++i) {
    for_runs(hc[i], count0);
    for_runs(hf[i], count1);
  }
  for (int c = tid + kListHold * kListThreads; c < nc; c += kListThreads)
    for_runs(core[(c / ky) * ccy + (c % ky)], count0);
  for (int c = tid + kListHold * kListThreads; c < nf; c += kListThreads)
    for_runs(own[(c / d.ey) * my + (c % d.ey)], count1);
  __syncthreads();
  if (tid < 128) {
    // wave w scans list w: lane l holds lengths 128 - 2l and 127 - 2l
    const int w = tid >> 6, lane = tid & 63;
    const int v0 = cnt[w][127 - 2 * lane], v1 = cnt[w][126 - 2 * lane];
    int incl = v0 + v1;
    for (int s2 = 1; s2 < 64; s2 <<= 1) {
      const int u = __shfl_up(incl, s2, 64);
      if (lane >= s2) incl += u;
    }
    const int excl = incl - v0 - v1;
    cnt[w][127 - 2 * lane] = excl;
    cnt[w][126 - 2 * lane] = excl + v0;
    if (lane == 63) ntot[w] = incl;
  }
  __syncthreads();
  unsigned* cr = runs + (size_t)o * kColRuns;
  unsigned* fr = fruns + (size_t)o * kFineRuns;
  const bool wc = ntot[0] <= kColRuns, wf = ntot[1] <= kFineRuns;
#pragma unroll
  for (int i = 0; i < kListHold; ++i) {
    const int c = tid + i * kListThreads;
    if (wc && c < nc) {
      const int X = c / ky, Y = c - X * ky;
      for_runs(hc[i], [&](int a, int L) { cr[atomicAdd(&cnt[0][L - 1], 1)] = pack_run(X, Y, a, L); });
    }
    if (wf && c < nf) {
      const int x = c / d.ey, y = c - x * d.ey;
      for_runs(hf[i], [&](int a, int L) { fr[atomicAdd(&cnt[1][L - 1], 1)] = pack_run(x, y, a, L); });
    }
  }
  if (wc)
    for (int c = tid + kListHold * kListThreads; c < nc; c += kListThreads) {
      const int X = c / ky, Y = c - X * ky;
      for_runs(core[X * ccy + Y],
               [&](int a, int L) { cr[atomicAdd(&cnt[0][L - 1], 1)] = pack_run(X, Y, a, L); });
    }
  if (wf)
    for (int c = tid + kListHold * kListThreads; c < nf; c += kListThreads) {
      const int x = c / d.ey, y = c - x * d.ey;
      for_runs(own[x * my + y],
               [&](int a, int L) { fr[atomicAdd(&cnt[1][L - 1], 1)] = pack_run(x, y, a, L); });
    }
  if (tid == 0) {
    nrun[2 * o] = clist ? ntot[0] : kColRuns + 1;
    nrun[2 * o + 1] = ntot[1];
  }
}

// Line masks of base items, for the column test's item masks: per item three
// tables of M128 at lines + line_off, bit k set when the voxel at coordinate k
// along the line is > 0 (lines of at most 128 voxels; a longer axis gets no
// table and is never a fitting orientation's z axis when the tray's z is at
// most 128): lines along z indexed u0 * by + u1, then along y indexed
// u0 * bz + u2, then along x indexed u1 * bz + u2.  Grid (ceil(max lines /
// 256), items), a thread per line.
struct ItemLines {
  int off;        // voxel offset of the base item in the item buffer
  int bx, by, bz;
  int line_off;   // first M128 of its tables
};
__global__ void __launch_bounds__(256)
k_item_lines(const int* __restrict__ items, const ItemLines* __restrict__ il, M128* __restrict__ lines) {
  const ItemLines L = il[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nz = L.bx * L.by, ny = L.bx * L.bz, nxl = L.by * L.bz;
  if (i >= nz + ny + nxl) return;
  const int* b = items + L.off;
  unsigned long long lo = 0ULL, hi = 0ULL;
  int start, step, len;
  if (i < nz) {
    const int u0 = i / L.by, u1 = i - u0 * L.by;
    start = (u0 * L.by + u1) * L.bz; step = 1; len = L.bz;
  } else if (i < nz + ny) {
    const int j = i - nz, u0 = j / L.bz, u2 = j - u0 * L.bz;
    start = u0 * L.by * L.bz + u2; step = L.bz; len = L.by;
  } else {
    const int j = i - nz - ny, u1 = j / L.bz, u2 = j - u1 * L.bz;
    start = u1 * L.bz + u2; step = L.by * L.bz; len = L.bx;
  }
  if (len <= 128)
    for (int k = 0; k < len; ++k) {
      const unsigned long long v = b[start + k * step] > 0 ? 1ULL : 0ULL;
      if (k < 64) lo |= v << k; else hi |= v << (k - 64);
    }
  lines[L.line_off + i] = M128{lo, hi};
}

// bit i of the result = bit n - 1 - i of a (0 < n <= 128; bits >= n of a are 0)
__device__ __forceinline__ M128 m_reverse(M128 a, int n) {
  const M128 r{__brevll(a.hi), __brevll(a.lo)};
  return m_shr(r, 128 - n);
}

// Item column masks, core masks and sorted run lists in one launch, for
// items whose footprint ex * ey is at most kMaskCols (the masks are built in
// LDS): one block per orientation.  Thread per item column: bit z of
// im[(o * mx + x) * my + y] is set when the item has a voxel > 0 at (x, y, z);
// core (X, Y) is the AND of the 3 x 3 columns around (2 X, 2 Y) (0 outside the
// item), written to cm[(o * ccx + X) * ccy + Y]; then the run lists of both,
// as k_run_lists writes them.  Same outputs as k_item_columns, k_core_columns
// and k_run_lists together.  With line masks (lines != nullptr, d.line_off >=
// 0) an item column is read as one line of the base item (reversed when the z
// axis is flipped), else gathered voxel by voxel.
constexpr int kMaskThreads = 1024;
constexpr int kMaskCols = 4096;
__global__ void __launch_bounds__(kMaskThreads)
k_item_masks(const int* __restrict__ items, const M128* __restrict__ lines,
             const OrientDesc* __restrict__ od, int m,
             M128* __restrict__ cm, int ccx, int ccy, M128* __restrict__ im, int mx, int my,
             unsigned* __restrict__ runs, unsigned* __restrict__ fruns, int* __restrict__ nrun,
             int* __restrict__ cfs, int* __restrict__ ctl, const int* __restrict__ pre,
             int* __restrict__ fs = nullptr, int begin = 0) {
  // begin: the first kernel of a greedy-loop item (in place of k_feas_begin
  // with the feasibility pass off): block 0 clears the item's control words
  // and fs, and only a halt stops the kernel (the skip word still belongs to
  // the previous item)
  if (begin ? (ctl && ctl[CTL_HALT]) : search_off(ctl)) return;
  const int o = blockIdx.x, tid = threadIdx.x;
  if (begin && o == 0) {
    if (tid == 0) {
      ctl[CTL_PRUNE] = 0;
      ctl[CTL_COARSE_DONE] = 0;
      ctl[CTL_SKIP_SEARCH] = 0;
      ctl[CTL_FEAS] = 0;
    }
    for (int j = tid; j < kFeasHeader + m && fs; j += blockDim.x) fs[j] = 0;
  }
  if (o >= m) return;
  if (cfs && tid == 0) {   // clears the column test's flags (cfs: null when already cleared)
    cfs[kFeasHeader + o] = 0;
    if (o == 0)
      for (int h = 0; h < kFeasHeader; ++h) cfs[h] = 0;
  }
  if (pre && pre[o] == 0) return;   // infeasible on an earlier tray: no test (k_column_test)
  const OrientDesc d = od[o];
  __shared__ M128 own[kMaskCols];
  __shared__ M128 core[kMaskCols / 2];   // ceil(ex/2) * ceil(ey/2) <= 2048 when ex * ey <= 4096
  __shared__ int cnt[2][128];
  __shared__ int ntot[2];
  if (tid < 256) cnt[tid >> 7][tid & 127] = 0;
  const int ne = d.ex * d.ey;
  const int* base = items + d.item_off;
  if (lines && d.line_off >= 0 && d.ez <= 128) {
    // the column is a line of the base item along base axis A = perm[2]
    const int A = d.perm[2], p = A == 0 ? 1 : 0, q = A == 2 ? 1 : 2;   // p < q: the others
    const int bd[3] = {d.bx, d.by, d.bz};
    const M128* tab = lines + d.line_off +
                      (A == 2 ? 0 : (A == 1 ? d.bx * d.by : d.bx * d.by + d.bx * d.bz));
    for (int c = tid; c < ne; c += kMaskThreads) {
      const int x = c / d.ey, y = c - x * d.ey;
      int u[3] = {0, 0, 0};
      u[d.perm[0]] = d.flip[0] ? bd[d.perm[0]] - 1 - x : x;
      u[d.perm[1]] = d.flip[1] ? bd[d.perm[1]] - 1 - y : y;
      M128 r = tab[u[p] * bd[q] + u[q]];
      if (d.flip[2]) r = m_reverse(r, d.ez);
      own[c] = r;
      im[((size_t)o * mx + x) * my + y] = r;
    }
  } else
  for (int c = tid; c < ne; c += kMaskThreads) {
    const int x = c / d.ey, y = c - x * d.ey;
    const int i0 = oriented_index(d, x, y, 0);
    const int dz = d.ez > 1 ? oriented_index(d, x, y, 1) - i0 : 0;
    unsigned long long lo = 0ULL, hi = 0ULL;
#pragma unroll 8
    for (int z = 0; z < d.ez; ++z) {
      const unsigned long long b = base[i0 + z * dz] > 0 ? 1ULL : 0ULL;
      if (z < 64) lo |= b << z; else hi |= b << (z - 64);
    }
    own[c] = M128{lo, hi};
    im[((size_t)o * mx + x) * my + y] = M128{lo, hi};
  }
  __syncthreads();
  const int kx = (d.ex + 1) / 2, ky = (d.ey + 1) / 2;
  for (int c = tid; c < kx * ky; c += kMaskThreads) {
    const int X = c / ky, Y = c - X * ky;
    M128 r{0ULL, 0ULL};
    if (2 * X + 1 < d.ex && 2 * Y + 1 < d.ey && X > 0 && Y > 0) {
      r = M128{~0ULL, ~0ULL};
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
          const M128 v = own[(2 * X + dx) * d.ey + (2 * Y + dy)];
          r.lo &= v.lo;
          r.hi &= v.hi;
        }
    }
    core[c] = r;
    cm[((size_t)o * ccx + X) * ccy + Y] = r;
  }
  __syncthreads();
  const bool clist = kx <= 128 && ky <= 256;   // the packed run's x and y fields
  const bool fine = d.ex <= 128 && d.ey <= 256;
  if (clist)
    for (int c = tid; c < kx * ky; c += kMaskThreads)
      for_runs(core[c], [&](int, int L) { atomicAdd(&cnt[0][L - 1], 1); });
  if (fine)
    for (int c = tid; c < ne; c += kMaskThreads)
      for_runs(own[c], [&](int, int L) { atomicAdd(&cnt[1][L - 1], 1); });
  __syncthreads();
  if (tid < 128) {
    // wave w scans list w: lane l holds lengths 128 - 2l and 127 - 2l
    const int w = tid >> 6, lane = tid & 63;
    const int v0 = cnt[w][127 - 2 * lane], v1 = cnt[w][126 - 2 * lane];
    int incl = v0 + v1;
    for (int s2 = 1; s2 < 64; s2 <<= 1) {
      const int u = __shfl_up(incl, s2, 64);
      if (lane >= s2) incl += u;
    }
    const int excl = incl - v0 - v1;
    cnt[w][127 - 2 * lane] = excl;
    cnt[w][126 - 2 * lane] = excl + v0;
    if (lane == 63) ntot[w] = incl;
  }
  __syncthreads();
  unsigned* cr = runs + (size_t)o * kColRuns;
  unsigned* fr = fruns + (size_t)o * kFineRuns;
  if (clist && ntot[0] <= kColRuns)
    for (int c = tid; c < kx * ky; c += kMaskThreads) {
      const int X = c / ky, Y = c - X * ky;
      for_runs(core[c], [&](int a, int L) { cr[atomicAdd(&cnt[0][L - 1], 1)] = pack_run(X, Y, a, L); });
    }
  if (fine && ntot[1] <= kFineRuns)
    for (int c = tid; c < ne; c += kMaskThreads) {
      const int x = c / d.ey, y = c - x * d.ey;
      for_runs(own[c], [&](int a, int L) { fr[atomicAdd(&cnt[1][L - 1], 1)] = pack_run(x, y, a, L); });
    }
  if (tid == 0) {
    nrun[2 * o] = clist ? ntot[0] : kColRuns + 1;
    nrun[2 * o + 1] = fine ? ntot[1] : 0;
  }
}

// OR over the G lanes of an aligned lane group (every lane of the group active)
template <int G>
__device__ __forceinline__ M128 group_or(M128 v) {
  for (int s = 1; s < G; s <<= 1) {
    v.lo |= __shfl_xor(v.lo, s, G);
    v.hi |= __shfl_xor(v.hi, s, G);
  }
  return v;
}
template <int G>
__device__ __forceinline__ int group_max(int v) {
  for (int s = 1; s < G; s <<= 1) v = max(v, __shfl_xor(v, s, G));
  return v;
}

template <int G>
__global__ void __launch_bounds__(256)
k_column_test(const M128* __restrict__ cols, int cx, int cy, const M128* __restrict__ cm,
              int ccx, int ccy, const M128* __restrict__ tf, int nx, int ny,
              const M128* __restrict__ im, int mx, int my, const OrientDesc* __restrict__ od,
              int m, const unsigned* __restrict__ run_lists, const unsigned* __restrict__ fine_lists,
              const int* __restrict__ nrun, const int* __restrict__ ht, int ty,
              int* __restrict__ cfs, const int* __restrict__ ctl, int tx,
              const int* __restrict__ pre, int* __restrict__ cnt,
              const int* __restrict__ od_item, int cap) {
  if (search_off(ctl)) return;
  const int o = blockIdx.y, tid = threadIdx.x;
  if (o >= m) return;
  // pre (nullable): flags of the same orientations on an earlier tray of the
  // run; one infeasible there is infeasible here (placements only add
  // occupied voxels) and keeps its flag 0 with no work
  if (pre && pre[o] == 0) return;
  const OrientDesc d = od[o];
  const int nqx = d.nvx > 0 ? (d.nvx - 1) / 2 + 1 : 0, nqy = d.nvy > 0 ? (d.nvy - 1) / 2 + 1 : 0;
  const int nq = nqx * nqy;
  const int nblk = (nq + 63) / 64;
  if ((int)blockIdx.x >= nblk || d.nvz <= 0) return;
  volatile int* flag = cfs + kFeasHeader + o;
  // A batched screen of several items (cnt != nullptr, cnt[od_item[o]]
  // counts the orientations of o's item proven feasible) needs no more tests
  // of an item once cap of its orientations are: an orientation left
  // untested then reads 2 (unknown), never 0 (infeasible).
  volatile int* icnt = cnt ? cnt + od_item[o] : nullptr;
  auto mark_feasible = [&]() {
    if (atomicExch((int*)flag, 1) != 1 && icnt) atomicAdd((int*)icnt, 1);
  };
  if (icnt && *icnt >= cap) {
    if (tid == 0) atomicCAS((int*)flag, 0, 2);
    return;
  }
  // whole-tray proof first: when no tray column is higher than H (the tray
  // top, ht[tx * ty]), the item's box at height H is empty at offset (0, 0),
  // a valid offset when H < nvz
  if (ht[tx * ty] < d.nvz) {
    if (tid == 0) mark_feasible();
    return;
  }
  __shared__ unsigned cand_list[256];
  __shared__ M128 cbad[64];
  __shared__ int ncand, skip;
  const int kx = (d.ex + 1) / 2, ky = (d.ey + 1) / 2;
  const M128* core = cm + (size_t)o * ccx * ccy;
  const bool fine = tf != nullptr && d.ex <= 128 && d.ey <= 256;
  const M128* own = im + (size_t)o * mx * my;
  // sorted run lists of this orientation (k_run_lists); the lanes of a quad
  // or group read consecutive entries, shared by every quad of the wave
  const unsigned* runs = run_lists + (size_t)o * kColRuns;
  const unsigned* fruns = fine_lists + (size_t)o * kFineRuns;
  const int nr = nrun[2 * o], nf = nrun[2 * o + 1];
  // the first kRunsLds entries of each list in LDS: every offset of the block
  // reads the same lists
  __shared__ unsigned sruns[kRunsLds], sfruns[kRunsLds];
  for (int i = tid; i < kRunsLds; i += 256) {
    if (i < nr && nr <= kColRuns) sruns[i] = runs[i];
    if (i < nf && nf <= kFineRuns) sfruns[i] = fruns[i];
  }
  __syncthreads();
  const int nv = d.nvz;
  const unsigned long long vlo = nv >= 64 ? ~0ULL : ((1ULL << nv) - 1ULL);
  const unsigned long long vhi = nv >= 128 ? ~0ULL : (nv > 64 ? ((1ULL << (nv - 64)) - 1ULL) : 0ULL);
  const size_t plane = (size_t)cx * cy;
  const size_t fplane = (size_t)nx * ny;
  const int j = tid & 3, gl = tid & (G - 1);
  // chunks of 64 offsets: blockIdx.x, blockIdx.x + gridDim.x, ...
  for (int qb = blockIdx.x; qb < nblk; qb += gridDim.x) {
    const int q0 = qb * 64;
    __syncthreads();   // the previous chunk's candidate list is consumed
    if (tid == 0) {
      ncand = 0;
      // one read, so that the whole block takes the same branch
      int sk = *flag == 1 ? 1 : 0;
      if (!sk && icnt && *icnt >= cap) {
        atomicCAS((int*)flag, 0, 2);
        sk = 1;
      }
      skip = sk;
    }
    __syncthreads();
    if (skip) return;
    const int q = q0 + (tid >> 2);
    const bool active = q < nq;
    const int Qx = active ? q / nqy : 0, Qy = active ? q - Qx * nqy : 0;
    M128 bad{0ULL, 0ULL};
    if (nr <= kColRuns) {
      // rounds of 4, 8, 16, 32, then 64 runs per offset; all four lanes of an
      // offset run the same rounds (the combine uses shuffles within the quad)
      bool done = !active;
      for (int r0 = 0, step = 4; r0 < nr; r0 += step, step = min(2 * step, 64)) {
        if (!done) {
          const int r1 = min(nr, r0 + step);
          for (int r = r0 + j; r < r1; r += 4) {
            const unsigned e = r < kRunsLds ? sruns[r] : runs[r];
            const int X = e & 127, Y = (e >> 7) & 255, ra = (e >> 15) & 127;
            const int L = ((e >> 22) & 127) + 1;
            bad = m_or(bad, run_blocks_at(cols, plane, (size_t)(Qx + X) * cy + (Qy + Y), ra, L));
          }
        }
        M128 u = bad;
        for (int s2 = 1; s2 < 4; s2 <<= 1) {
          u.lo |= __shfl_xor(u.lo, s2, 64);
          u.hi |= __shfl_xor(u.hi, s2, 64);
        }
        bad = u;
        if (!m_any_free(bad, vlo, vhi)) done = true;
        if (__all(done)) break;
      }
    } else if (active) {
      // no run list (too many runs, or extents beyond its fields): every lane
      // straight from the core masks
      for (int X = 0; X < kx; ++X)
        for (int Y = 0; Y < ky; ++Y)
          for_runs(core[X * ccy + Y], [&](int a, int L) {
            bad = m_or(bad, run_blocks_at(cols, plane, (size_t)(Qx + X) * cy + (Qy + Y), a, L));
          });
    }
    const bool cand = active && m_any_free(bad, vlo, vhi);
    if (!fine) {
      if (cand && j == 0) mark_feasible();
      continue;
    }
    // full-resolution stage: the block lists the full-resolution offsets
    // (2 Qx + rx, 2 Qy + ry) of every block offset with a height left; groups
    // of G lanes then take one candidate at a time, split the item's runs
    // between their lanes and combine in chunks of growing size, so that a
    // refuted candidate stops after its first blocked chunk and a proof costs
    // nf / G loads per lane
    if (cand) {
      if (j == 0) cbad[tid >> 2] = bad;
      const int qx = 2 * Qx + (j & 1), qy = 2 * Qy + (j >> 1);
      if (qx < d.nvx && qy < d.nvy) cand_list[atomicAdd(&ncand, 1)] = (unsigned)tid;
    }
    __syncthreads();
    const int nc = ncand;
    for (int c = tid / G; c < nc; c += 256 / G) {
      if (__shfl((int)*flag, 0, G)) break;   // the group leader's read, so the group stays together
      const int e = (int)cand_list[c];
      const int qq = q0 + (e >> 2), je = e & 3;
      const int QX = qq / nqy, QY = qq - QX * nqy;
      const int qx = 2 * QX + (je & 1), qy = 2 * QY + (je >> 1);
      // sufficient test first: every tray column under the footprint
      // [qx, qx + ex) x [qy, qy + ey) is at most h high (h = max of the tile
      // heights there), so the item's box at height h is empty, a valid
      // offset when h < nvz
      {
        const int tx0 = qx / kSkyTile, tx1 = (qx + d.ex - 1) / kSkyTile;
        const int ty0 = qy / kSkyTile, ty1 = (qy + d.ey - 1) / kSkyTile;
        const int nyt = ty1 - ty0 + 1, nt = (tx1 - tx0 + 1) * nyt;
        int h = 0;
        for (int i = gl; i < nt; i += G) {
          const int a = i / nyt;
          h = max(h, ht[(tx0 + a) * ty + ty0 + (i - a * nyt)]);
        }
        if (group_max<G>(h) < d.nvz) {
          if (gl == 0) mark_feasible();
          break;
        }
      }
      M128 fb = cbad[e >> 2];
      bool blocked = false;
      if (nf <= kFineRuns) {
        int step = G;
        for (int r0 = 0; r0 < nf && !blocked; r0 += step, step = min(step * 2, 64 * G)) {
          const int r1 = min(nf, r0 + step);
          for (int r = r0 + gl; r < r1; r += G) {
            const unsigned u = r < kRunsLds ? sfruns[r] : fruns[r];
            const int x = u & 127, y = (u >> 7) & 255, ra = (u >> 15) & 127, L = ((u >> 22) & 127) + 1;
            fb = m_or(fb, run_blocks_at(tf, fplane, (size_t)(qx + x) * ny + (qy + y), ra, L));
          }
          fb = group_or<G>(fb);
          blocked = !m_any_free(fb, vlo, vhi);
        }
      } else {
        // no run list: lanes over the item's columns
        for (int cc = gl; cc < d.ex * d.ey; cc += G) {
          const int x = cc / d.ey, y = cc - x * d.ey;
          for_runs(own[x * my + y], [&](int a, int L) {
            fb = m_or(fb, run_blocks_at(tf, fplane, (size_t)(qx + x) * ny + (qy + y), a, L));
          });
        }
        fb = group_or<G>(fb);
        blocked = !m_any_free(fb, vlo, vhi);
      }
      if (!blocked) {
        if (gl == 0) mark_feasible();
        break;
      }
    }
  }
}

// Column test decision, after k_column_test (cfs: its flags; its residual
// slot stays 0, the test being exact).  Some orientation left: the
// feasibility pass and the scored search skip the excluded ones (CTL_PRUNE).
// None left: every orientation is infeasible (CTL_COARSE_DONE); fs[0] = 0 so
// that a reader of the feasibility state sees a trusted empty result, and in
// the greedy loop (out != nullptr) the item is recorded as not found and its
// search and refresh are skipped, or, for a search over proposals
// (halt_not_found), the loop stops with CTL_HALT = 2 as k_decide would.
__global__ void k_coarse_decide(const int* __restrict__ cfs, int m, int* __restrict__ fs,
                                PackOut* __restrict__ out, int t, int halt_not_found,
                                int* __restrict__ ctl, int topk = 0,
                                int* __restrict__ gate = nullptr) {
  // gate (nullable): the learned loop's proposer gate (k_learn_gate), set
  // here: 1 (the proposer's kernels exit at once) unless the item is still to
  // be searched and more than topk of its orientations are feasible
  if (search_off(ctl)) {
    if (gate && threadIdx.x == 0) *gate = 1;
    return;
  }
  // one wavefront: a lane per flag, then a ballot
  int v = 0;
  for (int o = threadIdx.x; o < m; o += 64) v |= cfs[kFeasHeader + o];
  const bool any = __any(v != 0);
  if (gate) {   // m <= 64 here: one flag per lane
    const int f = threadIdx.x < m ? cfs[kFeasHeader + threadIdx.x] : 0;
    const int n = __popcll(__ballot(f != 0));
    if (threadIdx.x == 0) *gate = n <= topk ? 1 : 0;
  }
  if (threadIdx.x != 0) return;
  ctl[CTL_PRUNE] = any ? 1 : 0;
  ctl[CTL_COARSE_DONE] = any ? 0 : 1;
  if (any) return;
  fs[0] = 0;
  if (!out) return;
  PackOut r{};
  r.so.found = 0;
  r.so.idx = INT_MAX;
  out[t] = r;
  if (halt_not_found) {
    ctl[CTL_HALT] = 2;
    ctl[CTL_HALT_ITEM] = t;
  } else {
    ctl[CTL_SKIP_SEARCH] = 1;
  }
  ctl[CTL_SKIP_REFRESH] = 1;
}

// greedy loop: with the pass run and its residual within resid_limit, an
// item with no feasible orientation is decided here (not found) and its
// scored search, placement and refresh are skipped
//
// With cfs (the column-test flags the search prunes with), a trusted pass
// that left some orientation feasible narrows the search to exactly those.
__global__ void k_feas_decide(const int* __restrict__ fs, int m, PackOut* __restrict__ out, int t,
                              float resid_limit, int* __restrict__ cfs, int* __restrict__ ctl) {
  if (threadIdx.x != 0 || feas_off(ctl) || coarse_done(ctl)) return;
  const float res = __int_as_float(fs[0]);
  int any = 0;
  for (int o = 0; o < m; ++o) any |= fs[kFeasHeader + o];
  if (!(res <= resid_limit)) return;
  if (any) {
    if (cfs) {
      for (int o = 0; o < m; ++o) cfs[kFeasHeader + o] = fs[kFeasHeader + o];
      ctl[CTL_PRUNE] = 1;
    }
    return;
  }
  PackOut r{};
  r.so.found = 0;
  r.so.idx = INT_MAX;
  r.so.score = 0.0;
  r.so.resid = res;
  r.collision = 0;
  out[t] = r;
  ctl[CTL_SKIP_SEARCH] = 1;
  ctl[CTL_SKIP_REFRESH] = 1;
}

// ---------------------------------------------------------------------------
// selection: one block of kSelectThreads; wave w reduces orientations w,
// w+waves, ... over their x records, then thread 0 takes the first strictly
// smaller score across orientations (list order).
// ---------------------------------------------------------------------------
constexpr int kSelectThreads = 1024;
// the selection of one block of kSelectThreads; the winner is written to *out
// and, when sel != nullptr, to *sel (shared memory) for the same block
__device__ __forceinline__ void select_body(const WgBest* __restrict__ rec,
                                            const OrientDesc* __restrict__ od, int n_orient,
                                            int rec_stride, int rows_per_rec,
                                            WgBest* __restrict__ per_orient,
                                            SearchOut* __restrict__ out, SearchOut* sel) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int o = wave; o < n_orient; o += kSelectThreads / 64) {
    const OrientDesc d = od[o];
    // slab plan: one record per valid x slice; split plan: one per block of
    // rows_per_rec valid (x, y) rows
    const int nrec = rows_per_rec ? (d.nvx * d.nvy + rows_per_rec - 1) / rows_per_rec : d.nvx;
    double sc = 1e300;
    int idx = INT_MAX;
    long long p = 0;
    float res = 0.f;
    int uns = 0;
    for (int x = lane; x < nrec; x += 64) {
      const WgBest r = rec[(size_t)o * rec_stride + x];
      if (better(r.score, r.idx, sc, idx)) {
        sc = r.score; idx = r.idx;
        p = ((long long)r.p_hi << 32) | (unsigned int)r.p_lo;
      }
      res = fmaxf(res, r.resid);
      uns |= r.unsupported;
    }
    for (int s = 32; s > 0; s >>= 1) {
      const double s2 = __shfl_xor(sc, s);
      const int i2 = __shfl_xor(idx, s);
      const long long p2 = __shfl_xor(p, s);
      const float r2 = __shfl_xor(res, s);
      const int u2 = __shfl_xor(uns, s);
      if (better(s2, i2, sc, idx)) { sc = s2; idx = i2; p = p2; }
      res = fmaxf(res, r2);
      uns |= u2;
    }
    if (lane == 0) {
      WgBest b;
      b.score = sc; b.idx = idx;
      b.p_lo = (int)(p & 0xffffffffLL); b.p_hi = (int)(p >> 32);
      b.resid = res; b.unsupported = uns;
      per_orient[o] = b;
    }
  }
  __syncthreads();
  if (tid == 0) {
    SearchOut r;
    r.found = 0; r.orient = -1; r.idx = -1; r.pad = 0; r.score = 1e30; r.prox = 0;
    r.resid = 0.f; r.unsupported = 0;
    for (int o = 0; o < n_orient; ++o) {
      const WgBest b = per_orient[o];
      r.resid = fmaxf(r.resid, b.resid);
      r.unsupported |= b.unsupported;
      if (b.idx != INT_MAX && b.score < r.score) {
        r.found = 1; r.orient = o; r.idx = b.idx; r.score = b.score;
        r.prox = ((long long)b.p_hi << 32) | (unsigned int)b.p_lo;
      }
    }
    if (!r.found) r.score = 0.0;
    *out = r;
    if (sel) *sel = r;
  }
}

// keep (nullable): 1 when k_masked_argmin found the unmasked winner to be the
// masked one (the selection over its records then keeps out as it is)
__global__ void __launch_bounds__(kSelectThreads)
k_select(const WgBest* __restrict__ rec, const OrientDesc* __restrict__ od, int n_orient,
         int rec_stride, int rows_per_rec, WgBest* __restrict__ per_orient,
         SearchOut* __restrict__ out, long long* __restrict__ ver, const int* __restrict__ ctl,
         const int* __restrict__ keep = nullptr) {
  if (search_off(ctl) || (keep && *keep)) return;
  if (threadIdx.x < 3) ver[threadIdx.x] = 0;   // accumulator of the following k_verify
  select_body(rec, od, n_orient, rec_stride, rows_per_rec, per_orient, out, nullptr);
}

// ---------------------------------------------------------------------------
// exact verification of the selected offset: one block, direct sums.
// ver[0] = raw collision sum (tray * item), ver[1] = overlap count
// (tray != 0 & item > 0... item value weighted), ver[2] = proximity.
// ---------------------------------------------------------------------------
// grid-stride over the item voxels; ver[] must be zero on entry (k_select
// clears it) and is accumulated with 64-bit atomics.
__global__ void __launch_bounds__(256)
k_verify(const int* __restrict__ tray, const int* __restrict__ phi, const int* __restrict__ items,
         const OrientDesc* __restrict__ od, const SearchOut* __restrict__ so,
         long long* __restrict__ ver, int ny, int nz, const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  const SearchOut s = *so;
  const int tid = threadIdx.x;
  long long c = 0, ov = 0, p = 0;
  if (s.found) {
    const OrientDesc d = od[s.orient];
    const int qz = s.idx % nz, qy = (s.idx / nz) % ny, qx = s.idx / (ny * nz);
    const int* base = items + d.item_off;
    const int E = d.ex * d.ey * d.ez;
    for (int e = blockIdx.x * 256 + tid; e < E; e += gridDim.x * 256) {
      const int w2 = e % d.ez, w1 = (e / d.ez) % d.ey, w0 = e / (d.ey * d.ez);
      const int v = base[oriented_index(d, w0, w1, w2)];
      if (v == 0) continue;
      const size_t ti = ((size_t)(qx + w0) * ny + (qy + w1)) * nz + (qz + w2);
      const int t = tray[ti];
      c += (long long)t * v;
      ov += (t != 0) ? (long long)v : 0;
      p += (long long)phi[ti] * v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    c += __shfl_xor(c, o);
    ov += __shfl_xor(ov, o);
    p += __shfl_xor(p, o);
  }
  __shared__ long long r[3][4];
  if ((tid & 63) == 0) { r[0][tid >> 6] = c; r[1][tid >> 6] = ov; r[2][tid >> 6] = p; }
  __syncthreads();
  if (tid < 3) {
    const long long v = r[tid][0] + r[tid][1] + r[tid][2] + r[tid][3];
    if (v) atomicAdd((unsigned long long*)&ver[tid], (unsigned long long)v);
  }
}

// ---------------------------------------------------------------------------
// placement of an oriented item at an offset (from host args or SearchOut)
// phase 0: detect overlap (flag), phase 1: write id where tray == 0.
// ---------------------------------------------------------------------------
// In the greedy loop (ctl != nullptr) the offset comes from the SearchOut the
// device accepted and the id is id + the loop's placed count.
__global__ void __launch_bounds__(256)
k_place(int* __restrict__ tray, const int* __restrict__ items, const OrientDesc* __restrict__ od,
        const SearchOut* __restrict__ so, int orient, int qx, int qy, int qz, int id, int nx,
        int ny, int nz, int phase, int* __restrict__ flag, const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  if (ctl) id += ctl[CTL_PLACED];
  OrientDesc d;
  if (so) {
    const SearchOut s = *so;
    if (!s.found) return;
    d = od[s.orient];
    qz = s.idx % nz; qy = (s.idx / nz) % ny; qx = s.idx / (ny * nz);
  } else {
    d = od[orient];
  }
  if (phase == 1 && *flag) return;
  const int* base = items + d.item_off;
  const int E = d.ex * d.ey * d.ez;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < E; e += gridDim.x * 256) {
    const int w2 = e % d.ez, w1 = (e / d.ez) % d.ey, w0 = e / (d.ey * d.ez);
    const int X = qx + w0, Y = qy + w1, Z = qz + w2;
    if (X < 0 || Y < 0 || Z < 0 || X >= nx || Y >= ny || Z >= nz) continue;
    const int v = base[oriented_index(d, w0, w1, w2)];
    if (v <= 0) continue;
    int* t = tray + ((size_t)X * ny + Y) * nz + Z;
    if (phase == 0) {
      if (*t > 0) atomicOr(flag, 1);
    } else if (*t == 0) {
      *t = id;
    }
  }
}

// Greedy loop, one thread: accept the search of item t or hand it back to the
// host.  A search is accepted when every rounded FFT output was within
// resid_limit of its integer, the digit base covered the distance range and
// the exact re-check of the winner (k_verify) agrees; then the item is placed
// by the following kernels (when found) and CTL_PLACED counts it.
//
// With halt_not_found (a search over a proposed subset of orientations) a
// trusted search that found no feasible offset also stops the loop, with
// CTL_HALT = 2, so that the host can search every orientation of the item.
__global__ void k_decide(const SearchOut* __restrict__ so, const long long* __restrict__ ver,
                         PackOut* __restrict__ out, int t, float resid_limit,
                         int* __restrict__ ctl, int halt_not_found, int spec_max = 0,
                         const OrientDesc* __restrict__ od = nullptr, int id_offset = 0) {
  // an item the feasibility pass decided (nothing feasible) keeps its record
  if (threadIdx.x != 0 || ctl[CTL_HALT] || ctl[CTL_SKIP_SEARCH]) return;
  const SearchOut s = *so;
  const bool bad = s.unsupported || !(s.resid <= resid_limit) ||
                   (s.found && (ver[1] != 0 || ver[2] != s.prox));
  PackOut r;
  r.so = s;
  r.collision = ver[0];
  out[t] = r;
  if (bad || (halt_not_found && !s.found)) {
    ctl[CTL_HALT] = bad ? 1 : 2;
    ctl[CTL_HALT_ITEM] = t;
    ctl[CTL_SKIP_REFRESH] = 1;
    return;
  }
  ctl[CTL_SKIP_REFRESH] = s.found ? 0 : 1;
  if (!s.found) return;
  ctl[CTL_PLACED] += 1;
  // spec_max > 0 (sources "eq1", the field this library's own): a placement
  // with id != 1 adds no source, so the field stays and G1 is updated by the
  // placed voxels (k_spec_add), at most spec_max times in a row
  const int id = id_offset + ctl[CTL_PLACED];
  const bool inc = spec_max > 0 && id != 1 && od[s.orient].binary && ctl[CTL_SPEC_N] < spec_max;
  ctl[CTL_SPEC] = inc ? 1 : 0;
  ctl[CTL_SPEC_N] = inc ? ctl[CTL_SPEC_N] + 1 : 0;
}

// Up to 24 orientation indices passed by value.
struct OrientPick {
  int idx[24];
};

// dst[i] = src[pick.idx[i]] for i < k (the proposed orientations of one item)
__global__ void k_gather_od(const OrientDesc* __restrict__ src, OrientPick pick, int k,
                            OrientDesc* __restrict__ dst) {
  const int i = threadIdx.x;
  if (i < k) dst[i] = src[pick.idx[i]];
}

// ---------------------------------------------------------------------------
// Device-driven learned loop: the proposal pick as kernels.
// The column test has left per-orientation flags cfs[kFeasHeader + j] for the
// item's m fitting orientations (exact: 1 when orientation j has a
// collision-free valid offset); k_coarse_decide has already decided an item
// with none.
// ---------------------------------------------------------------------------
// gate = 1 (the proposer's inference kernels exit at once) unless the item is
// still to be searched and more than topk of its orientations are feasible
__global__ void k_learn_gate(const int* __restrict__ cfs, int m, int topk, int* __restrict__ gate,
                             const int* __restrict__ ctl) {
  const int j = threadIdx.x;   // one wavefront, m <= 64
  const int f = j < m ? cfs[kFeasHeader + j] : 0;
  const int n = __popcll(__ballot(f != 0));
  if (j == 0) *gate = (search_off(ctl) || n <= topk) ? 1 : 0;
}

// The orientations searched for an item: every feasible one when there are at
// most topk (no proposal needed), else the topk feasible ones with the
// highest logits (equal logits: the lower index first), in index order.  The
// descriptors go to od_pick[0, topk), padded with a copy of the first whose
// flag is 0 (the search skips it), and the flags of the pick list replace
// cfs[kFeasHeader + i], i < topk.  pick[i] = fitting index of slot i (-1 for
// padding); *used = 1 when logits chose, 0 when every feasible orientation
// is searched, -1 when the item was decided before (nothing searched).
// logits are indexed by orientation, map[j] is fitting j's orientation.
__global__ void k_learn_pick(const float* __restrict__ logits, const int* __restrict__ map,
                             const OrientDesc* __restrict__ od_fit, int m, int topk,
                             int* __restrict__ cfs, OrientDesc* __restrict__ od_pick,
                             int* __restrict__ pick, int* __restrict__ used,
                             const int* __restrict__ ctl) {
  const int j = threadIdx.x;   // one wavefront, m <= 64
  const bool off = search_off(ctl);
  const int f = (!off && j < m) ? cfs[kFeasHeader + j] : 0;
  const unsigned long long feas = __ballot(f != 0);
  const int n = __popcll(feas);
  bool sel = f != 0;
  if (n > topk && sel) {
    // rank among the feasible orientations: higher logit first, then index
    const float lj = logits[map[j]];
    int rank = 0;
    for (int i = 0; i < m; ++i) {
      if (!((feas >> i) & 1ULL) || i == j) continue;
      const float li = logits[map[i]];
      rank += (li > lj || (li == lj && i < j)) ? 1 : 0;
    }
    sel = rank < topk;
  }
  const unsigned long long chosen = __ballot(sel);
  const int k = __popcll(chosen);
  const int first = chosen ? __ffsll((long long)chosen) - 1 : 0;
  if (sel) {
    const int pos = __popcll(chosen & ((1ULL << j) - 1ULL));
    od_pick[pos] = od_fit[j];
    pick[pos] = j;
  }
  if (j < topk) {
    if (j >= k) {
      od_pick[j] = od_fit[first];
      pick[j] = -1;
    }
    cfs[kFeasHeader + j] = j < k ? 1 : 0;
  }
  if (j == 0) *used = off ? -1 : (n > topk ? 1 : 0);
}

// Per-orientation best scores of an accepted search: out[map[i]] = score of
// fitting orientation i (+inf when it has no feasible offset).
__global__ void k_record_scores(const WgBest* __restrict__ per, int m, const int* __restrict__ map,
                                double* __restrict__ out, const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  const int i = threadIdx.x;
  if (i >= m) return;
  const WgBest r = per[i];
  out[map[i]] = r.idx == INT_MAX ? (double)INFINITY : r.score;
}

// reset the distance-field statistics before a refresh
__global__ void k_stats_init(FieldStats* __restrict__ st, const int* __restrict__ ctl) {
  if (threadIdx.x != 0 || refresh_off(ctl)) return;
  st->sum = 0ULL;
  st->mn = INT_MAX;
  st->mx = INT_MIN;
  st->mxw = INT_MIN;
}

// ---------------------------------------------------------------------------
// split plan: the yz transforms as z-row kernels plus y-pencil kernels, each
// with a small LDS tile (used when a whole yz slab does not fit LDS, or when
// the autotuner finds it faster).
// ---------------------------------------------------------------------------
constexpr int kRowThreads = 256;
constexpr int kRowMaxEl = 16;
constexpr int kRowMax = kRowThreads * kRowMaxEl;   // complex elements per row block

enum { PENCIL_PLAIN = 0, PENCIL_SCORE = 1, PENCIL_ITEM = 2 };

// In-place FFT along n of complex data laid out [A][n][B]; block (b-tile, a)
// takes the T pencils b0..b0+T-1 of slab a.  Inputs at index >= n_in read as
// zero, outputs at index >= n_out are not stored.  PENCIL_SCORE: slab
// a = o*nx + x is skipped for x >= nvx(o) and stores only j < nvy(o);
// PENCIL_ITEM: slab a = o*nx + x is skipped for x >= ex(o) and reads j < ey(o).
template <bool FWD>
__global__ void __launch_bounds__(kPencilThreads)
k_pencil(float2* __restrict__ data, AxisPlan p_, int B, int T, int n_in, int n_out,
         const float2* __restrict__ tw_g, const OrientDesc* __restrict__ od, int nx, int mode,
         const int* __restrict__ ctl, int refresh_gate) {
  if (refresh_gate ? refresh_off(ctl) : search_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int n = p_.n;
  const int a = blockIdx.y, b0 = blockIdx.x * T, tid = threadIdx.x;
  if (b0 >= B) return;
  if (mode != PENCIL_PLAIN) {
    const int o = a / nx, x = a - o * nx;
    const OrientDesc d = od[o];
    if (mode == PENCIL_SCORE) {
      if (x >= d.nvx) return;
      n_out = min(n_out, d.nvy);
    } else {
      if (x >= d.ex) return;
      n_in = min(n_in, d.ey);
    }
  }
  float2* tw = smem;
  float2* buf = tw + n;
  load_table(tw, tw_g, n, tid, kPencilThreads);
  float2* base = data + (size_t)a * n * B + b0;
  {
    float2 v[kPencilMaxEl];
#pragma unroll
    for (int s = 0; s < kPencilMaxEl; ++s) {
      const int e = tid + s * kPencilThreads;
      const int y = e / T, t = e - y * T;
      v[s] = make_float2(0.f, 0.f);
      if (e < n * T && y < n_in && b0 + t < B) v[s] = base[(size_t)y * B + t];
    }
#pragma unroll
    for (int s = 0; s < kPencilMaxEl; ++s) {
      const int e = tid + s * kPencilThreads;
      if (e < n * T) buf[e] = v[s];
    }
  }
  __syncthreads();
  lds_fft<FWD, kPencilThreads, kPencilMaxEl, kPencilRadix>(buf, p_, T, T, 1, false, tw, tid);
  for (int e = tid; e < n_out * T; e += kPencilThreads) {
    const int y = e / T, t = e - y * T;
    if (b0 + t < B) base[(size_t)y * B + t] = buf[e];
  }
}

// forward z FFT of tray rows: grid (ceil(ny/R) * nx, 2 fields).  Field 0 rows
// are B + i*h (full pz), field 1 rows are l (first hz stored).
__global__ void __launch_bounds__(kRowThreads)
k_tray_rows_fwd(const int* __restrict__ tray, const int* __restrict__ phi,
                const FieldStats* __restrict__ st, float2* __restrict__ t1,
                float2* __restrict__ t2, int nx, int ny, int nz, int py, AxisPlan pz_, int R,
                const float2* __restrict__ tw_z, const int* __restrict__ ctl) {
  if (refresh_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int pz = pz_.n, hz = pz / 2 + 1, tid = threadIdx.x, field = blockIdx.y;
  const int jb = (ny + R - 1) / R;
  const int x = blockIdx.x / jb, j0 = (blockIdx.x - x * jb) * R;
  const int rows = min(R, ny - j0);
  float2* tw = smem;
  float2* buf = tw + pz;
  load_table(tw, tw_z, pz, tid, kRowThreads);
  const FieldParams fp = field_params(st, (long long)nx * ny * nz);
  const int half = (1 << fp.lb) >> 1;
  const float inv_pz = 1.0f / (float)pz;
  int dv[kRowMaxEl], tv[kRowMaxEl];
#pragma unroll
  for (int s = 0; s < kRowMaxEl; ++s) {
    const int e = tid + s * kRowThreads;
    const int r = fdiv(e, pz, inv_pz), k = e - r * pz;
    dv[s] = 0;
    tv[s] = 0;
    if (r < rows && k < nz) {
      const size_t g = ((size_t)x * ny + j0 + r) * nz + k;
      dv[s] = phi[g];
      if (field == 0) tv[s] = tray[g];
    }
  }
#pragma unroll
  for (int s = 0; s < kRowMaxEl; ++s) {
    const int e = tid + s * kRowThreads;
    const int r = fdiv(e, pz, inv_pz), k = e - r * pz;
    if (r >= rows) continue;
    float2 v = make_float2(0.f, 0.f);
    if (k < nz) {
      const int d = dv[s] - fp.mu;
      const int h = d >> fp.lb;
      if (field == 0) v = make_float2(tv[s] != 0 ? 1.f : 0.f, (float)h);
      else v.x = (float)(d - (h << fp.lb) - half);
    }
    buf[e] = v;
  }
  __syncthreads();
  lds_fft<true, kRowThreads, kRowMaxEl>(buf, pz_, 1, rows, pz, true, tw, tid);
  const int w = field == 0 ? pz : hz;
  float2* dst = field == 0 ? t1 : t2;
  for (int e = tid; e < rows * w; e += kRowThreads) {
    const int r = e / w, k = e - r * w;
    dst[((size_t)x * py + j0 + r) * w + k] = buf[r * pz + k];
  }
}

// forward z FFT of oriented item rows: grid (ceil(max_ey/R) * max_ex, m).
__global__ void __launch_bounds__(kRowThreads)
k_item_rows_fwd(const int* __restrict__ items, const OrientDesc* __restrict__ od,
                float2* __restrict__ ispec, int max_ex, int max_ey, int py, AxisPlan pz_, int R,
                const float2* __restrict__ tw_z, const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int pz = pz_.n, tid = threadIdx.x, o = blockIdx.y;
  const int jb = (max_ey + R - 1) / R;
  const int x = blockIdx.x / jb, j0 = (blockIdx.x - x * jb) * R;
  const OrientDesc d = od[o];
  if (x >= d.ex || j0 >= d.ey) return;
  const int rows = min(R, d.ey - j0);
  float2* tw = smem;
  float2* buf = tw + pz;
  load_table(tw, tw_z, pz, tid, kRowThreads);
  const int* base = items + d.item_off;
  const float inv_pz = 1.0f / (float)pz;
  int iv[kRowMaxEl];
#pragma unroll
  for (int s = 0; s < kRowMaxEl; ++s) {
    const int e = tid + s * kRowThreads;
    const int r = fdiv(e, pz, inv_pz), k = e - r * pz;
    iv[s] = (r < rows && k < d.ez) ? base[oriented_index(d, x, j0 + r, k)] : 0;
  }
#pragma unroll
  for (int s = 0; s < kRowMaxEl; ++s) {
    const int e = tid + s * kRowThreads;
    if (e < rows * pz) buf[e] = make_float2((float)iv[s], 0.f);
  }
  __syncthreads();
  lds_fft<true, kRowThreads, kRowMaxEl>(buf, pz_, 1, rows, pz, true, tw, tid);
  float2* dst = ispec + (((size_t)o * max_ex + x) * py + j0) * pz;
  for (int e = tid; e < rows * pz; e += kRowThreads) dst[e] = buf[e];
}

// pass B, split plan: grid (ceil(nx*ny/R), m).  Block takes R rows r of the
// orientation's valid rows (x = r / nvy, j = r % nvy): inverse z FFT of the
// Hermitian low-digit row and of the complex (c, h) row, round, score, argmin.
__global__ void __launch_bounds__(kRowThreads)
k_rows_score(const float2* __restrict__ y1, const float2* __restrict__ y2,
             const OrientDesc* __restrict__ od, const FieldStats* __restrict__ st,
             const double* __restrict__ pen, WgBest* __restrict__ rec, int rec_stride, int nx,
             int ny, int nz, int py, AxisPlan pz_, int R, const float2* __restrict__ tw_z,
             unsigned char* __restrict__ fgrid, long long* __restrict__ pgrid,
             const int* __restrict__ ctl) {
  if (search_off(ctl)) return;
  extern __shared__ float2 smem[];
  const int pz = pz_.n, hz = pz / 2 + 1, tid = threadIdx.x, o = blockIdx.y;
  const OrientDesc d = od[o];
  const int nrows = d.nvx * d.nvy;
  const int r0 = blockIdx.x * R;
  if (r0 >= nrows) return;
  const int rows = min(R, nrows - r0);
  // LDS rows [0, Rh): the low digit's rows paired, Z = X_2q + i X_2q+1 over
  // the full z spectrum (X at pz - k is conj X at k), so that the inverse
  // transform of row q holds row 2q's low-digit sums in its real part and row
  // 2q+1's in its imaginary part; rows [Rh, Rh + R): (c, h)
  const int Rh = (R + 1) / 2;
  float2* tw = smem;
  float2* buf = tw + pz;
  load_table(tw, tw_z, pz, tid, kRowThreads);
  const size_t so = (size_t)o * nx;
  const int nq = (rows + 1) / 2;
  for (int e = tid; e < nq * pz; e += kRowThreads) {
    const int q = e / pz, k = e - q * pz;
    const int kk = k < hz ? k : pz - k;
    const int ra = r0 + 2 * q, xa = ra / d.nvy, ja = ra - xa * d.nvy;
    const float2 a = y2[((so + xa) * py + ja) * hz + kk];
    float2 b = make_float2(0.f, 0.f);
    if (2 * q + 1 < rows) {
      const int rb = ra + 1, xb = rb / d.nvy, jb = rb - xb * d.nvy;
      b = y2[((so + xb) * py + jb) * hz + kk];
    }
    buf[q * pz + k] = k < hz ? make_float2(a.x - b.y, a.y + b.x)    // a + i b
                             : make_float2(a.x + b.y, b.x - a.y);   // conj(a) + i conj(b)
  }
  for (int e = tid; e < rows * pz; e += kRowThreads) {
    const int r = e / pz, k = e - r * pz;
    const int rr = r0 + r, x = rr / d.nvy, j = rr - x * d.nvy;
    buf[(Rh + r) * pz + k] = y1[((so + x) * py + j) * pz + k];
  }
  // rows past nq and past Rh + rows hold stale data and are not read
  __syncthreads();
  lds_fft<false, kRowThreads, kRowMaxEl>(buf, pz_, 1, Rh + R, pz, true, tw, tid);
  const FieldParams fp = field_params(st, (long long)nx * ny * nz);
  const long long b = 1LL << fp.lb;
  const long long addback = (b / 2 + fp.mu) * d.isum;
  double best = 1e300;
  int bidx = INT_MAX;
  long long bp = 0;
  float res = 0.f;
  for (int e = tid; e < rows * d.nvz; e += kRowThreads) {
    const int r = e / d.nvz, k = e - r * d.nvz;
    const int rr = r0 + r, x = rr / d.nvy, j = rr - x * d.nvy;
    const float2 lz = buf[(r >> 1) * pz + k];
    const float l = (r & 1) ? lz.y : lz.x;
    const float2 v = buf[(Rh + r) * pz + k];
    const float lr = rintf(l), c = rintf(v.x), h = rintf(v.y);
    res = fmaxf(res, fmaxf(fabsf(l - lr), fmaxf(fabsf(v.x - c), fabsf(v.y - h))));
    const long long p = b * (long long)h + (long long)lr + addback;
    const int idx = (x * ny + j) * nz + k;
    if (fgrid) {
      const size_t g = (size_t)o * nx * ny * nz + idx;
      fgrid[g] = c == 0.f ? 1 : 0;
      pgrid[g] = p;
    }
    if (c == 0.f) {
      const double sc = (double)p + pen[k];
      if (better(sc, idx, best, bidx)) { best = sc; bidx = idx; bp = p; }
    }
  }
  wg_argmin_write(best, bidx, bp, res, fp.ok ? 0 : 1, rec + (size_t)o * rec_stride + blockIdx.x,
                  tid, kRowThreads);
}

// ---------------------------------------------------------------------------
// exact direct-sum kernels (fallback search, standalone full-grid ops)
// item given as a list of (dx, dy, dz, value) non-zero voxels.
// ---------------------------------------------------------------------------
struct Tap { int dx, dy, dz, v; };

// mode 0: out[q] = sum a[q+y] v ; mode 1 (conv): out[q] = sum a[q-y] v
__global__ void __launch_bounds__(256)
k_direct_corr(const int* __restrict__ a, const Tap* __restrict__ taps, int ntap,
              long long* __restrict__ out, int nx, int ny, int nz, int conv) {
  const long long V = (long long)nx * ny * nz;
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < V; q += (long long)gridDim.x * 256) {
    const int k = (int)(q % nz), j = (int)((q / nz) % ny), i = (int)(q / ((long long)ny * nz));
    long long s = 0;
    for (int t = 0; t < ntap; ++t) {
      const Tap tp = taps[t];
      const int X = conv ? i - tp.dx : i + tp.dx;
      const int Y = conv ? j - tp.dy : j + tp.dy;
      const int Z = conv ? k - tp.dz : k + tp.dz;
      if (X < 0 || Y < 0 || Z < 0 || X >= nx || Y >= ny || Z >= nz) continue;
      s += (long long)a[((size_t)X * ny + Y) * nz + Z] * tp.v;
    }
    out[q] = s;
  }
}

// exact search over all in-bounds offsets of one orientation; per-block record
__global__ void __launch_bounds__(256)
k_direct_search(const int* __restrict__ tray, const int* __restrict__ phi,
                const Tap* __restrict__ taps, int ntap, int nvx, int nvy, int nvz, int ny,
                int nz, const double* __restrict__ pen, WgBest* __restrict__ rec) {
  const long long V = (long long)nvx * nvy * nvz;
  double best = 1e300;
  int bidx = INT_MAX;
  long long bp = 0;
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < V; q += (long long)gridDim.x * 256) {
    const int k = (int)(q % nvz), j = (int)((q / nvz) % nvy), i = (int)(q / ((long long)nvy * nvz));
    long long ov = 0, p = 0;
    for (int t = 0; t < ntap; ++t) {
      const Tap tp = taps[t];
      const size_t ti = ((size_t)(i + tp.dx) * ny + (j + tp.dy)) * nz + (k + tp.dz);
      const int tv = tray[ti];
      ov += (tv != 0) ? 1 : 0;
      p += (long long)phi[ti] * tp.v;
    }
    if (ov == 0) {
      const double sc = (double)p + pen[k];
      const int idx = (i * ny + j) * nz + k;
      if (better(sc, idx, best, bidx)) { best = sc; bidx = idx; bp = p; }
    }
  }
  wg_argmin_write(best, bidx, bp, 0.f, 0, rec + blockIdx.x, threadIdx.x, 256);
}

// exact feasibility / proximity grids of one orientation over its valid
// offsets (interlocking-free fallback); cells outside stay as cleared.
__global__ void __launch_bounds__(256)
k_direct_grid(const int* __restrict__ tray, const int* __restrict__ phi,
              const Tap* __restrict__ taps, int ntap, int nvx, int nvy, int nvz, int ny, int nz,
              unsigned char* __restrict__ fgrid, long long* __restrict__ pgrid) {
  const long long V = (long long)nvx * nvy * nvz;
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < V; q += (long long)gridDim.x * 256) {
    const int k = (int)(q % nvz), j = (int)((q / nvz) % nvy), i = (int)(q / ((long long)nvy * nvz));
    long long ov = 0, p = 0;
    for (int t = 0; t < ntap; ++t) {
      const Tap tp = taps[t];
      const size_t ti = ((size_t)(i + tp.dx) * ny + (j + tp.dy)) * nz + (k + tp.dz);
      ov += tray[ti] != 0 ? 1 : 0;
      p += (long long)phi[ti] * tp.v;
    }
    const size_t g = ((size_t)i * ny + j) * nz + k;
    fgrid[g] = ov == 0 ? 1 : 0;
    pgrid[g] = p;
  }
}

// ---------------------------------------------------------------------------
// interlocking-free flood fill on a feasibility grid
// free[q] = 1 for feasible offsets; reached grows from seed faces to fixpoint
// ---------------------------------------------------------------------------
// grid (blocks, m): seeds of grid o are its free cells on i|j|k == 0 or on the
// exit faces exits[3o..3o+2] (N-1-h, fft3.cu:276-287).
__global__ void __launch_bounds__(256)
k_ff_seed(const unsigned char* __restrict__ fr, unsigned char* __restrict__ reached,
          const int* __restrict__ exits, int nx, int ny, int nz, const int* __restrict__ skip) {
  if (skip && *skip) return;   // (k_if_direct: the fill is not needed)
  const long long V = (long long)nx * ny * nz;
  const int o = blockIdx.y;
  const int ex = exits[3 * o], ey = exits[3 * o + 1], ez = exits[3 * o + 2];
  const unsigned char* f = fr + (size_t)o * V;
  unsigned char* r = reached + (size_t)o * V;
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < V; q += (long long)gridDim.x * 256) {
    const int k = (int)(q % nz), j = (int)((q / nz) % ny), i = (int)(q / ((long long)ny * nz));
    const bool face = i == 0 || j == 0 || k == 0 || i == ex || j == ey || k == ez;
    r[q] = (f[q] && face) ? 1 : 0;
  }
}

// tile 4 x 8 x 32 with a one-cell halo; local fixpoint in LDS.
// grid (ceil(nz/32), ceil(ny/8), ceil(nx/4) * m): the z index also selects the grid o.
//
// Rounds run back to back on the device.  prev: nonzero when the previous
// round changed some cell (a round after one that changed nothing exits at
// once); changed: set when this round changes a cell.  act_prev / act_cur
// (one byte per tile, indexed like the grid): a tile is worked only when it
// or one of its six neighbouring tiles of the same grid changed in the
// previous round; every block writes its own act_cur entry.
constexpr int FFX = 4, FFY = 8, FFZ = 32;
__global__ void __launch_bounds__(FFX * FFY * FFZ)
k_ff_tile(const unsigned char* __restrict__ fr_all, unsigned char* __restrict__ reached_all,
          int nx, int ny, int nz, int xtiles, const int* __restrict__ prev,
          int* __restrict__ changed, const unsigned char* __restrict__ act_prev,
          unsigned char* __restrict__ act_cur) {
  if (*prev == 0) return;
  const int tid = threadIdx.x;
  const size_t tile = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  {
    const int xt_ = blockIdx.z % xtiles;
    const size_t sy = gridDim.x, sx = (size_t)gridDim.x * gridDim.y;
    bool act = act_prev[tile] != 0;
    act = act || (blockIdx.x > 0 && act_prev[tile - 1]);
    act = act || (blockIdx.x + 1 < gridDim.x && act_prev[tile + 1]);
    act = act || (blockIdx.y > 0 && act_prev[tile - sy]);
    act = act || (blockIdx.y + 1 < gridDim.y && act_prev[tile + sy]);
    act = act || (xt_ > 0 && act_prev[tile - sx]);
    act = act || (xt_ + 1 < xtiles && act_prev[tile + sx]);
    if (!act) {
      if (tid == 0) act_cur[tile] = 0;
      return;
    }
  }
  constexpr int HX = FFX + 2, HY = FFY + 2, HZ = FFZ + 2;
  __shared__ unsigned char sf[HX * HY * HZ], sr[HX * HY * HZ];
  __shared__ int gch;
  const int tz = tid % FFZ, ty = (tid / FFZ) % FFY, tx = tid / (FFZ * FFY);
  const int o = blockIdx.z / xtiles;
  const size_t V = (size_t)nx * ny * nz;
  const unsigned char* fr = fr_all + o * V;
  unsigned char* reached = reached_all + o * V;
  const int bx = (blockIdx.z - o * xtiles) * FFX, by = blockIdx.y * FFY, bz = blockIdx.x * FFZ;
  for (int e = tid; e < HX * HY * HZ; e += FFX * FFY * FFZ) {
    const int hz_ = e % HZ, hy_ = (e / HZ) % HY, hx_ = e / (HZ * HY);
    const int X = bx + hx_ - 1, Y = by + hy_ - 1, Z = bz + hz_ - 1;
    unsigned char f = 0, r = 0;
    if (X >= 0 && Y >= 0 && Z >= 0 && X < nx && Y < ny && Z < nz) {
      const size_t q = ((size_t)X * ny + Y) * nz + Z;
      f = fr[q];
      r = reached[q];
    }
    sf[e] = f;
    sr[e] = r;
  }
  if (tid == 0) gch = 0;
  __syncthreads();
  const int c = ((tx + 1) * HY + (ty + 1)) * HZ + (tz + 1);
  const bool inside = (bx + tx < nx) && (by + ty < ny) && (bz + tz < nz);
  // A z line of the tile is 32 consecutive lanes (half a wavefront): each
  // round takes one step from the 6 neighbours, then carries reached cells
  // along runs of free cells of the line with a segmented OR-scan over the
  // lanes (both directions, no barrier), so only x and y steps cost a round.
  const int zl = tid & (FFZ - 1);
  const bool free_c = inside && sf[c];
  bool mine = false;
  for (int it = 0; it < 4096; ++it) {
    int gen = 0;
    if (free_c) {
      gen = sr[c] | sr[c - 1] | sr[c + 1] | sr[c - HZ] | sr[c + HZ] | sr[c - HY * HZ] |
            sr[c + HY * HZ];
      gen = gen ? 1 : 0;
    }
    int gf = gen, pf = free_c ? 1 : 0, gb = gen, pb = pf;
#pragma unroll
    for (int d = 1; d < FFZ; d <<= 1) {
      const int g1 = __shfl_up(gf, d, FFZ), p1 = __shfl_up(pf, d, FFZ);
      const int g2 = __shfl_down(gb, d, FFZ), p2 = __shfl_down(pb, d, FFZ);
      if (zl >= d) {
        gf |= pf & g1;
        pf &= p1;
      }
      if (zl + d < FFZ) {
        gb |= pb & g2;
        pb &= p2;
      }
    }
    const int upd = (free_c && !sr[c] && (gf | gb)) ? 1 : 0;
    __syncthreads();
    if (upd) {
      sr[c] = 1;
      mine = true;
    }
    if (!__syncthreads_or(upd)) break;
  }
  if (inside && mine) {
    reached[((size_t)(bx + tx) * ny + (by + ty)) * nz + (bz + tz)] = 1;
    gch = 1;
  }
  __syncthreads();
  if (tid == 0) {
    act_cur[tile] = gch ? 1 : 0;
    // (the flag only goes from 0 to 1 here: once it reads 1 no atomic is needed)
    if (gch && !load_agent(changed)) atomicOr(changed, 1);
  }
}

// start of a batch of flood-fill rounds: flags[0] = go (1), flags[1..k] = 0;
// with init_act, every tile of act starts active
// ---- flood fill on bit columns (trays with nz <= 128) ---------------------
// The same fixpoint as k_ff_seed / k_ff_tile with each (x, y) column of an
// offset grid held as a 128-bit mask: bit z of column (o, x, y) is offset
// (x, y, z) of grid o.  Along z a column fills in one step by carry
// arithmetic: for free bits m and seeds s (s within m),
//     up(m, s) = (((s + m) ^ m) | s) & m
// holds every bit of a run of m at or above the lowest seed of that run:
// within a run [a, b] with seeds s1 < s2 < ..., m + s has bits a .. s1-1,
// s2, s3, ... and b+1 (the carry of s1 runs to the end of the run and stops
// at the clear bit b+1), so the xor with m leaves s1 .. b without s2, s3,
// ..., which the or puts back; runs do not interact.  down is up on the
// bit-reversed words; the union is every run of m that holds a seed.
struct U128 {
  unsigned long long lo, hi;
};
__device__ __forceinline__ U128 u_or(U128 a, U128 b) { return {a.lo | b.lo, a.hi | b.hi}; }
__device__ __forceinline__ U128 u_and(U128 a, U128 b) { return {a.lo & b.lo, a.hi & b.hi}; }
__device__ __forceinline__ bool u_eq(U128 a, U128 b) { return a.lo == b.lo && a.hi == b.hi; }
__device__ __forceinline__ U128 u_rev(U128 a) { return {__brevll(a.hi), __brevll(a.lo)}; }
__device__ __forceinline__ U128 fill_up(U128 m, U128 s) {
  const unsigned long long lo = s.lo + m.lo;
  const unsigned long long hi = s.hi + m.hi + (lo < s.lo ? 1ULL : 0ULL);
  return {((lo ^ m.lo) | s.lo) & m.lo, ((hi ^ m.hi) | s.hi) & m.hi};
}
// every run of m that holds a bit of s (s need not lie within m)
__device__ __forceinline__ U128 fill_runs(U128 m, U128 s) {
  s = u_and(s, m);
  return u_or(fill_up(m, s), u_rev(fill_up(u_rev(m), u_rev(s))));
}

// Free bits and seeds of every column of the m grids (one wavefront per
// column: lanes read z = lane and z = 64 + lane, then ballot).  Seeds: the
// free bits of the columns on an x or y face (x == 0 or the exit face x ==
// exits[3o], the same for y) and bits z == 0 and z == exits[3o + 2] of the
// others.  grid (ceil(nx ny / 4), m), 256 threads.
__global__ void __launch_bounds__(256)
k_ffb_pack(const unsigned char* __restrict__ fr, U128* __restrict__ fm, U128* __restrict__ rb,
           const int* __restrict__ exits, int nx, int ny, int nz, const int* __restrict__ skip) {
  if (skip && *skip) return;   // (k_if_direct: the fill is not needed)
  const long long ncol = (long long)nx * ny;
  const long long col = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, o = blockIdx.y;
  if (col >= ncol) return;   // whole wavefronts
  const unsigned char* f = fr + ((size_t)o * ncol + col) * nz;
  const unsigned long long lo = __ballot(lane < nz && f[lane] != 0);
  const unsigned long long hi = __ballot(64 + lane < nz && f[64 + lane] != 0);
  if (lane != 0) return;
  const int x = (int)(col / ny), y = (int)(col - (long long)x * ny);
  const int ex = exits[3 * o], ey = exits[3 * o + 1], ez = exits[3 * o + 2];
  const U128 m{lo, hi};
  U128 sd{1ULL, 0ULL};
  if (ez < 64) sd.lo |= 1ULL << ez;
  else sd.hi |= 1ULL << (ez - 64);
  const bool face = x == 0 || y == 0 || x == ex || y == ey;
  fm[(size_t)o * ncol + col] = m;
  rb[(size_t)o * ncol + col] = face ? m : u_and(m, sd);
}

// One round over 32 x 32-column tiles: load the tile's columns and a
// one-column halo, then iterate to the tile's local fixpoint (reached =
// runs of free bits holding a bit reached by the column or one of its four
// x / y neighbours), write back.  prev / changed / act_prev / act_cur as in
// k_ff_tile (a tile is worked when it or one of its four neighbouring tiles
// changed in the previous round).  grid (xt * yt, m), 1024 threads.
constexpr int FBT = 32;
__global__ void __launch_bounds__(FBT * FBT)
k_ffb_tile(const U128* __restrict__ fm_all, U128* __restrict__ rb_all, int nx, int ny, int xt,
           int yt, const int* __restrict__ prev, int* __restrict__ changed,
           const unsigned char* __restrict__ act_prev, unsigned char* __restrict__ act_cur) {
  if (*prev == 0) return;
  const int tid = threadIdx.x, o = blockIdx.y;
  const int tx0 = blockIdx.x % xt, ty0 = blockIdx.x / xt;
  const size_t tile = (size_t)o * xt * yt + blockIdx.x;
  {
    bool act = act_prev[tile] != 0;
    act = act || (tx0 > 0 && act_prev[tile - 1]);
    act = act || (tx0 + 1 < xt && act_prev[tile + 1]);
    act = act || (ty0 > 0 && act_prev[tile - xt]);
    act = act || (ty0 + 1 < yt && act_prev[tile + xt]);
    if (!act) {
      if (tid == 0) act_cur[tile] = 0;
      return;
    }
  }
  constexpr int H = FBT + 2;
  __shared__ U128 sr[H * H];
  __shared__ int gch;
  const size_t ncol = (size_t)nx * ny;
  const U128* fm = fm_all + (size_t)o * ncol;
  U128* rb = rb_all + (size_t)o * ncol;
  const int bx = tx0 * FBT, by = ty0 * FBT;
  // lanes along y (consecutive y are consecutive columns in memory)
  const int ly = tid % FBT, lx = tid / FBT;
  const int X = bx + lx, Y = by + ly;
  const bool inside = X < nx && Y < ny;
  for (int e = tid; e < H * H; e += FBT * FBT) {
    const int hy = e % H, hx = e / H;
    const int gx = bx + hx - 1, gy = by + hy - 1;
    U128 r{0ULL, 0ULL};
    if (gx >= 0 && gy >= 0 && gx < nx && gy < ny) r = rb[(size_t)gx * ny + gy];
    sr[e] = r;
  }
  const U128 m = inside ? fm[(size_t)X * ny + Y] : U128{0ULL, 0ULL};
  if (tid == 0) gch = 0;
  __syncthreads();
  const int c = (lx + 1) * H + (ly + 1);
  U128 r = sr[c];
  const U128 r0 = r;
  for (int it = 0; it < 4 * FBT * FBT; ++it) {
    const U128 nb = u_or(u_or(sr[c - 1], sr[c + 1]), u_or(sr[c - H], sr[c + H]));
    const U128 nr = fill_runs(m, u_or(r, nb));
    const bool upd = !u_eq(nr, r);
    __syncthreads();
    if (upd) {
      r = nr;
      sr[c] = r;
    }
    if (!__syncthreads_or(upd)) break;
  }
  const bool mine = inside && !u_eq(r, r0);
  if (mine) {
    rb[(size_t)X * ny + Y] = r;
    gch = 1;
  }
  __syncthreads();
  if (tid == 0) {
    act_cur[tile] = gch ? 1 : 0;
    // (the flag only goes from 0 to 1 here: once it reads 1 no atomic is needed)
    if (gch && !load_agent(changed)) atomicOr(changed, 1);
  }
}

// reached bits -> one byte per offset (the layout k_masked_argmin and the
// interlocking mask read); grid (ceil(nx ny / 4), m), 256 threads
__global__ void __launch_bounds__(256)
k_ffb_unpack(const U128* __restrict__ rb, unsigned char* __restrict__ reached, int nx, int ny,
             int nz) {
  const long long ncol = (long long)nx * ny;
  const long long col = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, o = blockIdx.y;
  if (col >= ncol) return;
  const U128 r = rb[(size_t)o * ncol + col];
  unsigned char* out = reached + ((size_t)o * ncol + col) * nz;
  if (lane < nz) out[lane] = (unsigned char)((r.lo >> lane) & 1ULL);
  if (64 + lane < nz) out[64 + lane] = (unsigned char)((r.hi >> lane) & 1ULL);
}

// dst[0, n) = v (a clear that a captured launch sequence may hold: no memset)
__global__ void k_fill_int(int* __restrict__ dst, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = v;
}

// skip (nullable): when set, the batch's first round already reads "nothing
// changed", so every round exits at once
__global__ void k_ff_batch_init(int* __restrict__ flags, int k, unsigned char* __restrict__ act,
                                long long n_act, const int* __restrict__ skip) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= k) flags[i] = i == 0 ? ((skip && *skip) ? 0 : 1) : 0;
  if (act)
    for (long long t = i; t < n_act; t += (long long)gridDim.x * blockDim.x) act[t] = 1;
}

// Interlocking-free search, only the winner wanted: before any fill, *keep =
// 1 when the unmasked winner is reachable on its own, so that the fill, the
// masked argmin and the selection after it are skipped.  It is when nothing
// is free (the masked search finds nothing either), when its offset is a
// seed of the fill (a free offset in an exit column, x = 0, y = 0, x = ex or
// y = ey), or when every offset straight below it to z = 0 or straight above
// it to z = ez is free (a path to a seed; z = 0 and z = ez are exit faces).
// One wavefront.
__global__ void k_if_direct(const SearchOut* __restrict__ win, const OrientDesc* __restrict__ od,
                            const unsigned char* __restrict__ ffree, int nx, int ny, int nz,
                            int* __restrict__ keep) {
  const int lane = threadIdx.x;
  const SearchOut s = *win;
  if (!s.found) {
    if (lane == 0) *keep = 1;
    return;
  }
  const OrientDesc d = od[s.orient];
  const int k = s.idx % nz, j = (s.idx / nz) % ny, x = s.idx / (ny * nz);
  const int ex = nx - 1 - d.hx, ey = ny - 1 - d.hy, ez = nz - 1 - d.hz;
  const bool face = x == 0 || j == 0 || x == ex || j == ey;
  const unsigned char* col = ffree + (size_t)s.orient * nx * ny * nz + ((size_t)x * ny + j) * nz;
  bool below = true, above = true;
  for (int z = lane; z < nz; z += 64) {
    const bool f = col[z] != 0;
    if (z <= k && !f) below = false;
    if (z >= k && z <= ez && !f) above = false;
  }
  below = __all(below);
  above = __all(above);
  if (lane == 0) *keep = (face || below || above) ? 1 : 0;
}

// grid (nx, m): argmin of double(p) + penalty[k] over reached offsets of slice x
// into rec[o*nx + x]; the residual/unsupported fields come from the
// orientation's unmasked summary per_o[o] (a k_select over the score records).
__global__ void __launch_bounds__(256)
k_masked_argmin(const unsigned char* __restrict__ reached, const long long* __restrict__ pgrid,
                const OrientDesc* __restrict__ od, const double* __restrict__ pen,
                const WgBest* __restrict__ per_o, WgBest* __restrict__ rec, int nx, int ny,
                int nz, const U128* __restrict__ rbits, const SearchOut* __restrict__ win,
                int* __restrict__ keep) {
  // rbits (nullable): the reached offsets as the bit fill left them, read in
  // place of the byte grid
  const int x = blockIdx.x, o = blockIdx.y, tid = threadIdx.x;
  // win (nullable): the unmasked search's winner.  When it is reachable (or
  // nothing is free at all) it is also the masked winner: every other
  // orientation's masked best is no better than its unmasked best, which lost
  // to it, and the winner's own orientation keeps it.  Every block then
  // returns, and *keep = 1 tells the following selection to keep it.
  if (win) {
    if (*keep) return;   // k_if_direct found the winner reachable (no fill ran)
    const SearchOut s = *win;
    bool same = !s.found;
    if (s.found) {
      const int k = s.idx % nz, j = (s.idx / nz) % ny, xw = s.idx / (ny * nz);
      if (rbits) {
        const U128 r = rbits[((size_t)s.orient * nx + xw) * ny + j];
        same = (((k < 64 ? r.lo : r.hi) >> (k & 63)) & 1ULL) != 0;
      } else {
        same = reached[(size_t)s.orient * nx * ny * nz + s.idx] != 0;
      }
    }
    if (x == 0 && o == 0 && tid == 0) *keep = same ? 1 : 0;
    if (same) return;
  }
  const OrientDesc d = od[o];
  if (x >= d.nvx) return;
  const size_t base = (size_t)o * nx * ny * nz + (size_t)x * ny * nz;
  double best = 1e300;
  int bidx = INT_MAX;
  long long bp = 0;
  const int S = d.nvy * d.nvz;
  for (int e = tid; e < S; e += 256) {
    const int j = e / d.nvz, k = e - j * d.nvz;
    const int q = j * nz + k;
    if (rbits) {
      const U128 r = rbits[((size_t)o * nx + x) * ny + j];
      if (!(((k < 64 ? r.lo : r.hi) >> (k & 63)) & 1ULL)) continue;
    } else if (!reached[base + q]) {
      continue;
    }
    const long long p = pgrid[base + q];
    const double sc = (double)p + pen[k];
    const int idx = x * ny * nz + q;
    if (better(sc, idx, best, bidx)) { best = sc; bidx = idx; bp = p; }
  }
  WgBest* out = rec + (size_t)o * nx + x;
  const WgBest s = per_o[o];
  wg_argmin_write(best, bidx, bp, s.resid, s.unsupported, out, tid, 256);
}

}  // namespace spk
