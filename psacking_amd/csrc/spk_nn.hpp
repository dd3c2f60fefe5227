// Kernels of the learned orientation proposer: a small 3-D CNN that scores
// the 24 orientations of an item against the current tray.  There is no
// counterpart in the reference (spectral_packer/packer.py:487-505 searches
// every orientation); the exact FFT search evaluates the proposals.
//
// Storage: activations, weights and activation gradients are bf16 (ushort
// bits, round-to-nearest-even from fp32), channels-last [batch][d0][d1][d2][c].
// Every product runs on v_mfma_f32_16x16x32_bf16 with fp32 accumulation;
// biases, master weights, weight gradients and the optimizer state are fp32.
//
// One kernel family serves all layers (3x3x3 stride-2 convolutions and the
// fully connected head, which is a 1x1x1 convolution over a 1x1x1 grid):
//   k_igemm<FWD>   Y[m][n]  = relu(sum_k im2col(X)[m][k] * W[n][k] + b[n])
//   k_igemm<BWD>   dX[m][c] = [X[m][c] > 0] * sum_(tap,n) dY[o(m,tap)][n] * Wt[c][tap,n]
//   k_wgrad        part[chunk][n][k] = sum_(m in chunk) dY[m][n] * im2col(X)[m][k]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spknn {

typedef unsigned short bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

// round to nearest even; inputs are finite
__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

__device__ __forceinline__ bool bf_pos(bf16_t h) { return h != 0 && !(h & 0x8000u); }

// Geometry of one layer.  ks = 3, st = 2, pad = 1 for the convolutions;
// ks = 1, st = 1, pad = 0 and all spatial sizes 1 for the dense layers.
// K = ks^3 * cin (forward reduction), Kp = K rounded up to 32; Kt = ks^3 * cout
// (input-gradient reduction), Ktp = Kt rounded up to 32.
struct ConvDesc {
  int B;
  int i0, i1, i2, cin;
  int o0, o1, o2, cout;
  int ks, st, pad;
  int K, Kp, Kt, Ktp;
};

enum { MODE_FWD = 0, MODE_BWD = 1 };

__device__ __forceinline__ void split3(int m, int d1, int d2, int d0, int& b, int& p0, int& p1,
                                       int& p2) {
  p2 = m % d2;
  m /= d2;
  p1 = m % d1;
  m /= d1;
  p0 = m % d0;
  b = m / d0;
}

// im2col element (row at output position (b,p0,p1,p2), column k)
__device__ __forceinline__ bf16_t col_fwd(const bf16_t* __restrict__ X, const ConvDesc& d, int b,
                                          int p0, int p1, int p2, int k) {
  if (k >= d.K) return 0;
  const int tap = k / d.cin, c = k - tap * d.cin;
  const int t0 = tap / (d.ks * d.ks), t1 = (tap / d.ks) % d.ks, t2 = tap % d.ks;
  const int x0 = p0 * d.st - d.pad + t0, x1 = p1 * d.st - d.pad + t1, x2 = p2 * d.st - d.pad + t2;
  if ((unsigned)x0 >= (unsigned)d.i0 || (unsigned)x1 >= (unsigned)d.i1 ||
      (unsigned)x2 >= (unsigned)d.i2)
    return 0;
  return X[((((size_t)b * d.i0 + x0) * d.i1 + x1) * d.i2 + x2) * d.cin + c];
}

// 8 consecutive im2col columns k..k+7; cin % 8 == 0 and k % 8 == 0, so they
// are 8 channels of one tap: one 16-byte load
__device__ __forceinline__ uint4 col_fwd8(const bf16_t* __restrict__ X, const ConvDesc& d, int b,
                                          int p0, int p1, int p2, int k) {
  uint4 z = make_uint4(0, 0, 0, 0);
  if (k >= d.K) return z;
  const int tap = k / d.cin, c = k - tap * d.cin;
  const int t0 = tap / (d.ks * d.ks), t1 = (tap / d.ks) % d.ks, t2 = tap % d.ks;
  const int x0 = p0 * d.st - d.pad + t0, x1 = p1 * d.st - d.pad + t1, x2 = p2 * d.st - d.pad + t2;
  if ((unsigned)x0 >= (unsigned)d.i0 || (unsigned)x1 >= (unsigned)d.i1 ||
      (unsigned)x2 >= (unsigned)d.i2)
    return z;
  return *reinterpret_cast<const uint4*>(
      X + ((((size_t)b * d.i0 + x0) * d.i1 + x1) * d.i2 + x2) * d.cin + c);
}

// 8 consecutive input-gradient columns k' = tap*cout + n .. +7 at input
// position (b,x0,x1,x2): the output-gradient row that tap maps the input to,
// or zeros when no output position covers it through this tap
__device__ __forceinline__ uint4 col_bwd8(const bf16_t* __restrict__ dY, const ConvDesc& d, int b,
                                          int x0, int x1, int x2, int k) {
  uint4 z = make_uint4(0, 0, 0, 0);
  if (k >= d.Kt) return z;
  const int tap = k / d.cout, n = k - tap * d.cout;
  const int t0 = tap / (d.ks * d.ks), t1 = (tap / d.ks) % d.ks, t2 = tap % d.ks;
  int q0 = x0 + d.pad - t0, q1 = x1 + d.pad - t1, q2 = x2 + d.pad - t2;
  if (q0 < 0 || q1 < 0 || q2 < 0) return z;
  if (q0 % d.st || q1 % d.st || q2 % d.st) return z;
  q0 /= d.st;
  q1 /= d.st;
  q2 /= d.st;
  if (q0 >= d.o0 || q1 >= d.o1 || q2 >= d.o2) return z;
  return *reinterpret_cast<const uint4*>(
      dY + ((((size_t)b * d.o0 + q0) * d.o1 + q1) * d.o2 + q2) * d.cout + n);
}

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// 256 threads = 4 waves; each wave owns 32 rows (two 16-row MFMA tiles) and
// NT 16-column tiles; the block covers 128 rows and 16*NT columns.
// FWD: X = layer input, Wm = W [N = cout][ldw = Kp], rows = output positions.
// BWD: X = output gradient dY, Wm = Wt [N = cin][ldw = Ktp], rows = input
//      positions; mask (layer input, post-ReLU) zeroes the gradient where the
//      activation was not positive.
//
// Split K (part != nullptr, forward only): block z reduces k in
// [z*kchunk, (z+1)*kchunk) and writes its fp32 partial tile to
// part[z][m][n]; k_split_finalize sums the slices in order and applies the
// epilogue.  Used when M is small (inference at batch 1), where the K loop of
// one or two blocks would otherwise be the whole kernel's latency.
template <int MODE, bool VEC, int NT>
__global__ void __launch_bounds__(256)
k_igemm(const bf16_t* __restrict__ X, const bf16_t* __restrict__ Wm, const float* __restrict__ bias,
        ConvDesc d, int M, int N, int Kloop, int ldw, int relu, const bf16_t* __restrict__ mask,
        bf16_t* __restrict__ Y, float* __restrict__ Yf, int kchunk, float* __restrict__ part,
        const int* __restrict__ gate) {
  if (gate && *gate) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mrow0 = (blockIdx.x * 4 + wave) * 32;
  const int nbase = blockIdx.y * 16 * NT;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  int rb[2], r0[2], r1[2], r2[2];
  bool rv[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int m = mrow0 + r * 16 + li;
    rv[r] = m < M;
    if (MODE == MODE_FWD)
      split3(rv[r] ? m : 0, d.o1, d.o2, d.o0, rb[r], r0[r], r1[r], r2[r]);
    else
      split3(rv[r] ? m : 0, d.i1, d.i2, d.i0, rb[r], r0[r], r1[r], r2[r]);
  }
  f32x4 acc[2][NT];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int k_begin = blockIdx.z * kchunk, k_end = min(Kloop, k_begin + kchunk);
  for (int k0 = k_begin; k0 < k_end; k0 += 32) {
    const int k = k0 + lk;
    bf16x8 a[2], b[NT];
    // element-wise im2col (cin % 8 != 0): the taps of the lane's 8 columns do
    // not depend on the row, so they are decomposed once for both rows
    int tp0[8], tp1[8], tp2[8], toff[8];
    if (MODE == MODE_FWD && !VEC) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int kk = k + j;
        const int tap = kk / d.cin, rest = tap % (d.ks * d.ks);
        tp0[j] = kk < d.K ? tap / (d.ks * d.ks) : -(1 << 20);   // out of range: never in bounds
        tp1[j] = rest / d.ks;
        tp2[j] = rest - tp1[j] * d.ks;
        // element offset from the row's corner (tap (0,0,0), channel 0)
        toff[j] = kk < d.K ? ((tp0[j] * d.i1 + tp1[j]) * d.i2 + tp2[j]) * d.cin + (kk - tap * d.cin)
                           : 0;
      }
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      uint4 v = make_uint4(0, 0, 0, 0);
      if (rv[r]) {
        if (MODE == MODE_BWD) {
          v = col_bwd8(X, d, rb[r], r0[r], r1[r], r2[r], k);
        } else if (VEC) {
          v = col_fwd8(X, d, rb[r], r0[r], r1[r], r2[r], k);
        } else {
          bf16_t e[8];
          const int c0 = r0[r] * d.st - d.pad, c1 = r1[r] * d.st - d.pad, c2 = r2[r] * d.st - d.pad;
          const long long corner =
              ((((long long)rb[r] * d.i0 + c0) * d.i1 + c1) * d.i2 + c2) * d.cin;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const bool in = (unsigned)(c0 + tp0[j]) < (unsigned)d.i0 &&
                            (unsigned)(c1 + tp1[j]) < (unsigned)d.i1 &&
                            (unsigned)(c2 + tp2[j]) < (unsigned)d.i2;
            e[j] = in ? X[corner + toff[j]] : (bf16_t)0;
          }
          v.x = e[0] | ((unsigned)e[1] << 16);
          v.y = e[2] | ((unsigned)e[3] << 16);
          v.z = e[4] | ((unsigned)e[5] << 16);
          v.w = e[6] | ((unsigned)e[7] << 16);
        }
      }
      a[r] = as_bf8(v);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n = nbase + t * 16 + li;
      uint4 w = make_uint4(0, 0, 0, 0);
      if (n < N) w = *reinterpret_cast<const uint4*>(Wm + (size_t)n * ldw + k);
      b[t] = as_bf8(w);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b[t], acc[r][t], 0, 0, 0);
  }
  // D tile: lane holds column li, rows 4*(lane>>4) + j
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n = nbase + t * 16 + li;
      if (n >= N) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int m = mrow0 + r * 16 + (lane >> 4) * 4 + j;
        if (m >= M) continue;
        float v = acc[r][t][j];
        const size_t o = (size_t)m * N + n;
        if (part) {
          part[(size_t)blockIdx.z * M * N + o] = v;
        } else if (MODE == MODE_FWD) {
          v += bias[n];
          if (relu) v = fmaxf(v, 0.f);
          if (Yf) Yf[o] = v;
          if (Y) Y[o] = f2bf(v);
        } else {
          if (mask && !bf_pos(mask[o])) v = 0.f;
          Y[o] = f2bf(v);
        }
      }
    }
}

// epilogue of a split-K forward: Y = relu(sum_z part[z] + bias)
// (Yc, nullable: column 0 of every row also at Yc[row], the scorer's logits)
__global__ void k_split_finalize(const float* __restrict__ part, int ks, int M, int N,
                                 const float* __restrict__ bias, int relu, bf16_t* __restrict__ Y,
                                 float* __restrict__ Yf, const int* __restrict__ gate,
                                 float* __restrict__ Yc = nullptr) {
  if (gate && *gate) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  float v = 0.f;
  for (int z = 0; z < ks; ++z) v += part[(size_t)z * M * N + i];
  v += bias[i % N];
  if (relu) v = fmaxf(v, 0.f);
  if (Yf) Yf[i] = v;
  if (Y) Y[i] = f2bf(v);
  if (Yc && i % N == 0) Yc[i / N] = v;
}

// Weight gradient partial sums over row chunks.  Block (kb, chunk) computes
// part[chunk][n][k] for all n < N and k in [64*kb, 64*kb + 64), staging 32
// rows at a time of dY and of im2col(X) in LDS, transposed so that each lane's
// 8 reduction elements are contiguous.
constexpr int kWgNMax = 128;   // largest N (output channels) handled
constexpr int kWgLd = 40;      // LDS row pitch in bf16 (32 rows + pad)

__global__ void __launch_bounds__(256)
k_wgrad(const bf16_t* __restrict__ X, const bf16_t* __restrict__ dY, ConvDesc d, int M, int N,
        int MC, float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) bf16_t sY[kWgNMax * kWgLd];
  __shared__ __attribute__((aligned(16))) bf16_t sA[64 * kWgLd];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  const int kc0 = blockIdx.x * 64;
  const int chunk = blockIdx.y;
  const int m_lo = chunk * MC, m_hi = min(M, m_lo + MC);
  const int NP = (N + 15) & ~15;
  const int ntn = NP / 16, ntiles = ntn * 4;
  f32x4 acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int ms = m_lo; ms < m_hi; ms += 32) {
    __syncthreads();
    for (int e = tid; e < 32 * NP; e += 256) {
      const int r = e / NP, n = e - r * NP;
      const int m = ms + r;
      sY[n * kWgLd + r] = (m < m_hi && n < N) ? dY[(size_t)m * N + n] : (bf16_t)0;
    }
    for (int e = tid; e < 32 * 64; e += 256) {
      const int r = e >> 6, kc = e & 63;
      const int m = ms + r, k = kc0 + kc;
      bf16_t v = 0;
      if (m < m_hi && k < d.K) {
        int b, p0, p1, p2;
        split3(m, d.o1, d.o2, d.o0, b, p0, p1, p2);
        v = col_fwd(X, d, b, p0, p1, p2, k);
      }
      sA[kc * kWgLd + r] = v;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int tile = wave + 4 * i;
      if (tile >= ntiles) break;
      const int tn = tile >> 2, tk = tile & 3;
      const bf16x8 a = as_bf8(*reinterpret_cast<const uint4*>(&sY[(tn * 16 + li) * kWgLd + lk]));
      const bf16x8 b = as_bf8(*reinterpret_cast<const uint4*>(&sA[(tk * 16 + li) * kWgLd + lk]));
      acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int tile = wave + 4 * i;
    if (tile >= ntiles) break;
    const int tn = tile >> 2, tk = tile & 3;
    const int k = kc0 + tk * 16 + li;
    if (k >= d.K) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = tn * 16 + (lane >> 4) * 4 + j;
      if (n < N) part[((size_t)chunk * N + n) * d.K + k] = acc[i][j];
    }
  }
}

// bias gradient partial sums: partb[chunk][n] = sum over the chunk's rows
__global__ void __launch_bounds__(256)
k_bgrad(const bf16_t* __restrict__ dY, int M, int N, int MC, float* __restrict__ partb) {
  __shared__ float red[256];
  const int chunk = blockIdx.x, n = blockIdx.y;
  const int m_lo = chunk * MC, m_hi = min(M, m_lo + MC);
  float s = 0.f;
  for (int m = m_lo + threadIdx.x; m < m_hi; m += 256) s += bf2f(dY[(size_t)m * N + n]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partb[(size_t)chunk * N + n] = red[0];
}

// out[i] = sum_c part[c*count + i]
__global__ void k_reduce_parts(const float* __restrict__ part, int nchunk, int count,
                               float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * count + i];
  out[i] = s;
}

// 2x2x2 average pool of X [B][d0][d1][d2][C] (even sizes) into rows of a wider
// buffer: out[b*ld + off + ((h0*h1n + h1)*h2n + h2)*C + c], each row written to
// rep consecutive rows (b*rep .. b*rep + rep-1) when rep > 1
__global__ void k_pool_fwd(const bf16_t* __restrict__ X, int B, int d0, int d1, int d2, int C,
                           bf16_t* __restrict__ out, int ld, int off, const int* __restrict__ gate,
                           int rep = 1) {
  if (gate && *gate) return;
  const int h0n = d0 / 2, h1n = d1 / 2, h2n = d2 / 2;
  const int per = h0n * h1n * h2n * C;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * per) return;
  const int b = i / per;
  int r = i - b * per;
  const int c = r % C;
  r /= C;
  const int h2 = r % h2n, h1 = (r / h2n) % h1n, h0 = r / (h2n * h1n);
  float s = 0.f;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int x0 = 2 * h0 + (e >> 2), x1 = 2 * h1 + ((e >> 1) & 1), x2 = 2 * h2 + (e & 1);
    s += bf2f(X[((((size_t)b * d0 + x0) * d1 + x1) * d2 + x2) * C + c]);
  }
  const bf16_t v = f2bf(s * 0.125f);
  for (int q = 0; q < rep; ++q) out[((size_t)b * rep + q) * ld + off + r * C + c] = v;
}

// gradient of the pool, masked by the pooled activation (post-ReLU)
__global__ void k_pool_bwd(const bf16_t* __restrict__ g, int ld, int off, const bf16_t* __restrict__ X,
                           int B, int d0, int d1, int d2, int C, bf16_t* __restrict__ dX) {
  const size_t n = (size_t)B * d0 * d1 * d2 * C;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % C);
  size_t r = i / C;
  const int x2 = (int)(r % d2), x1 = (int)((r / d2) % d1), x0 = (int)((r / ((size_t)d2 * d1)) % d0);
  const int b = (int)(r / ((size_t)d2 * d1 * d0));
  const int h1n = d1 / 2, h2n = d2 / 2;
  const int hr = ((x0 / 2) * h1n + x1 / 2) * h2n + x2 / 2;
  float v = bf2f(g[(size_t)b * ld + off + hr * C + c]) * 0.125f;
  if (!bf_pos(X[i])) v = 0.f;
  dX[i] = f2bf(v);
}

// Soft-target cross entropy per record over its valid orientations (mask
// bit o set).  The logit of orientation o of record b is z[(b*L + o)*zs]
// (channel 0 of the scorer's output row); dz = (softmax - q) / B on valid
// entries and 0 elsewhere, written to the scorer's output gradient (all zs
// channels of the row, the unused ones 0); logits also go compactly to
// zc[b*L + o].
__global__ void k_loss(const float* __restrict__ z, int zs, const float* __restrict__ q,
                       const int32_t* __restrict__ mask, int B, int L, bf16_t* __restrict__ dz,
                       float* __restrict__ loss_rows, float* __restrict__ zc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const unsigned mk = (unsigned)mask[b];
  const float* zr = z + (size_t)b * L * zs;
  const float* qr = q + (size_t)b * L;
  float mx = -INFINITY;
  for (int o = 0; o < L; ++o)
    if (mk >> o & 1u) mx = fmaxf(mx, zr[o * zs]);
  float se = 0.f;
  for (int o = 0; o < L; ++o)
    if (mk >> o & 1u) se += expf(zr[o * zs] - mx);
  const float lse = mx + logf(se);
  float loss = 0.f;
  const float inv = 1.f / (float)B;
  for (int o = 0; o < L; ++o) {
    float g = 0.f;
    const float zo = zr[o * zs];
    if (mk >> o & 1u) {
      const float p = expf(zo - lse);
      loss -= qr[o] * (zo - lse);
      g = (p - qr[o]) * inv;
    }
    bf16_t* d = dz + ((size_t)b * L + o) * zs;
    d[0] = f2bf(g);
    for (int c = 1; c < zs; ++c) d[c] = 0;
    zc[(size_t)b * L + o] = zo;
  }
  loss_rows[b] = mk ? loss : 0.f;
}

// compact logits: zc[r] = z[r*zs] (channel 0 of each scorer row)
__global__ void k_logits_compact(const float* __restrict__ z, int zs, int rows,
                                 float* __restrict__ zc, const int* __restrict__ gate) {
  if (gate && *gate) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) zc[r] = z[(size_t)r * zs];
}

// scorer input rows: concat[r] = [item row r (ni) | tray row r / R (nt)]
__global__ void k_concat_rows(const bf16_t* __restrict__ item, int ni, const bf16_t* __restrict__ tray,
                              int nt, int rows, int R, bf16_t* __restrict__ out,
                              const int* __restrict__ gate) {
  if (gate && *gate) return;
  const int w = ni + nt;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * w) return;
  const int r = (int)(i / w), c = (int)(i - (size_t)r * w);
  out[i] = c < ni ? item[(size_t)r * ni + c] : tray[(size_t)(r / R) * nt + (c - ni)];
}

// gradient of the tray rows: sum of the tray part of the R scorer rows of
// each record, g[b][j] = sum_o gc[(b*R + o)*w + ni + j]
__global__ void k_tray_gsum(const bf16_t* __restrict__ gc, int ni, int nt, int B, int R,
                            bf16_t* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * nt) return;
  const int b = i / nt, j = i - b * nt;
  const int w = ni + nt;
  float s = 0.f;
  for (int o = 0; o < R; ++o) s += bf2f(gc[((size_t)b * R + o) * w + ni + j]);
  g[i] = f2bf(s);
}

// Adam (bias-corrected) on the fp32 master parameters
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// bf16 working copies of one layer's weights: W [cout][Kp] (zero padded) and
// the input-gradient layout Wt [cin][Ktp], Wt[c][tap*cout + n] = W[n][tap*cin + c]
__global__ void k_cast_weights(const float* __restrict__ w, ConvDesc d, bf16_t* __restrict__ wb,
                               bf16_t* __restrict__ wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nW = d.cout * d.Kp, nT = d.cin * d.Ktp;
  if (i < nW) {
    const int n = i / d.Kp, k = i - n * d.Kp;
    wb[i] = k < d.K ? f2bf(w[(size_t)n * d.K + k]) : (bf16_t)0;
  } else if (i < nW + nT) {
    const int j = i - nW;
    const int c = j / d.Ktp, kt = j - c * d.Ktp;
    bf16_t v = 0;
    if (kt < d.Kt) {
      const int tap = kt / d.cout, n = kt - tap * d.cout;
      v = f2bf(w[(size_t)n * d.K + tap * d.cin + c]);
    }
    wt[j] = v;
  }
}

// dst row i = src row idx[i]; rows of `words` words of type T (uint4 or unsigned)
template <typename T>
__global__ void k_gather_rows(T* __restrict__ dst, const T* __restrict__ src,
                              long long words, const long long* __restrict__ idx, int n) {
  const int i = blockIdx.y;
  if (i >= n) return;
  const T* s = src + idx[i] * words;
  T* d = dst + (long long)i * words;
  for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < words;
       w += (long long)gridDim.x * blockDim.x)
    d[w] = s[w];
}

// ---------------------------------------------------------------------------
// input features
// ---------------------------------------------------------------------------
constexpr int kItemCells = 32;                 // item grid 32^3, pitch 2 voxels
constexpr int kTray0 = 64, kTray1 = 32, kTray2 = 32;

// An orientation of a base item as an index map: oriented axis a runs along
// base axis perm[a], reversed when flip[a] != 0; the base item's voxels start
// `off` ints past the item pointer handed to the kernel.
struct ItemMap {
  int bx, by, bz;
  int perm[3], flip[3];
  int off;
};

// occupancy of the oriented item: cell (a,b,c) = (# of oriented voxels > 0
// among x in {2a,2a+1}, y in {2b,2b+1}, z in {2c,2c+1}) / 8.  Grid (128, n):
// item map i -> out + i * 32^3
__global__ void k_item_feat(const int* __restrict__ item, const ItemMap* __restrict__ maps,
                            bf16_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kItemCells * kItemCells * kItemCells) return;
  const ItemMap mp = maps[blockIdx.y];
  const int bd[3] = {mp.bx, mp.by, mp.bz};
  int ext[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) ext[a] = bd[mp.perm[a]];
  const int c = i % kItemCells, b = (i / kItemCells) % kItemCells, a = i / (kItemCells * kItemCells);
  int cnt = 0;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int w[3] = {2 * a + (e >> 2), 2 * b + ((e >> 1) & 1), 2 * c + (e & 1)};
    if (w[0] >= ext[0] || w[1] >= ext[1] || w[2] >= ext[2]) continue;
    int u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int ax = mp.perm[d];
      u[ax] = mp.flip[d] ? bd[ax] - 1 - w[d] : w[d];
    }
    if (item[(size_t)mp.off + ((size_t)u[0] * mp.by + u[1]) * mp.bz + u[2]] > 0) ++cnt;
  }
  out[(size_t)blockIdx.y * (kItemCells * kItemCells * kItemCells) + i] = f2bf((float)cnt * 0.125f);
}

// column heights: h[x][y] = 1 + max{z : tray > 0}, 0 for an empty column.
// One wavefront per column: 64 heights per ballot, from the top down.
__global__ void k_heightmap(const int* __restrict__ tray, int nx, int ny, int nz,
                            int* __restrict__ h, const int* __restrict__ gate) {
  if (gate && *gate) return;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= nx * ny) return;
  const int* col = tray + (size_t)i * nz;
  int top = 0;
  for (int z0 = ((nz - 1) / 64) * 64; z0 >= 0; z0 -= 64) {
    const int z = z0 + lane;
    const unsigned long long b = __ballot(z < nz && col[z] > 0);
    if (b) {
      top = z0 + 64 - __clzll((long long)b);
      break;
    }
  }
  if (lane == 0) h[i] = top;
}

__device__ __forceinline__ void cell_range(int a, int n, int G, int& lo, int& hi) {
  lo = (int)((long long)a * n / G);
  hi = (int)((long long)(a + 1) * n / G);
  if (hi <= lo) hi = lo + 1;
}

// tray features [64][32][32][2]: channel 0 = occupied fraction of the cell,
// channel 1 = fraction of the cell below its columns' heights (z < h[x][y])
__global__ void k_tray_feat(const int* __restrict__ tray, const int* __restrict__ h, int nx, int ny,
                            int nz, bf16_t* __restrict__ out, const int* __restrict__ gate) {
  if (gate && *gate) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kTray0 * kTray1 * kTray2) return;
  const int c = i % kTray2, b = (i / kTray2) % kTray1, a = i / (kTray2 * kTray1);
  int x0, x1, y0, y1, z0, z1;
  cell_range(a, nx, kTray0, x0, x1);
  cell_range(b, ny, kTray1, y0, y1);
  cell_range(c, nz, kTray2, z0, z1);
  int occ = 0, below = 0;
  for (int x = x0; x < x1; ++x)
    for (int y = y0; y < y1; ++y) {
      const int hh = h[x * ny + y];
      const int* col = tray + ((size_t)x * ny + y) * nz;
      for (int z = z0; z < z1; ++z) {
        occ += col[z] > 0;
        below += z < hh;
      }
    }
  const float cnt = (float)((x1 - x0) * (y1 - y0) * (z1 - z0));
  out[2 * i] = f2bf((float)occ / cnt);
  out[2 * i + 1] = f2bf((float)below / cnt);
}

// k_heightmap and k_tray_feat in one kernel, from the tray's 128-bit column
// masks (bit z of tf[x ny + y] set when voxel (x, y, z) is nonzero; nz <= 128)
// instead of the int32 grid: the height of a column is one past its highest
// set bit and the occupied voxels of a cell are a popcount.  Equal to the two
// kernels when no tray voxel is negative (they count voxels > 0).
__device__ __forceinline__ int popc_range(unsigned long long lo, unsigned long long hi, int z0,
                                          int z1) {
  // set bits of (lo, hi) in [z0, z1), 0 <= z0 < z1 <= 128
  auto mask64 = [](int a, int b) -> unsigned long long {   // bits [a, b) of one word, 0..64
    if (b <= a) return 0ULL;
    const unsigned long long hi_m = b >= 64 ? ~0ULL : ((1ULL << b) - 1ULL);
    const unsigned long long lo_m = a >= 64 ? ~0ULL : ((1ULL << a) - 1ULL);
    return hi_m & ~lo_m;
  };
  const unsigned long long ml = mask64(min(z0, 64), min(z1, 64));
  const unsigned long long mh = mask64(max(z0 - 64, 0), max(z1 - 64, 0));
  return __popcll(lo & ml) + __popcll(hi & mh);
}
__global__ void __launch_bounds__(256)
k_tray_feat_m(const unsigned long long* __restrict__ tf, int nx, int ny, int nz,
              bf16_t* __restrict__ out, const int* __restrict__ gate) {
  if (gate && *gate) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kTray0 * kTray1 * kTray2) return;
  const int c = i % kTray2, b = (i / kTray2) % kTray1, a = i / (kTray2 * kTray1);
  int x0, x1, y0, y1, z0, z1;
  cell_range(a, nx, kTray0, x0, x1);
  cell_range(b, ny, kTray1, y0, y1);
  cell_range(c, nz, kTray2, z0, z1);
  int occ = 0, below = 0;
  for (int x = x0; x < x1; ++x)
    for (int y = y0; y < y1; ++y) {
      const size_t q = (size_t)x * ny + y;
      const unsigned long long lo = tf[2 * q], hi = tf[2 * q + 1];
      const int hh = hi ? 128 - __clzll((long long)hi) : (lo ? 64 - __clzll((long long)lo) : 0);
      occ += popc_range(lo, hi, z0, z1);
      below += max(0, min(hh, z1) - z0);
    }
  const float cnt = (float)((x1 - x0) * (y1 - y0) * (z1 - z0));
  out[2 * i] = f2bf((float)occ / cnt);
  out[2 * i + 1] = f2bf((float)below / cnt);
}

}  // namespace spknn
