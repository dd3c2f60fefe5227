// spk_fft.hpp -- workgroup-cooperative mixed-radix Stockham FFT in LDS (gfx950).
//
// One buffer in LDS holds `batch` transforms of length n, element e of
// transform b at buf[b*bs + e*es].  Each stage reads every butterfly's R
// inputs into registers, barriers, and writes the R outputs in Stockham
// (self-sorting) order, so a single LDS buffer suffices and results come out
// in natural order.  Radices 2,3,4,5,7,8; twiddles come from a table of the n
// roots exp(-2*pi*i*q/n), computed on the host in double and rounded once to
// float.  Stage s with radix R and span Ns (product of earlier radices):
//     j in [0, n/R), k = j mod Ns,
//     v_r = x[j + r*n/R] * w_n^{r*k*n/(Ns*R)},  V = DFT_R(v),
//     y[(j/Ns)*Ns*R + k + m*Ns] = V_m.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

constexpr int kMaxStages = 12;

constexpr int kRadixBits = 5;
constexpr int kMaxRadix = 25;       // largest butterfly (a 5x5 composite)
constexpr int kPencilRadix = 16;    // largest butterfly in the x-pencil kernels
constexpr unsigned long long kRadixMask = 31ULL;

struct AxisPlan {
  int n;
  int nst;
  unsigned long long rad;   // kRadixBits per stage, stage 0 in the low bits
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// i*sgn*z with sgn = -1 (forward) or +1 (inverse)
template <bool FWD>
__device__ __forceinline__ float2 isg(float2 z) {
  return FWD ? make_float2(z.y, -z.x) : make_float2(-z.y, z.x);
}

// cos/sin(2*pi*q/R) for the odd radices
template <int R> struct OddTab;
template <> struct OddTab<3> {
  static constexpr float c[3] = {1.0f, -0.5f, -0.5f};
  static constexpr float s[3] = {0.0f, 0.86602540378443865f, -0.86602540378443865f};
};
template <> struct OddTab<5> {
  static constexpr float c[5] = {1.0f, 0.30901699437494742f, -0.80901699437494742f,
                                 -0.80901699437494742f, 0.30901699437494742f};
  static constexpr float s[5] = {0.0f, 0.95105651629515357f, 0.58778525229247313f,
                                 -0.58778525229247313f, -0.95105651629515357f};
};
template <> struct OddTab<7> {
  static constexpr float c[7] = {1.0f, 0.62348980185873353f, -0.22252093395631440f,
                                 -0.90096886790241913f, -0.90096886790241913f,
                                 -0.22252093395631440f, 0.62348980185873353f};
  static constexpr float s[7] = {0.0f, 0.78183148246802981f, 0.97492791218182361f,
                                 0.43388373911755812f, -0.43388373911755812f,
                                 -0.97492791218182361f, -0.78183148246802981f};
};

template <int R, bool FWD> struct Bfly;

template <bool FWD> struct Bfly<2, FWD> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool FWD> struct Bfly<4, FWD> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    float2 s13 = cadd(v[1], v[3]), d13 = isg<FWD>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

template <bool FWD> struct Bfly<8, FWD> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<4, FWD>::run(e);
    Bfly<4, FWD>::run(o);
    const float h = 0.70710678118654752f;
    // o[m] *= w8^m, w8 = exp(sgn*i*pi/4)
    float2 o1 = o[1], o3 = o[3];
    if (FWD) {
      o[1] = make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
      o[3] = make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    } else {
      o[1] = make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y));
      o[3] = make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y));
    }
    o[2] = isg<FWD>(o[2]);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      v[m] = cadd(e[m], o[m]);
      v[m + 4] = csub(e[m], o[m]);
    }
  }
};

// odd radix: pair r with R-r
template <int R, bool FWD> struct BflyOdd {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr int H = (R - 1) / 2;
    float2 sp[H], dm[H];
    float2 v0 = v[0];
    float2 tot = v0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      sp[r - 1] = cadd(v[r], v[R - r]);
      dm[r - 1] = csub(v[r], v[R - r]);
      tot = cadd(tot, sp[r - 1]);
    }
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      float2 re = v0, im = make_float2(0.f, 0.f);
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const int q = (r * m) % R;
        re = cadd(re, cscale(sp[r - 1], OddTab<R>::c[q]));
        im = cadd(im, cscale(dm[r - 1], OddTab<R>::s[q]));
      }
      float2 t = isg<FWD>(im);
      v[m] = cadd(re, t);
      v[R - m] = csub(re, t);
    }
    v[0] = tot;
  }
};
template <bool FWD> struct Bfly<3, FWD> : BflyOdd<3, FWD> {};
template <bool FWD> struct Bfly<5, FWD> : BflyOdd<5, FWD> {};
template <bool FWD> struct Bfly<7, FWD> : BflyOdd<7, FWD> {};

// ---- composite radices in registers -------------------------------------
// cos/sin(2*pi*m/R) evaluated at compile time (range-reduced Taylor series in
// double, rounded once to float).
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 14; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
// exp(i*2*pi*m/R) for 0 <= m < R, via the octant of the angle
constexpr double root_cos(int m, int R) {
  m %= R;
  const double a = kTwoPi * m / R;
  return a <= kTwoPi / 2 ? (a <= kTwoPi / 4 ? taylor_cos(a) : -taylor_cos(kTwoPi / 2 - a))
                         : (a <= 3 * kTwoPi / 4 ? -taylor_cos(a - kTwoPi / 2) : taylor_cos(kTwoPi - a));
}
constexpr double root_sin(int m, int R) {
  m %= R;
  const double a = kTwoPi * m / R;
  return a <= kTwoPi / 2 ? (a <= kTwoPi / 4 ? taylor_sin(a) : taylor_sin(kTwoPi / 2 - a))
                         : (a <= 3 * kTwoPi / 4 ? -taylor_sin(a - kTwoPi / 2) : -taylor_sin(kTwoPi - a));
}

template <int R>
struct Roots {
  float c[R], s[R];
  constexpr Roots() : c(), s() {
    for (int m = 0; m < R; ++m) {
      c[m] = (float)root_cos(m, R);
      s[m] = (float)root_sin(m, R);
    }
  }
};
template <int R>
struct RootTab {
  static constexpr Roots<R> v{};
};

// X[k1 + A*k2] = sum_{n2} w_B^{n2 k2} w_R^{n2 k1} sum_{n1} x[B n1 + n2] w_A^{n1 k1},
// R = A*B, every index a compile-time constant after unrolling.
template <int A, int B, bool FWD>
struct BflyCT {
  static constexpr int R = A * B;
  __device__ __forceinline__ static void run(float2* v) {
    float2 y[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) y[n2][n1] = v[B * n1 + n2];
      Bfly<A, FWD>::run(y[n2]);
#pragma unroll
      for (int k1 = 1; k1 < A; ++k1) {
        if (n2 == 0) continue;
        const int q = (n2 * k1) % R;
        const float c = RootTab<R>::v.c[q];
        const float s = FWD ? -RootTab<R>::v.s[q] : RootTab<R>::v.s[q];
        y[n2][k1] = cmul(y[n2][k1], make_float2(c, s));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      float2 t[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n2] = y[n2][k1];
      Bfly<B, FWD>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
    }
  }
};
template <bool FWD> struct Bfly<6, FWD> : BflyCT<2, 3, FWD> {};
template <bool FWD> struct Bfly<9, FWD> : BflyCT<3, 3, FWD> {};
template <bool FWD> struct Bfly<10, FWD> : BflyCT<2, 5, FWD> {};
template <bool FWD> struct Bfly<12, FWD> : BflyCT<4, 3, FWD> {};
template <bool FWD> struct Bfly<14, FWD> : BflyCT<2, 7, FWD> {};
template <bool FWD> struct Bfly<15, FWD> : BflyCT<3, 5, FWD> {};
template <bool FWD> struct Bfly<16, FWD> : BflyCT<4, 4, FWD> {};
template <bool FWD> struct Bfly<25, FWD> : BflyCT<5, 5, FWD> {};

// n / d for 0 <= n < 2^22, d >= 1, with inv = 1/d: the float quotient is off
// by at most one, fixed by the remainder test.
__device__ __forceinline__ int fdiv(int n, int d, float inv) {
  int q = (int)((float)n * inv);
  const int r = n - q * d;
  q += (r >= d) ? 1 : 0;
  q -= (r < 0) ? 1 : 0;
  return q;
}

// One Stockham stage over `batch` transforms.  NT threads, at most MAXEL
// elements per thread (batch*n <= NT*MAXEL, checked by the host planner).
// jfast: consecutive lanes take consecutive j (use when es == 1), otherwise
// consecutive lanes take consecutive transforms (strided transforms).
template <int R, bool FWD, int NT, int MAXEL>
__device__ __forceinline__ void fft_stage(float2* buf, int n, int Ns, int es, int batch, int bs,
                                          bool jfast, const float2* tw, int tid) {
  constexpr int MS = (MAXEL + R - 1) / R;
  const int m = n / R;
  const int total = batch * m;
  const int tws = n / (Ns * R);
  const int dv = jfast ? m : batch;
  const float inv_dv = 1.0f / (float)dv;
  const float inv_ns = 1.0f / (float)Ns;
  // Two layouts only: rows (es == 1, transforms bs apart, lanes along j) and
  // columns (bs == 1, elements es apart, lanes along the transforms).  Every
  // per-r offset is then a uniform step times a compile-time r.
  const int in_step = jfast ? m : m * es;
  const int out_step = jfast ? Ns : Ns * es;
  float2 v[MS][R];
  int ob[MS];
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    const int b = tid + s * NT;
    ob[s] = -1;
    if (b < total) {
      const int qd = fdiv(b, dv, inv_dv);
      const int rd = b - qd * dv;
      const int j = jfast ? rd : qd;
      const int k = j - fdiv(j, Ns, inv_ns) * Ns;
      const int jo = (j - k) * R + k;
      int ib, obase;
      if (jfast) {
        const int off = qd * bs;
        ib = off + j;
        obase = off + jo;
      } else {
        ib = rd + j * es;
        obase = rd + jo * es;
      }
#pragma unroll
      for (int r = 0; r < R; ++r) v[s][r] = buf[ib + r * in_step];
      if (k != 0) {
        const int kt = k * tws;
        int ti = kt;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[ti];
          ti += kt;
          if (!FWD) w.y = -w.y;
          v[s][r] = cmul(v[s][r], w);
        }
      }
      Bfly<R, FWD>::run(v[s]);
      ob[s] = obase;
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    if (ob[s] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[ob[s] + r * out_step] = v[s][r];
    }
  }
  __syncthreads();
}

// Full transform of length p.n (all stages).  Caller has synchronised the
// buffer; returns synchronised.
template <bool FWD, int NT, int MAXEL, int RMAX = kMaxRadix>
__device__ __forceinline__ void lds_fft(float2* buf, const AxisPlan& p, int es, int batch, int bs,
                                        bool jfast,
                        const float2* tw, int tid) {
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int R = (int)((p.rad >> (kRadixBits * s)) & kRadixMask);
    // An opaque per-iteration copy of tid: without it the compiler hoists
    // every radix case's index set-up out of the loop and keeps all of them
    // live at once (the radix cases then spill at 128 VGPRs).
    int t = tid;
    asm volatile("" : "+v"(t));
    // radices above RMAX are not compiled in (the host plans such axes
    // without them): the largest butterfly sets the kernel's register need
#define SPK_STAGE(RR)                                                              \
  case RR:                                                                         \
    if constexpr (RR <= RMAX)                                                      \
      fft_stage<RR, FWD, NT, MAXEL>(buf, p.n, Ns, es, batch, bs, jfast, tw, t);  \
    break;
    switch (R) {
      SPK_STAGE(2) SPK_STAGE(3) SPK_STAGE(4) SPK_STAGE(5) SPK_STAGE(6) SPK_STAGE(7)
      SPK_STAGE(9) SPK_STAGE(10) SPK_STAGE(12) SPK_STAGE(14) SPK_STAGE(15) SPK_STAGE(16)
      SPK_STAGE(25)
      default: fft_stage<8, FWD, NT, MAXEL>(buf, p.n, Ns, es, batch, bs, jfast, tw, t); break;
    }
#undef SPK_STAGE
    Ns *= R;
  }
}

// ---- compile-time plans ----------------------------------------------------
// A transform length with its radices fixed at compile time (FPlan<N, R...>,
// applied in the order listed), or the runtime plan (DynPlan).  With a fixed
// plan every stride, count and divisor of the stages below is a constant:
// index arithmetic becomes shifts and multiplications, LDS addresses take
// immediate offsets and each thread's butterfly loop is fully unrolled.
struct DynPlan {
  static constexpr int N = 0;
};
template <int N_, int... Rs>
struct FPlan {
  static constexpr int N = N_;
};

template <int A, int B>
struct CeilDiv {
  static constexpr int v = (A + B - 1) / B;
};

// One Stockham stage (the same element mapping as fft_stage) with N, R, NS and
// the strides fixed.  Rows (JFAST): transforms BS apart, elements contiguous,
// batch (<= MAXB transforms) at run time.  Columns (!JFAST): transforms
// contiguous, elements ES apart, BATCH transforms fixed.
template <int R, bool FWD, int NT, int N, int NS, int ES, int BS, bool JFAST, int BATCH, int MAXB>
__device__ __forceinline__ void fstage(float2* buf, int batch_rt, const float2* tw, int tid) {
  constexpr int M = N / R;
  constexpr int TWS = N / (NS * R);
  constexpr int MAXT = (JFAST ? MAXB : BATCH) * M;
  constexpr int MS = CeilDiv<MAXT, NT>::v;
  constexpr int IN_STEP = JFAST ? M : M * ES;
  constexpr int OUT_STEP = JFAST ? NS : NS * ES;
  const int total = JFAST ? batch_rt * M : MAXT;
  float2 v[MS][R];
  int ob[MS];
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    const int b = tid + s * NT;
    ob[s] = -1;
    if (b < total) {
      int j, ib, obase;
      if (JFAST) {
        const int q = b / M;
        j = b - q * M;
        const int k = j % NS;
        ib = q * BS + j;
        obase = q * BS + (j - k) * R + k;
      } else {
        j = b / BATCH;
        const int col = b - j * BATCH;
        const int k = j % NS;
        ib = col + j * ES;
        obase = col + ((j - k) * R + k) * ES;
      }
#pragma unroll
      for (int r = 0; r < R; ++r) v[s][r] = buf[ib + r * IN_STEP];
      const int k = j % NS;
      if (NS > 1 && k != 0) {
        const int kt = k * TWS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * kt];
          if (!FWD) w.y = -w.y;
          v[s][r] = cmul(v[s][r], w);
        }
      }
      Bfly<R, FWD>::run(v[s]);
      ob[s] = obase;
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    if (ob[s] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[ob[s] + r * OUT_STEP] = v[s][r];
    }
  }
  __syncthreads();
}

template <bool FWD, int NT, int N, int NS, int ES, int BS, bool JFAST, int BATCH, int MAXB, int R,
          int... Rest>
__device__ __forceinline__ void fixed_fft_stages(float2* buf, int batch, const float2* tw,
                                                 int tid) {
  fstage<R, FWD, NT, N, NS, ES, BS, JFAST, BATCH, MAXB>(buf, batch, tw, tid);
  if constexpr (sizeof...(Rest) > 0)
    fixed_fft_stages<FWD, NT, N, NS * R, ES, BS, JFAST, BATCH, MAXB, Rest...>(buf, batch, tw, tid);
}

template <bool FWD, int NT, class P, int ES, int BS, bool JFAST, int BATCH, int MAXB>
struct FixedFft;
template <bool FWD, int NT, int N, int... Rs, int ES, int BS, bool JFAST, int BATCH, int MAXB>
struct FixedFft<FWD, NT, FPlan<N, Rs...>, ES, BS, JFAST, BATCH, MAXB> {
  __device__ __forceinline__ static void run(float2* buf, int batch, const float2* tw, int tid) {
    fixed_fft_stages<FWD, NT, N, 1, ES, BS, JFAST, BATCH, MAXB, Rs...>(buf, batch, tw, tid);
  }
};

// ---- in place, decimation in frequency (compile-time plans) ----------------
// The transform of FixedFft computed in place.  Stage s works on blocks of
// length L = N / (R_0 ... R_{s-1}); its butterfly (block, n), n < S = L / R_s,
// takes the R_s elements block*L + n + r*S, applies the R_s-point DFT,
// multiplies output m by w_L^(n*m) and writes it back to block*L + n + m*S.
// No two butterflies of a stage share an element, so a stage needs one
// barrier (after it) where Stockham's needs two.  The output comes out
// digit-reversed: X[k], k = m_0 + R_0 (m_1 + R_1 (m_2 + ...)), lands at
// position sum_s m_s * N / (R_0 ... R_s) (PlanPos<P>::at).  Rows (JFAST):
// transform q starts at rowmap(q) * BS, elements ES apart; columns (!JFAST):
// BATCH transforms at consecutive offsets, elements ES apart.
template <int N, int... Rs> struct IpPos;
template <int N, int R, int... Rest> struct IpPos<N, R, Rest...> {
  __device__ __forceinline__ static int at(int k) {
    if constexpr (sizeof...(Rest) == 0) return k;
    else return (k % R) * (N / R) + IpPos<N / R, Rest...>::at(k / R);
  }
};
template <class P> struct PlanPos;
template <int N, int... Rs> struct PlanPos<FPlan<N, Rs...>> {
  __device__ __forceinline__ static int at(int k) { return IpPos<N, Rs...>::at(k); }
};

struct IdentityMap {
  __device__ __forceinline__ int operator()(int q) const { return q; }
};

template <int R, bool FWD, int NT, int N, int L, int ES, int BS, bool JFAST, int BATCH, int MAXB,
          class RowMap>
__device__ __forceinline__ void ipstage(float2* buf, int batch_rt, const float2* tw, int tid,
                                        RowMap rowmap) {
  constexpr int S = L / R, M = N / R, TWS = N / L;
  constexpr int MAXT = (JFAST ? MAXB : BATCH) * M;
  constexpr int MS = CeilDiv<MAXT, NT>::v;
  const int total = JFAST ? batch_rt * M : MAXT;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    const int b = tid + s * NT;
    if (b < total) {
      int j, base;
      if (JFAST) {
        const int q = b / M;
        j = b - q * M;
        base = rowmap(q) * BS;
      } else {
        j = b / BATCH;
        base = b - j * BATCH;
      }
      const int blk = j / S, n = j - blk * S;
      const int i0 = base + (blk * L + n) * ES;
      float2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = buf[i0 + r * S * ES];
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
      for (int r = 0; r < R; ++r) buf[i0 + r * S * ES] = v[r];
    }
  }
  __syncthreads();
}

template <bool FWD, int NT, int N, int L, int ES, int BS, bool JFAST, int BATCH, int MAXB,
          class RowMap, int R, int... Rest>
__device__ __forceinline__ void ip_stages(float2* buf, int batch, const float2* tw, int tid,
                                          RowMap rm) {
  ipstage<R, FWD, NT, N, L, ES, BS, JFAST, BATCH, MAXB, RowMap>(buf, batch, tw, tid, rm);
  if constexpr (sizeof...(Rest) > 0)
    ip_stages<FWD, NT, N, L / R, ES, BS, JFAST, BATCH, MAXB, RowMap, Rest...>(buf, batch, tw, tid,
                                                                             rm);
}

template <bool FWD, int NT, class P, int ES, int BS, bool JFAST, int BATCH, int MAXB>
struct IpFft;
template <bool FWD, int NT, int N, int... Rs, int ES, int BS, bool JFAST, int BATCH, int MAXB>
struct IpFft<FWD, NT, FPlan<N, Rs...>, ES, BS, JFAST, BATCH, MAXB> {
  template <class RowMap>
  __device__ __forceinline__ static void run(float2* buf, int batch, const float2* tw, int tid,
                                             RowMap rm) {
    ip_stages<FWD, NT, N, N, ES, BS, JFAST, BATCH, MAXB, RowMap, Rs...>(buf, batch, tw, tid, rm);
  }
};

// A transform with the compile-time plan P when P is fixed (the strides and
// counts then come from the template arguments), else with the runtime plan p.
template <bool FWD, int NT, int MAXEL, int RMAX, class P, int ES, int BS, bool JFAST, int BATCH,
          int MAXB>
__device__ __forceinline__ void fft_axis(float2* buf, const AxisPlan& p, int es, int batch,
                                         int bs, bool jfast, const float2* tw, int tid) {
  if constexpr (P::N > 0)
    FixedFft<FWD, NT, P, ES, BS, JFAST, BATCH, MAXB>::run(buf, batch, tw, tid);
  else
    lds_fft<FWD, NT, MAXEL, RMAX>(buf, p, es, batch, bs, jfast, tw, tid);
}

// e / d and e % d with d = D when D > 0 (a constant), else the runtime d
template <int D>
__device__ __forceinline__ int div_by(int e, int d, float inv) {
  if constexpr (D > 0) return e / D;
  else return fdiv(e, d, inv);
}

}  // namespace spk
