// Host side of the learned orientation proposer (kernels in spk_nn.hpp).
// Included by spk_api.hip after its error, buffer and launch helpers.
//
// An orientation scorer shared by all orientations (260,696 parameters,
// channels-last, every conv 3x3x3 stride 2 pad 1, ReLU after every layer but
// the last):
//   oriented item [32,32,32,1] -> I1 [16^3,16] -> I2 [8^3,32] -> I3 [4^3,48] -> avg2 [2^3,48] = 384
//   tray [64,32,32,2] -> T1 [32,16,16,16] -> T2 [16,8,8,32] -> T3 [8,4,4,48] -> avg2 [4,2,2,48] = 768
//   [item 384 | tray 768] -> F1 128 -> F2 8 (channel 0 = the orientation's logit)
// A record is one tray and the 24 orientations of one item: the item branch
// runs on the 24 oriented items, the tray branch once, the head on 24 rows,
// and the 24 logits form one softmax.
// Parameters are stored layer by layer, W [cout][ks^3*cin] (tap-major,
// channel-minor columns; tap = (t0*3 + t1)*3 + t2) followed by b [cout].
#pragma once

#include "spk_nn.hpp"

namespace {

using spknn::bf16_t;
using spknn::ConvDesc;

enum { NN_I1, NN_I2, NN_I3, NN_T1, NN_T2, NN_T3, NN_F1, NN_F2, NN_LAYERS };
constexpr int kNnItemFeat = 32768;     // 32^3 x 1
constexpr int kNnTrayFeat = 131072;    // 64x32x32 x 2
constexpr int kNnItemPooled = 384;
constexpr int kNnTrayPooled = 768;
constexpr int kNnConcat = 1152;
constexpr int kNnLogits = 24;          // orientations per record
constexpr int kNnOut = 8;              // scorer outputs per row (channel 0 used)

int round32(int v) { return (v + 31) & ~31; }

ConvDesc conv_desc(int i0, int i1, int i2, int cin, int cout) {
  ConvDesc d{};
  d.i0 = i0; d.i1 = i1; d.i2 = i2; d.cin = cin;
  d.o0 = (i0 - 1) / 2 + 1; d.o1 = (i1 - 1) / 2 + 1; d.o2 = (i2 - 1) / 2 + 1; d.cout = cout;
  d.ks = 3; d.st = 2; d.pad = 1;
  d.K = 27 * cin; d.Kp = round32(d.K);
  d.Kt = 27 * cout; d.Ktp = round32(d.Kt);
  return d;
}

ConvDesc dense_desc(int cin, int cout) {
  ConvDesc d{};
  d.i0 = d.i1 = d.i2 = 1; d.cin = cin;
  d.o0 = d.o1 = d.o2 = 1; d.cout = cout;
  d.ks = 1; d.st = 1; d.pad = 0;
  d.K = cin; d.Kp = round32(cin);
  d.Kt = cout; d.Ktp = round32(cout);
  return d;
}

struct NnLayer {
  ConvDesc d;
  size_t w_off = 0, b_off = 0;     // master parameters (floats)
  size_t wb_off = 0, wt_off = 0;   // bf16 working copies
  int relu = 1;
  size_t in_size() const { return (size_t)d.i0 * d.i1 * d.i2 * d.cin; }
  size_t out_size() const { return (size_t)d.o0 * d.o1 * d.o2 * d.cout; }
};

size_t wgrad_chunks_max(const ConvDesc& d) {
  const int kb = (d.K + 63) / 64;
  return (size_t)((512 + kb - 1) / kb);
}

}  // namespace

struct spk_nn {
  int device = 0;
  hipStream_t stream = nullptr;
  int max_batch = 0;           // records per batch (24 scorer rows each)
  NnLayer L[NN_LAYERS];
  size_t n_params = 0, n_wbf = 0;
  DevBuf<float> params, grads, m1, m2, out, logits, loss_rows, part;
  DevBuf<bf16_t> wbf, ipool, tpool, concat, g_concat, g_tpool;
  DevBuf<bf16_t> act[NN_LAYERS], grad[NN_LAYERS];
  bool dirty = true;   // bf16 copies out of date
  // device word that, when nonzero, makes the inference kernels exit at once
  // (the device-driven learned loop skips items it does not propose for)
  const int* gate = nullptr;
  long long step = 0;
};

namespace {

template <typename K, typename... Args>
void nn_launch(hipStream_t s, K kernel, dim3 grid, dim3 block, size_t shmem, Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
  HIP_OK(hipGetLastError());
}

void nn_build(spk_nn* nn, int max_batch) {
  nn->max_batch = max_batch;
  nn->L[NN_I1].d = conv_desc(32, 32, 32, 1, 16);
  nn->L[NN_I2].d = conv_desc(16, 16, 16, 16, 32);
  nn->L[NN_I3].d = conv_desc(8, 8, 8, 32, 48);
  nn->L[NN_T1].d = conv_desc(64, 32, 32, 2, 16);
  nn->L[NN_T2].d = conv_desc(32, 16, 16, 16, 32);
  nn->L[NN_T3].d = conv_desc(16, 8, 8, 32, 48);
  nn->L[NN_F1].d = dense_desc(kNnConcat, 128);
  nn->L[NN_F2].d = dense_desc(128, kNnOut);
  nn->L[NN_F2].relu = 0;
  size_t p = 0, w = 0, part = 0;
  for (auto& l : nn->L) {
    l.w_off = p;
    p += (size_t)l.d.cout * l.d.K;
    l.b_off = p;
    p += l.d.cout;
    l.wb_off = w;
    w += (size_t)l.d.cout * l.d.Kp;
    l.wt_off = w;
    w += (size_t)l.d.cin * l.d.Ktp;
    part = std::max(part, wgrad_chunks_max(l.d) * l.d.cout * (l.d.K + 1));
    if (l.d.cout % 8) throw SpkError{SPK_EINVAL, "proposer layer width must be a multiple of 8"};
  }
  nn->n_params = p;
  nn->n_wbf = w;
  nn->params.ensure(p);
  nn->grads.ensure(p);
  nn->m1.ensure(p);
  nn->m2.ensure(p);
  nn->wbf.ensure(w);
  nn->part.ensure(part);
  const size_t B = (size_t)max_batch, R = B * kNnLogits;
  for (int i = 0; i < NN_LAYERS; ++i) {
    const bool tray = i == NN_T1 || i == NN_T2 || i == NN_T3;
    nn->act[i].ensure((tray ? B : R) * nn->L[i].out_size());
    nn->grad[i].ensure((tray ? B : R) * nn->L[i].out_size());
  }
  nn->ipool.ensure(R * kNnItemPooled);
  nn->tpool.ensure(B * kNnTrayPooled);
  nn->g_tpool.ensure(B * kNnTrayPooled);
  nn->concat.ensure(R * kNnConcat);
  nn->g_concat.ensure(R * kNnConcat);
  nn->out.ensure(R * kNnOut);
  nn->logits.ensure(R);
  nn->loss_rows.ensure(B);
  HIP_OK(hipMemsetAsync(nn->params.p, 0, p * sizeof(float), nn->stream));
  HIP_OK(hipMemsetAsync(nn->m1.p, 0, p * sizeof(float), nn->stream));
  HIP_OK(hipMemsetAsync(nn->m2.p, 0, p * sizeof(float), nn->stream));
  HIP_OK(hipStreamSynchronize(nn->stream));
  nn->dirty = true;
}

void nn_cast(spk_nn* nn, hipStream_t s) {
  if (!nn->dirty) return;
  for (auto& l : nn->L) {
    const int n = l.d.cout * l.d.Kp + l.d.cin * l.d.Ktp;
    nn_launch(s, spknn::k_cast_weights, dim3((n + 255) / 256), dim3(256), 0,
              (const float*)nn->params.p + l.w_off, l.d, nn->wbf.p + l.wb_off,
              nn->wbf.p + l.wt_off);
  }
  nn->dirty = false;
}

int nt_for(int n) { return n <= 16 ? 1 : n <= 32 ? 2 : n <= 48 ? 3 : 4; }

// forward of one layer over a batch of B rows of inputs; Yc (nullable): column
// 0 of every output row also written compactly to Yc[row] by the split-K
// epilogue; returns false when there was none (no split: Yc is not written)
bool nn_fwd(spk_nn* nn, hipStream_t s, const NnLayer& l, int B, const bf16_t* X, bf16_t* Y,
            float* Yf, float* Yc = nullptr) {
  ConvDesc d = l.d;
  d.B = B;
  const int M = B * d.o0 * d.o1 * d.o2, N = d.cout, NT = nt_for(N);
  const int gx = (M + 127) / 128, gy = (N + 16 * NT - 1) / (16 * NT);
  // split K when the row blocks alone cannot fill the chip: slices of >= 2
  // k-steps, at most 256 blocks, partials in nn->part
  const int ksteps = d.Kp / 32;
  int ks = 1;
  if (gx * gy < 64 && ksteps >= 4) {
    ks = std::min(ksteps / 2, std::max(1, 256 / (gx * gy)));
    if ((size_t)ks * M * N > nn->part.cap) ks = 1;
  }
  const int kchunk = ((ksteps + ks - 1) / ks) * 32;
  ks = (d.Kp + kchunk - 1) / kchunk;
  float* part = ks > 1 ? nn->part.p : nullptr;
  const dim3 grid(gx, gy, ks);
  const bf16_t* W = nn->wbf.p + l.wb_off;
  const float* bias = nn->params.p + l.b_off;
  const bool vec = d.cin % 8 == 0;
#define SPK_NN_FWD(V, T)                                                                         \
  nn_launch(s, spknn::k_igemm<spknn::MODE_FWD, V, T>, grid, dim3(256), 0, X, W, bias, d, M, N, \
            d.Kp, d.Kp, l.relu, (const bf16_t*)nullptr, Y, Yf, kchunk, part, nn->gate)
  if (vec) {
    switch (NT) {
      case 1: SPK_NN_FWD(true, 1); break;
      case 2: SPK_NN_FWD(true, 2); break;
      case 3: SPK_NN_FWD(true, 3); break;
      default: SPK_NN_FWD(true, 4); break;
    }
  } else {
    switch (NT) {
      case 1: SPK_NN_FWD(false, 1); break;
      case 2: SPK_NN_FWD(false, 2); break;
      case 3: SPK_NN_FWD(false, 3); break;
      default: SPK_NN_FWD(false, 4); break;
    }
  }
#undef SPK_NN_FWD
  if (part)
    nn_launch(s, spknn::k_split_finalize, dim3((M * N + 255) / 256), dim3(256), 0,
              (const float*)part, ks, M, N, bias, l.relu, Y, Yf, nn->gate, Yc);
  return part != nullptr;
}

// input gradient of one layer: dX = (dY * W) masked by the layer input > 0
void nn_bwd_data(spk_nn* nn, hipStream_t s, const NnLayer& l, int B, const bf16_t* dY,
                 const bf16_t* mask, bf16_t* dX) {
  ConvDesc d = l.d;
  d.B = B;
  const int M = B * d.i0 * d.i1 * d.i2, N = d.cin, NT = nt_for(N);
  const dim3 grid((M + 127) / 128, (N + 16 * NT - 1) / (16 * NT));
  const bf16_t* Wt = nn->wbf.p + l.wt_off;
#define SPK_NN_BWD(T)                                                                            \
  nn_launch(s, spknn::k_igemm<spknn::MODE_BWD, true, T>, grid, dim3(256), 0, dY, Wt,            \
            (const float*)nullptr, d, M, N, d.Ktp, d.Ktp, 0, mask, dX, (float*)nullptr, d.Ktp,   \
            (float*)nullptr, (const int*)nullptr)
  switch (NT) {
    case 1: SPK_NN_BWD(1); break;
    case 2: SPK_NN_BWD(2); break;
    case 3: SPK_NN_BWD(3); break;
    default: SPK_NN_BWD(4); break;
  }
#undef SPK_NN_BWD
}

// weight and bias gradients of one layer, written (not accumulated) into grads
void nn_wgrad(spk_nn* nn, hipStream_t s, const NnLayer& l, int B, const bf16_t* X,
              const bf16_t* dY) {
  ConvDesc d = l.d;
  d.B = B;
  const int M = B * d.o0 * d.o1 * d.o2, N = d.cout, K = d.K;
  const int kb = (K + 63) / 64;
  int chunks = std::max(1, std::min((M + 31) / 32, (512 + kb - 1) / kb));
  const int MC = (((M + chunks - 1) / chunks) + 31) & ~31;
  chunks = (M + MC - 1) / MC;
  float* part = nn->part.p;
  float* partb = part + (size_t)chunks * N * K;
  nn_launch(s, spknn::k_wgrad, dim3(kb, chunks), dim3(256), 0, X, dY, d, M, N, MC, part);
  nn_launch(s, spknn::k_reduce_parts, dim3((N * K + 255) / 256), dim3(256), 0,
            (const float*)part, chunks, N * K, nn->grads.p + l.w_off);
  nn_launch(s, spknn::k_bgrad, dim3(chunks, N), dim3(256), 0, dY, M, N, MC, partb);
  nn_launch(s, spknn::k_reduce_parts, dim3((N + 255) / 256), dim3(256), 0, (const float*)partb,
            chunks, N, nn->grads.p + l.b_off);
}

void nn_pool_fwd(hipStream_t s, const NnLayer& l, int B, const bf16_t* X, bf16_t* out, int ld,
                 int off, const int* gate, int rep = 1) {
  const ConvDesc& d = l.d;
  const int n = B * (d.o0 / 2) * (d.o1 / 2) * (d.o2 / 2) * d.cout;
  nn_launch(s, spknn::k_pool_fwd, dim3((n + 255) / 256), dim3(256), 0, X, B, d.o0, d.o1, d.o2,
            d.cout, out, ld, off, gate, rep);
}

void nn_pool_bwd(hipStream_t s, const NnLayer& l, int B, const bf16_t* g, int ld, int off,
                 const bf16_t* X, bf16_t* dX) {
  const ConvDesc& d = l.d;
  const size_t n = (size_t)B * d.o0 * d.o1 * d.o2 * d.cout;
  nn_launch(s, spknn::k_pool_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g, ld, off, X,
            B, d.o0, d.o1, d.o2, d.cout, dX);
}

// item branch: n oriented item grids -> pooled rows (pitch ld, default 384) in out
void nn_item_branch(spk_nn* nn, hipStream_t s, const bf16_t* item, int n, bf16_t* out,
                    int ld = kNnItemPooled) {
  nn_fwd(nn, s, nn->L[NN_I1], n, item, nn->act[NN_I1].p, nullptr);
  nn_fwd(nn, s, nn->L[NN_I2], n, nn->act[NN_I1].p, nn->act[NN_I2].p, nullptr);
  nn_fwd(nn, s, nn->L[NN_I3], n, nn->act[NN_I2].p, nn->act[NN_I3].p, nullptr);
  nn_pool_fwd(s, nn->L[NN_I3], n, nn->act[NN_I3].p, out, ld, 0, nn->gate);
}

// tray branch: B tray grids -> pooled rows (pitch 768) in nn->tpool
void nn_tray_branch(spk_nn* nn, hipStream_t s, const bf16_t* tray, int B) {
  nn_fwd(nn, s, nn->L[NN_T1], B, tray, nn->act[NN_T1].p, nullptr);
  nn_fwd(nn, s, nn->L[NN_T2], B, nn->act[NN_T1].p, nn->act[NN_T2].p, nullptr);
  nn_fwd(nn, s, nn->L[NN_T3], B, nn->act[NN_T2].p, nn->act[NN_T3].p, nullptr);
  nn_pool_fwd(s, nn->L[NN_T3], B, nn->act[NN_T3].p, nn->tpool.p, kNnTrayPooled, 0, nn->gate);
}

// One record's head input built in place: rows [item row r | tray row] of a
// buffer of kNnLogits rows of pitch kNnConcat whose item parts hold the
// record's pooled item rows (nn_item_branch with that pitch): the tray branch
// pools into the tray part of every row, then F1 and F2 read the rows and the
// logits come out of F2's split-K epilogue.
void nn_record_rows(spk_nn* nn, hipStream_t s, const bf16_t* tray, bf16_t* rows) {
  nn_fwd(nn, s, nn->L[NN_T1], 1, tray, nn->act[NN_T1].p, nullptr);
  nn_fwd(nn, s, nn->L[NN_T2], 1, nn->act[NN_T1].p, nn->act[NN_T2].p, nullptr);
  nn_fwd(nn, s, nn->L[NN_T3], 1, nn->act[NN_T2].p, nn->act[NN_T3].p, nullptr);
  nn_pool_fwd(s, nn->L[NN_T3], 1, nn->act[NN_T3].p, rows, kNnConcat, kNnItemPooled, nn->gate,
              kNnLogits);
  nn_fwd(nn, s, nn->L[NN_F1], kNnLogits, rows, nn->act[NN_F1].p, nullptr);
  if (!nn_fwd(nn, s, nn->L[NN_F2], kNnLogits, nn->act[NN_F1].p, nn->act[NN_F2].p, nn->out.p,
              nn->logits.p))
    nn_launch(s, spknn::k_logits_compact, dim3(1), dim3(256), 0, (const float*)nn->out.p, kNnOut,
              kNnLogits, nn->logits.p, nn->gate);
}

// head on B records: rows [item row r | tray row r/24] -> F1 -> F2 -> logits
void nn_head(spk_nn* nn, hipStream_t s, const bf16_t* ipool, int B) {
  const int rows = B * kNnLogits;
  const size_t n = (size_t)rows * kNnConcat;
  nn_launch(s, spknn::k_concat_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ipool,
            kNnItemPooled, (const bf16_t*)nn->tpool.p, kNnTrayPooled, rows, kNnLogits,
            nn->concat.p, nn->gate);
  nn_fwd(nn, s, nn->L[NN_F1], rows, nn->concat.p, nn->act[NN_F1].p, nullptr);
  nn_fwd(nn, s, nn->L[NN_F2], rows, nn->act[NN_F1].p, nn->act[NN_F2].p, nn->out.p);
  nn_launch(s, spknn::k_logits_compact, dim3((rows + 255) / 256), dim3(256), 0,
            (const float*)nn->out.p, kNnOut, rows, nn->logits.p, nn->gate);
}

// item [B*24 oriented item grids], tray [B grids] -> nn->logits [B][24]
void nn_forward(spk_nn* nn, hipStream_t s, const bf16_t* item, const bf16_t* tray, int B) {
  nn_cast(nn, s);
  nn_item_branch(nn, s, item, B * kNnLogits, nn->ipool.p);
  nn_tray_branch(nn, s, tray, B);
  nn_head(nn, s, nn->ipool.p, B);
}

// loss, then every parameter gradient into nn->grads
void nn_backward(spk_nn* nn, hipStream_t s, const bf16_t* item, const bf16_t* tray,
                 const float* target, const int32_t* mask, int B) {
  const int R = B * kNnLogits;
  nn_launch(s, spknn::k_loss, dim3((B + 63) / 64), dim3(64), 0, (const float*)nn->out.p, kNnOut,
            target, mask, B, kNnLogits, nn->grad[NN_F2].p, nn->loss_rows.p, nn->logits.p);
  const NnLayer* L = nn->L;
  nn_wgrad(nn, s, L[NN_F2], R, nn->act[NN_F1].p, nn->grad[NN_F2].p);
  nn_bwd_data(nn, s, L[NN_F2], R, nn->grad[NN_F2].p, nn->act[NN_F1].p, nn->grad[NN_F1].p);
  nn_wgrad(nn, s, L[NN_F1], R, nn->concat.p, nn->grad[NN_F1].p);
  nn_bwd_data(nn, s, L[NN_F1], R, nn->grad[NN_F1].p, nullptr, nn->g_concat.p);
  // item branch (R rows)
  nn_pool_bwd(s, L[NN_I3], R, nn->g_concat.p, kNnConcat, 0, nn->act[NN_I3].p, nn->grad[NN_I3].p);
  nn_wgrad(nn, s, L[NN_I3], R, nn->act[NN_I2].p, nn->grad[NN_I3].p);
  nn_bwd_data(nn, s, L[NN_I3], R, nn->grad[NN_I3].p, nn->act[NN_I2].p, nn->grad[NN_I2].p);
  nn_wgrad(nn, s, L[NN_I2], R, nn->act[NN_I1].p, nn->grad[NN_I2].p);
  nn_bwd_data(nn, s, L[NN_I2], R, nn->grad[NN_I2].p, nn->act[NN_I1].p, nn->grad[NN_I1].p);
  nn_wgrad(nn, s, L[NN_I1], R, item, nn->grad[NN_I1].p);
  // tray branch (B rows): the tray part of the 24 rows of a record summed
  nn_launch(s, spknn::k_tray_gsum, dim3((B * kNnTrayPooled + 255) / 256), dim3(256), 0,
            (const bf16_t*)nn->g_concat.p, kNnItemPooled, kNnTrayPooled, B, kNnLogits,
            nn->g_tpool.p);
  nn_pool_bwd(s, L[NN_T3], B, nn->g_tpool.p, kNnTrayPooled, 0, nn->act[NN_T3].p,
              nn->grad[NN_T3].p);
  nn_wgrad(nn, s, L[NN_T3], B, nn->act[NN_T2].p, nn->grad[NN_T3].p);
  nn_bwd_data(nn, s, L[NN_T3], B, nn->grad[NN_T3].p, nn->act[NN_T2].p, nn->grad[NN_T2].p);
  nn_wgrad(nn, s, L[NN_T2], B, nn->act[NN_T1].p, nn->grad[NN_T2].p);
  nn_bwd_data(nn, s, L[NN_T2], B, nn->grad[NN_T2].p, nn->act[NN_T1].p, nn->grad[NN_T1].p);
  nn_wgrad(nn, s, L[NN_T1], B, tray, nn->grad[NN_T1].p);
}

float nn_loss_sum(spk_nn* nn, hipStream_t s, int B) {
  std::vector<float> rows(B);
  HIP_OK(hipMemcpyAsync(rows.data(), nn->loss_rows.p, B * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  double t = 0.0;
  for (float r : rows) t += r;
  return (float)(t / B);
}

// feature grids of n oriented items (maps, device) of one item on the device
void nn_item_features(hipStream_t s, const int* item, const spknn::ItemMap* maps, int n,
                      bf16_t* out) {
  nn_launch(s, spknn::k_item_feat, dim3(kNnItemFeat / 256, n), dim3(256), 0, item, maps, out);
}

void nn_tray_features(hipStream_t s, const int* tray, int nx, int ny, int nz, int* heights,
                      bf16_t* out, const int* gate = nullptr) {
  nn_launch(s, spknn::k_heightmap, dim3((nx * ny + 3) / 4), dim3(256), 0, tray, nx, ny, nz,
            heights, gate);   // a wavefront per column
  nn_launch(s, spknn::k_tray_feat, dim3(kNnTrayFeat / 2 / 256), dim3(256), 0, tray,
            (const int*)heights, nx, ny, nz, out, gate);
}

// the same features from the tray's 128-bit column masks (level 0 of the
// column test's full-resolution masks), one kernel; equal when no tray voxel
// is negative
void nn_tray_features_masks(hipStream_t s, const void* masks, int nx, int ny, int nz, bf16_t* out,
                            const int* gate = nullptr) {
  nn_launch(s, spknn::k_tray_feat_m, dim3(kNnTrayFeat / 2 / 256), dim3(256), 0,
            (const unsigned long long*)masks, nx, ny, nz, out, gate);
}

template <typename F>
int nn_guarded(spk_nn* nn, F&& f) {
  if (!nn) return fail(SPK_EINVAL, "null proposer");
  try {
    DeviceGuard g(nn->device);
    return f();
  } catch (const HipError& e) {
    return fail(e.code, e.msg);
  } catch (const SpkError& e) {
    return fail(e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(SPK_ENOMEM, "host allocation failed");
  }
}

void nn_check_batch(const spk_nn* nn, int batch) {
  if (batch <= 0 || batch > nn->max_batch)
    throw SpkError{SPK_EINVAL, "batch must be in [1, max_batch]"};
}

}  // namespace
