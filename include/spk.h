/*
 * spk.h -- C-ABI of libspk.so, the MI355X (gfx950) spectral-packing library.
 *
 * Plain C types only: caller-owned host buffers, int status codes, an opaque
 * context per device.  Every entry point returns SPK_OK (0) or an error code;
 * spk_last_error() gives the message of the last failure on the calling
 * thread.  Grids are int32, row-major, index (i*ny + j)*nz + k, exactly the
 * layout of the reference FlatVoxelGrid (spectral_packing/types.h:24-62).
 *
 * Reference interfaces replaced (Vrroom/psacking, spectral_packing/):
 *   spk_tray_set          GPUTrayContext::initialize + gpu_tray_context_init
 *                         (fft3.cu:1106-1175, bindings.cpp:400-405) and
 *                         calculate_distance (fft3.cu:951-1004)
 *   spk_search            fft_search_placement (packing.cpp:590-642),
 *                         fft_search_with_gpu_context (fft3.cu:1506-1538),
 *                         fft_search_batch[_interlocking_free]
 *                         (fft3.cu:1549-1632)
 *   spk_search_oriented   the Python orientation loop packer.py:487-505 with
 *                         rotations.py:159-268 done as index maps on device
 *   spk_place             place_in_tray (packing.cpp:773-788)
 *   spk_pack              BinPacker.pack_voxels greedy loop (packer.py:447-529)
 *   spk_pack_gather       the same, items as one pointer per item array
 *   spk_distance          calculate_distance (fft3.cu:944-1004)
 *   spk_correlate         dft_conv3 / dft_corr3 (fft3.cu:643-846,
 *                         packing.cpp:225-239)
 *   spk_collision_grid    collision_grid (packing.cpp:241-251)
 *   spk_interlock_mask    gpu_interlocking_free_positions (fft3.cu:348-402)
 *                         behind interlocking_free_positions (bindings.cpp:561-648)
 *   spk_bounds            get_voxel_grid_bounds_flat (voxelGrid.cpp:279-299)
 *   spk_save_vox          saveAsVox (packing.cpp:275-305)
 *   spk_voxelize_stl      py_voxelize_stl (bindings.cpp:285-302): LibSL STL
 *                         reading, voxelize (packing.cpp:526-588) and
 *                         make_voxel_grid_tight (voxelGrid.cpp:166-178)
 *   spk_voxelize_mesh     voxelize (packing.cpp:526-588) on an in-memory mesh
 *   spk_feasible          (none: the loops' feasibility test as a query)
 */
#ifndef SPK_H
#define SPK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spk_ctx spk_ctx;

enum {
  SPK_OK = 0,
  SPK_EINVAL = 1,       /* bad argument (shape, pointer, flag)            */
  SPK_EDEVICE = 2,      /* HIP runtime / kernel failure                   */
  SPK_ENOMEM = 3,       /* device allocation failed                       */
  SPK_EOVERLAP = 4,     /* place: item intersects occupied tray voxels     */
  SPK_ESTATE = 5,       /* context has no tray                            */
  SPK_EUNSUPPORTED = 6  /* size or value range outside what is built       */
};

enum { SPK_METRIC_L1 = 0, SPK_METRIC_EDT = 1 };          /* distance metric  */
enum { SPK_SOURCES_EQ1 = 0, SPK_SOURCES_NONZERO = 1 };   /* distance sources */
enum { SPK_PENALTY_POW = 0, SPK_PENALTY_CUBE = 1 };      /* height penalty   */

/* An axis-aligned orientation as an index map: output axis a runs along base
 * axis perm[a], reversed when flip[a] != 0. */
typedef struct {
  int32_t perm[3];
  int32_t flip[3];
} spk_orient;

typedef struct {
  int32_t pos[3];      /* (x,y,z) offset, (-1,-1,-1) when not found         */
  int32_t found;       /* 1 if a feasible offset exists                     */
  int32_t orient;      /* index into the orientation list, -1 if none      */
  int32_t path;        /* 0 = FFT path, 1 = exact direct-sum path          */
  double score;        /* proximity + height penalty, 0.0 when not found   */
  int64_t collision;   /* exact overlap count at pos (0 when found)         */
  int64_t proximity;   /* exact proximity sum at pos                        */
  float residual;      /* max distance of an FFT output to its integer      */
  int32_t reserved;
} spk_result;

const char* spk_last_error(void);
int spk_version(int* major, int* minor, int* patch);

int spk_ctx_create(int device, spk_ctx** out);
int spk_ctx_destroy(spk_ctx* ctx);

/* Upload a tray, compute (or take, when dist != NULL) its distance field and
 * build the cached spectra.  generation is stored for cache keying. */
int spk_tray_set(spk_ctx* ctx, const int32_t* tray, int nx, int ny, int nz,
                 const int32_t* dist, int metric, int sources, uint64_t generation);
int spk_tray_get(spk_ctx* ctx, int32_t* tray_out, int32_t* dist_out);
/* Statistics of the tray's distance field as the scored search uses them
 * (no counterpart in the reference; for tests and diagnostics): the sum,
 * minimum and maximum of the field, and the largest min(field, distance to
 * the outside of the grid) over its voxels, the distance to the outside of
 * voxel (x, y, z) being min(x+1, nx-x, y+1, ny-y, z+1, nz-z).  Any pointer
 * may be NULL. */
int spk_tray_stats(spk_ctx* ctx, int64_t* sum, int32_t* mn, int32_t* mx, int32_t* mxw);
int spk_tray_info(spk_ctx* ctx, int* nx, int* ny, int* nz, uint64_t* generation,
                  int* valid);
int spk_tray_clear(spk_ctx* ctx);

/* Search n explicit item grids (one per orientation) against the cached tray.
 * items[o] has dims[3*o .. 3*o+2].  best gets the winner (strictly smaller
 * score wins, list order), per_orient (nullable) every orientation's best. */
int spk_search(spk_ctx* ctx, const int32_t* const* items, const int32_t* dims, int n,
               int penalty, int interlocking_free, spk_result* best,
               spk_result* per_orient);

/* Same, for one base item and n orientation index maps applied on device. */
int spk_search_oriented(spk_ctx* ctx, const int32_t* item, int ix, int iy, int iz,
                        const spk_orient* orients, int n, int penalty,
                        int interlocking_free, spk_result* best, spk_result* per_orient);

/* Which orientations of an item fit the context's tray without collision
 * (no reference counterpart: the feasibility test the greedy loops use to
 * skip items that fit nowhere).  full[o] = 1 when orientation o has a valid
 * offset with no occupied tray voxel under the item (the full-resolution
 * collision pass, FP32 FFT rounded to integers; *residual = its largest
 * rounding distance, to be compared with option "resid_limit").
 * column[o] (may be NULL): the same answer from the column test, an exact
 * bit-mask test (tray z <= 128): the tray reduced to 2x2 column blocks at
 * full height against the item's x-y-eroded core, then full-resolution
 * columns on the offsets the blocks leave open (1 everywhere when it does not
 * apply).  n <= 24; needs the slab plan. */
int spk_feasible(spk_ctx* ctx, const int32_t* item, int ix, int iy, int iz,
                 const spk_orient* orients, int n, int32_t* column, int32_t* full,
                 float* residual);

/* Write id into the cached tray where the (oriented) item is > 0 and the tray
 * is 0, then refresh the distance field and spectra.  SPK_EOVERLAP (tray left
 * unchanged) if an item voxel lands on an occupied voxel. */
int spk_place(spk_ctx* ctx, const int32_t* item, int ix, int iy, int iz,
              const spk_orient* orient, int x, int y, int z, int32_t id);

/* Greedy loop over n_items items (concatenated in items, dims n_items*3), in
 * the given order, each trying n_orient orientations; results[t] is the
 * outcome of the t-th processed item.  The tray must have been set. */
int spk_pack(spk_ctx* ctx, const int32_t* items, const int32_t* dims, int n_items,
             const int32_t* order, const spk_orient* orients, int n_orient,
             int penalty, int interlocking_free, int32_t id_offset,
             spk_result* results);

/* spk_pack with the items given as one pointer per item instead of one
 * concatenated buffer: the bindings' list of item arrays (bindings.cpp
 * pack_voxels' vector of grids, packer.py:447-529) goes in without a copy.
 * items[t] holds dims[3t] * dims[3t+1] * dims[3t+2] values. */
int spk_pack_gather(spk_ctx* ctx, const int32_t* const* items, const int32_t* dims,
                    int n_items, const int32_t* order, const spk_orient* orients,
                    int n_orient, int penalty, int interlocking_free, int32_t id_offset,
                    spk_result* results);

/* Arguments of one spk_pack call, for spk_pack_many (and the learned and
 * records loops).  The items are either concatenated in items, or, when
 * item_ptrs is not null, one pointer per item there (items is then unused). */
typedef struct {
  const int32_t* items;
  const int32_t* dims;
  int n_items;
  const int32_t* order;
  const spk_orient* orients;
  int n_orient;
  int penalty;
  int interlocking_free;
  int32_t id_offset;
  spk_result* results;
  const int32_t* const* item_ptrs;
} spk_pack_args;

/* Many independent packing problems at once (no reference counterpart: the
 * multi-instance batch of uneven trays).  ctxs[i] must hold instance i's tray
 * (spk_tray_set); every context's work runs on its own stream.  A pool of
 * host workers (option "many_workers" of ctxs[0], default 8) takes the
 * instances largest first (tray voxels x items) from a shared queue, so
 * their kernels overlap on the device.  Returns the first failing
 * instance's status (its message in spk_last_error()); the other instances
 * still complete. */
int spk_pack_many(spk_ctx* const* ctxs, int n, const spk_pack_args* args);

/* Standalone full-grid operations (host in, host out). */
int spk_distance(spk_ctx* ctx, const int32_t* grid, int nx, int ny, int nz, int metric,
                 int sources, int32_t* out);
/* spk_distance with no device allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory -- serving loops and
 * harnesses that forbid hipMalloc).  The grid, the intermediate field and the
 * result live in a static device array of the library (32 MiB + 256 bytes
 * per GPU, held until the process ends); grids of more than 4,194,304 voxels
 * fail with SPK_EUNSUPPORTED.  Calls on one GPU share that array: the library
 * serialises them with a lock per device, so overlapping calls from several
 * threads wait for each other.  Same results as spk_distance. */
int spk_distance_noalloc(spk_ctx* ctx, const int32_t* grid, int nx, int ny, int nz, int metric,
                         int sources, int32_t* out);
int spk_correlate(spk_ctx* ctx, const int32_t* a, const int32_t* b, int nx, int ny,
                  int nz, int convolve, int32_t* out);
int spk_collision_grid(spk_ctx* ctx, const int32_t* tray, int nx, int ny, int nz,
                       const int32_t* item, int ix, int iy, int iz, int32_t* out);
int spk_interlock_mask(spk_ctx* ctx, const int32_t* tray, int nx, int ny, int nz,
                       const int32_t* obj, int ox, int oy, int oz, int32_t* out);

/* place_in_tray on caller grids (packing.cpp:773-788 via bindings.cpp:259-276):
 * out = tray with id written where item > 0 and tray == 0; SPK_EOVERLAP (out
 * untouched) if an item > 0 voxel meets a tray > 0 voxel.  Item parts outside
 * the tray are ignored. */
int spk_place_grid(spk_ctx* ctx, const int32_t* tray, int nx, int ny, int nz,
                   const int32_t* item, int ix, int iy, int iz, int x, int y, int z,
                   int32_t id, int32_t* out);

/* Host grid utilities. */
int spk_bounds(const int32_t* grid, int nx, int ny, int nz, int32_t lo[3], int32_t hi[3]);
int spk_save_vox(const int32_t* grid, int nx, int ny, int nz, const char* path);

/* Surface voxelization of a triangle mesh on the GPU, bit-exact with the
 * reference: exact vertex merge, fixed-point transform into a box of
 * `resolution` voxels along the longest axis with a 2.5% margin, three-view
 * rasterization with one parity bit per view and voxel, binarisation, tight
 * crop.  dims receives the cropped shape (x, y, z).  With out == NULL only dims
 * are computed; otherwise out (x-major int32 0/1) must hold capacity >=
 * dims[0]*dims[1]*dims[2] values.  resolution lies in [1, 1024].  A mesh that
 * marks no voxel fails with SPK_EINVAL and the reference's message.
 * spk_voxelize_mesh: positions are n_vertices float32 triples; triangles are
 * n_triangles int32 index triples.  spk_voxelize_stl reads a binary or ASCII
 * STL file with the reference's rules (binary exactly when the file size is
 * 84 + 50 * facet count). */
int spk_voxelize_mesh(int device, const float* positions, int64_t n_vertices,
                      const int32_t* triangles, int64_t n_triangles, int resolution,
                      int32_t* dims, int32_t* out, int64_t capacity);
int spk_voxelize_stl(int device, const char* path, int resolution, int32_t* dims, int32_t* out,
                     int64_t capacity);

/* Per-kernel HIP-event timing on the context's stream (off by default). */
int spk_timing_enable(spk_ctx* ctx, int on);
int spk_timing_get(spk_ctx* ctx, const char* kernel, double* total_ms, int64_t* launches);
/* units of work a kernel did while timing was on, to turn launch times into
 * per-unit rates: for "score_yz" and "corr_x" the orientations the score pass
 * searched, for the feasibility pass the orientation pairs it transformed
 * (both counted on the device: orientations the column test excluded do not
 * count), for "spec_add" the incremental tray-spectrum updates that ran; for
 * other kernels the orientations handed to them */
int spk_timing_units(spk_ctx* ctx, const char* kernel, int64_t* units);
int spk_timing_reset(spk_ctx* ctx);
/* Tuning/diagnostic options of a context (no reference counterpart):
 *   "resid_limit"  largest accepted distance of an FFT output from its
 *                  integer before a search is redone on the exact direct-sum
 *                  path (default 0.25; a negative value forces the exact path
 *                  for every search, which the tests use to check that path).
 *   "plan"         FFT plan of the tray: 0 auto, 1 yz slabs in one block,
 *                  2 z rows + y pencils (large y*z planes); applies at the
 *                  next spk_tray_set.
 *   "feas"         feasibility pass of the greedy loops (slab plan): 0 off,
 *                  1 (default) for an item after one that placed nothing,
 *                  2 every item.  The pass tests every orientation for a
 *                  zero-collision offset at about a third of the scored
 *                  search's cost; an item with none is recorded as not found
 *                  without the scored search.  Results are unchanged.
 *   "autotune"     1 (default): on the first spk_tray_set of a tray shape in
 *                  the process, time each admissible x-pencil tile on a
 *                  24-orientation search and feasibility pass and keep the
 *                  fastest; 0: fixed tile.
 *   "tile"         0 (default) tuned or default tile; 4, 8 or 16 force the
 *                  x-pencil tile (applies at the next spk_tray_set).
 *   "certificate"  1: before an item's search, compare the tray field's
 *                  maximum with the item's inscribed-ball bound (the largest
 *                  distance from an item voxel to a voxel outside it); when
 *                  the field (distance to every occupied voxel, sources
 *                  "nonzero") is smaller everywhere, no orientation fits and
 *                  the item is skipped.  Exact; pays off when free space is
 *                  scattered in small gaps.  Default 0.
 *   "coarse"       1 (default): the exact column test before every search
 *                  (slab plan, tray z <= 128), in place of the feasibility
 *                  pass; 0 off.  Applies at the next spk_tray_set.
 *   "fine_group"   lanes per candidate offset in the column test's
 *                  full-resolution stage: 4, 8, 16, 32 or 64 (default).
 *   "col_chunks"   64-offset chunks per column-test block (default 1).
 *   "screen_all"   1 (default): the learned loop screens every item with the
 *                  column test before proposing; 0: only after an item that
 *                  placed nothing.
 *   "score_radix"  largest y/z butterfly of the score pass: 10 (default) or
 *                  25; "yz_radix" the same for the other y/z passes (2..25),
 *                  "pad_y"/"pad_z" padded FFT sizes (0: smallest 2^a3^b5^c7^d).
 *   "corr_chunk"   orientations per pass-A block (0: automatic).
 *   "ff_bits"      1 (default): the interlocking flood fill of trays with
 *                  z <= 128 holds each (x, y) column of an offset grid as a
 *                  128-bit mask (runs of free offsets fill along z in one
 *                  step); 0: one byte per offset.
 *   "spec_inc"     after a placement in a device-driven loop that leaves the
 *                  distance field unchanged (sources "eq1", the library's own
 *                  field, a placed id other than 1, an item of 0/1 values,
 *                  slab plan), the tray spectrum is updated by the placed
 *                  voxels alone, G1 += e^{-2 pi i k.q / P} FFT(item), at most
 *                  this many times in a row before a full transform (default
 *                  16; 0: always the full transform).
 *   "inc_split"    blocks of the incremental distance x pass per group of
 *                  64 (y, z) lines, each over one share of x (1 to 16,
 *                  default 8).
 *   "score2"       1 (default): the score pass holds both digit slabs in LDS
 *                  at once and transforms them in shared stages (compile-time
 *                  plans, when they fit the device's LDS); 0: one after the
 *                  other.
 *   "reg_x"        pass A for padded x lengths 16 R (R = 6, 8, 9, 10, 12,
 *                  14, 15, 16; a tray's x length is padded to one when that
 *                  adds at most an eighth): the x transforms in registers
 *                  with one LDS exchange per transform, 2 (default) two LDS
 *                  buffers per block, 1 three for the length 240 (two fewer
 *                  barriers, a third more LDS); 0: the LDS FFT kernel for
 *                  every length (set before spk_tray_set).
 *   "feas_group"   orientation pairs per feasibility launch pair (0: all).
 *   "graph"        1: the device-driven loops (spk_pack, spk_pack_gather,
 *                  spk_pack_learned) capture each enqueue of their items as
 *                  one hipGraph and launch it (off while timing is on);
 *                  0 (default): the launches go to the stream directly.
 *   "learn_dev"    1 (default): spk_pack_learned runs device-driven (the
 *                  pick of the proposals as a kernel, no host round trip per
 *                  item) when the column test screens every item, min_conf
 *                  is 0 and the certificate is off; 0: host-driven.
 *   "prescreen"    items per round of the device-driven loops (spk_pack,
 *                  spk_pack_gather, spk_pack_learned; default 16): at the
 *                  start of a round one batched column test over its items
 *                  decides those that fit nowhere on the tray as it is (they
 *                  fit nowhere later either: placements only add occupied
 *                  voxels), and only the others are enqueued; 0: one round,
 *                  no batched test.
 *   "many_workers" host workers of spk_pack_many (set on its first
 *                  context; default 8).
 *   "inc_dist"     1 (default): after a placement the distance field is
 *                  updated from the placed voxels alone (the field is a
 *                  minimum over sources: Phi = min(Phi, transform of the new
 *                  sources), computed on the item's box and one pass over the
 *                  tray), when the field is this library's own and the item's
 *                  x extent is at most 128; 0: the full transform every time.
 *   "inc_masks"    1 (default): after a placement the column test's tray masks
 *                  are rebuilt for the columns under the placed box only (a
 *                  placement changes no other column); 0: the whole tray.
 *   "dist_y"       distance y pass: 0 outward search, 1 Felzenszwalb
 *                  envelope, 2 full scan (default); "dist_x": 0 outward
 *                  search (default), 2 full scan.
 * None of them changes a result. */
int spk_set_option(spk_ctx* ctx, const char* name, double value);

/* Plan actually used for the cached tray: FFT sizes and digit base. */
int spk_plan_info(spk_ctx* ctx, int* px, int* py, int* pz, int* digit_base);
/* x-pencil tile and plan in use (split = 1 for z rows + y pencils); the tile
 * is the autotuner's choice when option "autotune" is on (the default). */
int spk_tile_info(spk_ctx* ctx, int* tile, int* split);

/* ---- Device memory (no reference counterpart) -------------------------------
 * For callers that hold no HIP runtime of their own (ctypes, other
 * languages): device buffers for the proposer's features, targets and masks.
 * Every call synchronises the device.  kind: 0 host->device, 1 device->host,
 * 2 device->device.  spk_dev_gather copies rows idx[0..n) (host array, each
 * in [0, src_rows)) of src into consecutive rows of dst; row_bytes must be a
 * multiple of 4. */
int spk_dev_alloc(int device, int64_t bytes, void** out);
int spk_dev_free(int device, void* ptr);
int spk_dev_copy(int device, void* dst, const void* src, int64_t bytes, int kind);
int spk_dev_gather(int device, void* dst, const void* src, int64_t row_bytes, int64_t src_rows,
                   const int64_t* idx, int n);

/* ---- Learned orientation proposer (no reference counterpart) --------------
 * The reference searches every orientation (spectral_packer/packer.py:487-505,
 * fft3.cu:1549-1632).  Here a 3-D CNN orientation scorer (260,696 parameters,
 * bf16 weights and activations, fp32 accumulation) rates each orientation of
 * an item against the tray: an item branch on the oriented item's occupancy
 * at pitch 2 (32^3 cells), a tray branch on the tray's occupancy and
 * below-surface fraction on a 64x32x32 grid, and a head on both; the 24
 * orientation logits of a record form one softmax, and the exact search
 * evaluates the top-k.  Feature buffers, targets and masks are DEVICE pointers
 * (bf16 bits as uint16); logits, parameters and gradients are host buffers.
 * A batch of B records holds B*24 oriented item grids (record b, orientation
 * o at row b*24 + o) and B tray grids.  One proposer is used by one thread at
 * a time. */
typedef struct spk_nn spk_nn;
#define SPK_NN_ITEM_FEATURES 32768    /* [32][32][32][1] per oriented item */
#define SPK_NN_TRAY_FEATURES 131072   /* [64][32][32][2] per tray */
#define SPK_NN_LOGITS 24              /* orientations per record */

int spk_nn_create(int device, int max_batch, spk_nn** out);
int spk_nn_destroy(spk_nn* nn);
int spk_nn_num_params(spk_nn* nn, int64_t* n);
/* fp32 master parameters, layer by layer: W [cout][27*cin or cin] then b [cout];
 * setting them clears the optimizer state. */
int spk_nn_set_params(spk_nn* nn, const float* params, int64_t n);
int spk_nn_get_params(spk_nn* nn, float* params, int64_t n);
/* logits [batch][24] */
int spk_nn_forward(spk_nn* nn, const uint16_t* item_feat, const uint16_t* tray_feat, int batch,
                   float* logits);
/* soft-target cross entropy over the orientations whose mask bit is set:
 * mean loss and d(loss)/d(params) for one batch, parameters unchanged */
int spk_nn_gradient(spk_nn* nn, const uint16_t* item_feat, const uint16_t* tray_feat,
                    const float* target, const int32_t* mask, int batch, float* grads,
                    float* loss);
/* the same, followed by one Adam step (beta 0.9/0.999, eps 1e-8) */
int spk_nn_train_step(spk_nn* nn, const uint16_t* item_feat, const uint16_t* tray_feat,
                      const float* target, const int32_t* mask, int batch, float lr,
                      float* loss);
/* features of a host item in one orientation (NULL = as given) into a
 * device buffer of SPK_NN_ITEM_FEATURES bf16 */
int spk_nn_item_features(spk_nn* nn, const int32_t* item, int ix, int iy, int iz,
                         const spk_orient* orient, uint16_t* out);
/* features of the context's cached tray into a device buffer */
int spk_tray_features(spk_ctx* ctx, uint16_t* out);

/* spk_pack (plain, not interlocking-free) that also records, for item t, the
 * tray features it saw (tray_feat + t*SPK_NN_TRAY_FEATURES, device) and each
 * orientation's best score (scores[t*n_orient + o], host; +inf when the
 * orientation does not fit or has no feasible offset). */
int spk_pack_records(spk_ctx* ctx, const spk_pack_args* args, uint16_t* tray_feat,
                     double* scores);

/* The greedy loop with learned proposals (plain, not interlocking-free; at
 * most 24 orientations).  Candidates are an item's fitting orientations,
 * narrowed to the feasible ones when the feasibility pass ran first (it does
 * after an item that placed nothing, unless option "feas" is 0).  With more
 * than topk candidates whose topk best-scored ones hold a softmax mass >=
 * min_conf, only those are searched and the best placement among them is
 * taken; all candidates are searched otherwise, and the other candidates
 * also when no proposal is feasible.  proposals[t*topk + j] (optional) lists
 * the searched proposals (-1 when the item was searched over all its
 * candidates); counts[5] (optional): items decided on proposals, searches of
 * all candidates, searches after infeasible proposals, exact host-path
 * decisions, items the feasibility pass found infeasible. */
int spk_pack_learned(spk_ctx* ctx, spk_nn* nn, const spk_pack_args* args, int topk,
                     float min_conf, int32_t* proposals, int32_t* counts);

#ifdef __cplusplus
}
#endif

#endif /* SPK_H */
