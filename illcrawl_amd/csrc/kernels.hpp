// Host launchers of the device kernels (defined in the .hip sources).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace ilc {

// ---- grid.hip
void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
void launch_bbox(const float4* p4, const float* h, uint64_t n, float* partial, int nparts,
                 uint32_t* bad, bbox_result* out, hipStream_t s);
void launch_keys(const float4* p4, uint32_t n, const grid_params& g, uint32_t* keys,
                 uint32_t* vals, uint32_t* err, hipStream_t s);
// Per-cell max h (float bits, atomic max; maxh_bits cleared by the caller)
// of n particles, without sorting them; out-of-grid particles count in err.
void launch_cell_maxh(const float4* p4, const float* h, uint32_t n, const grid_params& g,
                      uint32_t* maxh_bits, uint32_t* err, hipStream_t s);
uint64_t scan_temp_words(uint64_t n);
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* temp,
                           hipStream_t s);
uint64_t radix_temp_words(uint64_t n);
void launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t n, int key_bits, uint32_t* temp,
                       hipStream_t s);
// The same sort with a second value array moved alongside; temp holds
// radix2_temp_words(n).
uint64_t radix2_temp_words(uint64_t n);
void launch_radix_sort2(uint32_t* keys, uint32_t* vals, uint32_t* vals2, uint32_t n, int key_bits,
                        uint32_t* temp, hipStream_t s);
// coef = w norm / h^2 (pow04: coef^0.4, the projected quartic's form); h
// NULL: each particle's h is the z of its p4
void launch_build_soa(const float4* p4, const float* h, const uint32_t* perm,
                      const uint32_t* skeys, uint32_t n, float norm, bool pow04, float4* part,
                      float* hs, uint32_t* count, uint32_t* maxh_bits, hipStream_t s,
                      const float2* zc_in = nullptr, float2* zc_out = nullptr);
// Reorder (key, perm) pairs, n of them, by descending h[perm] (top 16 bits,
// stable); horder holds 2n words, temp radix_temp_words(n).
void launch_h_order(const float* h, uint32_t* key, uint32_t* perm, uint32_t n, uint32_t* horder,
                    uint32_t* temp, hipStream_t s);
// The reference's within-cell order (stable by input index) of a grid whose
// sorted keys and perm are given; by_index holds n words.
void launch_reference_perm(const uint32_t* key, const uint32_t* perm, uint32_t n, int key_bits,
                           uint32_t* out_perm, uint32_t* by_index, uint32_t* temp, hipStream_t s);
void launch_export_cells(const uint32_t* start, const uint32_t* maxh_bits, uint32_t ncells,
                         int32_t* cells2, float* maxh, hipStream_t s);
void launch_bbox_final(const float* partial, int nparts, const uint32_t* bad, bbox_result* out,
                       hipStream_t s);

// ---- batch.hip
// Per-view particle pass: quantity functor (mass or bolometric X-ray), camera
// projection with depth mask, and either bounding-box partials (bbox mode) or
// the projected float4 particles (write mode).
struct view_proj {
  float m[9];
  float c[3];
  float depth;
  int quantity;  // 0 mass, 1 X-ray
  double scale;  // quantity scaling applied to the mass (w = float(scale * m))
};
struct particle_soa {
  const float* pos3;  // x, y, z interleaved
  const float* mass;
  const float* hsml;
  const float* rho;
  const float* temp;
  const float* xe;
  uint64_t n;
};
struct gaunt_table {
  const float* y;  // device
  int n;
  float x0, dx;
};
// Crawler block body (periodic correction + cutout + compaction).
constexpr int kMaxCrawlColumns = 8;
struct crawl_columns {
  const float* in[kMaxCrawlColumns];
  float* out[kMaxCrawlColumns];
  int n;
};
uint64_t crawl_tiles(uint64_t n);
// Largest block of the allocation-free block body (ILC_CRAWL_STATIC_MAX_ROWS).
constexpr uint64_t kCrawlStaticMaxRows = uint64_t(1) << 29;
// Device address, on the current device, of the block body's static scratch
// (`words` uint32 at most; fails above its size).  Calls on one device share it.
uint32_t* crawl_static_scratch(uint64_t words);
// counts/offsets: crawl_tiles(n) + 1 words each; scan_temp: scan_temp_words(tiles + 1)
void launch_crawl(const double period[3], const double gmin[3], const double gmax[3],
                  const double vmin[3], const double vmax[3], const float* xyz, uint64_t n,
                  const crawl_columns& cols, uint32_t* counts, uint32_t* offsets,
                  uint32_t* scan_temp, float4* out4, uint64_t* index_out, hipStream_t s);

// Row filter of the batched grids: camera-frame (x', y') inside a window
// around the view's screen, and/or the view's depth cut (z' in [0, depth]).
struct window_f {
  float x0, x1, y0, y1;
  int use_window, use_depth;
  // the batch's shared build: rows whose support cannot reach the screen
  // rectangle [sx0, sx1] x [sy0, sy1] (d^2 / h^2 >= cull2 by the gather's own
  // cull arithmetic) only raise their cell's max h
  int use_screen;
  float sx0, sx1, sy0, sy1, cull2;
  // the grid cells that can hold a pixel point (the screen's cell range with
  // a margin): only their max h is read by the reference's search radius, so
  // only rows in them raise it
  int scx0, scx1, scy0, scy1;
  // line-of-sight kernel clip: a row counts for the depth cut when its
  // support [z' - h, z' + h] overlaps [0, depth] (instead of its centre lying
  // inside), and the scatter writes its clip parameters
  int clip;
};
// One view of a multi-view pass over the snapshot.
struct view_build_desc {
  view_proj vp;
  window_f win;
  grid_params g;  // the view's grid (cell keys of the scatter)
};
// Bounding boxes of every view in one read of the snapshot: out[v] as
// launch_view_bbox would give it (every view's vp.scale must be 1).  partial
// holds multi_bbox_parts(n) * 8 * nv floats, bad nv words.  one_group: every
// view is a slab of one slab group (the same x', y' projection bit for bit,
// the same quantity).
int multi_bbox_parts(uint64_t n);
void launch_multi_bbox(const particle_soa& ps, const view_build_desc* vd_dev, int nv,
                       const gaunt_table& gt, float* partial, uint32_t* bad, bbox_result* out,
                       bool one_group, hipStream_t s);
// A slab stack: the listed views are slabs of one stack (one slab group:
// the same x', y' projection; depth row (0, 0, m8) with m8 = +-1; centre
// masks, no window or screen test) whose depth ranges [s_k, s_k + depth_k],
// s_k = m8 * cz_k, follow one another without overlap, with room for FP32
// rounding.  A particle at t = m8 * z can then only be kept by the last slab
// with s_k <= t or the one before it: the multi-view passes test those two
// with the views' own arithmetic instead of every slab.  n = 0: no stack.
constexpr int kStackMax = 64;
struct slab_stack {
  int n = 0;
  float m8 = 1.0f;
  float s[kStackMax];    // s_k in ascending order
  int32_t view[kStackMax];  // the view of the k-th slab
};

// Kept particles per window tile (window_tiles(n) of them) of every listed view
// in one read: counts[v * ntiles + t], totals[v] += the same (not cleared).
// counts / totals: kept rows (window, depth cut and screen test);
// counts_plain / totals_plain: the same without the screen test (the rows a
// single view's window scatter keeps)
// stack (n > 0): the listed views are that stack's slabs, every one of them.
void launch_multi_window_count(const float* pos3, const float* hsml, uint64_t n,
                               const view_build_desc* vd_dev, const int32_t* views_dev,
                               int nviews, uint32_t* counts, uint32_t* totals,
                               uint32_t* counts_plain, uint32_t* totals_plain,
                               const slab_stack& stack, hipStream_t s);
uint32_t window_tiles(uint64_t n);
// X-ray weight of every particle (the quantity of particle_of, view-free).
void launch_particle_xray(const particle_soa& ps, const gaunt_table& gt, float* wx,
                          hipStream_t s);
// The kept particles of every listed view, in input order, at the global
// positions offs[v * ntiles + t] (exclusive scan of the multi-window counts):
// projected float4 (x', y', h, w with the depth mask), the descending-h sort
// key (as launch_h_order) and the row index as its value, the global cell key
// (the view's first cell cbase[v] + the cell key in the view's grid, so that
// it orders rows by view, then cell) and the snapshot index.
// wx: launch_particle_xray.  maxh_bits: the per-cell max h of the shared cell
// arrays (rows kept out by the screen test go there only)
// zc (NULL unless clip mode): per kept row, (0 - z') / h and (depth - z') / h
void launch_multi_window_scatter(const particle_soa& ps, const float* wx,
                                 const view_build_desc* vd_dev, const int32_t* views_dev,
                                 int nviews, const uint32_t* offs, float4* p4c, uint32_t* hqkey,
                                 uint32_t* hqval, uint32_t* gkey, uint32_t* orig, uint32_t* err,
                                 const uint32_t* cbase, uint32_t* maxh_bits, float2* zc,
                                 const slab_stack& stack, hipStream_t s);
// out[i] = src[idx[i]]
void launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* out,
                       hipStream_t s);
// lkey[i] = gkey[i] - cbase[v], v the view whose cells hold gkey[i] (cbase:
// the nv views' first cells, rising strictly)
void launch_local_keys(const uint32_t* gkey, const uint32_t* cbase, int nv, uint32_t n,
                       uint32_t* lkey, hipStream_t s);
void launch_window_scatter(const particle_soa& ps, const view_proj& vp, const gaunt_table& gt,
                           const window_f& win, const uint32_t* offsets, const grid_params& g,
                           float4* p4c, float* hc, uint32_t* keys, uint32_t* vals, uint32_t* orig,
                           uint32_t* err, hipStream_t s);
// one view's adjoint (sorted order, projected frame) into snapshot gradients;
// orig (may be NULL): the snapshot index of each grid input row
void launch_scatter_grads(const uint32_t* perm, const uint32_t* orig, uint32_t n,
                          const float* gw, const float* gh, const float* gx, const float* gy,
                          const float* pos3, const view_proj& vp, float* out_w, float* out_h,
                          float* out_pos3, hipStream_t s);
int view_bbox_parts(uint64_t n);
void launch_view_bbox(const particle_soa& ps, const view_proj& vp, const gaunt_table& gt,
                      float* partial, uint32_t* bad, bbox_result* out, hipStream_t s);
void launch_view_project(const particle_soa& ps, const view_proj& vp, const gaunt_table& gt,
                         float4* out, hipStream_t s);

// ---- gather.hip
// Device view of one built grid (sorted SoA + cells).
struct grid_view {
  const float4* part;         // (x, y, coef, 1/h^2) sorted by cell
  // line-of-sight clip (OP_CLIP): per sorted particle, (0 - z') / h and
  // (depth - z') / h, the view's depth range relative to the particle
  const float2* clip;
  const uint32_t* key;        // sorted cell keys
  int sparse;                 // fewer than 0.4 particles per cell: stream row ranges
  const float* hs;            // sorted h
  const uint32_t* start;      // ncells + 1 offsets
  const uint32_t* maxh_bits;  // per-cell max h (float bits)
  int32_t nc[3];
  float minc[3];
  float cs[3];
  float max_h;
  float cell_radius;
  int32_t kz0;  // z layer of the particles (single layer in reference mode)
  uint32_t ncells;
};

// Pixel-corner points of a camera (pixel.hpp): projection rows and centre,
// the FP32 corner and screen axes, the pixel size.
struct pixel_gen {
  float m[9], c[3];
  float mn[3], b0[3], b1[3];
  float ps;
};
pixel_gen pixel_gen_of(const camera_frame& f);

// Evaluation points laid out as an image (lnx x lny, row-major): loaded from
// pts, or (gen != 0) the pixel corners of the camera pg, computed where they
// are used.
struct point_view {
  const float4* pts;
  uint32_t n;
  uint32_t lnx, lny;
  int gen = 0;
  pixel_gen pg{};
};

// OP_CLIP: exact mode with the line-of-sight kernel clip (grid_view::clip)
enum gather_op { OP_REFERENCE = 0, OP_EXACT = 1, OP_COUNT = 2, OP_ADJOINT = 3, OP_CLIP = 4 };

struct tile_shape {
  int tw;         // tile width (points)
  int ppt;        // points per thread (tile height = nb/tw * ppt)
  int nb = 256;   // threads per workgroup (one tile per workgroup)
};

void launch_gather(const grid_view& g, const point_view& p, int op, int kernel,
                   const tile_shape& ts, double* acc, unsigned long long* counts,
                   const float* cubic_table, hipStream_t s);

// Adjoint of the exact-mode gather: adds into per-sorted-particle gradient
// arrays (dL/dw, dL/dh, dL/dx, dL/dy) for dL/d(point value) = grad, point
// value = scale * accumulator, W = norm * shape / h^2 (quartic: s^2.5; cubic:
// the tabulated F with its tabulated derivative dtab).
void launch_adjoint(const grid_view& g, const point_view& p, int kernel, const float* grad,
                    float scale, float norm, const float* ftab, const float* dtab, float* gw,
                    float* gh, float* gx, float* gy, hipStream_t s);
void launch_unpermute4(const uint32_t* perm, uint32_t n, const float* a, const float* b,
                       const float* c, const float* d, float* oa, float* ob, float* oc, float* od,
                       hipStream_t s);

// One view of a batch: its grid, its points, its accumulator, its tiling.
// fused = 1: the gather writes the view's accumulator (acc = sum, without
// reading it) and its image (img = float(sum * double(scale)), as
// launch_retrieve) instead of adding to a cleared accumulator; every
// (view, tile) item is visited once per launch.
struct view_desc {
  grid_view g;
  point_view p;
  double* acc;
  uint32_t tiles_x;
  uint32_t ntiles;
  float* img = nullptr;
  float scale = 1.0f;
  int fused = 0;
  // the view's mean particles per tile (kept rows / tiles, at least 1): the
  // unit of the work-item order's cost classes
  uint32_t cost_unit = 1;
};

// tiles of a point image for a tile shape
void tile_grid(const tile_shape& ts, uint32_t lnx, uint32_t lny, uint32_t* tiles_x,
               uint32_t* ntiles);
// workgroups of the persistent batch gather that fit on the device at once
int gather_batch_grid(int op, int kernel, const tile_shape& ts);
// Work items split into kQueues queues (queue q: items[seg[q] .. seg[q+1])):
// a workgroup drains the queue of its XCD (blockIdx % kQueues) first, then
// the others.  counters: kQueues counters kQueueStride words apart, zeroed.
constexpr int kQueues = 8;
constexpr int kQueueStride = 64;
struct queue_segs {
  uint32_t seg[kQueues + 1];
};
// The persistent gather's items reordered inside every queue: tiles with the
// most particles under them first (stable inside a cost class).  keys, idx:
// n words each; radix_temp: radix_temp_words(n); out: n items.
void launch_order_items(const view_desc* views, const uint2* items, const queue_segs& segs,
                        uint32_t n, const tile_shape& ts, uint32_t* keys, uint32_t* idx,
                        uint32_t* radix_temp, uint2* out, hipStream_t s);
void launch_gather_batch(const view_desc* views, const uint2* items, const queue_segs& segs,
                         uint32_t* counters, int op, int kernel, const tile_shape& ts,
                         const float* cubic_table, int grid_blocks, hipStream_t s);

// ---- project.hip
void launch_project_particles(const camera_frame& f, float depth, const float4* in, uint64_t n,
                              float4* out, hipStream_t s);
void launch_pixel_points(const camera_frame& f, uint32_t nx, uint32_t ny, float4* out,
                         hipStream_t s);
void launch_world_pixels(const camera_frame& f, uint32_t nx, uint32_t ny, float4* out,
                         hipStream_t s);
void launch_project_points(const camera_frame& f, const float4* in, uint64_t n, float4* out,
                           hipStream_t s);
void launch_vector_scale(float* x, float alpha, uint64_t n, hipStream_t s);
void launch_set_w(float4* p4, const float* w, uint64_t n, hipStream_t s);
void launch_retrieve(const double* acc, uint64_t n, float scale, float* out, hipStream_t s);

// ---- smooth3d.hip: the volumetric smoothing gather (grid with g.pz set)
//
// The level index of a volumetric grid: particles binned by h into levels
// (level L holds h <= max_h 2^(L - 23)), each level sorted by a fine cell key
// (L << 57 | z << 38 | y << 19 | x) of cell size cs[L] ~ the level's h.
constexpr int kLevels3 = 24;
constexpr uint32_t kFineMax3 = (1u << 19) - 1;
// meta layout: count[24], max h bits[24], start[25]
constexpr int kLvMetaWords = 3 * kLevels3 + 1;
struct level_view {
  const unsigned long long* key;  // sorted fine keys
  const float4* pos;              // (x, y, z, 1/h^2)
  const float* w;                 // w K / h^3
  const int2* cell;               // reference cell (cx | cy << 16, cz)
  const uint32_t* meta;
  // plane directory: pdir[poff[L] + z] = first index of level L whose fine z
  // is >= z, for z = 0 .. poff[L + 1] - poff[L] - 1 (the last: the level's end)
  const uint32_t* pdir;
  uint32_t poff[kLevels3 + 1];
  float cs[kLevels3];
};
// Fine cell sizes of the levels for a grid of that extent and max h (host).
void level_cell_sizes(const grid_params& g, float max_h, float cs[kLevels3]);
// Offsets of the levels' plane directories (host): every fine z a particle
// can have, plus one entry for the level's end.
void level_planes(const grid_params& g, const float cs[kLevels3], uint32_t poff[kLevels3 + 1]);
// Builds the level index of particles p4 (x, y, z, w) with smoothing lengths
// h on grid g, and the grid's per-cell max h (maxh_bits, ncells words); err
// counts particles outside the grid.  norm: the kernel's 3-D normalisation
// (coef = w norm / h^3).  tmp: 11 n words, 16-byte aligned; radix_temp:
// radix_temp_words(n).
void launch_volumetric_index(const float4* p4, const float* h, uint32_t n, const grid_params& g,
                             float max_h, float norm, const level_view& lv,
                             unsigned long long* key, float4* pos, float* w, int2* cell,
                             uint32_t* meta, uint32_t* pdir, uint32_t* maxh_bits, uint32_t* err,
                             uint32_t* tmp, uint32_t* radix_temp, hipStream_t s);
// op OP_REFERENCE / OP_EXACT add to acc, OP_COUNT adds evaluation counts to cnt
void launch_smooth3d(const grid_view& g, const level_view& lv, const point_view& p, int op,
                     double* acc, hipStream_t s, unsigned long long* cnt = nullptr);
void launch_quantity_scaled(const float* in, double scale, uint64_t n, float* w, hipStream_t s);
void launch_quantity_stellar_mass(const float* sft, const float* mass, uint64_t n, float* w,
                                  hipStream_t s);
// The reference's spectral X-ray emissivity (2-D Gaunt table on the device,
// t[i + nx j], i along ln(gamma^2) from -6, j along ln(u) from -16, step 0.2)
void launch_xray_flux(const float* rho, const float* u, const float* xe, uint64_t n,
                      const float* gaunt_dev, int nx, int ny, float z, float dl, float e_min,
                      float e_max, int num_samples, float* out, int32_t* steps, hipStream_t s);
void launch_xray_spectral_flux(const float* rho, const float* u, const float* xe, uint64_t n,
                               const float* gaunt_dev, int nx, int ny, float z, float dl, float E,
                               float dE, float* out, hipStream_t s);
void launch_xray_bolometric(const float* rho, const float* temp, const float* xe, uint64_t n,
                            const float* table_dev, int32_t ntab, float x0, float dx, float* w,
                            hipStream_t s);
void launch_fill_f64(double* p, double v, uint64_t n, hipStream_t s);
unsigned long long launch_sum_u64(const unsigned long long* p, uint64_t n,
                                  unsigned long long* scratch, hipStream_t s);

}  // namespace ilc
