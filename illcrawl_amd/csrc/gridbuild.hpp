// Device-resident particle grid and its construction, shared by the
// single-view backend and the batched renderer.
#pragma once

#include "common.hpp"
#include "kernels.hpp"

namespace ilc {

// One built grid: the reference's smoothing_particle_grid
// (smoothing_particle_grid.hpp:29-46) as sorted SoA + cell offsets.  Cells,
// their ranges and max h are the reference's; within a cell the particles are
// ordered by descending h (the reference keeps input order, which
// launch_reference_perm recovers).
struct grid_store {
  dbuf<float4> part;          // (x, y, coef, 1/h^2), cell-sorted
  dbuf<uint32_t> key;         // sorted cell keys
  dbuf<float> hs;             // sorted h
  dbuf<uint32_t> perm;        // sorted position -> input index
  dbuf<uint32_t> start;       // ncells + 1
  dbuf<uint32_t> maxh_bits;   // ncells
  // volumetric grids: the level index the gather searches (part, key, hs,
  // perm and start are not built for them)
  dbuf<unsigned long long> lv_key;
  dbuf<float4> lv_pos;
  dbuf<float> lv_w;
  dbuf<int2> lv_cell;
  dbuf<uint32_t> lv_meta;
  dbuf<uint32_t> lv_pdir;
  uint32_t lv_poff[kLevels3 + 1] = {};
  float lv_cs[kLevels3] = {};
  grid_params gp{};
  float max_h = 0.f, cell_radius = 0.f;
  int32_t kz0 = 0;
  uint64_t n = 0;
  // particles per cell that decides how the gather enumerates candidates;
  // 0 = n (a windowed grid sets the count of the full set it stands for, so
  // that it is enumerated like the full grid)
  uint64_t density_n = 0;
  bool volumetric = false;  // 3-D smoothing: the level index and maxh_bits only
  bool ready = false;
  grid_view view() const;
  level_view levels() const;
};

// Scratch shared by successive builds on one stream.
struct grid_scratch {
  dbuf<uint32_t> temp, counts, flags, horder, lv_tmp;
  dbuf<float> bbox_partial;
  dbuf<bbox_result> bbox_dev;
};

// Grid parameters from a bounding box (host, double; particle_grid.cpp:43-100)
// plus the z layer of the particles; reference mode requires one layer.
grid_params plan_grid(const bbox_result& bb, uint64_t n, uint64_t target, int mode,
                      int32_t* kz0);

// Enqueue keys, stable sort, sorted SoA, per-cell counts / max h and cell
// offsets for n float4 particles p4 with smoothing lengths h.  Out-of-grid
// keys are counted into err (device uint32, not cleared here).
void enqueue_grid_build(grid_store& gs, grid_scratch& sc, const float4* p4, const float* h,
                        uint64_t n, const grid_params& g, int32_t kz0, float max_h, int kernel,
                        uint32_t* err, hipStream_t s);

// Only the per-cell max h (gs.maxh_bits) and the grid parameters of n
// particles, no sorted arrays (gs.part, key, hs, perm and start stay as they
// were): for grids whose users read nothing else.
void enqueue_grid_maxh(grid_store& gs, const float4* p4, const float* h, uint64_t n,
                       const grid_params& g, int32_t kz0, float max_h, uint32_t* err,
                       hipStream_t s);

// The same from keys already in gs.key (cell keys) and gs.perm (indices into
// p4 / h), n of them, in the order the within-cell order must follow.
void enqueue_grid_from_keys(grid_store& gs, grid_scratch& sc, const float4* p4, const float* h,
                            uint64_t n, const grid_params& g, int32_t kz0, float max_h,
                            int kernel, hipStream_t s);

// The grid of the 3-D smoothing: the reference grid's parameters and per-cell
// max h, and the level index (coef = w K / h^3) the volumetric gather
// searches.  Out-of-grid particles are counted into err.
void enqueue_volumetric_grid(grid_store& gs, grid_scratch& sc, const float4* p4, const float* h,
                             uint64_t n, const grid_params& g, float max_h, int kernel,
                             uint32_t* err, hipStream_t s);

float kernel_norm(int kernel);
// the 3-D quartic's normalisation: the reference's 1.04445f (ILC_KERNEL_QUARTIC_REF)
// or 105 / (32 pi) (ILC_KERNEL_QUARTIC)
float kernel_norm_3d(int kernel);

// Bits of a radix key that holds values up to maxval.
int key_bits_for(uint64_t maxval);

}  // namespace ilc
