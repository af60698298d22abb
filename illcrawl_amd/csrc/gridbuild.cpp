// Grid construction on the device (see gridbuild.hpp).
#pragma clang fp contract(off)

#include "gridbuild.hpp"

#include <algorithm>
#include <cmath>

namespace ilc {

int key_bits_for(uint64_t maxval) {
  int b = 0;
  while (b < 64 && (maxval >> b)) ++b;
  return b;
}

float kernel_norm(int kernel) {
  switch (kernel) {
    case ILC_KERNEL_QUARTIC_REF:
      // 2 * norm * integral of smoothing_kernel.cl:57-60, with the h powers
      // folded into 1/h^2 (h^5 of the s2^2.5 factor over h^3 * h^4)
      return static_cast<float>(2.0 * double(1.04445f) * double(1.066667f));
    case ILC_KERNEL_QUARTIC:
      return static_cast<float>(3.5 / M_PI);
    default:
      return 1.0f;
  }
}

float kernel_norm_3d(int kernel) {
  // smoothing_kernel.cl:27-37: norm = 105 / (32 pi h^3)
  return kernel == ILC_KERNEL_QUARTIC_REF ? 1.04445f : static_cast<float>(105.0 / (32.0 * M_PI));
}

level_view grid_store::levels() const {
  level_view lv;
  lv.key = lv_key.p;
  lv.pos = lv_pos.p;
  lv.w = lv_w.p;
  lv.cell = lv_cell.p;
  lv.meta = lv_meta.p;
  lv.pdir = lv_pdir.p;
  for (int l = 0; l <= kLevels3; ++l) lv.poff[l] = lv_poff[l];
  for (int l = 0; l < kLevels3; ++l) lv.cs[l] = lv_cs[l];
  return lv;
}

void level_cell_sizes(const grid_params& g, float max_h, float cs[kLevels3]) {
  // level L: h <= H_L = max_h 2^(L - 23), cells of H_L, at least the grid
  // extent over the 19-bit fine coordinate range
  double ext = 0.0;
  for (int a = 0; a < 3; ++a) ext = std::max(ext, double(g.nc[a]) * double(g.cs[a]));
  const float floor_cs = static_cast<float>(ext / 524000.0);
  for (int l = 0; l < kLevels3; ++l) {
    const float hl = std::ldexp(max_h, l - (kLevels3 - 1));
    cs[l] = std::max(hl, floor_cs);
    if (!(cs[l] > 0.0f) || !std::isfinite(cs[l])) cs[l] = std::max(max_h, 1e-30f);
  }
}

void level_planes(const grid_params& g, const float cs[kLevels3], uint32_t poff[kLevels3 + 1]) {
  // a particle's fine z is at most floor(extent / cs) (one more by rounding)
  const double ext = double(g.nc[2]) * double(g.cs[2]);
  uint32_t off = 0;
  for (int l = 0; l < kLevels3; ++l) {
    poff[l] = off;
    const double np = std::floor(ext / double(cs[l])) + 2.0;
    off += static_cast<uint32_t>(std::min(np, double(kFineMax3) + 1.0)) + 1u;
  }
  poff[kLevels3] = off;
}

grid_view grid_store::view() const {
  grid_view g{};
  g.part = part.p;
  g.key = key.p;
  g.hs = hs.p;
  g.start = start.p;
  g.maxh_bits = maxh_bits.p;
  for (int a = 0; a < 3; ++a) {
    g.nc[a] = gp.nc[a];
    g.minc[a] = gp.minc[a];
    g.cs[a] = gp.cs[a];
  }
  g.max_h = max_h;
  g.cell_radius = cell_radius;
  g.kz0 = kz0;
  g.ncells = static_cast<uint32_t>(gp.ncells());
  const uint64_t dn = density_n ? density_n : n;
  g.sparse = dn * 5 < uint64_t(g.ncells) * 2 ? 1 : 0;
  return g;
}

grid_params plan_grid(const bbox_result& bb, uint64_t n, uint64_t target, int mode,
                      int32_t* kz0) {
  if (bb.bad)
    fail(ILC_EINVAL, "%u particles have non-finite coordinates/w or h outside [1.1e-19, 1.8e19]",
         bb.bad);
  const grid_params g = compute_grid_params(bb, n, target);
  const float zlo = (bb.mn[2] - g.minc[2]) / g.cs[2];
  const float zhi = (bb.mx[2] - g.minc[2]) / g.cs[2];
  *kz0 = static_cast<int32_t>(std::floor(zlo));
  if (mode == ILC_MODE_REFERENCE && std::floor(zlo) != std::floor(zhi))
    fail(ILC_EINVAL, "reference mode needs projected particles (one z layer); z spans [%g, %g]",
         double(bb.mn[2]), double(bb.mx[2]));
  // the gather's masked list packs a cell's x and y into 16 bits each
  if (mode == ILC_MODE_REFERENCE && (g.nc[0] > 65536 || g.nc[1] > 65536))
    fail(ILC_ERANGE, "grid of %d x %d cells in x, y (at most 65536 each)", g.nc[0], g.nc[1]);
  return g;
}

void enqueue_grid_build(grid_store& gs, grid_scratch& sc, const float4* p4, const float* h,
                        uint64_t n, const grid_params& g, int32_t kz0, float max_h, int kernel,
                        uint32_t* err, hipStream_t s) {
  gs.key.ensure(n);
  gs.perm.ensure(n);
  launch_keys(p4, uint32_t(n), g, gs.key.p, gs.perm.p, err, s);
  enqueue_grid_from_keys(gs, sc, p4, h, n, g, kz0, max_h, kernel, s);
}

void enqueue_grid_maxh(grid_store& gs, const float4* p4, const float* h, uint64_t n,
                       const grid_params& g, int32_t kz0, float max_h, uint32_t* err,
                       hipStream_t s) {
  const uint64_t ncells = g.ncells();
  gs.ready = false;
  gs.maxh_bits.ensure(ncells);
  launch_fill_u32(gs.maxh_bits.p, 0u, ncells, s);
  launch_cell_maxh(p4, h, uint32_t(n), g, gs.maxh_bits.p, err, s);
  gs.volumetric = false;
  gs.gp = g;
  gs.n = n;
  gs.max_h = max_h;
  float dd = g.cs[0] * g.cs[0];
  dd = dd + g.cs[1] * g.cs[1];
  dd = dd + g.cs[2] * g.cs[2];
  gs.cell_radius = 0.5f * std::sqrt(dd);
  gs.kz0 = kz0;
  gs.ready = true;
}

void enqueue_grid_from_keys(grid_store& gs, grid_scratch& sc, const float4* p4, const float* h,
                            uint64_t n, const grid_params& g, int32_t kz0, float max_h,
                            int kernel, hipStream_t s) {
  const uint64_t ncells = g.ncells();
  gs.ready = false;
  gs.key.ensure(n);
  gs.perm.ensure(n);
  gs.start.ensure(ncells + 1);
  gs.maxh_bits.ensure(ncells);
  gs.part.ensure(n);
  gs.hs.ensure(n);
  sc.counts.ensure(ncells + 1);
  sc.temp.ensure(std::max(radix_temp_words(n), scan_temp_words(ncells + 1)));
  sc.horder.ensure(2 * std::max<uint64_t>(n, 1));
  launch_h_order(h, gs.key.p, gs.perm.p, uint32_t(n), sc.horder.p, sc.temp.p, s);
  launch_radix_sort(gs.key.p, gs.perm.p, uint32_t(n), key_bits_for(ncells - 1), sc.temp.p, s);
  launch_fill_u32(sc.counts.p, 0u, ncells + 1, s);
  launch_fill_u32(gs.maxh_bits.p, 0u, ncells, s);
  gs.volumetric = false;
  launch_build_soa(p4, h, gs.perm.p, gs.key.p, uint32_t(n), kernel_norm(kernel),
                   kernel != ILC_KERNEL_CUBIC, gs.part.p,
                   gs.hs.p, sc.counts.p, gs.maxh_bits.p, s);
  launch_exclusive_scan(sc.counts.p, gs.start.p, ncells + 1, sc.temp.p, s);
  gs.gp = g;
  gs.n = n;
  gs.max_h = max_h;
  float dd = g.cs[0] * g.cs[0];
  dd = dd + g.cs[1] * g.cs[1];
  dd = dd + g.cs[2] * g.cs[2];
  gs.cell_radius = 0.5f * std::sqrt(dd);
  gs.kz0 = kz0;
  gs.ready = true;
}

void enqueue_volumetric_grid(grid_store& gs, grid_scratch& sc, const float4* p4, const float* h,
                             uint64_t n, const grid_params& g, float max_h, int kernel,
                             uint32_t* err, hipStream_t s) {
  gs.ready = false;
  gs.maxh_bits.ensure(g.ncells());
  gs.lv_key.ensure(n);
  gs.lv_pos.ensure(n);
  gs.lv_w.ensure(n);
  gs.lv_cell.ensure(n);
  gs.lv_meta.ensure(kLvMetaWords);
  sc.lv_tmp.ensure(11 * std::max<uint64_t>(n, 1));
  sc.temp.ensure(radix_temp_words(n));
  level_cell_sizes(g, max_h, gs.lv_cs);
  level_planes(g, gs.lv_cs, gs.lv_poff);
  gs.lv_pdir.ensure(gs.lv_poff[kLevels3]);
  launch_volumetric_index(p4, h, uint32_t(n), g, max_h, kernel_norm_3d(kernel), gs.levels(),
                          gs.lv_key.p, gs.lv_pos.p, gs.lv_w.p, gs.lv_cell.p, gs.lv_meta.p,
                          gs.lv_pdir.p, gs.maxh_bits.p, err, sc.lv_tmp.p, sc.temp.p, s);
  gs.gp = g;
  gs.n = n;
  gs.max_h = max_h;
  float dd = g.cs[0] * g.cs[0];
  dd = dd + g.cs[1] * g.cs[1];
  dd = dd + g.cs[2] * g.cs[2];
  gs.cell_radius = 0.5f * std::sqrt(dd);
  gs.kz0 = 0;
  gs.volumetric = true;
  gs.ready = true;
}

}  // namespace ilc
