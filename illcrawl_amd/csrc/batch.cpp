// Batched multi-view rendering (C ABI ilc_batch_*).
//
// Many projections of one resident snapshot (animation frames, rotated
// cameras, slabs) in one persistent gather launch.  This replaces the
// reference's frame loop distributed over MPI ranks (animation.cpp:44-100,
// uniform_work_partitioner.cpp:36-52) on a single GPU.
//
// A render is
//   1. plan: for every view, one fused pass (quantity, projection, depth
//      mask) reduces the projected bounding box; one synchronisation; the
//      host derives each view's reference grid (particle_grid.cpp:43-100);
//   2. device work: per view, the projected particles and the grid build
//      (keys, stable radix sort, sorted SoA, cell offsets), all resident;
//      then ONE persistent gather over every (view, tile) work item, the
//      views with the coarsest pixels first, pulled from per-XCD device
//      counters; then per-view
//      retrieval (FP64 -> FP32 times the view's dA scale).
// Step 2 can be captured in a hipGraph and replayed while the plan is
// unchanged.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "gridbuild.hpp"
#include "kernels.hpp"

namespace ilc {
std::vector<float> cubic_projected_table(int n);
std::vector<float> cubic_projected_dtable(int n);
}

using namespace ilc;

namespace {

// Scratch of one stream building grids.
struct build_lane {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  grid_scratch scratch;
  dbuf<float4> p4tmp;
  dbuf<float> htmp;
  dbuf<uint32_t> wcounts, woffs;
  ~build_lane() {
    if (done) (void)hipEventDestroy(done);
    if (st) (void)hipStreamDestroy(st);
  }
};
constexpr int kLanes = 4;

struct view_state {
  ilc_view v{};
  camera_frame f{};
  grid_store grid;
  dbuf<float4> pts;
  dbuf<double> acc;
  dbuf<float> img;
  uint32_t tiles_x = 0, ntiles = 0;
  grid_params gp{};
  int32_t kz0 = 0;
  float max_h = 0.f;
  bool windowed = false;  // grid built from the particles in `win` only
  bool compact = false;   // particle lists from the depth cut only; max h from the group
  int group = -1;         // slab group (views with identical projected x', y')
  window_f win{};
  uint32_t nwin = 0;      // particles kept by the filter (from the plan)
  uint32_t nwin_plain = 0;  // the same without the screen test (single-view builds)
  dbuf<uint32_t> orig;    // filtered grids: snapshot index of each grid input row
  // the view's grid as the last render built it: its own grid, or a slice of
  // the batch's shared grid arrays
  grid_view gv{};
  const uint32_t* gperm = nullptr;  // sorted position -> grid input row
  const uint32_t* gorig = nullptr;  // grid input row -> snapshot index (NULL: identity)
  uint64_t gbase = 0;               // first sorted position of the view
  uint64_t gn = 0;                  // particles in the grid
  bool gready = false;
  uint64_t sbase = 0, cbase = 0;    // shared build: first kept row, first cell
};

template <class T>
struct pinned {
  T* p = nullptr;
  size_t cap = 0;
  ~pinned() {
    if (p) (void)hipHostFree(p);
  }
  T* ensure(size_t n) {
    if (n <= cap && p) return p;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    ILC_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (n ? n : 1)));
    cap = n;
    return p;
  }
};

}  // namespace

struct ilc_batch {
  int device = 0;
  uint64_t target = 1000;
  int mode = ILC_MODE_REFERENCE;
  int kernel = ILC_KERNEL_QUARTIC_REF;
  tile_shape ts{32, 4};

  particle_soa ps{};
  bool have_particles = false;
  dbuf<float> gaunt;
  int gaunt_n = 0;
  float gx0 = 0.f, gdx = 1.f;
  dbuf<float> cubic, dcubic;
  bool dcubic_ready = false;
  bool cubic_ready = false;

  std::vector<std::unique_ptr<view_state>> views;
  grid_scratch scratch;
  dbuf<float4> p4tmp;
  dbuf<float> htmp;
  // extra streams (and their scratch) for the per-view grid builds of a
  // persistent render; the builds of different views are independent
  std::vector<std::unique_ptr<build_lane>> lanes;
  hipEvent_t ev_fork = nullptr;
  std::vector<std::unique_ptr<grid_store>> group_grids;  // per-cell max h per slab group
  std::vector<size_t> group_leader;
  dbuf<uint32_t> wcounts, woffs, wtotal;
  pinned<uint32_t> wtotal_host;
  dbuf<float> bbox_parts;
  dbuf<bbox_result> bbox_dev;
  dbuf<uint32_t> bad;
  pinned<bbox_result> bbox_host;
  dbuf<view_desc> desc_dev;
  pinned<view_desc> desc_host;
  // the multi-view passes of the plan: per-view projection and window, the
  // filtered views, and their kept counts per window tile (view-major)
  dbuf<view_build_desc> bdesc_dev;
  pinned<view_build_desc> bdesc_host;
  dbuf<int32_t> fviews_dev;
  pinned<int32_t> fviews_host;
  dbuf<uint32_t> wcounts_all;  // per (view, window tile); the plain counts follow
  // shared build of every view's grid (persistent renders): one scatter of
  // every view's kept particles, one sort by (view, cell, descending h), one
  // SoA pass; a view's grid is a slice
  bool shared = false;
  uint64_t sh_ntot = 0, sh_ncells = 0;
  // line-of-sight kernel clip of every view's depth range (exact mode,
  // quartic kernels, persistent renders): per kept row and per sorted row,
  // the range relative to the particle in units of h
  int clip = 0;
  dbuf<float2> sh_zc, sh_zcs;
  dbuf<uint32_t> sh_offs, sh_gkey, sh_orig, sh_key, sh_perm, sh_sk, sh_counts, sh_start, sh_maxh,
      sh_temp, sh_cbase;
  pinned<uint32_t> sh_cbase_host;
  dbuf<float> sh_wx, sh_hs;
  dbuf<float4> sh_p4c, sh_part;
  // every view a slab of one stack (the multi-view passes test two slabs per
  // particle instead of every view)
  slab_stack stack;
  dbuf<uint2> items_dev;
  pinned<uint2> items_host;
  uint32_t nitems = 0;
  // the items as the gather takes them: every queue heaviest tiles first
  dbuf<uint2> items_ord;
  dbuf<uint32_t> item_keys, item_idx, item_temp;
  queue_segs segs{};
  dbuf<uint32_t> qcounters;  // per-queue work counters (kQueueStride apart)
  dbuf<uint32_t> counter;  // [0] unused, [1] out-of-grid particles
  pinned<uint32_t> flags_host;

  bool planned = false;
  std::vector<uint64_t> plan_sig;
  int grid_blocks = 0;

  hipStream_t own = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // timing of the gather launches of the last eager render
  std::vector<hipEvent_t> tev;
  int tev_used = 0;
  float last_gather_ms = 0.f;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  bool ran_eagerly = false;

  ~ilc_batch() {
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    lanes.clear();
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (own) (void)hipStreamDestroy(own);
  }
};

namespace {

template <class F>
ilc_status guarded(F&& f) {
  try {
    f();
    set_last_error("");
    return ILC_OK;
  } catch (const ilc::error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return ILC_ENOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return ILC_EINVAL;
  }
}

struct device_guard {
  int prev = -1;
  explicit device_guard(int dev) {
    ILC_HIP(hipGetDevice(&prev));
    if (prev != dev) ILC_HIP(hipSetDevice(dev));
  }
  ~device_guard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
  }
};

// The views as one slab stack (slab_stack), when they are one: a single slab
// group of centre-masked views (no window, no screen test, no clip), the
// depth row (0, 0, +-1) the same for all, depth ranges [s, s + depth],
// s = m8 * cz, that follow one another without overlap, each thicker than
// FP32 rounding of its coordinates.  Else n = 0.
slab_stack find_stack(const ilc_batch* b) {
  slab_stack st;
  const size_t nv = b->views.size();
  if (nv < 2 || nv > size_t(kStackMax) || b->clip) return st;
  const float m8 = b->views[0]->f.minv[8];
  if (!(m8 == 1.0f || m8 == -1.0f)) return st;
  std::vector<std::pair<double, size_t>> order;
  for (size_t v = 0; v < nv; ++v) {
    const view_state& vs = *b->views[v];
    if (!vs.compact || vs.group != 0 || vs.win.use_window || vs.win.use_screen ||
        !vs.win.use_depth || vs.win.clip || vs.f.minv[8] != m8)
      return st;
    order.emplace_back(double(m8 * vs.f.center[2]), v);
  }
  std::sort(order.begin(), order.end());
  for (size_t k = 0; k < nv; ++k) {
    const double s0 = order[k].first;
    const double d = double(static_cast<float>(b->views[order[k].second]->v.cam.depth));
    if (!(std::isfinite(s0) && d > 1e-5 * (std::fabs(s0) + d + 1.0))) return st;
    if (k + 1 < nv && order[k + 1].first < s0 + d) return st;
  }
  for (size_t k = 0; k < nv; ++k) {
    st.s[k] = static_cast<float>(order[k].first);
    st.view[k] = static_cast<int32_t>(order[k].second);
  }
  st.m8 = m8;
  st.n = static_cast<int>(nv);
  return st;
}

view_proj proj_of(const ilc_batch* b, const view_state& vs) {
  view_proj vp;
  for (int k = 0; k < 9; ++k) vp.m[k] = vs.f.minv[k];
  for (int k = 0; k < 3; ++k) vp.c[k] = vs.f.center[k];
  vp.depth = static_cast<float>(vs.v.cam.depth);
  vp.quantity = vs.v.quantity;
  vp.scale = 1.0;
  (void)b;
  return vp;
}

// A view's grid as the gather sees it (set by the build of the last render).
grid_view grid_of(const ilc_batch* b, const view_state& vs) {
  (void)b;
  return vs.gv;
}

float cell_radius_of(const grid_params& g) {
  float dd = g.cs[0] * g.cs[0];
  dd = dd + g.cs[1] * g.cs[1];
  dd = dd + g.cs[2] * g.cs[2];
  return 0.5f * std::sqrt(dd);
}

// the view's own grid (per-view build); a depth-compacted view takes the
// per-cell max h of its group's full grid
void own_grid(const ilc_batch* b, view_state& vs) {
  vs.gv = vs.grid.view();
  if (vs.compact) vs.gv.maxh_bits = b->group_grids[size_t(vs.group)]->maxh_bits.p;
  vs.gperm = vs.grid.perm.p;
  vs.gorig = (vs.windowed || vs.compact) ? vs.orig.p : nullptr;
  vs.gbase = 0;
  vs.gn = vs.grid.n;
  vs.gready = true;
}

gaunt_table gaunt_of(const ilc_batch* b) {
  return gaunt_table{b->gaunt.p, b->gaunt_n, b->gx0, b->gdx};
}

void require_ready(const ilc_batch* b) {
  ILC_REQUIRE(b != nullptr, "null batch");
  if (!b->have_particles) fail(ILC_ENOTREADY, "no particles set");
  if (b->views.empty()) fail(ILC_ENOTREADY, "no views set");
}

void ensure_cubic(ilc_batch* b) {
  if (b->kernel != ILC_KERNEL_CUBIC || b->cubic_ready) return;
  std::vector<float> t = cubic_projected_table(4096);
  b->cubic.ensure(t.size());
  ILC_HIP(hipMemcpy(b->cubic.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
  b->cubic_ready = true;
}

// 1. plan: bounding boxes of every view, one sync, grid parameters.
void plan(ilc_batch* b, hipStream_t s) {
  const size_t nv = b->views.size();
  const int nparts = multi_bbox_parts(b->ps.n);
  b->bbox_parts.ensure(nv * size_t(nparts) * 8);
  b->bbox_dev.ensure(nv);
  b->bad.ensure(nv);
  b->bbox_host.ensure(nv);
  b->bdesc_host.ensure(nv);
  b->bdesc_dev.ensure(nv);
  for (size_t v = 0; v < nv; ++v) {
    b->bdesc_host.p[v].vp = proj_of(b, *b->views[v]);
    b->bdesc_host.p[v].win = window_f{0.f, 0.f, 0.f, 0.f, 0, 0};
  }
  // slab groups: views whose projection rows for x' and y' and centre x, y are
  // bitwise equal, with no z term in those rows and no x, y term in the depth
  // row, project every particle to the same (x', y') and differ only in the
  // depth cut; they share one full grid (same bounding box, cells and per-cell
  // max h) and keep only their depth-cut particles in their lists.
  b->group_leader.clear();
  for (size_t v = 0; v < nv; ++v) {
    b->views[v]->group = -1;
    b->views[v]->compact = false;
  }
  for (size_t v = 0; v < nv; ++v) {
    view_state& vs = *b->views[v];
    const float* m = vs.f.minv;
    if (vs.group >= 0 || m[2] != 0.0f || m[5] != 0.0f || m[6] != 0.0f || m[7] != 0.0f) continue;
    for (size_t u = v + 1; u < nv; ++u) {
      view_state& us = *b->views[u];
      if (us.group >= 0 || us.v.quantity != vs.v.quantity) continue;
      if (std::memcmp(us.f.minv, m, 6 * sizeof(float)) != 0 ||
          std::memcmp(us.f.center, vs.f.center, 2 * sizeof(float)) != 0 ||
          us.f.minv[6] != 0.0f || us.f.minv[7] != 0.0f)
        continue;
      if (vs.group < 0) {
        vs.group = static_cast<int>(b->group_leader.size());
        b->group_leader.push_back(v);
      }
      us.group = vs.group;
    }
  }
  for (size_t v = 0; v < nv; ++v) b->views[v]->compact = b->views[v]->group >= 0;
  while (b->group_grids.size() < b->group_leader.size())
    b->group_grids.push_back(std::make_unique<grid_store>());
  bool one_group = !b->group_leader.empty();
  for (size_t v = 0; v < nv; ++v) one_group = one_group && b->views[v]->group == 0;
  ILC_HIP(hipMemcpyAsync(b->bdesc_dev.p, b->bdesc_host.p, sizeof(view_build_desc) * nv,
                         hipMemcpyHostToDevice, s));
  launch_multi_bbox(b->ps, b->bdesc_dev.p, static_cast<int>(nv), gaunt_of(b), b->bbox_parts.p,
                    b->bad.p, b->bbox_dev.p, one_group, s);
  ILC_HIP(hipMemcpyAsync(b->bbox_host.p, b->bbox_dev.p, sizeof(bbox_result) * nv,
                         hipMemcpyDeviceToHost, s));
  ILC_HIP(hipStreamSynchronize(s));
  std::vector<uint64_t> sig;
  for (size_t v = 0; v < nv; ++v) {
    view_state& vs = *b->views[v];
    vs.gp = plan_grid(b->bbox_host.p[v], b->ps.n, b->target, b->mode, &vs.kz0);
    vs.max_h = b->bbox_host.p[v].max_h;
    uint32_t bits[10];
    std::memcpy(bits, vs.gp.nc, 12);
    std::memcpy(bits + 3, vs.gp.minc, 12);
    std::memcpy(bits + 6, vs.gp.cs, 12);
    std::memcpy(bits + 9, &vs.max_h, 4);
    for (uint32_t x : bits) sig.push_back(x);
    sig.push_back(uint64_t(vs.kz0));
  }

  // windows: the screen grown by the largest reach of any point (F3 radius
  // <= max h; support up to 2 h) plus whole cells.  Reference mode needs every
  // point's own cell inside the grid (an out-of-grid point's unclamped key
  // can alias any cell), else the view keeps all particles.
  for (size_t v = 0; v < nv; ++v) {
    view_state& vs = *b->views[v];
    const double ps = vs.v.cam.width / double(vs.v.cam.nx);
    const double hx = (0.5 * vs.v.cam.nx + 2.0) * ps, hy = (0.5 * vs.v.cam.ny + 2.0) * ps;
    const double cs = std::max(vs.gp.cs[0], vs.gp.cs[1]);
    const double reach = 2.0 * double(vs.max_h) + 4.0 * cs;
    bool ok = std::isfinite(reach) && std::isfinite(hx) && std::isfinite(hy);
    if (ok && b->mode == ILC_MODE_REFERENCE) {
      for (int a = 0; a < 2; ++a) {
        const double e = a == 0 ? hx : hy;
        const double lo = std::floor((-e - double(vs.gp.minc[a])) / double(vs.gp.cs[a]));
        const double hi = std::floor((e - double(vs.gp.minc[a])) / double(vs.gp.cs[a]));
        ok = ok && lo >= 1.0 && hi <= double(vs.gp.nc[a]) - 2.0;
      }
    }
    const bbox_result& bb = b->bbox_host.p[v];
    // no gain when the window holds the whole bounding box
    if (ok && -hx - reach <= bb.mn[0] && hx + reach >= bb.mx[0] && -hy - reach <= bb.mn[1] &&
        hy + reach >= bb.mx[1])
      ok = false;
    vs.windowed = ok && !vs.compact;
    if (vs.windowed) {
      vs.win = window_f{std::nextafter(float(-hx - reach), -INFINITY),
                        std::nextafter(float(hx + reach), INFINITY),
                        std::nextafter(float(-hy - reach), -INFINITY),
                        std::nextafter(float(hy + reach), INFINITY), 1, 0};
      // the screen: every pixel point's camera-frame coordinates, with a
      // margin for their FP32 rounding; rows whose support disk (the
      // gather's cull, with the count pass's slack) misses it are not kept
      double cmax = 0.0;
      for (int a = 0; a < 3; ++a) cmax = std::max(cmax, std::fabs(vs.v.cam.position[a]));
      const double mg = 1e-5 * (cmax + 2.0 * std::max(hx, hy) + double(vs.max_h));
      const float sup = b->kernel == ILC_KERNEL_CUBIC ? 2.0f : 1.0f;
      vs.win.use_screen = 1;
      vs.win.sx0 = std::nextafter(float(-hx - mg), -INFINITY);
      vs.win.sx1 = std::nextafter(float(hx + mg), INFINITY);
      vs.win.sy0 = std::nextafter(float(-hy - mg), -INFINITY);
      vs.win.sy1 = std::nextafter(float(hy + mg), INFINITY);
      vs.win.cull2 = sup * sup * 1.0001f;
      // two cells of margin absorb the FP32 cell arithmetic of the device
      auto cell_at = [&](float x, int a) {
        return std::floor((double(x) - double(vs.gp.minc[a])) / double(vs.gp.cs[a]));
      };
      vs.win.scx0 = int(std::max(cell_at(vs.win.sx0, 0) - 2.0, -1.0));
      vs.win.scx1 = int(std::min(cell_at(vs.win.sx1, 0) + 2.0, double(vs.gp.nc[0])));
      vs.win.scy0 = int(std::max(cell_at(vs.win.sy0, 1) - 2.0, -1.0));
      vs.win.scy1 = int(std::min(cell_at(vs.win.sy1, 1) + 2.0, double(vs.gp.nc[1])));
    } else if (vs.compact) {
      vs.win = window_f{0.f, 0.f, 0.f, 0.f, 0, 1};
    } else {
      vs.win = window_f{0.f, 0.f, 0.f, 0.f, 0, 0};  // keeps every particle
    }
    vs.win.clip = b->clip;
  }
  {
    // kept counts per window tile of every view (unfiltered views keep
    // all), in one pass; the render scans and scatters from them
    const uint32_t nt = window_tiles(b->ps.n);
    b->wcounts_all.ensure(2 * nv * size_t(nt));
    b->wtotal.ensure(2 * nv);
    b->wtotal_host.ensure(2 * nv);
    b->fviews_host.ensure(nv);
    b->fviews_dev.ensure(nv);
    int nf = 0;
    for (size_t v = 0; v < nv; ++v) {
      b->bdesc_host.p[v].win = b->views[v]->win;
      b->bdesc_host.p[v].g = b->views[v]->gp;
      b->fviews_host.p[nf++] = static_cast<int32_t>(v);
    }
    ILC_HIP(hipMemcpyAsync(b->bdesc_dev.p, b->bdesc_host.p, sizeof(view_build_desc) * nv,
                           hipMemcpyHostToDevice, s));
    ILC_HIP(hipMemcpyAsync(b->fviews_dev.p, b->fviews_host.p, sizeof(int32_t) * size_t(nf),
                           hipMemcpyHostToDevice, s));
    launch_fill_u32(b->wtotal.p, 0u, 2 * nv, s);
    b->stack = find_stack(b);
    launch_multi_window_count(b->ps.pos3, b->ps.hsml, b->ps.n, b->bdesc_dev.p, b->fviews_dev.p,
                              nf, b->wcounts_all.p, b->wtotal.p,
                              b->wcounts_all.p + nv * size_t(nt), b->wtotal.p + nv, b->stack, s);
    ILC_HIP(hipMemcpyAsync(b->wtotal_host.p, b->wtotal.p, sizeof(uint32_t) * 2 * nv,
                           hipMemcpyDeviceToHost, s));
    ILC_HIP(hipStreamSynchronize(s));
  }
  uint64_t ntot = 0, ncells = 0;
  b->sh_cbase_host.ensure(nv);
  for (size_t v = 0; v < nv; ++v) {
    view_state& vs = *b->views[v];
    const bool filtered = vs.windowed || vs.compact;
    vs.nwin = filtered ? b->wtotal_host.p[v] : 0u;
    vs.nwin_plain = filtered ? b->wtotal_host.p[nv + v] : 0u;
    sig.push_back(filtered ? uint64_t(vs.nwin) : ~uint64_t(0));
    sig.push_back(uint64_t(vs.nwin_plain));
    sig.push_back(uint64_t(int64_t(vs.group)));
    vs.sbase = ntot;
    vs.cbase = ncells;
    b->sh_cbase_host.p[v] = static_cast<uint32_t>(ncells);
    ntot += b->wtotal_host.p[v];
    ncells += vs.gp.ncells();
  }
  // shared build when every index fits in 31 bits.  ILC_BATCH_SHARED_MAX_ROWS
  // lowers the row limit, so that tests reach the per-view builds (taken
  // above 2^31 rows) with small snapshots.
  uint64_t max_rows = uint64_t(1) << 31;
  if (const char* e = std::getenv("ILC_BATCH_SHARED_MAX_ROWS"))
    max_rows = std::min<uint64_t>(max_rows, std::strtoull(e, nullptr, 10));
  b->shared = ntot > 0 && ntot < max_rows && ncells < (uint64_t(1) << 31);
  b->sh_ntot = ntot;
  b->sh_ncells = ncells;
  sig.push_back(b->shared ? ntot : 0);
  sig.push_back(uint64_t(b->clip));
  sig.push_back(uint64_t(b->ts.tw) << 40 | uint64_t(b->ts.nb) << 20 | uint64_t(b->ts.ppt));
  // work items: views with the coarsest pixels first (their tiles cover the
  // most area, so over dense regions they stream the most particles and run
  // longest: started first, they finish while the finer views' many light
  // tiles keep the GPU busy), then larger views first; tiles in image order
  std::vector<uint32_t> order(nv);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
    const ilc_camera& ca = b->views[a]->v.cam;
    const ilc_camera& cc = b->views[c]->v.cam;
    const double psa = ca.width / double(ca.nx), psc = cc.width / double(cc.nx);
    if (psa != psc) return psa > psc;
    return uint64_t(ca.nx) * ca.ny > uint64_t(cc.nx) * cc.ny;
  });
  uint64_t nitems = 0;
  for (size_t v = 0; v < nv; ++v) {
    view_state& vs = *b->views[v];
    tile_grid(b->ts, vs.v.cam.nx, vs.v.cam.ny, &vs.tiles_x, &vs.ntiles);
    nitems += vs.ntiles;
  }
  ILC_REQUIRE(nitems < (uint64_t(1) << 32), "too many work items");
  // queue q: the q-th eighth (contiguous tiles, image order) of every view,
  // largest views first
  uint2* it = b->items_host.ensure(nitems);
  uint64_t k = 0;
  for (int q = 0; q < kQueues; ++q) {
    b->segs.seg[q] = static_cast<uint32_t>(k);
    for (uint32_t v : order) {
      const uint64_t nt = b->views[v]->ntiles;
      const uint64_t t0 = nt * uint64_t(q) / kQueues, t1 = nt * uint64_t(q + 1) / kQueues;
      for (uint64_t t = t0; t < t1; ++t) it[k++] = make_uint2(v, static_cast<uint32_t>(t));
    }
  }
  b->segs.seg[kQueues] = static_cast<uint32_t>(k);
  b->nitems = static_cast<uint32_t>(nitems);
  b->items_dev.ensure(nitems);
  b->items_ord.ensure(nitems);
  b->item_keys.ensure(nitems);
  b->item_idx.ensure(nitems);
  b->item_temp.ensure(radix_temp_words(nitems));
  b->desc_dev.ensure(nv);
  b->desc_host.ensure(nv);
  b->counter.ensure(2);
  b->flags_host.ensure(2);
  const bool same = b->planned && sig == b->plan_sig;
  if (!same) b->graph_valid = false;
  b->plan_sig = sig;
  b->planned = true;
  b->grid_blocks = gather_batch_grid(
      b->clip ? OP_CLIP : (b->mode == ILC_MODE_REFERENCE ? OP_REFERENCE : OP_EXACT), b->kernel,
      b->ts);
}

// An event for timing the gather launch(es) of an eager render.
hipEvent_t timing_event(ilc_batch* b) {
  if (b->tev_used == int(b->tev.size())) {
    hipEvent_t e;
    ILC_HIP(hipEventCreate(&e));
    b->tev.push_back(e);
  }
  return b->tev[b->tev_used++];
}

// Everything a grid build writes besides the view's own grid.
struct build_ctx {
  grid_scratch* sc;
  dbuf<float4>* p4tmp;
  dbuf<float>* htmp;
  dbuf<uint32_t>* wcounts;
  dbuf<uint32_t>* woffs;
  hipStream_t st;
};

size_t view_index(const ilc_batch* b, const view_state& vs) {
  for (size_t v = 0; v < b->views.size(); ++v)
    if (b->views[v].get() == &vs) return v;
  fail(ILC_EINVAL, "view not in batch");
}

// the per-cell max h of slab group gi's full grid (its depth-compacted views
// search their own cells with it; nothing else of that grid is read)
void build_group(ilc_batch* b, size_t gi, const build_ctx& c) {
  view_state& ls = *b->views[b->group_leader[gi]];
  c.p4tmp->ensure(b->ps.n);
  launch_view_project(b->ps, proj_of(b, ls), gaunt_of(b), c.p4tmp->p, c.st);
  enqueue_grid_maxh(*b->group_grids[gi], c.p4tmp->p, b->ps.hsml, b->ps.n, ls.gp, ls.kz0,
                    ls.max_h, b->counter.p + 1, c.st);
}

// view v's grid and cleared accumulator
// clear: clear the view's accumulator (renders whose gather adds to it; the
// persistent gather writes it whole)
void build_view(ilc_batch* b, view_state& vs, const build_ctx& c, bool clear) {
  const hipStream_t st = c.st;
  c.p4tmp->ensure(b->ps.n);
  if (vs.windowed || vs.compact) {
    // kept particles in input order, then the grid from their keys
    const uint32_t nt = window_tiles(b->ps.n);
    const uint64_t m = vs.nwin_plain;
    c.woffs->ensure(size_t(nt) + 1);
    c.htmp->ensure(std::max<uint64_t>(m, 1));
    vs.grid.key.ensure(std::max<uint64_t>(m, 1));
    vs.grid.perm.ensure(std::max<uint64_t>(m, 1));
    c.sc->temp.ensure(scan_temp_words(uint64_t(nt) + 1));
    // the plan counted the kept particles of every window tile of this view
    launch_exclusive_scan(b->wcounts_all.p + (b->views.size() + view_index(b, vs)) * size_t(nt),
                          c.woffs->p, nt, c.sc->temp.p, st);
    vs.orig.ensure(std::max<uint64_t>(m, 1));
    launch_window_scatter(b->ps, proj_of(b, vs), gaunt_of(b), vs.win, c.woffs->p, vs.gp,
                          c.p4tmp->p, c.htmp->p, vs.grid.key.p, vs.grid.perm.p, vs.orig.p,
                          b->counter.p + 1, st);
    enqueue_grid_from_keys(vs.grid, *c.sc, c.p4tmp->p, c.htmp->p, m, vs.gp, vs.kz0, vs.max_h,
                           b->kernel, st);
    vs.grid.density_n = vs.windowed ? b->ps.n : 0;
  } else {
    launch_view_project(b->ps, proj_of(b, vs), gaunt_of(b), c.p4tmp->p, st);
    enqueue_grid_build(vs.grid, *c.sc, c.p4tmp->p, b->ps.hsml, b->ps.n, vs.gp, vs.kz0, vs.max_h,
                       b->kernel, b->counter.p + 1, st);
    vs.grid.density_n = 0;
  }
  own_grid(b, vs);
  if (clear) launch_fill_f64(vs.acc.p, 0.0, uint64_t(vs.v.cam.nx) * vs.v.cam.ny, st);
}

// every view's grid in one pipeline on stream s (persistent renders)
void build_shared(ilc_batch* b, hipStream_t s) {
  const size_t nv = b->views.size();
  const uint32_t nt = window_tiles(b->ps.n);
  const uint64_t ntot = b->sh_ntot, nc = b->sh_ncells;
  const uint32_t n32 = static_cast<uint32_t>(ntot);
  b->sh_offs.ensure(nv * size_t(nt) + 1);
  b->sh_temp.ensure(std::max({radix2_temp_words(ntot), scan_temp_words(nv * uint64_t(nt) + 1),
                              scan_temp_words(nc + 1)}));
  for (auto* d : {&b->sh_gkey, &b->sh_orig, &b->sh_key, &b->sh_perm, &b->sh_sk})
    d->ensure(ntot);
  b->sh_hs.ensure(ntot);
  b->sh_p4c.ensure(ntot);
  b->sh_part.ensure(ntot);
  b->sh_counts.ensure(nc + 1);
  b->sh_start.ensure(nc + 1);
  b->sh_maxh.ensure(nc);
  b->sh_cbase.ensure(nv);
  ILC_HIP(hipMemcpyAsync(b->sh_cbase.p, b->sh_cbase_host.p, sizeof(uint32_t) * nv,
                         hipMemcpyHostToDevice, s));
  // kept particles of every view at global positions
  launch_exclusive_scan(b->wcounts_all.p, b->sh_offs.p, nv * uint64_t(nt), b->sh_temp.p, s);
  bool any_xray = false;
  for (auto& v : b->views) any_xray |= v->v.quantity == ILC_QUANTITY_XRAY;
  if (any_xray) {
    b->sh_wx.ensure(b->ps.n);
    launch_particle_xray(b->ps, gaunt_of(b), b->sh_wx.p, s);
  }
  // per-cell max h starts at zero: the scatter raises it for the rows its
  // screen test leaves out, the SoA pass for the kept rows.  A batch of
  // depth-compacted views only (slabs of groups) reads every max h from its
  // groups' full grids and needs none of it.
  bool all_compact = true;
  for (auto& v : b->views) all_compact = all_compact && v->compact;
  uint32_t* const maxh = all_compact ? nullptr : b->sh_maxh.p;
  if (maxh) launch_fill_u32(maxh, 0u, nc, s);
  if (b->clip) {
    b->sh_zc.ensure(ntot);
    b->sh_zcs.ensure(ntot);
  }
  launch_multi_window_scatter(b->ps, any_xray ? b->sh_wx.p : nullptr, b->bdesc_dev.p,
                              b->fviews_dev.p, static_cast<int>(nv), b->sh_offs.p, b->sh_p4c.p,
                              b->sh_sk.p, b->sh_perm.p, b->sh_gkey.p, b->sh_orig.p,
                              b->counter.p + 1, b->sh_cbase.p, maxh,
                              b->clip ? b->sh_zc.p : nullptr, b->stack, s);
  // stable sorts by descending h (each row's global cell key moving with
  // it), then by global cell key (view, then cell): a view's rows end up in
  // one range, by cell, biggest h first, ties in input order
  launch_radix_sort2(b->sh_sk.p, b->sh_perm.p, b->sh_gkey.p, n32, 16, b->sh_temp.p, s);
  launch_radix_sort(b->sh_gkey.p, b->sh_perm.p, n32, key_bits_for(b->sh_ncells - 1), b->sh_temp.p,
                    s);
  // view-local cell keys for the gather; the global ones (sh_gkey) give the
  // cell ranges
  launch_local_keys(b->sh_gkey.p, b->sh_cbase.p, static_cast<int>(nv), n32, b->sh_key.p, s);
  launch_fill_u32(b->sh_counts.p, 0u, nc + 1, s);
  launch_build_soa(b->sh_p4c.p, nullptr, b->sh_perm.p, b->sh_gkey.p, n32, kernel_norm(b->kernel),
                   b->kernel != ILC_KERNEL_CUBIC,
                   b->sh_part.p, b->sh_hs.p, b->sh_counts.p, maxh, s,
                   b->clip ? b->sh_zc.p : nullptr, b->clip ? b->sh_zcs.p : nullptr);
  launch_exclusive_scan(b->sh_counts.p, b->sh_start.p, nc + 1, b->sh_temp.p, s);
  for (size_t v = 0; v < nv; ++v) {
    view_state& vs = *b->views[v];
    grid_view g{};
    g.part = b->sh_part.p;
    g.clip = b->clip ? b->sh_zcs.p : nullptr;
    g.key = b->sh_key.p;
    g.hs = b->sh_hs.p;
    g.start = b->sh_start.p + vs.cbase;
    g.maxh_bits = vs.compact ? b->group_grids[size_t(vs.group)]->maxh_bits.p
                             : b->sh_maxh.p + vs.cbase;
    for (int a = 0; a < 3; ++a) {
      g.nc[a] = vs.gp.nc[a];
      g.minc[a] = vs.gp.minc[a];
      g.cs[a] = vs.gp.cs[a];
    }
    g.max_h = vs.max_h;
    g.cell_radius = cell_radius_of(vs.gp);
    g.kz0 = vs.kz0;
    g.ncells = static_cast<uint32_t>(vs.gp.ncells());
    const uint64_t kept = vs.compact ? uint64_t(vs.nwin) : (vs.windowed ? vs.nwin : b->ps.n);
    // enumerate like the grid the view stands for (grid_store::view)
    const uint64_t dn = vs.compact ? kept : b->ps.n;
    g.sparse = dn * 5 < uint64_t(g.ncells) * 2 ? 1 : 0;
    vs.gv = g;
    vs.gperm = b->sh_perm.p;
    vs.gorig = b->sh_orig.p;
    vs.gbase = vs.sbase;
    vs.gn = kept;
    vs.gready = true;
  }
}

view_desc make_desc(const ilc_batch* b, size_t v) {
  const view_state& vs = *b->views[v];
  const uint64_t P = uint64_t(vs.v.cam.nx) * vs.v.cam.ny;
  view_desc d{};
  d.g = grid_of(b, vs);
  // the pixel corners, computed by the gather as the point kernel does
  d.p = point_view{vs.pts.p, static_cast<uint32_t>(P), vs.v.cam.nx, vs.v.cam.ny, 1,
                   pixel_gen_of(vs.f)};
  d.acc = vs.acc.p;
  d.tiles_x = vs.tiles_x;
  d.ntiles = vs.ntiles;
  return d;
}

// The persistent gather's descriptor of view v: it writes the accumulator
// and the image itself.  desc_host is the source of the captured graph's
// descriptor copy, so only persistent renders write it.
void fill_desc(ilc_batch* b, size_t v) {
  view_desc d = make_desc(b, v);
  const view_state& vs = *b->views[v];
  d.img = vs.img.p;
  d.scale = vs.v.scale;
  d.fused = 1;
  const uint64_t rows = (vs.windowed || vs.compact) ? uint64_t(vs.nwin) : b->ps.n;
  d.cost_unit = static_cast<uint32_t>(
      std::min<uint64_t>(0xffffffu, std::max<uint64_t>(1, rows / std::max<uint32_t>(vs.ntiles, 1u))));
  b->desc_host.p[v] = d;
}

void ensure_lanes(ilc_batch* b) {
  if (!b->ev_fork) ILC_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  while (b->lanes.size() < size_t(kLanes)) {
    auto l = std::make_unique<build_lane>();
    ILC_HIP(hipStreamCreateWithFlags(&l->st, hipStreamNonBlocking));
    ILC_HIP(hipEventCreateWithFlags(&l->done, hipEventDisableTiming));
    b->lanes.push_back(std::move(l));
  }
}

// 2. device work of a planned render: persistent (the grid builds of all
// views spread over kLanes streams, then one gather launch over every
// (view, tile)) or plain (build, gather and retrieve one view after the
// other on `s`).
void device_work(ilc_batch* b, bool persistent, hipStream_t s, bool timed = false) {
  const size_t nv = b->views.size();
  b->tev_used = 0;
  const int op = b->clip ? OP_CLIP : (b->mode == ILC_MODE_REFERENCE ? OP_REFERENCE : OP_EXACT);
  if (b->clip && (!persistent || !b->shared))
    fail(ILC_EINVAL, "the line-of-sight clip needs a persistent render with the shared build");
  launch_fill_u32(b->counter.p, 0u, 2, s);
  b->qcounters.ensure(size_t(kQueues) * kQueueStride);
  launch_fill_u32(b->qcounters.p, 0u, uint64_t(kQueues) * kQueueStride, s);
  const build_ctx main{&b->scratch, &b->p4tmp, &b->htmp, &b->wcounts, &b->woffs, s};
  if (!persistent) {
    for (size_t gi = 0; gi < b->group_leader.size(); ++gi) build_group(b, gi, main);
    for (size_t v = 0; v < nv; ++v) {
      view_state& vs = *b->views[v];
      build_view(b, vs, main, true);
      const view_desc d = make_desc(b, v);
      if (timed) ILC_HIP(hipEventRecord(timing_event(b), s));
      launch_gather(d.g, d.p, op, b->kernel, b->ts, vs.acc.p, nullptr, b->cubic.p, s);
      if (timed) ILC_HIP(hipEventRecord(timing_event(b), s));
      launch_retrieve(vs.acc.p, d.p.n, vs.v.scale, vs.img.p, s);
    }
  } else {
    ensure_lanes(b);
    ILC_HIP(hipEventRecord(b->ev_fork, s));
    for (auto& l : b->lanes) ILC_HIP(hipStreamWaitEvent(l->st, b->ev_fork, 0));
    size_t job = 0;
    auto lane_ctx = [&](size_t j) {
      build_lane& l = *b->lanes[j % b->lanes.size()];
      return build_ctx{&l.scratch, &l.p4tmp, &l.htmp, &l.wcounts, &l.woffs, l.st};
    };
    for (size_t gi = 0; gi < b->group_leader.size(); ++gi) build_group(b, gi, lane_ctx(job++));
    if (b->shared) {
      build_shared(b, s);
      for (size_t v = 0; v < nv; ++v) fill_desc(b, v);
    } else {
      for (size_t v = 0; v < nv; ++v) {
        build_view(b, *b->views[v], lane_ctx(job++), false);
        fill_desc(b, v);
      }
    }
    for (auto& l : b->lanes) {
      ILC_HIP(hipEventRecord(l->done, l->st));
      ILC_HIP(hipStreamWaitEvent(s, l->done, 0));
    }
    ILC_HIP(hipMemcpyAsync(b->desc_dev.p, b->desc_host.p, sizeof(view_desc) * nv,
                           hipMemcpyHostToDevice, s));
    ILC_HIP(hipMemcpyAsync(b->items_dev.p, b->items_host.p, sizeof(uint2) * b->nitems,
                           hipMemcpyHostToDevice, s));
    launch_order_items(b->desc_dev.p, b->items_dev.p, b->segs, b->nitems, b->ts, b->item_keys.p,
                       b->item_idx.p, b->item_temp.p, b->items_ord.p, s);
    if (timed) ILC_HIP(hipEventRecord(timing_event(b), s));
    launch_gather_batch(b->desc_dev.p, b->items_ord.p, b->segs, b->qcounters.p, op, b->kernel,
                        b->ts, b->cubic.p, b->grid_blocks, s);
    if (timed) ILC_HIP(hipEventRecord(timing_event(b), s));
    // the gather wrote every view's accumulator and image
  }
  ILC_HIP(hipMemcpyAsync(b->flags_host.p, b->counter.p, sizeof(uint32_t) * 2,
                         hipMemcpyDeviceToHost, s));
}

// sum of the timed gather launches (pairs of events) of the last eager render
void collect_timing(ilc_batch* b) {
  float tot = 0.f;
  for (int k = 0; k + 1 < b->tev_used; k += 2) {
    float ms = 0.f;
    ILC_HIP(hipEventElapsedTime(&ms, b->tev[k], b->tev[k + 1]));
    tot += ms;
  }
  b->last_gather_ms = tot;
}

void check_flags(ilc_batch* b) {
  if (b->flags_host.p[1])
    fail(ILC_ERANGE, "%u particles fall outside their view's grid", b->flags_host.p[1]);
}

}  // namespace

extern "C" {

ilc_status ilc_batch_create(const ilc_psg_config* cfg, ilc_batch** out) {
  return guarded([&] {
    ILC_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    ilc_psg_config c{};
    if (cfg) c = *cfg;
    ILC_REQUIRE(c.mode == ILC_MODE_REFERENCE || c.mode == ILC_MODE_EXACT, "bad mode %d", c.mode);
    ILC_REQUIRE(c.kernel >= ILC_KERNEL_QUARTIC_REF && c.kernel <= ILC_KERNEL_CUBIC,
                "bad kernel %d", c.kernel);
    ILC_REQUIRE(!(c.mode == ILC_MODE_REFERENCE && c.kernel == ILC_KERNEL_CUBIC),
                "reference mode uses a quartic kernel");
    int ndev = 0;
    ILC_HIP(hipGetDeviceCount(&ndev));
    ILC_REQUIRE(c.device >= 0 && c.device < ndev, "device %d not present", c.device);
    std::unique_ptr<ilc_batch> b(new ilc_batch());
    b->device = c.device;
    b->target = c.target_particles_per_cell ? c.target_particles_per_cell : 1000;
    b->mode = c.mode;
    b->kernel = c.kernel;
    if (c.tile_w || c.tile_h) {
      ILC_REQUIRE(c.tile_w > 0 && c.tile_h > 0 && (c.tile_h * c.tile_w) % kBlock == 0,
                  "tile %dx%d is not a multiple of %d points", c.tile_w, c.tile_h, kBlock);
      b->ts = tile_shape{c.tile_w, c.tile_w * c.tile_h / kBlock};
    }
    device_guard dg(b->device);
    ILC_HIP(hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking));
    ILC_HIP(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
    ILC_HIP(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
    *out = b.release();
  });
}

ilc_status ilc_batch_destroy(ilc_batch* b) {
  return guarded([&] {
    if (!b) return;
    device_guard dg(b->device);
    (void)hipStreamSynchronize(b->own);
    delete b;
  });
}

ilc_status ilc_batch_set_particles(ilc_batch* b, const float* pos3_dev, const float* mass_dev,
                                   const float* hsml_dev, const float* rho_dev,
                                   const float* temp_dev, const float* xe_dev, uint64_t n) {
  return guarded([&] {
    ILC_REQUIRE(b != nullptr, "null batch");
    ILC_REQUIRE(n > 0 && n < (uint64_t(1) << 31), "particle count %llu not in [1, 2^31)",
                (unsigned long long)n);
    ILC_REQUIRE(pos3_dev && hsml_dev, "positions and smoothing lengths are required");
    b->ps = particle_soa{pos3_dev, mass_dev, hsml_dev, rho_dev, temp_dev, xe_dev, n};
    b->have_particles = true;
    b->graph_valid = false;
    b->planned = false;
    // a new snapshot may need larger buffers: the next graph render first
    // runs eagerly, so that no allocation happens inside a capture
    b->ran_eagerly = false;
  });
}

ilc_status ilc_batch_set_gaunt_table(ilc_batch* b, const float* table_host, int32_t ntab,
                                     float x0, float dx) {
  return guarded([&] {
    ILC_REQUIRE(b && table_host, "null argument");
    ILC_REQUIRE(ntab >= 1 && ntab <= 4096, "table size %d not in [1, 4096]", ntab);
    ILC_REQUIRE(dx > 0.0f, "table spacing must be positive");
    device_guard dg(b->device);
    b->gaunt.ensure(size_t(ntab));
    ILC_HIP(hipMemcpy(b->gaunt.p, table_host, sizeof(float) * ntab, hipMemcpyHostToDevice));
    b->gaunt_n = ntab;
    b->gx0 = x0;
    b->gdx = dx;
    b->graph_valid = false;
  });
}

ilc_status ilc_batch_set_views(ilc_batch* b, const ilc_view* views, int32_t nviews,
                               void* stream) {
  return guarded([&] {
    ILC_REQUIRE(b && views && nviews > 0, "null or empty view list");
    device_guard dg(b->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::vector<std::unique_ptr<view_state>> vs;
    for (int32_t k = 0; k < nviews; ++k) {
      const ilc_view& v = views[k];
      ILC_REQUIRE(v.quantity == ILC_QUANTITY_MASS || v.quantity == ILC_QUANTITY_XRAY,
                  "view %d: bad quantity %d", k, v.quantity);
      std::unique_ptr<view_state> st(new view_state());
      st->v = v;
      st->f = make_camera_frame(v.cam);
      const uint64_t P = uint64_t(v.cam.nx) * v.cam.ny;
      ILC_REQUIRE(P < (uint64_t(1) << 32), "view %d: too many pixels", k);
      st->pts.ensure(P);
      st->acc.ensure(P);
      st->img.ensure(P);
      launch_pixel_points(st->f, v.cam.nx, v.cam.ny, st->pts.p, s);
      vs.push_back(std::move(st));
    }
    b->views = std::move(vs);
    b->planned = false;
    b->graph_valid = false;
    b->ran_eagerly = false;
    ILC_HIP(hipStreamSynchronize(s));
  });
}

ilc_status ilc_batch_render(ilc_batch* b, int32_t flags, void* stream) {
  return guarded([&] {
    require_ready(b);
    device_guard dg(b->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const bool persistent = (flags & ILC_RENDER_PLAIN) == 0;
    const bool use_graph = persistent && (flags & ILC_RENDER_GRAPH);
    for (auto& v : b->views)
      if (v->v.quantity == ILC_QUANTITY_XRAY && !b->gaunt_n)
        fail(ILC_ENOTREADY, "X-ray view without a Gaunt table");
    ensure_cubic(b);
    plan(b, s);
    // A plain render points every view at its own grid; a graph replay
    // re-runs no host code, so the next graph render captures afresh and the
    // views' grids (read by backward and view_info) are the shared ones again.
    if (!persistent) b->graph_valid = false;
    if (!use_graph || !b->ran_eagerly) {
      // eager (also sizes every buffer before a later capture)
      device_work(b, persistent, s, true);
      ILC_HIP(hipStreamSynchronize(s));
      collect_timing(b);
      if (persistent) b->ran_eagerly = true;
      check_flags(b);
      return;
    }
    if (!b->graph_valid) {
      if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->gexec = nullptr;
      b->graph = nullptr;
      ILC_HIP(hipStreamBeginCapture(b->own, hipStreamCaptureModeThreadLocal));
      try {
        device_work(b, true, b->own);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->own, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      ILC_HIP(hipStreamEndCapture(b->own, &b->graph));
      ILC_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->graph_valid = true;
    }
    ILC_HIP(hipEventRecord(b->ev_in, s));
    ILC_HIP(hipStreamWaitEvent(b->own, b->ev_in, 0));
    ILC_HIP(hipGraphLaunch(b->gexec, b->own));
    ILC_HIP(hipEventRecord(b->ev_out, b->own));
    ILC_HIP(hipStreamWaitEvent(s, b->ev_out, 0));
    ILC_HIP(hipStreamSynchronize(s));
    check_flags(b);
  });
}

ilc_status ilc_batch_view_info(const ilc_batch* b, int32_t view, ilc_grid_info* out) {
  return guarded([&] {
    ILC_REQUIRE(b && out, "null argument");
    ILC_REQUIRE(view >= 0 && size_t(view) < b->views.size(), "view %d out of range", view);
    const view_state& vs = *b->views[view];
    if (!vs.gready) fail(ILC_ENOTREADY, "view %d has no grid (render first)", view);
    for (int a = 0; a < 3; ++a) {
      out->num_cells[a] = vs.gv.nc[a];
      out->min_corner[a] = vs.gv.minc[a];
      out->cell_size[a] = vs.gv.cs[a];
    }
    out->max_h = vs.gv.max_h;
    out->num_particles = vs.gn;
  });
}

ilc_status ilc_batch_image(const ilc_batch* b, int32_t view, float* out_dev, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(b && out_dev, "null argument");
    ILC_REQUIRE(view >= 0 && size_t(view) < b->views.size(), "view %d out of range", view);
    device_guard dg(b->device);
    const view_state& vs = *b->views[view];
    ILC_HIP(hipMemcpyAsync(out_dev, vs.img.p, sizeof(float) * vs.v.cam.nx * vs.v.cam.ny,
                           hipMemcpyDeviceToDevice, reinterpret_cast<hipStream_t>(stream)));
  });
}

ilc_status ilc_batch_image_f64(const ilc_batch* b, int32_t view, double* out_dev, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(b && out_dev, "null argument");
    ILC_REQUIRE(view >= 0 && size_t(view) < b->views.size(), "view %d out of range", view);
    device_guard dg(b->device);
    const view_state& vs = *b->views[view];
    ILC_HIP(hipMemcpyAsync(out_dev, vs.acc.p, sizeof(double) * vs.v.cam.nx * vs.v.cam.ny,
                           hipMemcpyDeviceToDevice, reinterpret_cast<hipStream_t>(stream)));
  });
}

ilc_status ilc_batch_count_evals(ilc_batch* b, uint64_t* per_view_host, void* stream) {
  return guarded([&] {
    require_ready(b);
    ILC_REQUIRE(per_view_host, "null output");
    device_guard dg(b->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!b->ran_eagerly) fail(ILC_ENOTREADY, "render once before counting");
    dbuf<unsigned long long> cnt, tmp;
    tmp.ensure(1);
    for (size_t v = 0; v < b->views.size(); ++v) {
      view_state& vs = *b->views[v];
      const uint64_t P = uint64_t(vs.v.cam.nx) * vs.v.cam.ny;
      cnt.ensure(P);
      launch_fill_u32(reinterpret_cast<uint32_t*>(cnt.p), 0u, 2 * P, s);
      point_view p{vs.pts.p, static_cast<uint32_t>(P), vs.v.cam.nx, vs.v.cam.ny};
      launch_gather(grid_of(b, vs), p, OP_COUNT, b->kernel, b->ts, nullptr, cnt.p, nullptr, s);
      per_view_host[v] = launch_sum_u64(cnt.p, P, tmp.p, s);
    }
  });
}

ilc_status ilc_batch_backward(ilc_batch* b, const float* const* dl_dimage_dev, float* dl_dw_dev,
                              float* dl_dh_dev, float* dl_dpos3_dev, void* stream) {
  return guarded([&] {
    require_ready(b);
    ILC_REQUIRE(dl_dimage_dev && dl_dw_dev && dl_dh_dev && dl_dpos3_dev, "null argument");
    if (b->mode != ILC_MODE_EXACT)
      fail(ILC_EINVAL, "backward needs exact mode (the reference-mode search box is not smooth)");
    if (b->clip) fail(ILC_EINVAL, "backward of the line-of-sight clip is not provided");
    if (!b->planned) fail(ILC_ENOTREADY, "render before backward");
    device_guard dg(b->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint64_t n = b->ps.n;
    launch_fill_u32(reinterpret_cast<uint32_t*>(dl_dw_dev), 0u, n, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(dl_dh_dev), 0u, n, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(dl_dpos3_dev), 0u, 3 * n, s);
    ensure_cubic(b);
    if (b->kernel == ILC_KERNEL_CUBIC && !b->dcubic_ready) {
      const std::vector<float> t = cubic_projected_dtable(4096);
      b->dcubic.ensure(t.size());
      ILC_HIP(hipMemcpy(b->dcubic.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
      b->dcubic_ready = true;
    }
    // per-particle gradients in the grids' sorted order: one array of each
    // component spanning every view's range (the adjoint indexes a view's
    // particles by their sorted position)
    uint64_t span = 0;
    for (size_t v = 0; v < b->views.size(); ++v) {
      const view_state& vs = *b->views[v];
      ILC_REQUIRE(dl_dimage_dev[v], "null gradient image for view %zu", v);
      if (!vs.gready) fail(ILC_ENOTREADY, "view %zu has no grid (render first)", v);
      span = std::max(span, vs.gbase + vs.gn);
    }
    dbuf<float> g4;
    g4.ensure(4 * std::max<uint64_t>(span, 1));
    float* gw = g4.p;
    float* gh = g4.p + span;
    float* gx = g4.p + 2 * span;
    float* gy = g4.p + 3 * span;
    for (size_t v = 0; v < b->views.size(); ++v) {
      view_state& vs = *b->views[v];
      const uint64_t m = vs.gn, o = vs.gbase;
      if (!m) continue;
      for (float* c : {gw, gh, gx, gy})
        launch_fill_u32(reinterpret_cast<uint32_t*>(c + o), 0u, m, s);
      const uint64_t P = uint64_t(vs.v.cam.nx) * vs.v.cam.ny;
      const point_view p{vs.pts.p, static_cast<uint32_t>(P), vs.v.cam.nx, vs.v.cam.ny};
      launch_adjoint(grid_of(b, vs), p, b->kernel, dl_dimage_dev[v], vs.v.scale,
                     kernel_norm(b->kernel), b->cubic.p, b->dcubic.p, gw, gh, gx, gy, s);
      launch_scatter_grads(vs.gperm + o, vs.gorig, static_cast<uint32_t>(m), gw + o, gh + o,
                           gx + o, gy + o, b->ps.pos3, proj_of(b, vs), dl_dw_dev, dl_dh_dev,
                           dl_dpos3_dev, s);
    }
    ILC_HIP(hipStreamSynchronize(s));
  });
}

ilc_status ilc_batch_last_gather_ms(const ilc_batch* b, float* ms) {
  return guarded([&] {
    ILC_REQUIRE(b && ms, "null argument");
    *ms = b->last_gather_ms;
  });
}

ilc_status ilc_batch_set_depth_clip(ilc_batch* b, int32_t clip) {
  return guarded([&] {
    ILC_REQUIRE(b, "null batch");
    if (clip) {
      ILC_REQUIRE(b->mode == ILC_MODE_EXACT, "the line-of-sight clip needs exact mode");
      ILC_REQUIRE(b->kernel != ILC_KERNEL_CUBIC, "the line-of-sight clip needs a quartic kernel");
    }
    b->clip = clip ? 1 : 0;
    b->graph_valid = false;
    b->planned = false;
    b->ran_eagerly = false;
  });
}

ilc_status ilc_batch_set_tile(ilc_batch* b, int32_t tile_w, int32_t tile_h) {
  return ilc_batch_set_tile_wg(b, tile_w, tile_h, kBlock);
}

ilc_status ilc_batch_set_tile_wg(ilc_batch* b, int32_t tile_w, int32_t tile_h,
                                 int32_t workgroup) {
  return guarded([&] {
    ILC_REQUIRE(b, "null batch");
    ILC_REQUIRE(workgroup == 64 || workgroup == 256, "workgroup of %d threads (64 or 256)",
                workgroup);
    ILC_REQUIRE(tile_w > 0 && tile_h > 0 && (tile_w * tile_h) % workgroup == 0,
                "tile %dx%d is not a multiple of %d points", tile_w, tile_h, workgroup);
    tile_shape ts{tile_w, tile_w * tile_h / workgroup, workgroup};
    (void)tile_grid;  // validated on use
    b->ts = ts;
    b->graph_valid = false;
    b->ran_eagerly = false;
  });
}

ilc_status ilc_batch_get_tile(const ilc_batch* b, int32_t* tile_w, int32_t* tile_h,
                              int32_t* workgroup) {
  return guarded([&] {
    ILC_REQUIRE(b, "null batch");
    if (tile_w) *tile_w = b->ts.tw;
    if (tile_h) *tile_h = (b->ts.nb / b->ts.tw) * b->ts.ppt;
    if (workgroup) *workgroup = b->ts.nb;
  });
}

ilc_status ilc_batch_autotune(ilc_batch* b, int32_t* tile_w, int32_t* tile_h, double* best_ms,
                              void* stream) {
  return guarded([&] {
    require_ready(b);
    device_guard dg(b->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // one tile per workgroup of 256 threads, and 16x16 tiles of one wave
    static const tile_shape cand[] = {{16, 1, 256}, {32, 2, 256}, {32, 4, 256}, {64, 4, 256},
                                      {16, 4, 64},  {8, 1, 64}};
    ensure_cubic(b);  // the cubic gather reads its table from the first launch
    struct events {
      hipEvent_t e[2] = {nullptr, nullptr};
      ~events() {
        for (hipEvent_t x : e)
          if (x) (void)hipEventDestroy(x);
      }
    } ev;
    ILC_HIP(hipEventCreate(&ev.e[0]));
    ILC_HIP(hipEventCreate(&ev.e[1]));
    // on an error the batch keeps its previous tile and renders eagerly next
    struct restore {
      ilc_batch* b;
      tile_shape ts;
      bool armed = true;
      ~restore() {
        if (!armed) return;
        b->ts = ts;
        b->graph_valid = false;
        b->ran_eagerly = false;
      }
    } rs{b, b->ts};
    tile_shape best = b->ts;
    float best_t = INFINITY;
    for (const tile_shape& ts : cand) {
      b->ts = ts;
      b->graph_valid = false;
      plan(b, s);
      device_work(b, true, s);  // warm (allocations, caches)
      ILC_HIP(hipEventRecord(ev.e[0], s));
      device_work(b, true, s);
      ILC_HIP(hipEventRecord(ev.e[1], s));
      ILC_HIP(hipEventSynchronize(ev.e[1]));
      float ms = 0.f;
      ILC_HIP(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
      if (ms < best_t) {
        best_t = ms;
        best = ts;
      }
    }
    rs.armed = false;
    b->ts = best;
    b->graph_valid = false;
    b->ran_eagerly = false;
    if (tile_w) *tile_w = best.tw;
    if (tile_h) *tile_h = (best.nb / best.tw) * best.ppt;
    if (best_ms) *best_ms = best_t;
  });
}

}  // extern "C"
