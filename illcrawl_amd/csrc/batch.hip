// Per-view particle pass of the batched renderer: the quantity functor
// (mass, or bolometric X-ray emissivity from a Gaunt table in LDS), the
// camera projection with depth mask, and either bounding-box partials or
// the projected float4 particles.
//
// Fuses, per view and in one read of the resident snapshot, what the
// reference does in the quantity transformation (quantity.cpp:84-131,
// quantities.cl:97-136, 701-726), projective_smoothing::setup_particles
// (projective_smoothing_reconstruction.cpp:98-114, projection.cl:56-95) and
// the bounding-box scan of the grid constructor (particle_grid.cpp:43-67).
// The FP32 operation order is the one of project.hip, so the projected
// particles are bit-identical to the single-view path's.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace ilc {

namespace {

constexpr int kMaxGaunt = 4096;

__device__ inline float dot3(const float* r, float x, float y, float z) {
#pragma clang fp contract(off)
  const float a = r[0] * x;
  const float b = r[1] * y;
  const float s = a + b;
  const float d = r[2] * z;
  return s + d;
}

__device__ inline float sample_1d(const float* tab, int n, float x0, float dx, float x) {
#pragma clang fp contract(off)
  const float d = x - x0;
  const float u = d / dx;
  const float us = u - 0.5f;
  const float fl = floorf(us);
  const float a = us - fl;
  int i0 = 0;
  if (fl >= 2147483647.0f) i0 = INT32_MAX;
  else if (fl <= -2147483648.0f) i0 = INT32_MIN;
  else if (fl == fl) i0 = int(fl);
  const float t0 = (i0 >= 0 && i0 < n) ? tab[i0] : 0.0f;
  const float t1 = (i0 >= -1 && i0 < n - 1) ? tab[i0 + 1] : 0.0f;
  const float wa = 1.0f - a;
  const float p0 = wa * t0;
  const float p1 = a * t1;
  return p0 + p1;
}

__device__ inline float wmin(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline float wmax(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// w and camera-frame (x, y, z) of particle i
__device__ inline void particle_of(const particle_soa& ps, const view_proj& vp, const float* tab,
                                   int ntab, float gx0, float gdx, uint64_t i, float* out,
                                   float* w) {
#pragma clang fp contract(off)
  const float x = ps.pos3[3 * i], y = ps.pos3[3 * i + 1], z = ps.pos3[3 * i + 2];
  if (vp.quantity == 1) {
    const float T = ps.temp[i];
    float ion = 4.046263f * ps.rho[i];
    ion = ion * ps.xe[i];
    const float g = sample_1d(tab, ntab, gx0, gdx, logf(T));
    float e = ion * ion;
    e = e * sqrtf(T);
    *w = e * g;
  } else {
    *w = static_cast<float>(vp.scale * static_cast<double>(ps.mass[i]));
  }
  const float vx = x - vp.c[0], vy = y - vp.c[1], vz = z - vp.c[2];
  out[0] = dot3(vp.m + 0, vx, vy, vz);
  out[1] = dot3(vp.m + 3, vx, vy, vz);
  out[2] = dot3(vp.m + 6, vx, vy, vz);
}

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_view_pass(particle_soa ps, view_proj vp,
                                                       gaunt_table gt, float4* __restrict__ out,
                                                       float* __restrict__ part,
                                                       uint32_t* __restrict__ bad) {
  __shared__ float tab[kMaxGaunt];
  if (vp.quantity == 1) {
    for (int i = threadIdx.x; i < gt.n; i += kBlock) tab[i] = gt.y[i];
    __syncthreads();
  }
  float mn0 = INFINITY, mn1 = INFINITY, mx0 = -INFINITY, mx1 = -INFINITY, mh = 0.0f;
  uint32_t nbad = 0;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < ps.n;
       i += uint64_t(gridDim.x) * kBlock) {
    float pc[3], w;
    particle_of(ps, vp, tab, gt.n, gt.x0, gt.dx, i, pc, &w);
    if (pc[2] < 0.0f || pc[2] > vp.depth) w = 0.0f;
    if constexpr (WRITE) {
      out[i] = make_float4(pc[0], pc[1], 0.0f, w);
    } else {
      const float hh = ps.hsml[i];
      const bool ok = isfinite(pc[0]) && isfinite(pc[1]) && isfinite(w) && hh >= 1.1e-19f &&
                      hh <= 1.8e19f;
      if (!ok) {
        ++nbad;
        continue;
      }
      mn0 = fminf(mn0, pc[0]);
      mx0 = fmaxf(mx0, pc[0]);
      mn1 = fminf(mn1, pc[1]);
      mx1 = fmaxf(mx1, pc[1]);
      mh = fmaxf(mh, hh);
    }
  }
  if constexpr (!WRITE) {
    __shared__ float red[4][5];
    __shared__ uint32_t redb[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float v[5] = {wmin(mn0), wmin(mn1), wmax(mx0), wmax(mx1), wmax(mh)};
    uint32_t b = nbad;
    for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o, 64);
    if (lane == 0) {
      for (int k = 0; k < 5; ++k) red[wave][k] = v[k];
      redb[wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float r[5];
      for (int k = 0; k < 5; ++k) r[k] = red[0][k];
      uint32_t bb = redb[0];
      for (int w = 1; w < 4; ++w) {
        r[0] = fminf(r[0], red[w][0]);
        r[1] = fminf(r[1], red[w][1]);
        r[2] = fmaxf(r[2], red[w][2]);
        r[3] = fmaxf(r[3], red[w][3]);
        r[4] = fmaxf(r[4], red[w][4]);
        bb += redb[w];
      }
      // layout of the bbox partials: min x, y, z, max x, y, z, max h
      float* o = part + blockIdx.x * 8;
      o[0] = r[0];
      o[1] = r[1];
      o[2] = 0.0f;
      o[3] = r[2];
      o[4] = r[3];
      o[5] = 0.0f;
      o[6] = r[4];
      if (bb) atomicAdd(bad, bb);
    }
  }
}

// ---- every view in one pass
//
// A workgroup holds kMultiItems particles per thread in registers and loops
// over the views; per view it reduces what the single-view kernels reduce
// (bounding box, max h and bad count; or kept counts per window tile), so each
// snapshot read serves every view.

constexpr int kMultiItems = 8;
constexpr int kMultiChunk = kBlock * kMultiItems;

// one_group: every view is a slab of one slab group (the same x', y'
// projection, bit for bit, and the same quantity).  A particle's bounding-box
// terms are then the same for every view, unless its weight is not finite
// inside some slab's depth cut (that slab counts it as bad): a workgroup
// without such a particle projects once and writes the same partial for
// every view; the others take the per-view loop.
__global__ __launch_bounds__(kBlock) void k_multi_bbox(particle_soa ps,
                                                        const view_build_desc* __restrict__ vd,
                                                        int nv, gaunt_table gt,
                                                        float* __restrict__ part,
                                                        uint32_t* __restrict__ bad,
                                                        int one_group) {
#pragma clang fp contract(off)
  __shared__ float tab[kMaxGaunt];
  __shared__ float red[2][4][5];
  __shared__ uint32_t redb[2][4];
  bool any_xray = false, any_mass = false;
  for (int v = 0; v < nv; ++v) {
    any_xray |= vd[v].vp.quantity == 1;
    any_mass |= vd[v].vp.quantity != 1;
  }
  if (any_xray)
    for (int i = threadIdx.x; i < gt.n; i += kBlock) tab[i] = gt.y[i];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nparts = gridDim.x;
  float px[kMultiItems], py[kMultiItems], pz[kMultiItems], hh[kMultiItems];
  float wm[kMultiItems], wx[kMultiItems];
  bool in[kMultiItems];
#pragma unroll
  for (int r = 0; r < kMultiItems; ++r) {
    const uint64_t i = uint64_t(blockIdx.x) * kMultiChunk + uint64_t(r) * kBlock + threadIdx.x;
    in[r] = i < ps.n;
    px[r] = py[r] = pz[r] = hh[r] = wm[r] = wx[r] = 0.0f;
    if (in[r]) {
      px[r] = ps.pos3[3 * i];
      py[r] = ps.pos3[3 * i + 1];
      pz[r] = ps.pos3[3 * i + 2];
      hh[r] = ps.hsml[i];
      // the two quantities of particle_of (views of this pass have scale 1,
      // where float(1.0 * double(m)) is m)
      if (any_mass) wm[r] = ps.mass[i];
      if (any_xray) {
        const float T = ps.temp[i];
        float ion = 4.046263f * ps.rho[i];
        ion = ion * ps.xe[i];
        const float g = sample_1d(tab, gt.n, gt.x0, gt.dx, logf(T));
        float e = ion * ion;
        e = e * sqrtf(T);
        wx[r] = e * g;
      }
    }
  }
  if (one_group) {
    const view_proj vp = vd[0].vp;
    float mn0 = INFINITY, mn1 = INFINITY, mx0 = -INFINITY, mx1 = -INFINITY, mh = 0.0f;
    uint32_t nbad = 0;
    int wbad = 0;
#pragma unroll
    for (int r = 0; r < kMultiItems; ++r) {
      if (!in[r]) continue;
      const float vx = px[r] - vp.c[0], vy = py[r] - vp.c[1], vz = pz[r] - vp.c[2];
      const float ax = dot3(vp.m + 0, vx, vy, vz);
      const float ay = dot3(vp.m + 3, vx, vy, vz);
      const float w = vp.quantity == 1 ? wx[r] : wm[r];
      const bool ok = isfinite(ax) && isfinite(ay) && hh[r] >= 1.1e-19f && hh[r] <= 1.8e19f;
      if (ok && !isfinite(w)) wbad = 1;
      if (!ok) {
        ++nbad;
        continue;
      }
      mn0 = fminf(mn0, ax);
      mx0 = fmaxf(mx0, ax);
      mn1 = fminf(mn1, ay);
      mx1 = fmaxf(mx1, ay);
      mh = fmaxf(mh, hh[r]);
    }
    if (!__syncthreads_or(wbad)) {
      const float rv[5] = {wmin(mn0), wmin(mn1), wmax(mx0), wmax(mx1), wmax(mh)};
      uint32_t bsum = nbad;
      for (int o = 32; o > 0; o >>= 1) bsum += __shfl_xor(bsum, o, 64);
      if (lane == 0) {
        for (int k = 0; k < 5; ++k) red[0][wave][k] = rv[k];
        redb[0][wave] = bsum;
      }
      __syncthreads();
      float q[5];
      for (int k = 0; k < 5; ++k) q[k] = red[0][0][k];
      uint32_t bb = redb[0][0];
      for (int w = 1; w < 4; ++w) {
        q[0] = fminf(q[0], red[0][w][0]);
        q[1] = fminf(q[1], red[0][w][1]);
        q[2] = fmaxf(q[2], red[0][w][2]);
        q[3] = fmaxf(q[3], red[0][w][3]);
        q[4] = fmaxf(q[4], red[0][w][4]);
        bb += redb[0][w];
      }
      for (int v = threadIdx.x; v < nv; v += kBlock) {
        float* o = part + (uint64_t(v) * nparts + blockIdx.x) * 8;
        o[0] = q[0];
        o[1] = q[1];
        o[2] = 0.0f;
        o[3] = q[2];
        o[4] = q[3];
        o[5] = 0.0f;
        o[6] = q[4];
        if (bb) atomicAdd(bad + v, bb);
      }
      return;
    }
  }
  // a weight's finiteness decides a particle's validity only inside a view's
  // depth range: with every weight of the tile finite, the depth coordinate
  // is not needed
  int wnan = 0;
#pragma unroll
  for (int r = 0; r < kMultiItems; ++r)
    wnan |= (in[r] && ((any_mass && !isfinite(wm[r])) || (any_xray && !isfinite(wx[r])))) ? 1 : 0;
  const bool wfinite = !__syncthreads_or(wnan);
  for (int v = 0; v < nv; ++v) {
    const view_proj vp = vd[v].vp;
    float mn0 = INFINITY, mn1 = INFINITY, mx0 = -INFINITY, mx1 = -INFINITY, mh = 0.0f;
    uint32_t nbad = 0;
#pragma unroll
    for (int r = 0; r < kMultiItems; ++r) {
      if (!in[r]) continue;
      const float vx = px[r] - vp.c[0], vy = py[r] - vp.c[1], vz = pz[r] - vp.c[2];
      const float ax = dot3(vp.m + 0, vx, vy, vz);
      const float ay = dot3(vp.m + 3, vx, vy, vz);
      bool wok = true;
      if (!wfinite) {
        const float az = dot3(vp.m + 6, vx, vy, vz);
        float w = vp.quantity == 1 ? wx[r] : wm[r];
        if (az < 0.0f || az > vp.depth) w = 0.0f;
        wok = isfinite(w);
      }
      const bool ok = isfinite(ax) && isfinite(ay) && wok && hh[r] >= 1.1e-19f &&
                      hh[r] <= 1.8e19f;
      if (!ok) {
        ++nbad;
        continue;
      }
      mn0 = fminf(mn0, ax);
      mx0 = fmaxf(mx0, ax);
      mn1 = fminf(mn1, ay);
      mx1 = fmaxf(mx1, ay);
      mh = fmaxf(mh, hh[r]);
    }
    const float rv[5] = {wmin(mn0), wmin(mn1), wmax(mx0), wmax(mx1), wmax(mh)};
    uint32_t bsum = nbad;
    for (int o = 32; o > 0; o >>= 1) bsum += __shfl_xor(bsum, o, 64);
    const int slot = v & 1;
    if (lane == 0) {
      for (int k = 0; k < 5; ++k) red[slot][wave][k] = rv[k];
      redb[slot][wave] = bsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float q[5];
      for (int k = 0; k < 5; ++k) q[k] = red[slot][0][k];
      uint32_t bb = redb[slot][0];
      for (int w = 1; w < 4; ++w) {
        q[0] = fminf(q[0], red[slot][w][0]);
        q[1] = fminf(q[1], red[slot][w][1]);
        q[2] = fmaxf(q[2], red[slot][w][2]);
        q[3] = fmaxf(q[3], red[slot][w][3]);
        q[4] = fmaxf(q[4], red[slot][w][4]);
        bb += redb[slot][w];
      }
      float* o = part + (uint64_t(v) * nparts + blockIdx.x) * 8;
      o[0] = q[0];
      o[1] = q[1];
      o[2] = 0.0f;
      o[3] = q[2];
      o[4] = q[3];
      o[5] = 0.0f;
      o[6] = q[4];
      if (bb) atomicAdd(bad + v, bb);
    }
  }
}

// one workgroup per view: the reduction of k_bbox_final
__global__ __launch_bounds__(kBlock) void k_multi_bbox_final(const float* __restrict__ part,
                                                              int nparts,
                                                              const uint32_t* __restrict__ bad,
                                                              bbox_result* __restrict__ out) {
  __shared__ float red[4][7];
  const int v = blockIdx.x;
  const float* pv = part + uint64_t(v) * nparts * 8;
  float r[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], pv[i * 8 + k]);
    for (int k = 3; k < 7; ++k) r[k] = fmaxf(r[k], pv[i * 8 + k]);
  }
  for (int k = 0; k < 3; ++k) r[k] = wmin(r[k]);
  for (int k = 3; k < 7; ++k) r[k] = wmax(r[k]);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int k = 0; k < 7; ++k) red[wave][k] = r[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], red[w][k]);
      for (int k = 3; k < 7; ++k) r[k] = fmaxf(r[k], red[w][k]);
    }
    bbox_result& o = out[v];
    for (int k = 0; k < 3; ++k) {
      o.mn[k] = r[k];
      o.mx[k] = r[3 + k];
    }
    o.max_h = r[6];
    o.bad = bad[v];
  }
}

// ---- filtered grid input
//
// Window: only the cells that some point of the view can reach matter to its
// image (the gather never visits the others), so the batch can build a view's
// grid from the particles whose projection falls in a window around the
// screen: the screen grown by more than the largest reach plus whole cells,
// so that every cell a point can reach keeps all of its particles.  Kept
// particles stay in input order (order-preserving compaction), so the
// within-cell order of the reachable cells, their counts and max h, and hence
// the images, are the ones of the full grid.
//
// Depth: views that project every particle to the same (x', y') (slabs of one
// stack) share one full grid for the per-cell max h and the bounding box, and
// each keeps for its particle lists only the particles inside its depth cut
// (the others have w = 0 and add nothing).

// as grid.hip's (grid3d_get_cell_indices)
__device__ inline int32_t floor_to_int(float f) {
  const float fl = floorf(f);
  if (!(fl == fl)) return 0;
  if (fl >= 2147483647.0f) return INT32_MAX;
  if (fl <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(fl);
}

constexpr int kWinItems = 8;                    // particles per thread
constexpr int kWinTile = kBlock * kWinItems;    // particles per workgroup

__device__ inline bool in_window(const float* pos3, const view_proj& vp, const window_f& win,
                                 uint64_t i, float* px, float* py) {
#pragma clang fp contract(off)
  const float x = pos3[3 * i], y = pos3[3 * i + 1], z = pos3[3 * i + 2];
  const float vx = x - vp.c[0], vy = y - vp.c[1], vz = z - vp.c[2];
  *px = dot3(vp.m + 0, vx, vy, vz);
  *py = dot3(vp.m + 3, vx, vy, vz);
  bool keep = true;
  if (win.use_window)
    keep = *px >= win.x0 && *px <= win.x1 && *py >= win.y0 && *py <= win.y1;
  if (win.use_depth) {
    const float pz = dot3(vp.m + 6, vx, vy, vz);
    keep = keep && !(pz < 0.0f || pz > vp.depth);
  }
  return keep;
}

// the depth cut of a row at camera depth az: the reference's centre mask
// (projection.cl:80-91), or with the line-of-sight clip, its support
// [az - h, az + h] overlapping [0, depth]
__device__ inline bool depth_keep(float az, float h, float depth, int clip) {
  return clip ? !(az + h < 0.0f || az - h > depth) : !(az < 0.0f || az > depth);
}

// the support disk (radius^2 = cull2 h^2) of a particle at (ax, ay) meets
// the screen rectangle: the gather's tile cull with the screen for the tile
// box (every tile box lies inside the rectangle, so a particle any tile keeps
// passes here)
__device__ inline bool reaches_screen(const window_f& win, float ax, float ay, float ih2) {
  const float dx = fmaxf(fmaxf(win.sx0 - ax, ax - win.sx1), 0.0f);
  const float dy = fmaxf(fmaxf(win.sy0 - ay, ay - win.sy1), 0.0f);
  const float d2 = fmaf(dy, dy, dx * dx);
  return d2 * ih2 < win.cull2;
}

// views per workgroup barrier in the multi-view passes: a tile's per-wave
// counts of kViewChunk views are exchanged at once
constexpr int kViewChunk = 16;

// The slabs of a stack that can keep a particle at depth coordinate
// t = m8 * z: k, the last slab with s_k <= t, and k - 1 (-1: none).  Each is
// then tested with its view's own arithmetic.  ss: the s_k, n of them.
__device__ inline int last_slab_below(const float* ss, int n, float t) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ss[mid] <= t)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

// What a slab's centre mask reads of its view: the depth row, the centre and
// the depth (kept in LDS by the stack passes).
struct slab_par {
  float m[3], c[3], depth;
};

__device__ inline slab_par slab_par_of(const view_proj& vp) {
  return slab_par{{vp.m[6], vp.m[7], vp.m[8]}, {vp.c[0], vp.c[1], vp.c[2]}, vp.depth};
}

// the centre mask of a slab at (x, y, z), in the arithmetic of the per-view
// passes
__device__ inline bool slab_keeps(const slab_par& sp, float x, float y, float z, float h) {
#pragma clang fp contract(off)
  const float vx = x - sp.c[0], vy = y - sp.c[1], vz = z - sp.c[2];
  return depth_keep(dot3(sp.m, vx, vy, vz), h, sp.depth, 0);
}

// every listed view's kept rows per tile, with and without the screen test,
// from one read of the tile (k_count_totals sums them per view)
__global__ __launch_bounds__(kBlock) void k_multi_window_count(
    const float* __restrict__ pos3, const float* __restrict__ hsml, uint64_t n,
    const view_build_desc* __restrict__ vd, const int32_t* __restrict__ views, int nviews,
    uint32_t ntiles, uint32_t* __restrict__ counts, uint32_t* __restrict__ counts_plain,
    slab_stack stack) {
#pragma clang fp contract(off)
  __shared__ uint32_t wc[2][kViewChunk][8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = uint64_t(blockIdx.x) * kWinTile;
  float px[kWinItems], py[kWinItems], pz[kWinItems], ih2[kWinItems], hr[kWinItems];
  bool in[kWinItems];
  int bad = 0;
#pragma unroll
  for (int r = 0; r < kWinItems; ++r) {
    const uint64_t i = base + uint64_t(r) * kBlock + threadIdx.x;
    in[r] = i < n;
    px[r] = in[r] ? pos3[3 * i] : 0.0f;
    py[r] = in[r] ? pos3[3 * i + 1] : 0.0f;
    pz[r] = in[r] ? pos3[3 * i + 2] : 0.0f;
    const float h = in[r] ? hsml[i] : 1.0f;
    hr[r] = h;
    ih2[r] = 1.0f / (h * h);  // as the grid's SoA holds it
    bad |= (isfinite(px[r]) && isfinite(py[r]) && isfinite(pz[r])) ? 0 : 1;
  }
  if (stack.n > 0) {
    // a slab stack: two slabs to test per particle.  A tile with a
    // non-finite coordinate takes the per-view loop (a NaN depth passes
    // every centre mask there).
    __shared__ float ss[kStackMax];
    __shared__ int32_t sv[kStackMax];
    __shared__ uint32_t sc[kStackMax];
    __shared__ slab_par sp[kStackMax];
    if (int(threadIdx.x) < stack.n) {
      ss[threadIdx.x] = stack.s[threadIdx.x];
      sv[threadIdx.x] = stack.view[threadIdx.x];
      sc[threadIdx.x] = 0;
      sp[threadIdx.x] = slab_par_of(vd[stack.view[threadIdx.x]].vp);
    }
    if (!__syncthreads_or(bad)) {
#pragma unroll
      for (int r = 0; r < kWinItems; ++r) {
        if (!in[r]) continue;
        const int k = last_slab_below(ss, stack.n, stack.m8 * pz[r]);
        for (int u = max(k - 1, 0); u <= k; ++u)
          if (slab_keeps(sp[u], px[r], py[r], pz[r], hr[r])) atomicAdd(&sc[u], 1u);
      }
      __syncthreads();
      if (int(threadIdx.x) < stack.n) {
        const uint64_t o = uint64_t(sv[threadIdx.x]) * ntiles + blockIdx.x;
        counts[o] = sc[threadIdx.x];
        counts_plain[o] = sc[threadIdx.x];
      }
      return;
    }
  }
  int slot = 0;
  for (int k0 = 0; k0 < nviews; k0 += kViewChunk, slot ^= 1) {
    const int nk = min(kViewChunk, nviews - k0);
    for (int kk = 0; kk < nk; ++kk) {
      const int v = views[k0 + kk];
      const view_proj vp = vd[v].vp;
      const window_f win = vd[v].win;
      uint32_t c = 0, cp = 0;
      if (!win.use_window && !win.use_depth) {
#pragma unroll
        for (int r = 0; r < kWinItems; ++r) c += __popcll(__ballot(in[r]));
        cp = c;
      } else {
#pragma unroll
        for (int r = 0; r < kWinItems; ++r) {
          const float vx = px[r] - vp.c[0], vy = py[r] - vp.c[1], vz = pz[r] - vp.c[2];
          bool keep = in[r];
          float ax = 0.0f, ay = 0.0f;
          if (win.use_window) {
            ax = dot3(vp.m + 0, vx, vy, vz);
            ay = dot3(vp.m + 3, vx, vy, vz);
            keep = keep && ax >= win.x0 && ax <= win.x1 && ay >= win.y0 && ay <= win.y1;
          }
          if (win.use_depth) {
            const float az = dot3(vp.m + 6, vx, vy, vz);
            keep = keep && depth_keep(az, hr[r], vp.depth, win.clip);
          }
          cp += __popcll(__ballot(keep));
          if (win.use_screen) keep = keep && reaches_screen(win, ax, ay, ih2[r]);
          c += __popcll(__ballot(keep));
        }
      }
      if (lane == 0) {
        wc[slot][kk][wave] = c;
        wc[slot][kk][4 + wave] = cp;
      }
    }
    // chunk j writes wc[j & 1]; its readers are done before barrier j + 1,
    // which every writer of chunk j + 2 passes first
    __syncthreads();
    if (int(threadIdx.x) < nk) {
      const uint32_t* w = wc[slot][threadIdx.x];
      const uint64_t o = uint64_t(views[k0 + int(threadIdx.x)]) * ntiles + blockIdx.x;
      counts[o] = w[0] + w[1] + w[2] + w[3];
      counts_plain[o] = w[4] + w[5] + w[6] + w[7];
    }
  }
}

// totals[v] += the sum of view v's tile counts, one workgroup per listed view
// (per-tile atomics would all land on the few lines of the totals)
__global__ __launch_bounds__(kBlock) void k_count_totals(const int32_t* __restrict__ views,
                                                         uint32_t ntiles,
                                                         const uint32_t* __restrict__ counts,
                                                         uint32_t* __restrict__ totals,
                                                         const uint32_t* __restrict__ counts_plain,
                                                         uint32_t* __restrict__ totals_plain) {
  __shared__ uint32_t red[2][kBlock / 64];
  const int v = views[blockIdx.x];
  const uint32_t* c = counts + uint64_t(v) * ntiles;
  const uint32_t* cp = counts_plain + uint64_t(v) * ntiles;
  uint32_t a = 0, b = 0;
  for (uint32_t t = threadIdx.x; t < ntiles; t += kBlock) {
    a += c[t];
    b += cp[t];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += uint32_t(__shfl_xor(int(a), o));
    b += uint32_t(__shfl_xor(int(b), o));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s0 = 0, s1 = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      s0 += red[0][w];
      s1 += red[1][w];
    }
    // atomic: a view listed twice adds twice, as per-tile adds would
    if (s0) atomicAdd(totals + v, s0);
    if (s1) atomicAdd(totals_plain + v, s1);
  }
}

__global__ __launch_bounds__(kBlock) void k_particle_xray(particle_soa ps, gaunt_table gt,
                                                           float* __restrict__ wx) {
#pragma clang fp contract(off)
  __shared__ float tab[kMaxGaunt];
  for (int i = threadIdx.x; i < gt.n; i += kBlock) tab[i] = gt.y[i];
  __syncthreads();
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < ps.n;
       i += uint64_t(gridDim.x) * kBlock) {
    const float T = ps.temp[i];
    float ion = 4.046263f * ps.rho[i];
    ion = ion * ps.xe[i];
    const float g = sample_1d(tab, gt.n, gt.x0, gt.dx, logf(T));
    float e = ion * ion;
    e = e * sqrtf(T);
    wx[i] = e * g;
  }
}

// descending-h sort key: the top 16 bits of h's FP32 pattern, reversed
__device__ inline uint32_t hq_key(float h) { return 0x7fffu - (__float_as_uint(h) >> 16); }

// k_window_scatter for every listed view from one read of the tile.  For
// kViewChunk views at a time, the kept counts of the kWinItems rows (rows of
// kBlock particles) are taken per wave first (a thread's keep bits in two
// registers), so one barrier gives every row's offset in all of them; each
// view's rows are then written one at a time, skipping those with nothing
// kept, in the order of k_window_scatter (rows, then threads).
__global__ __launch_bounds__(kBlock) void k_multi_window_scatter(
    particle_soa ps, const float* __restrict__ wx, const view_build_desc* __restrict__ vd,
    const int32_t* __restrict__ views, int nviews, uint32_t ntiles,
    const uint32_t* __restrict__ offs, float4* __restrict__ p4c, uint32_t* __restrict__ hqkey,
    uint32_t* __restrict__ hqval, uint32_t* __restrict__ gkey, uint32_t* __restrict__ orig,
    uint32_t* __restrict__ err,
    const uint32_t* __restrict__ cbase, uint32_t* __restrict__ maxh_bits,
    float2* __restrict__ zc, slab_stack stack) {
#pragma clang fp contract(off)
  static_assert(kViewChunk <= 16 && kWinItems == 8, "keep bits: 8 views per 64-bit word");
  // per-wave kept counts: wc[2][kViewChunk][kWinItems][4] for the per-view
  // loop, or [kWinItems][4][kStackMax] for a slab stack
  constexpr int kCountWords = kWinItems * 4 * kStackMax;
  static_assert(2 * kViewChunk * kWinItems * 4 <= kCountWords, "count table");
  __shared__ __align__(16) uint32_t cwords[kCountWords];
  auto& wc = *reinterpret_cast<uint32_t(*)[2][kViewChunk][kWinItems][4]>(cwords);
  // the tile's positions, read once (SoA); h and 1/h^2 in registers
  __shared__ float sx[kWinTile], sy[kWinTile], sz[kWinTile];
  // one view's kept tile positions (the per-view loop)
  __shared__ uint16_t klist[kWinTile];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = uint64_t(blockIdx.x) * kWinTile;
  float hr[kWinItems], ih2[kWinItems];
  int bad = 0;
#pragma unroll
  for (int r = 0; r < kWinItems; ++r) {
    const uint64_t i = base + uint64_t(r) * kBlock + threadIdx.x;
    const bool ok = i < ps.n;
    const float x = ok ? ps.pos3[3 * i] : 0.0f, y = ok ? ps.pos3[3 * i + 1] : 0.0f;
    const float z = ok ? ps.pos3[3 * i + 2] : 0.0f;
    sx[r * kBlock + threadIdx.x] = x;
    sy[r * kBlock + threadIdx.x] = y;
    sz[r * kBlock + threadIdx.x] = z;
    bad |= (isfinite(x) && isfinite(y) && isfinite(z)) ? 0 : 1;
    const float h = ok ? ps.hsml[i] : 1.0f;
    hr[r] = h;
    ih2[r] = 1.0f / (h * h);
  }
  __syncthreads();
  // cell key of (x', y', 0) in a view's grid: grid3d_get_cell_indices,
  // particle_grid.cl:45-53
  auto cell_key = [](const grid_params& g, float ax, float ay, bool* in_grid,
                     int32_t* cxy = nullptr) {
    const float c3[3] = {ax, ay, 0.0f};
    int32_t idx[3];
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
      const float d = c3[a] - g.minc[a];
      idx[a] = floor_to_int(d / g.cs[a]);
      ok = ok && idx[a] >= 0 && idx[a] < g.nc[a];
    }
    *in_grid = ok;
    if (cxy) {
      cxy[0] = idx[0];
      cxy[1] = idx[1];
    }
    return ok ? uint32_t(idx[0]) + uint32_t(idx[1]) * uint32_t(g.nc[0]) +
                    uint32_t(idx[2]) * uint32_t(g.nc[0]) * uint32_t(g.nc[1])
              : 0u;
  };
  // the filter computes only the coordinates it tests (the flags are
  // uniform); the kept path recomputes all three.  A row inside the window
  // and depth cut whose support cannot reach the screen is not kept; it
  // raises its cell's max h, which the search box of the reference reads.
  auto kept = [&](const view_proj& vp, const window_f& win, const grid_params& g, uint32_t cb,
                  float x, float y, float z, float h, float ih2r, bool ok) {
    const float vx = x - vp.c[0], vy = y - vp.c[1], vz = z - vp.c[2];
    bool keep = ok;
    float ax = 0.0f, ay = 0.0f;
    if (win.use_window) {
      ax = dot3(vp.m + 0, vx, vy, vz);
      ay = dot3(vp.m + 3, vx, vy, vz);
      keep = keep && ax >= win.x0 && ax <= win.x1 && ay >= win.y0 && ay <= win.y1;
    }
    if (win.use_depth) {
      const float az = dot3(vp.m + 6, vx, vy, vz);
      keep = keep && depth_keep(az, h, vp.depth, win.clip);
    }
    if (win.use_screen && keep) {
      if (!reaches_screen(win, ax, ay, ih2r)) {
        keep = false;
        bool in_grid;
        int32_t cxy[2];
        const uint32_t key = cell_key(g, ax, ay, &in_grid, cxy);
        if (!in_grid) {
          atomicAdd(err, 1u);
        } else if (cxy[0] >= win.scx0 && cxy[0] <= win.scx1 && cxy[1] >= win.scy0 &&
                   cxy[1] <= win.scy1) {
          uint32_t* m = maxh_bits + cb + key;
          const uint32_t hb = __float_as_uint(h);
          if (hb > __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(m, hb);
        }
      }
    }
    return keep;
  };
  // kept row j of view v: particle i, at tile position t
  auto emit = [&](int v, uint32_t j, int t, uint64_t i) {
    const view_proj vp = vd[v].vp;
    const window_f win = vd[v].win;
    const float vx = sx[t] - vp.c[0], vy = sy[t] - vp.c[1], vz = sz[t] - vp.c[2];
    const float ax = dot3(vp.m + 0, vx, vy, vz);
    const float ay = dot3(vp.m + 3, vx, vy, vz);
    const float az = dot3(vp.m + 6, vx, vy, vz);
    const grid_params& g = vd[v].g;
    float w = vp.quantity == 1 ? wx[i] : static_cast<float>(vp.scale * double(ps.mass[i]));
    // h rides in the z slot, so the SoA pass reads one line per row; the
    // descending-h sort starts from (hq key, row)
    const float hh = ps.hsml[i];
    if (!depth_keep(az, hh, vp.depth, win.clip)) w = 0.0f;
    if (zc) zc[j] = make_float2((0.0f - az) / hh, (vp.depth - az) / hh);
    p4c[j] = make_float4(ax, ay, hh, w);
    hqkey[j] = hq_key(hh);
    hqval[j] = j;
    bool in_grid;
    const uint32_t key = cell_key(g, ax, ay, &in_grid);
    if (!in_grid) atomicAdd(err, 1u);
    gkey[j] = cbase[v] + key;
    orig[j] = uint32_t(i);
  };
  if (stack.n > 0 && !__syncthreads_or(bad)) {
    // A slab stack: each particle's two candidate slabs, tested with their
    // views' arithmetic.  Kept rows of one slab go in the order of the
    // per-view loop (item r, then wave, then lane): per (r, wave) each slab
    // present gets its count and every lane its rank among the wave's lanes
    // that keep the particle for that slab (a lane's two slabs differ).  The
    // first lane of each slab writes the count.  (A tile with a non-finite
    // coordinate takes the per-view loop, as in the count pass.)
    __shared__ float ss[kStackMax];
    __shared__ int32_t sv[kStackMax];
    __shared__ slab_par sp[kStackMax];
    auto& cnt = *reinterpret_cast<uint32_t(*)[kWinItems][4][kStackMax]>(cwords);
    if (int(threadIdx.x) < stack.n) {
      ss[threadIdx.x] = stack.s[threadIdx.x];
      sv[threadIdx.x] = stack.view[threadIdx.x];
      sp[threadIdx.x] = slab_par_of(vd[stack.view[threadIdx.x]].vp);
    }
    for (int e = threadIdx.x; e < kCountWords; e += kBlock) cwords[e] = 0u;
    __syncthreads();
    const uint64_t below = (uint64_t(1) << lane) - 1u;
    // per item: slab + 1 (7 bits) and rank (6 bits) of the two slots
    uint32_t packed[kWinItems];
#pragma unroll
    for (int r = 0; r < kWinItems; ++r) {
      const int t = r * kBlock + threadIdx.x;
      const bool ok = base + uint64_t(t) < ps.n;
      int a = -1, b = -1;
      if (ok) {
        const int k = last_slab_below(ss, stack.n, stack.m8 * sz[t]);
        if (k >= 0 && slab_keeps(sp[k], sx[t], sy[t], sz[t], hr[r])) a = k;
        if (k >= 1 && slab_keeps(sp[k - 1], sx[t], sy[t], sz[t], hr[r])) b = k - 1;
      }
      // multi-split on the slab index: the ballots of its six bits in each
      // slot give, for any slab x, the lanes holding x in that slot
      uint64_t va = __ballot(a >= 0), vb = __ballot(b >= 0);
      uint64_t ba[6], bb[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        ba[q] = __ballot(a >= 0 && ((a >> q) & 1));
        bb[q] = __ballot(b >= 0 && ((b >> q) & 1));
      }
      auto peers_of = [&](int x) {
        uint64_t pa = va, pb = vb;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          pa &= ((x >> q) & 1) ? ba[q] : ~ba[q];
          pb &= ((x >> q) & 1) ? bb[q] : ~bb[q];
        }
        return pa | pb;
      };
      uint32_t ra = 0, rb = 0;
      if (a >= 0) {
        const uint64_t pe = peers_of(a);
        ra = uint32_t(__popcll(pe & below));
        if (ra == 0) cnt[r][wave][a] = uint32_t(__popcll(pe));
      }
      if (b >= 0) {
        const uint64_t pe = peers_of(b);
        rb = uint32_t(__popcll(pe & below));
        if (rb == 0) cnt[r][wave][b] = uint32_t(__popcll(pe));
      }
      packed[r] = uint32_t(a + 1) | (ra << 7) | (uint32_t(b + 1) << 13) | (rb << 20);
    }
    __syncthreads();
    // each slab's first position per (r, wave): its tile offset plus the rows
    // before, in (r, wave) order
    if (int(threadIdx.x) < stack.n) {
      uint32_t run = offs[uint64_t(sv[threadIdx.x]) * ntiles + blockIdx.x];
      for (int r = 0; r < kWinItems; ++r)
        for (int w = 0; w < 4; ++w) {
          const uint32_t c = cnt[r][w][threadIdx.x];
          cnt[r][w][threadIdx.x] = run;
          run += c;
        }
    }
    __syncthreads();
#pragma unroll 1
    for (int r = 0; r < kWinItems; ++r) {
      const int t = r * kBlock + threadIdx.x;
      const uint64_t i = base + uint64_t(t);
      const int a = int(packed[r] & 127u) - 1, b = int((packed[r] >> 13) & 127u) - 1;
      if (a >= 0) emit(sv[a], cnt[r][wave][a] + ((packed[r] >> 7) & 63u), t, i);
      if (b >= 0) emit(sv[b], cnt[r][wave][b] + ((packed[r] >> 20) & 63u), t, i);
    }
    return;
  }
  int slot = 0;
  for (int k0 = 0; k0 < nviews; k0 += kViewChunk, slot ^= 1) {
    const int nk = min(kViewChunk, nviews - k0);
    // this thread's kept rows: view kk of the chunk in byte kk % 8 of
    // bits[kk / 8]
    uint64_t bits_lo = 0, bits_hi = 0;
    for (int kk = 0; kk < nk; ++kk) {
      const int v = views[k0 + kk];
      const view_proj vp = vd[v].vp;
      const window_f win = vd[v].win;
      const uint32_t cb = cbase[v];
      uint32_t rb = 0;
#pragma unroll
      for (int r = 0; r < kWinItems; ++r) {
        const int t = r * kBlock + threadIdx.x;
        const bool ok = base + uint64_t(t) < ps.n;
        const bool keep = kept(vp, win, vd[v].g, cb, sx[t], sy[t], sz[t], hr[r], ih2[r], ok);
        rb |= keep ? 1u << r : 0u;
        const uint32_t c = __popcll(__ballot(keep));
        if (lane == 0) wc[slot][kk][r][wave] = c;
      }
      const uint64_t sh = uint64_t(rb) << (8 * (kk & 7));
      if (kk < 8)
        bits_lo |= sh;
      else
        bits_hi |= sh;
    }
    // chunk j writes wc[j & 1]; every read of it comes before barrier j + 1,
    // which every writer of chunk j + 2 passes first
    __syncthreads();
    for (int kk = 0; kk < nk; ++kk) {
      const int v = views[k0 + kk];
      const uint32_t bits = uint32_t((kk < 8 ? bits_lo : bits_hi) >> (8 * (kk & 7))) & 0xffu;
      // the view's kept tile positions in row order (item r, then wave,
      // then lane), listed in LDS; then every thread emits consecutive rows
      uint32_t total = 0;
#pragma unroll
      for (int r = 0; r < kWinItems; ++r) {
        const uint4 cw = *reinterpret_cast<const uint4*>(wc[slot][kk][r]);
        const uint32_t before = (wave > 0 ? cw.x : 0u) + (wave > 1 ? cw.y : 0u) +
                                (wave > 2 ? cw.z : 0u);
        const bool keep = (bits >> r) & 1u;
        const uint64_t bal = __ballot(keep);
        if (keep)
          klist[total + before +
                __builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32),
                                          __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u))] =
              uint16_t(r * kBlock + threadIdx.x);
        total += cw.x + cw.y + cw.z + cw.w;
      }
      if (total == 0) continue;
      __syncthreads();
      const uint32_t out = offs[uint64_t(v) * ntiles + blockIdx.x];
      for (uint32_t e = threadIdx.x; e < total; e += kBlock) {
        const int t = klist[e];
        emit(v, out + e, t, base + uint64_t(t));
      }
      // the list is rewritten for the next view
      __syncthreads();
    }
  }
}


__global__ __launch_bounds__(kBlock) void k_gather_u32(const uint32_t* __restrict__ src,
                                                       const uint32_t* __restrict__ idx,
                                                       uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

// Work-item order of the persistent gather.  A tile's cost is dominated by
// the particles of the grid cells under its points: a tile over a dense core
// streams and evaluates the most, and started last it runs on alone while
// the rest of the GPU idles.  Each item gets the key queue << 5 | class.  A
// heavy tile (at least 16 times the view's mean particles per tile) has
// class 31 - bit length of its count, every other tile class 31.  A stable
// sort by key keeps every queue's items together, the heavy tiles first and
// heaviest first, and the light majority in its original order (neighbours
// in one queue share their XCD's L2).  The count visits at most 32 rows (a
// stride over taller boxes) and 8 layers: it is an estimate.
__global__ __launch_bounds__(kBlock) void k_item_cost(const view_desc* __restrict__ views,
                                                      const uint2* __restrict__ items,
                                                      queue_segs segs, uint32_t n, int tw, int th,
                                                      uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t q = 0;
  while (q + 1 < uint32_t(kQueues) && i >= segs.seg[q + 1]) ++q;
  const uint2 it = items[i];
  const view_desc& vd = views[it.x];
  const grid_view& g = vd.g;
  uint64_t cost = 0;
  const uint32_t tx = it.y % vd.tiles_x, ty = it.y / vd.tiles_x;
  const uint32_t px0 = tx * uint32_t(tw), py0 = ty * uint32_t(th);
  if (px0 < vd.p.lnx && py0 < vd.p.lny && g.nc[0] > 0 && g.nc[1] > 0 && g.nc[2] > 0) {
    const uint32_t px1 = min(px0 + uint32_t(tw), vd.p.lnx) - 1u;
    const uint32_t py1 = min(py0 + uint32_t(th), vd.p.lny) - 1u;
    const uint64_t ia = uint64_t(py0) * vd.p.lnx + px0, ib = uint64_t(py1) * vd.p.lnx + px1;
    if (ib < vd.p.n) {
      const float4 a = vd.p.pts[ia], b = vd.p.pts[ib];
      const float x0 = fminf(a.x, b.x), x1 = fmaxf(a.x, b.x);
      const float y0 = fminf(a.y, b.y), y1 = fmaxf(a.y, b.y);
      const float fx0 = floorf((x0 - g.minc[0]) / g.cs[0]), fx1 = floorf((x1 - g.minc[0]) / g.cs[0]);
      const float fy0 = floorf((y0 - g.minc[1]) / g.cs[1]), fy1 = floorf((y1 - g.minc[1]) / g.cs[1]);
      // the whole box off the grid (or not a number): no cells
      if (fx1 >= 0.0f && fy1 >= 0.0f && fx0 < float(g.nc[0]) && fy0 < float(g.nc[1])) {
        const int cx0 = int(fmaxf(fx0, 0.0f)), cx1 = int(fminf(fx1, float(g.nc[0] - 1)));
        const int cy0 = int(fmaxf(fy0, 0.0f)), cy1 = int(fminf(fy1, float(g.nc[1] - 1)));
        const int nrows = cy1 - cy0 + 1;
        const int step = (nrows + 31) / 32;
        const int nl = min(g.nc[2], 8);
        const uint32_t nx = uint32_t(g.nc[0]), nxny = nx * uint32_t(g.nc[1]);
        for (int cz = 0; cz < nl; ++cz)
          for (int cy = cy0; cy <= cy1; cy += step) {
            const uint32_t kb = uint32_t(cy) * nx + uint32_t(cz) * nxny;
            cost += uint64_t(g.start[kb + uint32_t(cx1) + 1u] - g.start[kb + uint32_t(cx0)]) *
                    uint64_t(step);
          }
      }
    }
  }
  const uint32_t c32 = cost > 0xffffffffull ? 0xffffffffu : uint32_t(cost);
  const int bits = c32 ? 32 - __clz(c32) : 0;
  const bool heavy = cost >= 16u * uint64_t(vd.cost_unit);
  keys[i] = (q << 5) | (heavy ? uint32_t(31 - min(bits, 31)) : 31u);
  idx[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_permute_items(const uint2* __restrict__ items,
                                                          const uint32_t* __restrict__ idx,
                                                          uint32_t n, uint2* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = items[idx[i]];
}

// a row's view is the last v with cbase[v] <= its global key (every view has
// at least one cell, so the bases rise strictly)
__global__ __launch_bounds__(kBlock) void k_local_keys(const uint32_t* __restrict__ gkey,
                                                       const uint32_t* __restrict__ cbase, int nv,
                                                       uint32_t n, uint32_t* __restrict__ lkey) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = gkey[i];
  int a = 0, b = nv;  // cbase[a] <= k < cbase[b] (cbase[nv] taken as infinite)
  while (b - a > 1) {
    const int m = (a + b) >> 1;
    if (cbase[m] <= k)
      a = m;
    else
      b = m;
  }
  lkey[i] = k - cbase[a];
}

// Kept particles in input order: offsets = exclusive scan of the counts;
// writes the projected float4 (x', y', 0, w with the depth mask), h, the cell
// key (as k_keys computes it) and the compacted index.
__global__ __launch_bounds__(kBlock) void k_window_scatter(
    particle_soa ps, view_proj vp, gaunt_table gt, window_f win,
    const uint32_t* __restrict__ offsets, grid_params g, float4* __restrict__ p4c,
    float* __restrict__ hc, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
    uint32_t* __restrict__ orig, uint32_t* __restrict__ err) {
#pragma clang fp contract(off)
  __shared__ float tab[kMaxGaunt];
  __shared__ uint32_t wc[2][4];
  if (vp.quantity == 1) {
    for (int i = threadIdx.x; i < gt.n; i += kBlock) tab[i] = gt.y[i];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = uint64_t(blockIdx.x) * kWinTile;
  uint32_t out = offsets[blockIdx.x];
  __syncthreads();
  for (int r = 0; r < kWinItems; ++r) {
    const uint64_t i = base + uint64_t(r) * kBlock + threadIdx.x;
    float px = 0.f, py = 0.f;
    const bool keep = i < ps.n && in_window(ps.pos3, vp, win, i, &px, &py);
    const uint64_t bal = __ballot(keep);
    const int slot = r & 1;
    if (lane == 0) wc[slot][wave] = __popcll(bal);
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
      const uint32_t cw = wc[slot][w];
      if (w < wave) before += cw;
      tot += cw;
    }
    if (keep) {
      const uint32_t j = out + before +
                         __builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32),
                                                   __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u));
      float pc[3], w;
      particle_of(ps, vp, tab, gt.n, gt.x0, gt.dx, i, pc, &w);
      if (pc[2] < 0.0f || pc[2] > vp.depth) w = 0.0f;
      p4c[j] = make_float4(pc[0], pc[1], 0.0f, w);
      hc[j] = ps.hsml[i];
      // cell key of (x', y', 0): grid3d_get_cell_indices, particle_grid.cl:45-53
      const float c3[3] = {pc[0], pc[1], 0.0f};
      int32_t idx[3];
      bool ok = true;
      for (int a = 0; a < 3; ++a) {
        const float d = c3[a] - g.minc[a];
        idx[a] = floor_to_int(d / g.cs[a]);
        ok = ok && idx[a] >= 0 && idx[a] < g.nc[a];
      }
      uint32_t key = 0;
      if (ok)
        key = uint32_t(idx[0]) + uint32_t(idx[1]) * uint32_t(g.nc[0]) +
              uint32_t(idx[2]) * uint32_t(g.nc[0]) * uint32_t(g.nc[1]);
      else
        atomicAdd(err, 1u);
      keys[j] = key;
      vals[j] = j;
      orig[j] = uint32_t(i);
    }
    out += tot;
  }
}

// Adjoint of one view of a batch: gradients of the view's grid (sorted
// order, projected frame) added to the snapshot's per-particle gradients:
// dL/dw (0 for particles the view's depth cut removes, whose w does not
// enter the image), dL/dh, and dL/dpos = J^T (dL/dx', dL/dy') with J the
// first two rows of the view's projection.
__global__ __launch_bounds__(kBlock) void k_scatter_grads(
    const uint32_t* __restrict__ perm, const uint32_t* __restrict__ orig, uint32_t n,
    const float* __restrict__ gw, const float* __restrict__ gh, const float* __restrict__ gx,
    const float* __restrict__ gy, const float* __restrict__ pos3, view_proj vp,
    float* __restrict__ out_w, float* __restrict__ out_h, float* __restrict__ out_pos3) {
#pragma clang fp contract(off)
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t c = perm[s];
  const uint32_t i = orig ? orig[c] : c;
  const float x = pos3[3 * uint64_t(i)], y = pos3[3 * uint64_t(i) + 1],
              z = pos3[3 * uint64_t(i) + 2];
  const float pz = dot3(vp.m + 6, x - vp.c[0], y - vp.c[1], z - vp.c[2]);
  const bool kept = !(pz < 0.0f || pz > vp.depth);
  if (kept) atomicAdd(out_w + i, gw[s]);
  atomicAdd(out_h + i, gh[s]);
  const float ax = gx[s], ay = gy[s];
  for (int k = 0; k < 3; ++k) {
    const float v = fmaf(ax, vp.m[k], ay * vp.m[3 + k]);
    if (v != 0.0f) atomicAdd(out_pos3 + 3 * uint64_t(i) + k, v);
  }
}

}  // namespace

void launch_scatter_grads(const uint32_t* perm, const uint32_t* orig, uint32_t n,
                          const float* gw, const float* gh, const float* gx, const float* gy,
                          const float* pos3, const view_proj& vp, float* out_w, float* out_h,
                          float* out_pos3, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_scatter_grads, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, perm,
                     orig, n, gw, gh, gx, gy, pos3, vp, out_w, out_h, out_pos3);
  ILC_HIP(hipGetLastError());
}

uint32_t window_tiles(uint64_t n) { return static_cast<uint32_t>((n + kWinTile - 1) / kWinTile); }

int multi_bbox_parts(uint64_t n) {
  return static_cast<int>(std::max<uint64_t>(1, (n + kMultiChunk - 1) / kMultiChunk));
}

void launch_particle_xray(const particle_soa& ps, const gaunt_table& gt, float* wx,
                          hipStream_t s) {
  if (!ps.n) return;
  if (gt.n < 1 || gt.n > kMaxGaunt || !gt.y)
    fail(ILC_EINVAL, "X-ray view without a Gaunt table (1..%d entries)", kMaxGaunt);
  uint64_t nb = (ps.n + kBlock - 1) / kBlock;
  if (nb > 16384) nb = 16384;
  hipLaunchKernelGGL(k_particle_xray, dim3(uint32_t(nb)), dim3(kBlock), 0, s, ps, gt, wx);
  ILC_HIP(hipGetLastError());
}

void launch_multi_window_scatter(const particle_soa& ps, const float* wx,
                                 const view_build_desc* vd_dev, const int32_t* views_dev,
                                 int nviews, const uint32_t* offs, float4* p4c, uint32_t* hqkey,
                                 uint32_t* hqval,
                                 uint32_t* gkey, uint32_t* orig, uint32_t* err,
                                 const uint32_t* cbase, uint32_t* maxh_bits, float2* zc,
                                 const slab_stack& stack, hipStream_t s) {
  const uint32_t nb = window_tiles(ps.n);
  if (!nb || nviews <= 0) return;
  ILC_REQUIRE(cbase != nullptr, "the scatter needs the views' first cells");
  ILC_REQUIRE(stack.n == 0 || (stack.n == nviews && stack.n <= kStackMax && zc == nullptr),
              "a slab stack lists every view, at most %d, centre masks only", kStackMax);
  hipLaunchKernelGGL(k_multi_window_scatter, dim3(nb), dim3(kBlock), 0, s, ps, wx, vd_dev,
                     views_dev, nviews, nb, offs, p4c, hqkey, hqval, gkey, orig, err, cbase,
                     maxh_bits, zc, stack);
  ILC_HIP(hipGetLastError());
}

void launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* out,
                       hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather_u32, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, src, idx,
                     n, out);
  ILC_HIP(hipGetLastError());
}

void launch_local_keys(const uint32_t* gkey, const uint32_t* cbase, int nv, uint32_t n,
                       uint32_t* lkey, hipStream_t s) {
  if (!n) return;
  ILC_REQUIRE(nv >= 1, "no views");
  hipLaunchKernelGGL(k_local_keys, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, gkey,
                     cbase, nv, n, lkey);
  ILC_HIP(hipGetLastError());
}

void launch_order_items(const view_desc* views, const uint2* items, const queue_segs& segs,
                        uint32_t n, const tile_shape& ts, uint32_t* keys, uint32_t* idx,
                        uint32_t* radix_temp, uint2* out, hipStream_t s) {
  if (!n) return;
  const int th = (kBlock / ts.tw) * ts.ppt;
  hipLaunchKernelGGL(k_item_cost, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, views,
                     items, segs, n, ts.tw, th, keys, idx);
  ILC_HIP(hipGetLastError());
  launch_radix_sort(keys, idx, n, 8, radix_temp, s);
  hipLaunchKernelGGL(k_permute_items, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, items,
                     idx, n, out);
  ILC_HIP(hipGetLastError());
}

void launch_multi_bbox(const particle_soa& ps, const view_build_desc* vd_dev, int nv,
                       const gaunt_table& gt, float* partial, uint32_t* bad, bbox_result* out,
                       bool one_group, hipStream_t s) {
  if (nv <= 0) return;
  if (gt.n > kMaxGaunt || (gt.n > 0 && !gt.y))
    fail(ILC_EINVAL, "Gaunt table of %d entries (at most %d)", gt.n, kMaxGaunt);
  const int nparts = multi_bbox_parts(ps.n);
  launch_fill_u32(bad, 0u, uint64_t(nv), s);
  hipLaunchKernelGGL(k_multi_bbox, dim3(nparts), dim3(kBlock), 0, s, ps, vd_dev, nv, gt, partial,
                     bad, one_group ? 1 : 0);
  ILC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_multi_bbox_final, dim3(nv), dim3(kBlock), 0, s, partial, nparts, bad, out);
  ILC_HIP(hipGetLastError());
}

void launch_multi_window_count(const float* pos3, const float* hsml, uint64_t n,
                               const view_build_desc* vd_dev, const int32_t* views_dev,
                               int nviews, uint32_t* counts, uint32_t* totals,
                               uint32_t* counts_plain, uint32_t* totals_plain,
                               const slab_stack& stack, hipStream_t s) {
  const uint32_t nb = window_tiles(n);
  if (!nb || nviews <= 0) return;
  ILC_REQUIRE(stack.n == 0 || (stack.n == nviews && stack.n <= kStackMax),
              "a slab stack lists every view, at most %d", kStackMax);
  hipLaunchKernelGGL(k_multi_window_count, dim3(nb), dim3(kBlock), 0, s, pos3, hsml, n, vd_dev,
                     views_dev, nviews, nb, counts, counts_plain, stack);
  ILC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_count_totals, dim3(uint32_t(nviews)), dim3(kBlock), 0, s, views_dev, nb,
                     counts, totals, counts_plain, totals_plain);
  ILC_HIP(hipGetLastError());
}

void launch_window_scatter(const particle_soa& ps, const view_proj& vp, const gaunt_table& gt,
                           const window_f& win, const uint32_t* offsets, const grid_params& g,
                           float4* p4c, float* hc, uint32_t* keys, uint32_t* vals, uint32_t* orig,
                           uint32_t* err, hipStream_t s) {
  if (vp.quantity == 1 && (gt.n < 1 || gt.n > kMaxGaunt || !gt.y))
    fail(ILC_EINVAL, "X-ray view without a Gaunt table (1..%d entries)", kMaxGaunt);
  const uint32_t nb = window_tiles(ps.n);
  if (!nb) return;
  hipLaunchKernelGGL(k_window_scatter, dim3(nb), dim3(kBlock), 0, s, ps, vp, gt, win, offsets, g,
                     p4c, hc, keys, vals, orig, err);
  ILC_HIP(hipGetLastError());
}

int view_bbox_parts(uint64_t n) {
  return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(1024, (n + kBlock - 1) / kBlock)));
}

void launch_view_bbox(const particle_soa& ps, const view_proj& vp, const gaunt_table& gt,
                      float* partial, uint32_t* bad, bbox_result* out, hipStream_t s) {
  if (vp.quantity == 1 && (gt.n < 1 || gt.n > kMaxGaunt || !gt.y))
    fail(ILC_EINVAL, "X-ray view without a Gaunt table (1..%d entries)", kMaxGaunt);
  const int nparts = view_bbox_parts(ps.n);
  launch_fill_u32(bad, 0u, 1, s);
  hipLaunchKernelGGL(k_view_pass<false>, dim3(nparts), dim3(kBlock), 0, s, ps, vp, gt, nullptr,
                     partial, bad);
  ILC_HIP(hipGetLastError());
  launch_bbox_final(partial, nparts, bad, out, s);
}

void launch_view_project(const particle_soa& ps, const view_proj& vp, const gaunt_table& gt,
                         float4* out, hipStream_t s) {
  uint64_t nb = (ps.n + kBlock - 1) / kBlock;
  if (nb > 16384) nb = 16384;
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL(k_view_pass<true>, dim3(uint32_t(nb)), dim3(kBlock), 0, s, ps, vp, gt, out,
                     nullptr, nullptr);
  ILC_HIP(hipGetLastError());
}

}  // namespace ilc
