// Per-point projected-kernel gather on gfx950.
//
// Replaces projective_smoothing_grid -> spg_run_reconstruction
// (projective_smoothing_backend_grid.cl:30-56, smoothing_particle_grid.cl:31-127).
//
// One 256-thread workgroup (4 waves of 64) owns a tile of TW x TH points of
// the point array seen as an image; each thread owns PPT points of one column.
// The workgroup
//   1. loads its points and, in reference mode, each point's search box
//      exactly as the reference computes it (cell of the point, radius =
//      min(max_h[cell] + cell_radius, max_h), index box of p -/+ radius, with
//      the unclamped 64-bit key of the reference);
//   2. reduces the tile's bounding box and the union / intersection of the
//      points' index boxes;
//   3. walks the candidate cells row by row: a row of cells with consecutive
//      x is one contiguous span of the cell-sorted particle array, streamed
//      256 particles at a time with coalesced 16-byte loads;
//   4. culls each streamed particle against the tile (support disk vs tile
//      box) and compacts the survivors into LDS with wave ballots.  In
//      reference mode a survivor whose cell lies inside every point's index
//      box goes to the "full" list, the others to the "masked" list, which
//      carries the cell coordinates for a per-point integer test;
//   5. when a list holds FLUSH particles, every thread evaluates all listed
//      particles for its points (LDS broadcast reads), accumulating in FP32
//      over 32 particles and then into a per-point FP64 accumulator.
// The cell-distance cull of the reference (smoothing_particle_grid.cl:95-100)
// never removes a particle whose support covers the point (a cell's centre
// is within the cell half-diagonal of every particle in it), so it is not
// repeated; particles outside their support contribute exactly 0 in the
// reference and are culled here.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

#include <type_traits>

namespace ilc {

namespace {

constexpr int kCap = 512;    // list capacity (entries)
constexpr int kFlush = 256;  // evaluate when a list reaches this size
constexpr int kCubicTable = 4096;

__device__ inline int32_t floor_to_int(float f) {
  const float fl = floorf(f);
  if (!(fl == fl)) return 0;
  if (fl >= 2147483647.0f) return INT32_MAX;
  if (fl <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(fl);
}

// grid3d_get_cell_indices of p - r (or p + r) on one axis, FP32, no FMA.
__device__ inline int32_t cell_of(float p, float r, float mn, float cs) {
#pragma clang fp contract(off)
  const float a = p + r;
  const float d = a - mn;
  return floor_to_int(d / cs);
}

__device__ inline float wmin(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline float wmax(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wmini(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wmaxi(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ inline uint32_t prefix_count(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Exclusive scan over the 256 threads of the workgroup; *total = sum.
// Contains two workgroup barriers.
__device__ inline uint32_t block_scan_excl(uint32_t v, uint32_t* total, uint32_t* ws) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) ws[wave] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int w = 0; w < 4; ++w) {
    if (w < wave) off += ws[w];
    tot += ws[w];
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Kernel shape times coefficient.  KERNEL 0: (1 - r^2/h^2)^2.5 (both
// quartic normalisations); KERNEL 2: tabulated projected cubic spline on
// u = r^2/h^2 in [0, 4), linear interpolation.
template <int KERNEL>
__device__ inline float shape(float r2, float ih2, const float* tab) {
  if constexpr (KERNEL == 2) {
    const float u = r2 * ih2;
    const float x = fminf(u * (float(kCubicTable - 1) / 4.0f), float(kCubicTable - 1));
    const int i = min(int(x), kCubicTable - 2);
    const float f = x - float(i);
    const float a = tab[i], b = tab[i + 1];
    return fmaf(f, b - a, a);
  } else {
    float s = fmaf(-r2, ih2, 1.0f);
    s = fmaxf(s, 0.0f);
    return (s * s) * __builtin_amdgcn_sqrtf(s);
  }
}

template <bool REF, int KERNEL>
struct gather_smem {
  float4 Lf[kCap];
  float4 Lm[REF ? kCap : 1];
  int2 Lmc[REF ? kCap : 1];
  uint32_t wcnt[2][8];
  float redf[4][4];
  int redi[4][10];
  float tab[KERNEL == 2 ? kCubicTable : 1];
  uint32_t cell_b[kBlock];
  uint32_t cell_off[kBlock];
  int2 cell_info[kBlock];
  uint32_t scan_ws[4];
};

// One tile (tx, ty) of the point image: everything between the point loads
// and the accumulator update.  All 256 threads of the workgroup call it with
// the same arguments.
template <int TW, int PPT, int OP, int KERNEL>
__device__ inline void gather_tile(const grid_view& g, const point_view& p, uint32_t tx,
                                   uint32_t ty, double* __restrict__ acc_out,
                                   unsigned long long* __restrict__ counts, float support,
                                   gather_smem<OP == OP_REFERENCE, KERNEL>& sm) {
  constexpr int ROWS = kBlock / TW;
  constexpr int TH = ROWS * PPT;
  constexpr bool REF = OP == OP_REFERENCE;
  constexpr bool COUNT = OP == OP_COUNT;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;


  const int col = t % TW, rg = t / TW;

  float ex[PPT], ey[PPT];
  uint32_t pid[PPT];
  bool pv[PPT];
  int lox[PPT], hix[PPT], loy[PPT], hiy[PPT];
  bool zin[PPT];
  double a64[PPT];
  uint32_t cnt_acc[PPT];

  float bminx = INFINITY, bminy = INFINITY, bmaxx = -INFINITY, bmaxy = -INFINITY;
  int ulx = INT32_MAX, uly = INT32_MAX, uhx = INT32_MIN, uhy = INT32_MIN;
  int ilx = INT32_MIN, ily = INT32_MIN, ihx = INT32_MAX, ihy = INT32_MAX;
  int anyz = 0, allz = 1, anyv = 0;

#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t px = tx * TW + col;
    const uint32_t py = ty * TH + rg * PPT + k;
    pid[k] = py * p.lnx + px;
    pv[k] = px < p.lnx && py < p.lny && pid[k] < p.n;
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (pv[k]) q = p.pts[pid[k]];
    ex[k] = q.x;
    ey[k] = q.y;
    a64[k] = 0.0;
    cnt_acc[k] = 0;
    lox[k] = hix[k] = loy[k] = hiy[k] = 0;
    zin[k] = false;
    if (pv[k]) {
      anyv = 1;
      bminx = fminf(bminx, q.x);
      bmaxx = fmaxf(bmaxx, q.x);
      bminy = fminf(bminy, q.y);
      bmaxy = fmaxf(bmaxy, q.y);
      if constexpr (REF) {
        // search radius of spg_run_reconstruction (smoothing_particle_grid.cl:54-75)
        const int32_t c0 = cell_of(q.x, 0.0f, g.minc[0], g.cs[0]);
        const int32_t c1 = cell_of(q.y, 0.0f, g.minc[1], g.cs[1]);
        const int32_t c2 = cell_of(q.z, 0.0f, g.minc[2], g.cs[2]);
        const uint64_t nx = uint64_t(int64_t(g.nc[0])), ny = uint64_t(int64_t(g.nc[1]));
        const uint64_t key = uint64_t(int64_t(c0)) + uint64_t(int64_t(c1)) * nx +
                             uint64_t(int64_t(c2)) * nx * ny;
        float R = g.max_h;
        if (key < uint64_t(g.ncells)) {
#pragma clang fp contract(off)
          R = fminf(__uint_as_float(g.maxh_bits[key]) + g.cell_radius, g.max_h);
        }
        lox[k] = cell_of(q.x, -R, g.minc[0], g.cs[0]);
        hix[k] = cell_of(q.x, R, g.minc[0], g.cs[0]);
        loy[k] = cell_of(q.y, -R, g.minc[1], g.cs[1]);
        hiy[k] = cell_of(q.y, R, g.minc[1], g.cs[1]);
        const int32_t loz = cell_of(q.z, -R, g.minc[2], g.cs[2]);
        const int32_t hiz = cell_of(q.z, R, g.minc[2], g.cs[2]);
        zin[k] = loz <= g.kz0 && g.kz0 <= hiz;
        ilx = max(ilx, lox[k]);
        ihx = min(ihx, hix[k]);
        ily = max(ily, loy[k]);
        ihy = min(ihy, hiy[k]);
        allz &= int(zin[k]);
        if (zin[k]) {
          anyz = 1;
          ulx = min(ulx, lox[k]);
          uhx = max(uhx, hix[k]);
          uly = min(uly, loy[k]);
          uhy = max(uhy, hiy[k]);
        }
      }
    }
  }

  // ---- tile reductions
  {
    const float v0 = wmin(bminx), v1 = wmin(bminy), v2 = wmax(bmaxx), v3 = wmax(bmaxy);
    int iv[10] = {wmini(ulx), wmini(uly), wmaxi(uhx), wmaxi(uhy), wmaxi(ilx),
                  wmaxi(ily), wmini(ihx), wmini(ihy), wmaxi(anyz), wmini(allz)};
    const int av = wmaxi(anyv);
    if (lane == 0) {
      sm.redf[wave][0] = v0;
      sm.redf[wave][1] = v1;
      sm.redf[wave][2] = v2;
      sm.redf[wave][3] = v3;
      for (int j = 0; j < 10; ++j) sm.redi[wave][j] = iv[j];
      // waves without a valid point must not widen the intersection
      if (!av) {
        sm.redi[wave][4] = INT32_MIN;
        sm.redi[wave][5] = INT32_MIN;
        sm.redi[wave][6] = INT32_MAX;
        sm.redi[wave][7] = INT32_MAX;
        sm.redi[wave][9] = 1;
      }
    }
    if (lane == 0) sm.wcnt[0][4 + wave] = av;
  }
  __syncthreads();
  float tbx0 = sm.redf[0][0], tby0 = sm.redf[0][1], tbx1 = sm.redf[0][2], tby1 = sm.redf[0][3];
  int r_ulx = sm.redi[0][0], r_uly = sm.redi[0][1], r_uhx = sm.redi[0][2], r_uhy = sm.redi[0][3];
  int r_ilx = sm.redi[0][4], r_ily = sm.redi[0][5], r_ihx = sm.redi[0][6], r_ihy = sm.redi[0][7];
  int r_anyz = sm.redi[0][8], r_allz = sm.redi[0][9];
  int any_valid = sm.wcnt[0][4];
  for (int w = 1; w < 4; ++w) {
    tbx0 = fminf(tbx0, sm.redf[w][0]);
    tby0 = fminf(tby0, sm.redf[w][1]);
    tbx1 = fmaxf(tbx1, sm.redf[w][2]);
    tby1 = fmaxf(tby1, sm.redf[w][3]);
    r_ulx = min(r_ulx, sm.redi[w][0]);
    r_uly = min(r_uly, sm.redi[w][1]);
    r_uhx = max(r_uhx, sm.redi[w][2]);
    r_uhy = max(r_uhy, sm.redi[w][3]);
    r_ilx = max(r_ilx, sm.redi[w][4]);
    r_ily = max(r_ily, sm.redi[w][5]);
    r_ihx = min(r_ihx, sm.redi[w][6]);
    r_ihy = min(r_ihy, sm.redi[w][7]);
    r_anyz = max(r_anyz, sm.redi[w][8]);
    r_allz = min(r_allz, sm.redi[w][9]);
    any_valid |= sm.wcnt[0][4 + w];
  }
  __syncthreads();
  if (!uniform(any_valid)) return;

  // ---- candidate cell box
  int cx0, cx1, cy0, cy1, cz0, cz1;
  if constexpr (REF) {
    if (!uniform(r_anyz) || g.kz0 < 0 || g.kz0 >= g.nc[2]) {
      cx0 = 1;
      cx1 = 0;
    } else {
      cx0 = max(r_ulx, 0);
      cx1 = min(r_uhx, g.nc[0] - 1);
    }
    cy0 = max(r_uly, 0);
    cy1 = min(r_uhy, g.nc[1] - 1);
    cz0 = cz1 = g.kz0;
  } else {
    // one extra cell on each side absorbs rounding in p -/+ reach
    const float reach = support * g.max_h;
    cx0 = max(max(cell_of(tbx0, -reach, g.minc[0], g.cs[0]), INT32_MIN + 1) - 1, 0);
    cx1 = min(min(cell_of(tbx1, reach, g.minc[0], g.cs[0]), INT32_MAX - 1) + 1, g.nc[0] - 1);
    cy0 = max(max(cell_of(tby0, -reach, g.minc[1], g.cs[1]), INT32_MIN + 1) - 1, 0);
    cy1 = min(min(cell_of(tby1, reach, g.minc[1], g.cs[1]), INT32_MAX - 1) + 1, g.nc[1] - 1);
    cz0 = 0;
    cz1 = g.nc[2] - 1;
  }
  cx0 = uniform(cx0);
  cx1 = uniform(cx1);
  cy0 = uniform(cy0);
  cy1 = uniform(cy1);
  cz0 = uniform(cz0);
  cz1 = uniform(cz1);
  const bool allz_u = uniform(r_allz) != 0;
  const float cull2 = COUNT ? support * support * 1.0001f : support * support;

  int nf = 0, nm = 0;  // list sizes (identical in every thread)
  int slot = 0;

  // Evaluate every listed particle for this thread's points: FP32 partial
  // sums over 32 particles, then FP64.  Loop bounds are wave-uniform.
  auto eval_full = [&](int cnt) {
    cnt = uniform(cnt);
    for (int j0 = 0; j0 < cnt; j0 += 32) {
      const int jn = min(32, cnt - j0);
      float a[PPT];
#pragma unroll
      for (int k = 0; k < PPT; ++k) a[k] = 0.0f;
      int j = 0;
      if constexpr (!COUNT) {
        for (; j + 4 <= jn; j += 4) {
          float4 q[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) q[u] = sm.Lf[j0 + j + u];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
              const float dx = q[u].x - ex[k], dy = q[u].y - ey[k];
              const float r2 = fmaf(dy, dy, dx * dx);
              a[k] = fmaf(q[u].z, shape<KERNEL>(r2, q[u].w, sm.tab), a[k]);
            }
          }
        }
      }
      for (; j < jn; ++j) {
        const float4 q = sm.Lf[j0 + j];
        if constexpr (COUNT) {
#pragma clang fp contract(off)
          const double hs = double(support) * double(q.z);
          const double hs2 = hs * hs;
#pragma unroll
          for (int k = 0; k < PPT; ++k) {
            const double dx = double(q.x) - double(ex[k]);
            const double dy = double(q.y) - double(ey[k]);
            const double r2 = dx * dx + dy * dy;
            cnt_acc[k] += (r2 < hs2) ? 1u : 0u;
          }
        } else {
#pragma unroll
          for (int k = 0; k < PPT; ++k) {
            const float dx = q.x - ex[k], dy = q.y - ey[k];
            const float r2 = fmaf(dy, dy, dx * dx);
            a[k] = fmaf(q.z, shape<KERNEL>(r2, q.w, sm.tab), a[k]);
          }
        }
      }
      if constexpr (!COUNT) {
#pragma unroll
        for (int k = 0; k < PPT; ++k) a64[k] += double(a[k]);
      }
    }
  };
  auto eval_masked = [&](int cnt) {
    if constexpr (REF) {
      cnt = uniform(cnt);
      for (int j0 = 0; j0 < cnt; j0 += 32) {
        const int jn = min(32, cnt - j0);
        float a[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) a[k] = 0.0f;
        for (int j = 0; j < jn; ++j) {
          const float4 q = sm.Lm[j0 + j];
          const int2 c = sm.Lmc[j0 + j];
#pragma unroll
          for (int k = 0; k < PPT; ++k) {
            const bool in = zin[k] & (c.x >= lox[k]) & (c.x <= hix[k]) & (c.y >= loy[k]) &
                            (c.y <= hiy[k]);
            const float dx = q.x - ex[k], dy = q.y - ey[k];
            const float r2 = fmaf(dy, dy, dx * dx);
            const float v = shape<KERNEL>(r2, q.w, sm.tab);
            a[k] = fmaf(q.z, in ? v : 0.0f, a[k]);
          }
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k) a64[k] += double(a[k]);
      }
    }
  };

  // Candidate cells are the box [cx0,cx1] x [cy0,cy1] x [cz0,cz1], visited
  // 256 at a time (one per thread, x fastest).  A cell is skipped unless
  // its max h can reach the tile (the conservative form of the reference's
  // cell cull); the particles of the reached cells are then streamed with a
  // block-wide offset table, culled one by one against the tile, and
  // appended to the LDS lists.
  const int ncols = cx1 >= cx0 ? cx1 - cx0 + 1 : 0;
  const int nrows = cy1 >= cy0 ? cy1 - cy0 + 1 : 0;
  const int nlays = cz1 >= cz0 ? cz1 - cz0 + 1 : 0;
  const int64_t ncand = int64_t(ncols) * nrows * nlays;
  const uint32_t nxny = uint32_t(g.nc[0]) * uint32_t(g.nc[1]);
  // reached cells must overlap the tile box grown by the support; float
  // rounding of particle coordinates vs cell edges is absorbed by a margin
  const float sup_h2 = support * support * 1.00002f;
  for (int64_t c0 = 0; c0 < ncand; c0 += kBlock) {
    const int64_t ci = c0 + t;
    uint32_t cb = 0, cn = 0;
    int ccx = 0, ccy = 0, cfull = 0;
    if (ci < ncand) {
      const int col = int(ci % ncols);
      const int64_t rest = ci / ncols;
      ccx = cx0 + col;
      ccy = cy0 + int(rest % nrows);
      const int ccz = cz0 + int(rest / nrows);
      const uint32_t key = uint32_t(ccx) + uint32_t(ccy) * uint32_t(g.nc[0]) + uint32_t(ccz) * nxny;
      cb = g.start[key];
      const uint32_t ce = g.start[key + 1];
      if (ce > cb) {
        const float mh = __uint_as_float(g.maxh_bits[key]);
        const float lo_x = g.minc[0] + float(ccx) * g.cs[0];
        const float lo_y = g.minc[1] + float(ccy) * g.cs[1];
        // a particle's cell comes from fl((x - min) / w): its coordinate can
        // sit a few ulps of |x| + |min| outside the cell's float edges
        const float mgx = 1e-4f * g.cs[0] + 4e-7f * (fabsf(g.minc[0]) + fabsf(lo_x) + g.cs[0]);
        const float mgy = 1e-4f * g.cs[1] + 4e-7f * (fabsf(g.minc[1]) + fabsf(lo_y) + g.cs[1]);
        const float dx = fmaxf(fmaxf((lo_x - mgx) - tbx1, tbx0 - (lo_x + g.cs[0] + mgx)), 0.0f);
        const float dy = fmaxf(fmaxf((lo_y - mgy) - tby1, tby0 - (lo_y + g.cs[1] + mgy)), 0.0f);
        if (dx * dx + dy * dy < mh * mh * sup_h2) cn = ce - cb;
        if constexpr (REF)
          cfull = (allz_u && ccy >= r_ily && ccy <= r_ihy && ccx >= r_ilx && ccx <= r_ihx) ? 1 : 0;
      }
    }
    // block exclusive scan of the reached cells' particle counts
    uint32_t tot;
    const uint32_t coff = block_scan_excl(cn, &tot, sm.scan_ws);
    sm.cell_b[t] = cb;
    sm.cell_off[t] = coff;
    sm.cell_info[t] = make_int2(ccx, (ccy << 1) | cfull);
    __syncthreads();
    tot = uint32_t(uniform(int(tot)));
    for (uint32_t v0 = 0; v0 < tot; v0 += kBlock) {
      const uint32_t v = v0 + uint32_t(t);
      const bool valid = v < tot;
      float4 q = make_float4(0.f, 0.f, 0.f, 1.f);
      bool keep = false, full = false;
      int2 cinf = make_int2(0, 0);
      if (valid) {
        // last cell whose offset is <= v (it has a nonzero count)
        int lo = 0;
#pragma unroll
        for (int step = kBlock / 2; step > 0; step >>= 1)
          if (sm.cell_off[lo + step] <= v) lo += step;
        const uint32_t sidx = sm.cell_b[lo] + (v - sm.cell_off[lo]);
        cinf = sm.cell_info[lo];
        q = g.part[sidx];
        if (q.z != 0.0f) {
          const float dx = fmaxf(fmaxf(tbx0 - q.x, q.x - tbx1), 0.0f);
          const float dy = fmaxf(fmaxf(tby0 - q.y, q.y - tby1), 0.0f);
          const float d2 = fmaf(dy, dy, dx * dx);
          keep = d2 * q.w < cull2;
        }
        full = keep && (!REF || (cinf.y & 1));
        if constexpr (COUNT) {
          if (keep) q.z = g.hs[sidx];
        }
      }
      const uint64_t bf = __ballot(full);
      const uint64_t bm = __ballot(keep && !full);
      if (lane == 0) {
        sm.wcnt[slot][wave] = __popcll(bf);
        sm.wcnt[slot][4 + wave] = __popcll(bm);
      }
      __syncthreads();
      int basef = nf, basem = nm, totf = 0, totm = 0;
      for (int w = 0; w < 4; ++w) {
        const int cf = sm.wcnt[slot][w], cm = sm.wcnt[slot][4 + w];
        if (w < wave) {
          basef += cf;
          basem += cm;
        }
        totf += cf;
        totm += cm;
      }
      if (full) sm.Lf[basef + prefix_count(bf)] = q;
      if constexpr (REF) {
        if (keep && !full) {
          const int pos = basem + prefix_count(bm);
          sm.Lm[pos] = q;
          sm.Lmc[pos] = make_int2(cinf.x, cinf.y >> 1);
        }
      }
      nf += totf;
      nm += totm;
      slot ^= 1;
      if (nf >= kFlush || nm >= kFlush) {
        __syncthreads();
        eval_full(nf);
        eval_masked(nm);
        __syncthreads();
        nf = 0;
        nm = 0;
      }
    }
    __syncthreads();  // the cell table is rewritten by the next chunk
  }
  if (nf > 0 || nm > 0) {
    __syncthreads();
    eval_full(nf);
    eval_masked(nm);
  }

#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (!pv[k]) continue;
    if constexpr (COUNT) {
      if (counts) counts[pid[k]] += cnt_acc[k];
    } else {
      acc_out[pid[k]] += a64[k];
    }
  }
}

template <int TW, int PPT, int OP, int KERNEL>
__global__ __launch_bounds__(kBlock) void k_gather(grid_view g, point_view p, uint32_t tiles_x,
                                                   uint32_t ntiles, double* __restrict__ acc_out,
                                                   unsigned long long* __restrict__ counts,
                                                   const float* __restrict__ cubic_table,
                                                   float support) {
  __shared__ gather_smem<OP == OP_REFERENCE, KERNEL> sm;
  if constexpr (KERNEL == 2) {
    for (int i = threadIdx.x; i < kCubicTable; i += kBlock) sm.tab[i] = cubic_table[i];
  }
  // XCD-contiguous tile order: workgroups are dealt round-robin to the 8
  // XCDs, so give XCD x the x-th eighth of the tiles (shared L2 reuse of the
  // neighbouring tiles' particle spans).
  const uint32_t bid = blockIdx.x;
  const uint32_t per = ntiles / 8u;
  const uint32_t tile = bid < per * 8u ? (bid % 8u) * per + bid / 8u : bid;
  const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
  gather_tile<TW, PPT, OP, KERNEL>(g, p, tx, ty, acc_out, counts, support, sm);
}

// Persistent form over many views: every workgroup takes the next
// (view, tile) item from a device counter until the list is exhausted.
// Items are ordered by estimated cost, largest first.
template <int TW, int PPT, int OP, int KERNEL>
__global__ __launch_bounds__(kBlock) void k_gather_batch(const view_desc* __restrict__ views,
                                                         const uint2* __restrict__ items,
                                                         uint32_t nitems,
                                                         uint32_t* __restrict__ counter,
                                                         const float* __restrict__ cubic_table,
                                                         float support) {
  __shared__ gather_smem<OP == OP_REFERENCE, KERNEL> sm;
  __shared__ uint32_t next;
  if constexpr (KERNEL == 2) {
    for (int i = threadIdx.x; i < kCubicTable; i += kBlock) sm.tab[i] = cubic_table[i];
  }
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t w = uint32_t(uniform(int(next)));
    if (w >= nitems) break;
    const uint2 it = items[w];
    const uint32_t vi = uint32_t(uniform(int(it.x))), ti = uint32_t(uniform(int(it.y)));
    const view_desc& vd = views[vi];
    const uint32_t tiles_x = vd.tiles_x;
    gather_tile<TW, PPT, OP, KERNEL>(vd.g, vd.p, ti % tiles_x, ti / tiles_x, vd.acc, nullptr,
                                     support, sm);
  }
}

template <int TW, int PPT, int OP, int KERNEL>
void launch_one(const grid_view& g, const point_view& p, double* acc, unsigned long long* counts,
                const float* cubic_table, float support, hipStream_t s) {
  constexpr int TH = (kBlock / TW) * PPT;
  const uint32_t tiles_x = (p.lnx + TW - 1) / TW;
  const uint32_t tiles_y = (p.lny + TH - 1) / TH;
  const uint32_t ntiles = tiles_x * tiles_y;
  if (!ntiles) return;
  hipLaunchKernelGGL((k_gather<TW, PPT, OP, KERNEL>), dim3(ntiles), dim3(kBlock), 0, s, g, p,
                     tiles_x, ntiles, acc, counts, cubic_table, support);
  ILC_HIP(hipGetLastError());
}

template <int OP, int KERNEL>
void launch_shape(const grid_view& g, const point_view& p, const tile_shape& ts, double* acc,
                  unsigned long long* counts, const float* tab, float support, hipStream_t s) {
  if (ts.tw == 16 && ts.ppt == 1)
    launch_one<16, 1, OP, KERNEL>(g, p, acc, counts, tab, support, s);
  else if (ts.tw == 32 && ts.ppt == 2)
    launch_one<32, 2, OP, KERNEL>(g, p, acc, counts, tab, support, s);
  else if (ts.tw == 32 && ts.ppt == 4)
    launch_one<32, 4, OP, KERNEL>(g, p, acc, counts, tab, support, s);
  else if (ts.tw == 64 && ts.ppt == 4)
    launch_one<64, 4, OP, KERNEL>(g, p, acc, counts, tab, support, s);
  else
    fail(ILC_EINVAL, "gather: unsupported tile shape tw=%d ppt=%d", ts.tw, ts.ppt);
}

template <int TW, int PPT, int OP, int KERNEL>
void launch_batch_one(const view_desc* views, const uint2* items, uint32_t nitems,
                      uint32_t* counter, const float* tab, float support, int grid,
                      hipStream_t s) {
  hipLaunchKernelGGL((k_gather_batch<TW, PPT, OP, KERNEL>), dim3(grid), dim3(kBlock), 0, s, views,
                     items, nitems, counter, tab, support);
  ILC_HIP(hipGetLastError());
}

template <int TW, int PPT, int OP, int KERNEL>
int batch_occupancy() {
  int per_cu = 0;
  ILC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gather_batch<TW, PPT, OP, KERNEL>,
                                                      kBlock, 0));
  return per_cu;
}

// dispatch over (tile shape, op, kernel) to F<TW, PPT, OP, KERNEL>
template <template <int, int, int, int> class F, class... A>
auto dispatch(const tile_shape& ts, int op, int kernel, A&&... a) {
  auto by_kernel = [&](auto tw, auto ppt) {
    constexpr int TW = decltype(tw)::value, PPT = decltype(ppt)::value;
    if (op == OP_REFERENCE) {
      if (kernel == ILC_KERNEL_CUBIC) fail(ILC_EINVAL, "reference mode uses the quartic kernel");
      return F<TW, PPT, OP_REFERENCE, 0>::run(a...);
    }
    if (kernel == ILC_KERNEL_CUBIC) return F<TW, PPT, OP_EXACT, 2>::run(a...);
    return F<TW, PPT, OP_EXACT, 0>::run(a...);
  };
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I4 = std::integral_constant<int, 4>;
  if (ts.tw == 16 && ts.ppt == 1) return by_kernel(I16{}, I1{});
  if (ts.tw == 32 && ts.ppt == 2) return by_kernel(I32{}, I2{});
  if (ts.tw == 32 && ts.ppt == 4) return by_kernel(I32{}, I4{});
  if (ts.tw == 64 && ts.ppt == 4) return by_kernel(I64{}, I4{});
  fail(ILC_EINVAL, "gather: unsupported tile shape tw=%d ppt=%d", ts.tw, ts.ppt);
}

template <int TW, int PPT, int OP, int KERNEL>
struct batch_launcher {
  static int run(const view_desc* v, const uint2* it, uint32_t ni, uint32_t* c, const float* tab,
                 float sup, int grid, hipStream_t s) {
    launch_batch_one<TW, PPT, OP, KERNEL>(v, it, ni, c, tab, sup, grid, s);
    return 0;
  }
};

template <int TW, int PPT, int OP, int KERNEL>
struct batch_occ {
  static int run() { return batch_occupancy<TW, PPT, OP, KERNEL>(); }
};

}  // namespace

void tile_grid(const tile_shape& ts, uint32_t lnx, uint32_t lny, uint32_t* tiles_x,
               uint32_t* ntiles) {
  const uint32_t th = uint32_t(kBlock / ts.tw) * uint32_t(ts.ppt);
  *tiles_x = (lnx + ts.tw - 1) / ts.tw;
  *ntiles = *tiles_x * ((lny + th - 1) / th);
}

int gather_batch_grid(int op, int kernel, const tile_shape& ts) {
  int dev = 0, cus = 0;
  ILC_HIP(hipGetDevice(&dev));
  ILC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int per_cu = dispatch<batch_occ>(ts, op, kernel);
  return std::max(1, per_cu) * cus;
}

void launch_gather_batch(const view_desc* views, const uint2* items, uint32_t nitems,
                         uint32_t* counter, int op, int kernel, const tile_shape& ts,
                         const float* cubic_table, int grid_blocks, hipStream_t s) {
  const float support = kernel == ILC_KERNEL_CUBIC ? 2.0f : 1.0f;
  dispatch<batch_launcher>(ts, op, kernel, views, items, nitems, counter, cubic_table, support,
                           grid_blocks, s);
}

void launch_gather(const grid_view& g, const point_view& p, int op, int kernel,
                   const tile_shape& ts, double* acc, unsigned long long* counts,
                   const float* cubic_table, hipStream_t s) {
  const float support = kernel == ILC_KERNEL_CUBIC ? 2.0f : 1.0f;
  if (op == OP_COUNT) {
    launch_one<32, 4, OP_COUNT, 0>(g, p, nullptr, counts, nullptr, support, s);
    return;
  }
  if (op == OP_REFERENCE) {
    if (kernel == ILC_KERNEL_CUBIC) fail(ILC_EINVAL, "reference mode uses the quartic kernel");
    launch_shape<OP_REFERENCE, 0>(g, p, ts, acc, counts, cubic_table, support, s);
  } else if (kernel == ILC_KERNEL_CUBIC) {
    launch_shape<OP_EXACT, 2>(g, p, ts, acc, counts, cubic_table, support, s);
  } else {
    launch_shape<OP_EXACT, 0>(g, p, ts, acc, counts, cubic_table, support, s);
  }
}

}  // namespace ilc
