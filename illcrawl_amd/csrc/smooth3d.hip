// Volumetric smoothing on gfx950: the 3-D grid reconstruction of dm::grid
// (dm_reconstruction_backend_grid.cl:26-52 running spg_run_reconstruction,
// smoothing_particle_grid.cl:31-127, without SPG_PROJECTION).  Every
// evaluation point (x, y, z) sums w * quartic_polynomial3d(|x_p - x|, h)
// (smoothing_kernel.cl:27-37) over the particles of the cells in its search
// box.
//
// Candidates come from the level index of the grid, not from the reference
// cells: particles binned by h into levels (level L holds h <= max_h 2^(L-23)),
// each level sorted by a fine cell key (z, y, x) whose cell size is about the
// level's h.  The reference's cells are coarse (64 particles each on average,
// far more in clustered regions), and a tile inside a cell would stream all
// of them; a level row holds only the particles near the tile.
// The result is the reference's, since the search box test is made on each
// particle's reference cell.
//
// One 256-thread workgroup owns a 16 x 16 tile of the point array seen as an
// image (one point per thread, e.g. a volumetric slice).  It
//   1. loads its points and, in reference mode, each point's search box as
//      the reference computes it (radius = min(max_h[cell] + cell_radius,
//      max_h) from the unclamped 64-bit key of the point's cell, index box of
//      x -/+ radius on each axis);
//   2. reduces the tile's 3-D bounding box and the union and intersection of
//      the search boxes;
//   3. per level, takes the fine cells within the level's max h of the tile
//      box: one key range per (y, z) row of them (per z plane, or one range,
//      when the rows are too many), found by binary search in the level's
//      sorted keys;
//   4. streams the ranges, culls each particle against the tile box (support
//      sphere vs box) and, in reference mode, drops it when its reference cell
//      is in no point's search box; the survivors go by wave ballot to two LDS
//      lists: full (the cell is in every point's box) and masked (in some);
//   5. evaluates each list for its point in FP32 over 32 particles, then into
//      an FP64 accumulator.
//
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace ilc {

namespace {

constexpr int kCap3 = 384;
constexpr int kFlush3 = 128;
constexpr uint32_t kRowCap3 = 64;  // key rows per level and tile before coarser ranges
static_assert(kFlush3 - 1 + kBlock <= kCap3, "list capacity");
constexpr int kTile3 = 16;  // 16 x 16 points per workgroup

__device__ inline int32_t floor_to_int(float f) {
  const float fl = floorf(f);
  if (!(fl == fl)) return 0;
  if (fl >= 2147483647.0f) return INT32_MAX;
  if (fl <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(fl);
}

// grid3d_get_cell_indices of p + r on one axis, FP32, no FMA
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

struct smem3 {
  float4 Ls[kCap3];  // full: (x, y, z, 1/h^2)
  float Lw[kCap3];   // full: coef
  float4 Lm[kCap3];  // masked: (x, y, z, 1/h^2)
  float Lmw[kCap3];  // masked: coef
  int2 Lmc[kCap3];   // masked: reference cell (cx | cy << 16, cz)
  uint32_t row_b[kBlock];
  uint32_t row_off[kBlock];
  uint32_t scan_ws[4];
  uint32_t wcnt[2][8];
  float redf[4][6];
  int redi[4][13];
  // the tile's key rows per level: fine box, row mode, first row, start
  int lv_f0[kLevels3][3];
  int lv_f1[kLevels3][3];
  int lv_mode[kLevels3];
  uint32_t lv_row0[kLevels3 + 1];
  uint32_t lv_s[kLevels3 + 1];
  uint32_t lv_p[kLevels3 + 1];  // plane directory offsets
};

// level of a particle: the smallest L with h <= max_h 2^(L - 23) (by the
// exponent of max_h / h; any assignment is correct, the gather searches
// each level with its particles' largest h)
__device__ inline int level_of(float h, float max_h) {
  const float r = max_h / h;
  const int e = int((__float_as_uint(r) >> 23) & 0xffu) - 127;
  return min(max(kLevels3 - 1 - e, 0), kLevels3 - 1);
}

// fine cell coordinate, clamped to [0, kFineMax3]; monotone in x
__device__ inline uint32_t fine_of(float x, float mn, float cs) {
#pragma clang fp contract(off)
  const float f = floorf((x - mn) / cs);
  if (!(f >= 0.0f)) return 0u;
  if (f >= float(kFineMax3)) return kFineMax3;
  return uint32_t(f);
}

__device__ inline unsigned long long fine_key(int l, uint32_t fx, uint32_t fy, uint32_t fz) {
  return (static_cast<unsigned long long>(l) << 57) | (static_cast<unsigned long long>(fz) << 38) |
         (static_cast<unsigned long long>(fy) << 19) | static_cast<unsigned long long>(fx);
}

// first index in [a, b) whose key is >= k (keys sorted ascending)
__device__ inline uint32_t lower_bound_u64(const unsigned long long* __restrict__ key, uint32_t a,
                                           uint32_t b, unsigned long long k) {
  while (a < b) {
    const uint32_t m = (a + b) >> 1;
    if (key[m] < k)
      a = m + 1;
    else
      b = m;
  }
  return a;
}

// the same for a k expected close after a: steps of 1, 2, 4, ... from a
// bracket it, then a binary search inside the bracket (a row of fine cells
// holds few particles, so this takes a few dependent loads, not log2 n)
__device__ inline uint32_t gallop_lower_bound_u64(const unsigned long long* __restrict__ key,
                                                  uint32_t a, uint32_t b, unsigned long long k) {
  uint32_t lo = a, step = 1;
  while (lo < b && key[lo] < k) {
    a = lo + 1;
    lo = (b - lo > step) ? lo + step : b;
    step <<= 1;
  }
  return lower_bound_u64(key, a, lo, k);
}

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

// 7 waves per SIMD (the LDS allows 7 workgroups per CU): the row searches
// and candidate loads are latency-bound, and more waves hide them (6 -> 7
// waves: 2.85 -> 2.64 ms on 8 slices of 2048^2 through 16M particles, with 4
// registers spilled at 7)
template <bool REF, bool COUNT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_smooth3d(
    grid_view g, level_view lv, point_view p, uint32_t tiles_x, double* __restrict__ acc,
    unsigned long long* __restrict__ cnt) {
  static_assert(!(REF && COUNT), "evaluations are counted with the exact search");
  __shared__ smem3 sm;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const uint32_t px = tx * kTile3 + uint32_t(t % kTile3), py = ty * kTile3 + uint32_t(t / kTile3);
  const uint32_t pid = py * p.lnx + px;
  const bool pv = px < p.lnx && py < p.lny && pid < p.n;
  const float4 q = pv ? p.pts[pid] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float qc[3] = {q.x, q.y, q.z};

  // the point's search box (reference mode)
  int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
  if (REF && pv) {
    int32_t c[3];
    for (int a = 0; a < 3; ++a) c[a] = cell_of(qc[a], 0.0f, g.minc[a], g.cs[a]);
    const uint64_t nx = uint64_t(int64_t(g.nc[0])), ny = uint64_t(int64_t(g.nc[1]));
    const uint64_t key = uint64_t(int64_t(c[0])) + uint64_t(int64_t(c[1])) * nx +
                         uint64_t(int64_t(c[2])) * nx * ny;
    float R = g.max_h;
    if (key < uint64_t(g.ncells)) {
#pragma clang fp contract(off)
      R = fminf(__uint_as_float(g.maxh_bits[key]) + g.cell_radius, g.max_h);
    }
    for (int a = 0; a < 3; ++a) {
      lo[a] = cell_of(qc[a], -R, g.minc[a], g.cs[a]);
      hi[a] = cell_of(qc[a], R, g.minc[a], g.cs[a]);
    }
  }

  // ---- tile reductions: bounding box; union and intersection of the boxes
  {
    float f[6];
    for (int a = 0; a < 3; ++a) {
      f[a] = wmin(pv ? qc[a] : INFINITY);
      f[3 + a] = wmax(pv ? qc[a] : -INFINITY);
    }
    int iv[13];
    for (int a = 0; a < 3; ++a) {
      iv[a] = wmini(pv ? lo[a] : INT32_MAX);       // union
      iv[3 + a] = wmaxi(pv ? hi[a] : INT32_MIN);
      iv[6 + a] = wmaxi(pv ? lo[a] : INT32_MIN);   // intersection
      iv[9 + a] = wmini(pv ? hi[a] : INT32_MAX);
    }
    iv[12] = wmaxi(pv ? 1 : 0);
    if (lane == 0) {
      for (int k = 0; k < 6; ++k) sm.redf[wave][k] = f[k];
      for (int k = 0; k < 13; ++k) sm.redi[wave][k] = iv[k];
    }
  }
  __syncthreads();
  float bmin[3], bmax[3];
  int ulo[3], uhi[3], ilo[3], ihi[3];
  int any = 0;
  for (int a = 0; a < 3; ++a) {
    bmin[a] = INFINITY;
    bmax[a] = -INFINITY;
    ulo[a] = INT32_MAX;
    uhi[a] = INT32_MIN;
    ilo[a] = INT32_MIN;
    ihi[a] = INT32_MAX;
  }
  for (int w = 0; w < 4; ++w) {
    for (int a = 0; a < 3; ++a) {
      bmin[a] = fminf(bmin[a], sm.redf[w][a]);
      bmax[a] = fmaxf(bmax[a], sm.redf[w][3 + a]);
      ulo[a] = min(ulo[a], sm.redi[w][a]);
      uhi[a] = max(uhi[a], sm.redi[w][3 + a]);
      ilo[a] = max(ilo[a], sm.redi[w][6 + a]);
      ihi[a] = min(ihi[a], sm.redi[w][9 + a]);
    }
    any |= sm.redi[w][12];
  }
  // the boxes are the workgroup's: scalar registers, not vector ones
  for (int a = 0; a < 3; ++a) {
    bmin[a] = __int_as_float(uniform(__float_as_int(bmin[a])));
    bmax[a] = __int_as_float(uniform(__float_as_int(bmax[a])));
    ulo[a] = uniform(ulo[a]);
    uhi[a] = uniform(uhi[a]);
    ilo[a] = uniform(ilo[a]);
    ihi[a] = uniform(ihi[a]);
  }
  __syncthreads();
  if (!uniform(any)) return;

  // per point: its box clamped to the cells a list entry can carry
  // (0..65535 per axis); an empty box fails every cell
  int bx0 = INT32_MAX, by0 = 0, bz0 = 0;
  uint32_t bxw = 0, byw = 0, bzw = 0;
  if (REF && pv) {
    const int x0 = max(lo[0], 0), x1 = min(hi[0], 65535);
    const int y0 = max(lo[1], 0), y1 = min(hi[1], 65535);
    const int z0 = max(lo[2], 0), z1 = min(hi[2], INT32_MAX - 1);
    if (x0 <= x1 && y0 <= y1 && z0 <= z1) {
      bx0 = x0;
      by0 = y0;
      bz0 = z0;
      bxw = uint32_t(x1 - x0);
      byw = uint32_t(y1 - y0);
      bzw = uint32_t(z1 - z0);
    }
  }

  double a64 = 0.0;
  int ns = 0, nm = 0, slot = 0;
  auto eval_full = [&](int cnt) {
    cnt = uniform(cnt);
    for (int j0 = 0; j0 < cnt; j0 += 32) {
      const int jn = min(32, cnt - j0);
      float a = 0.0f;
      for (int j = 0; j < jn; ++j) {
        const float4 e = sm.Ls[j0 + j];
        const float w = sm.Lw[j0 + j];
        const float dx = e.x - q.x, dy = e.y - q.y, dz = e.z - q.z;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const float s = fmaxf(fmaf(-r2, e.w, 1.0f), 0.0f);
        if constexpr (COUNT)
          a += (s > 0.0f && w != 0.0f) ? 1.0f : 0.0f;
        else
          a = fmaf(w, s * s, a);
      }
      a64 += double(a);
    }
  };
  auto eval_masked = [&](int cnt) {
    if constexpr (REF) {
      cnt = uniform(cnt);
      for (int j0 = 0; j0 < cnt; j0 += 32) {
        const int jn = min(32, cnt - j0);
        float a = 0.0f;
        for (int j = 0; j < jn; ++j) {
          const float4 e = sm.Lm[j0 + j];
          const float w = sm.Lmw[j0 + j];
          const int2 c = sm.Lmc[j0 + j];
          const int cx = uniform(c.x) & 0xffff, cy = int(uint32_t(uniform(c.x)) >> 16);
          const int cz = uniform(c.y);
          const bool in = uint32_t(cx - bx0) <= bxw && uint32_t(cy - by0) <= byw &&
                          uint32_t(cz - bz0) <= bzw;
          const float dx = e.x - q.x, dy = e.y - q.y, dz = e.z - q.z;
          const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
          const float s = fmaxf(fmaf(-r2, e.w, 1.0f), 0.0f);
          a = fmaf(in ? w : 0.0f, s * s, a);
        }
        a64 += double(a);
      }
    }
  };

  const float sup2 = 1.00002f;  // support 1 h, slack for rounding

  // ---- the tile's key rows, level by level
  if (t < kLevels3) {
    const int l = t;
    int f0[3] = {0, 0, 0}, f1[3] = {-1, -1, -1};
    int mode = 0;
    uint32_t rows = 0;
    if (lv.meta[l] != 0u) {
      // every particle whose support reaches a point of the tile lies within
      // the level's largest h of the tile box; the margin covers rounding
      const float R = __uint_as_float(lv.meta[kLevels3 + l]);
      const float c = lv.cs[l];
      for (int a = 0; a < 3; ++a) {
        const float m = R + 1e-5f * (fabsf(bmin[a]) + fabsf(bmax[a]) + R + fabsf(g.minc[a]));
        f0[a] = int(fine_of(bmin[a] - m, g.minc[a], c));
        f1[a] = int(fine_of(bmax[a] + m, g.minc[a], c));
      }
      const uint32_t ny = uint32_t(f1[1] - f0[1] + 1), nz = uint32_t(f1[2] - f0[2] + 1);
      if (ny <= kRowCap3 && nz <= kRowCap3 && ny * nz <= kRowCap3) {
        mode = 0;
        rows = ny * nz;
      } else if (nz <= kRowCap3) {
        mode = 1;  // one range per z plane over rows f0y..f1y
        rows = nz;
      } else {
        mode = 2;  // one range over the whole box
        rows = 1;
      }
    }
    for (int a = 0; a < 3; ++a) {
      sm.lv_f0[l][a] = f0[a];
      sm.lv_f1[l][a] = f1[a];
    }
    sm.lv_mode[l] = mode;
    sm.lv_row0[l] = rows;
    sm.lv_s[l] = lv.meta[2 * kLevels3 + l];
    sm.lv_p[l] = lv.poff[l];
    if (l == kLevels3 - 1) {
      sm.lv_s[kLevels3] = lv.meta[3 * kLevels3];
      sm.lv_p[kLevels3] = lv.poff[kLevels3];
    }
  }
  __syncthreads();
  if (t == 0) {
    uint32_t r0 = 0;
    for (int l = 0; l < kLevels3; ++l) {
      const uint32_t r = sm.lv_row0[l];
      sm.lv_row0[l] = r0;
      r0 += r;
    }
    sm.lv_row0[kLevels3] = r0;
  }
  __syncthreads();
  const uint32_t nrows = uint32_t(uniform(int(sm.lv_row0[kLevels3])));

  for (uint32_t rbase = 0; rbase < nrows; rbase += kBlock) {
    const uint32_t r = rbase + uint32_t(t);
    uint32_t rb = 0, rn = 0;
    if (r < nrows) {
      int l = 0;
      while (sm.lv_row0[l + 1] <= r) ++l;  // r < lv_row0[kLevels3] ends the walk
      const uint32_t k = r - sm.lv_row0[l];
      const int mode = sm.lv_mode[l];
      const int* f0 = sm.lv_f0[l];
      const int* f1 = sm.lv_f1[l];
      int y0 = f0[1], y1 = f1[1], z0 = f0[2], z1 = f1[2];
      if (mode == 0) {
        const uint32_t ny = uint32_t(f1[1] - f0[1] + 1);
        y0 = y1 = f0[1] + int(k % ny);
        z0 = z1 = f0[2] + int(k / ny);
      } else if (mode == 1) {
        z0 = z1 = f0[2] + int(k);
      }
      const unsigned long long klo = fine_key(l, uint32_t(f0[0]), uint32_t(y0), uint32_t(z0));
      const unsigned long long khi = fine_key(l, uint32_t(f1[0]), uint32_t(y1), uint32_t(z1));
      // the keys of planes z0 .. z1 from the plane directory (a plane
      // beyond the level's last holds no particle)
      const uint32_t planes = sm.lv_p[l + 1] - sm.lv_p[l] - 1;
      const uint32_t za = min(uint32_t(z0), planes - 1), zb = min(uint32_t(z1), planes - 1);
      const uint32_t s_ = lv.pdir[sm.lv_p[l] + za];
      const uint32_t e_ = lv.pdir[sm.lv_p[l] + zb + 1];
      const uint32_t a_ = lower_bound_u64(lv.key, s_, e_, klo);
      rb = a_;
      rn = gallop_lower_bound_u64(lv.key, a_, e_, khi + 1ull) - a_;
    }
    uint32_t tot;
    const uint32_t roff = block_scan_excl(rn, &tot, sm.scan_ws);
    sm.row_b[t] = rb;
    sm.row_off[t] = roff;
    __syncthreads();
    tot = uint32_t(uniform(int(tot)));
    for (uint32_t v0 = 0; v0 < tot; v0 += kBlock) {
      const uint32_t v = v0 + uint32_t(t);
      bool keep = false, full = false;
      float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
      float w = 0.0f;
      int2 cinf = make_int2(0, 0);
      if (v < tot) {
        // last row whose offset is <= v (it has a nonzero count)
        int l = 0;
#pragma unroll
        for (int step = kBlock / 2; step > 0; step >>= 1)
          if (sm.row_off[l + step] <= v) l += step;
        const uint32_t sidx = sm.row_b[l] + (v - sm.row_off[l]);
        e = lv.pos[sidx];  // (x, y, z, 1/h^2)
        w = lv.w[sidx];
        if (w != 0.0f) {
          const float pc[3] = {e.x, e.y, e.z};
          float d2 = 0.0f;
          for (int a = 0; a < 3; ++a) {
            const float d = fmaxf(fmaxf(bmin[a] - pc[a], pc[a] - bmax[a]), 0.0f);
            d2 = fmaf(d, d, d2);
          }
          keep = d2 * e.w < sup2;
        }
        full = keep;
        if (REF && keep) {
          // the reference sums a particle only when its cell is in the
          // point's search box: none (drop), every (full) or some (masked)
          cinf = lv.cell[sidx];
          const int cc[3] = {cinf.x & 0xffff, int(uint32_t(cinf.x) >> 16), cinf.y};
          bool inu = true, ini = true;
          for (int a = 0; a < 3; ++a) {
            inu = inu && cc[a] >= ulo[a] && cc[a] <= uhi[a];
            ini = ini && cc[a] >= ilo[a] && cc[a] <= ihi[a];
          }
          keep = inu;
          full = inu && ini;
        }
      }
      const uint64_t bs = __ballot(full);
      const uint64_t bm = __ballot(keep && !full);
      if (lane == 0) {
        sm.wcnt[slot][wave] = __popcll(bs);
        sm.wcnt[slot][4 + wave] = __popcll(bm);
      }
      __syncthreads();
      int bases = ns, basem = nm, tots = 0, totm = 0;
      for (int ww = 0; ww < 4; ++ww) {
        const int cs_ = sm.wcnt[slot][ww], cm_ = sm.wcnt[slot][4 + ww];
        if (ww < wave) {
          bases += cs_;
          basem += cm_;
        }
        tots += cs_;
        totm += cm_;
      }
      if (full) {
        const int pos = bases + prefix_count(bs);
        sm.Ls[pos] = e;
        sm.Lw[pos] = w;
      }
      if (REF && keep && !full) {
        const int pos = basem + prefix_count(bm);
        sm.Lm[pos] = e;
        sm.Lmw[pos] = w;
        sm.Lmc[pos] = cinf;
      }
      ns += tots;
      nm += totm;
      slot ^= 1;
      if (ns >= kFlush3 || nm >= kFlush3) {
        __syncthreads();
        eval_full(ns);
        eval_masked(nm);
        __syncthreads();
        ns = nm = 0;
      }
    }
    __syncthreads();  // the row table is rewritten by the next chunk
  }
  if (ns > 0 || nm > 0) {
    __syncthreads();
    eval_full(ns);
    eval_masked(nm);
  }
  if (pv) {
    if constexpr (COUNT)
      cnt[pid] += static_cast<unsigned long long>(a64);
    else
      acc[pid] += a64;
  }
}

// ---- level index build, from the particles in input order

// grid3d_get_cell_indices of a particle (particle_grid.cl) as the grid build
// computes it; false when the particle lies outside the grid
__device__ inline bool cell_index_3d(const float4& q, const grid_params& g, int32_t idx[3]) {
#pragma clang fp contract(off)
  const float c[3] = {q.x, q.y, q.z};
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    const float d = c[a] - g.minc[a];
    idx[a] = floor_to_int(d / g.cs[a]);
    ok = ok && idx[a] >= 0 && idx[a] < g.nc[a];
  }
  return ok;
}

// per particle: the reference cell's max h (atomic max), the level and its
// fine key (lo, hi words), the level counts and max h, and a 32-byte record
// (x, y, z, h | w, cell xy, cell z, 0) in input order, so that the sorted
// gather reads one line per particle instead of one from p4 and one from h
__global__ __launch_bounds__(kBlock) void k_vol_keys(const float4* __restrict__ p4,
                                                      const float* __restrict__ hsml, uint32_t n,
                                                      grid_params g, float max_h, level_view lv,
                                                      float4* __restrict__ rec,
                                                      uint32_t* __restrict__ lo,
                                                      uint32_t* __restrict__ hi,
                                                      uint32_t* __restrict__ idx,
                                                      uint32_t* __restrict__ meta,
                                                      uint32_t* __restrict__ maxh_bits,
                                                      uint32_t* __restrict__ err) {
  __shared__ uint32_t cnt[kLevels3], mh[kLevels3];
  __shared__ float scs[kLevels3];
  __shared__ float4 srec[2 * kBlock];
  const int t = threadIdx.x;
  if (t == 0) {
#pragma unroll
    for (int l = 0; l < kLevels3; ++l) {
      cnt[l] = 0u;
      mh[l] = 0u;
      scs[l] = lv.cs[l];
    }
  }
  __syncthreads();
  const uint32_t nx = uint32_t(g.nc[0]), nxy = nx * uint32_t(g.nc[1]);
  // the whole block runs every round (the records go out through LDS)
  for (uint64_t base = uint64_t(blockIdx.x) * kBlock; base < n;
       base += uint64_t(gridDim.x) * kBlock) {
    const uint32_t i = uint32_t(base) + uint32_t(t);
    if (i < n) {
      const float4 q = p4[i];
      const float h = hsml[i];
      int32_t c[3];
      if (cell_index_3d(q, g, c)) {
        // clustered cells hold millions of particles: read the running max
        // first and go to the atomic unit only with a larger h (a stale read
        // is never above the true max, so nothing is skipped wrongly)
        uint32_t* m = &maxh_bits[uint32_t(c[0]) + uint32_t(c[1]) * nx + uint32_t(c[2]) * nxy];
        const uint32_t hb = __float_as_uint(h);
        if (hb > __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
          atomicMax(m, hb);
      } else {
        atomicAdd(err, 1u);
        c[0] = c[1] = c[2] = 0;
      }
      srec[2 * t] = make_float4(q.x, q.y, q.z, h);
      srec[2 * t + 1] = make_float4(q.w, __uint_as_float(uint32_t(c[0]) | (uint32_t(c[1]) << 16)),
                                    __uint_as_float(uint32_t(c[2])), 0.0f);
      const int l = level_of(h, max_h);
      const float cs = scs[l];
      const unsigned long long k = fine_key(l, fine_of(q.x, g.minc[0], cs),
                                            fine_of(q.y, g.minc[1], cs),
                                            fine_of(q.z, g.minc[2], cs));
      lo[i] = uint32_t(k);
      hi[i] = uint32_t(k >> 32);
      idx[i] = i;
      atomicAdd(&cnt[l], 1u);
      if (__float_as_uint(h) > mh[l]) atomicMax(&mh[l], __float_as_uint(h));
    }
    __syncthreads();
    // the round's records are contiguous: whole 16-byte words per lane
    const uint32_t m2 = 2u * uint32_t(min(uint64_t(kBlock), uint64_t(n) - base));
    for (uint32_t k = uint32_t(t); k < m2; k += kBlock) rec[2 * base + k] = srec[k];
    __syncthreads();
  }
  if (t < kLevels3 && cnt[t]) {
    atomicAdd(&meta[t], cnt[t]);
    atomicMax(&meta[kLevels3 + t], mh[t]);
  }
}

// pdir[poff[L] + z]: first index of level L whose fine z is >= z; the last
// entry of a level is its end
__global__ __launch_bounds__(kBlock) void k_lv_planes(level_view lv, uint32_t* __restrict__ pdir) {
  __shared__ uint32_t sp[kLevels3 + 1], ss[kLevels3 + 1];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int l = 0; l <= kLevels3; ++l) {
      sp[l] = lv.poff[l];
      ss[l] = lv.meta[2 * kLevels3 + l];
    }
  }
  __syncthreads();
  const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= sp[kLevels3]) return;
  int l = 0;
  while (sp[l + 1] <= e) ++l;  // e < sp[kLevels3] ends the walk
  const uint32_t z = e - sp[l];
  const uint32_t last = sp[l + 1] - sp[l] - 1;
  pdir[e] = z == last ? ss[l + 1]
                      : lower_bound_u64(lv.key, ss[l], ss[l + 1],
                                        fine_key(l, 0u, 0u, min(z, kFineMax3)));
}

__global__ void k_lv_starts(uint32_t* __restrict__ meta) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t s = 0;
  for (int l = 0; l < kLevels3; ++l) {
    meta[2 * kLevels3 + l] = s;
    s += meta[l];
  }
  meta[3 * kLevels3] = s;
}

// the level-ordered arrays: key, (x, y, z, 1/h^2), coef = w norm / h^3 and
// the reference cell
__global__ __launch_bounds__(kBlock) void k_vol_soa(const float4* __restrict__ rec, uint32_t n,
                                                     grid_params g, float max_h, float norm,
                                                     level_view lv,
                                                     const uint32_t* __restrict__ order,
                                                     unsigned long long* __restrict__ key,
                                                     float4* __restrict__ pos,
                                                     float* __restrict__ w,
                                                     int2* __restrict__ cell) {
  __shared__ float scs[kLevels3];
  const int t = threadIdx.x;
  if (t == 0) {
#pragma unroll
    for (int l = 0; l < kLevels3; ++l) scs[l] = lv.cs[l];
  }
  __syncthreads();
  const uint32_t j = blockIdx.x * kBlock + uint32_t(t);
  if (j >= n) return;
  const uint32_t i = order[j];
  const float4 r0 = rec[2 * uint64_t(i)];
  const float4 r1 = rec[2 * uint64_t(i) + 1];
  const float4 q = make_float4(r0.x, r0.y, r0.z, r1.x);
  const float h = r0.w;
  const int l = level_of(h, max_h);
  const float cs = scs[l];
  key[j] = fine_key(l, fine_of(q.x, g.minc[0], cs), fine_of(q.y, g.minc[1], cs),
                    fine_of(q.z, g.minc[2], cs));
  const float ih2 = 1.0f / (h * h);
  pos[j] = make_float4(q.x, q.y, q.z, ih2);
  w[j] = q.w * (norm * (ih2 / h));
  cell[j] = make_int2(int(__float_as_uint(r1.y)), int(__float_as_uint(r1.z)));
}

}  // namespace

void launch_volumetric_index(const float4* p4, const float* h, uint32_t n, const grid_params& g,
                             float max_h, float norm, const level_view& lv,
                             unsigned long long* key, float4* pos, float* w, int2* cell,
                             uint32_t* meta, uint32_t* pdir, uint32_t* maxh_bits, uint32_t* err,
                             uint32_t* tmp, uint32_t* radix_temp, hipStream_t s) {
  launch_fill_u32(meta, 0u, kLvMetaWords, s);
  launch_fill_u32(maxh_bits, 0u, g.ncells(), s);
  if (n) {
    float4* rec = reinterpret_cast<float4*>(tmp);
    uint32_t* lo = tmp + 8 * uint64_t(n);
    uint32_t* hi = lo + n;
    uint32_t* idx = hi + n;
    // a capped grid: each workgroup adds its level counts to the same few
    // global words once
    const uint32_t nb = std::min<uint32_t>((n + kBlock - 1) / kBlock, 2048u);
    hipLaunchKernelGGL(k_vol_keys, dim3(nb), dim3(kBlock), 0, s, p4, h, n, g, max_h, lv, rec, lo,
                       hi, idx, meta, maxh_bits, err);
    ILC_HIP(hipGetLastError());
    // (hi, lo) order by two stable sorts: low word, then high word
    launch_radix_sort(lo, idx, n, 32, radix_temp, s);
    launch_gather_u32(hi, idx, n, lo, s);
    launch_radix_sort(lo, idx, n, 30, radix_temp, s);
    const dim3 grid{(n + kBlock - 1) / kBlock}, block{kBlock};
    hipLaunchKernelGGL(k_vol_soa, grid, block, 0, s, rec, n, g, max_h, norm, lv, idx, key, pos, w,
                       cell);
    ILC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_lv_starts, dim3(1), dim3(64), 0, s, meta);
  ILC_HIP(hipGetLastError());
  // the directory searches the sorted keys (lv.key is key, lv.meta is meta)
  const uint32_t nd = lv.poff[kLevels3];
  ILC_REQUIRE(lv.key == key && lv.meta == meta && lv.pdir == pdir, "level index buffers");
  if (nd) {
    hipLaunchKernelGGL(k_lv_planes, dim3((nd + kBlock - 1) / kBlock), dim3(kBlock), 0, s, lv, pdir);
    ILC_HIP(hipGetLastError());
  }
}

void launch_smooth3d(const grid_view& g, const level_view& lv, const point_view& p, int op,
                     double* acc, hipStream_t s, unsigned long long* cnt) {
  if (!p.n) return;
  ILC_REQUIRE(lv.key != nullptr, "the grid is not a volumetric grid");
  ILC_REQUIRE(op == OP_REFERENCE || op == OP_EXACT || op == OP_COUNT, "bad op %d", op);
  ILC_REQUIRE(op == OP_COUNT ? cnt != nullptr : acc != nullptr, "null output");
  const uint32_t tiles_x = (p.lnx + kTile3 - 1) / kTile3;
  const uint64_t tiles = uint64_t(tiles_x) * ((p.lny + kTile3 - 1) / kTile3);
  ILC_REQUIRE(tiles < (uint64_t(1) << 31), "too many tiles");
  const dim3 grid{static_cast<uint32_t>(tiles)}, block{kBlock};
  if (op == OP_REFERENCE)
    hipLaunchKernelGGL((k_smooth3d<true, false>), grid, block, 0, s, g, lv, p, tiles_x, acc, cnt);
  else if (op == OP_EXACT)
    hipLaunchKernelGGL((k_smooth3d<false, false>), grid, block, 0, s, g, lv, p, tiles_x, acc, cnt);
  else
    hipLaunchKernelGGL((k_smooth3d<false, true>), grid, block, 0, s, g, lv, p, tiles_x, acc, cnt);
  ILC_HIP(hipGetLastError());
}

}  // namespace ilc
