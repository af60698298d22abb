// Volumetric smoothing on gfx950: the 3-D grid reconstruction of dm::grid
// (dm_reconstruction_backend_grid.cl:26-52 running spg_run_reconstruction,
// smoothing_particle_grid.cl:31-127, without SPG_PROJECTION).  Every
// evaluation point (x, y, z) sums w * quartic_polynomial3d(|x_p - x|, h)
// (smoothing_kernel.cl:27-37) over the particles of the cells in its search
// box.
//
// One 256-thread workgroup owns a 16 x 16 tile of the point array seen as an
// image (one point per thread, e.g. a volumetric slice).  It
//   1. loads its points and, in reference mode, each point's search box as
//      the reference computes it (radius = min(max_h[cell] + cell_radius,
//      max_h) from the unclamped 64-bit key of the point's cell, index box of
//      x -/+ radius on each axis);
//   2. reduces the tile's 3-D bounding box and the union and intersection of
//      the search boxes;
//   3. visits the candidate cells 256 at a time, drops those whose max h
//      cannot reach the tile box, and reads only the prefix of each cell
//      (particles by descending h) that can;
//   4. culls each particle against the tile box (support sphere vs box) and
//      compacts the survivors into two LDS lists by wave ballot: full (the
//      cell is in every point's search box) and masked (in some);
//   5. evaluates each list for its point in FP32 over 32 particles, then into
//      an FP64 accumulator.
//
// The count variant (exact search) adds 1 per pair with w != 0 and s > 0
// instead of the kernel value: the number of particle-kernel evaluations.
//
// Kernel arithmetic: W = K / h^3 (1 - q^2)^2 for q = r / h < 1 and 0 beyond
// (the reference's 1 + q^4 - 2 q^2 with q = min(r / h, 1)); the grid stores
// coef = w K / h^3 and 1 / h^2, so a pair costs s = max(1 - r^2 / h^2, 0) and
// coef * s^2.  The cell-distance cull of the reference never removes a
// particle whose support reaches the point (its cell centre is within the
// cell half-diagonal of it), so it is not repeated.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace ilc {

namespace {

constexpr int kCap3 = 384;
constexpr int kFlush3 = 128;
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
  int2 Lmc[kCap3];   // masked: cell (cx | cy << 16, cz)
  uint32_t cell_b[kBlock];
  uint32_t cell_off[kBlock];
  int cell_xy[kBlock];  // cx | cy << 16
  int cell_zf[kBlock];  // cz << 1 | full
  uint32_t scan_ws[4];
  uint32_t wcnt[2][8];
  float redf[4][6];
  int redi[4][13];
};

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

template <bool REF, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_smooth3d(grid_view g, point_view p, uint32_t tiles_x,
                                                      double* __restrict__ acc,
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
  __syncthreads();
  if (!uniform(any)) return;

  // ---- candidate cell box
  int c0[3], c1[3];
  for (int a = 0; a < 3; ++a) {
    if (REF) {
      c0[a] = max(ulo[a], 0);
      c1[a] = min(uhi[a], g.nc[a] - 1);
    } else {
      // one extra cell on each side absorbs rounding in x -/+ reach
      c0[a] = max(max(cell_of(bmin[a], -g.max_h, g.minc[a], g.cs[a]), INT32_MIN + 1) - 1, 0);
      c1[a] = min(min(cell_of(bmax[a], g.max_h, g.minc[a], g.cs[a]), INT32_MAX - 1) + 1,
                  g.nc[a] - 1);
    }
    c0[a] = uniform(c0[a]);
    c1[a] = uniform(c1[a]);
  }
  const int64_t ncx = c1[0] >= c0[0] ? c1[0] - c0[0] + 1 : 0;
  const int64_t ncy = c1[1] >= c0[1] ? c1[1] - c0[1] + 1 : 0;
  const int64_t ncz = c1[2] >= c0[2] ? c1[2] - c0[2] + 1 : 0;
  const int64_t ncand = ncx * ncy * ncz;
  const uint32_t nxny = uint32_t(g.nc[0]) * uint32_t(g.nc[1]);

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
          const int cx = uniform(c.x) & 0xffff, cy = uniform(c.x) >> 16, cz = uniform(c.y);
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
  for (int64_t cbase = 0; cbase < ncand; cbase += kBlock) {
    const int64_t ci = cbase + t;
    uint32_t cb = 0, cn = 0;
    int cxy = 0, czf = 0;
    if (ci < ncand) {
      const int cx = c0[0] + int(ci % ncx);
      const int64_t rest = ci / ncx;
      const int cy = c0[1] + int(rest % ncy);
      const int cz = c0[2] + int(rest / ncy);
      const uint32_t key = uint32_t(cx) + uint32_t(cy) * uint32_t(g.nc[0]) + uint32_t(cz) * nxny;
      cb = g.start[key];
      const uint32_t ce = g.start[key + 1];
      if (ce > cb) {
        const float mh = __uint_as_float(g.maxh_bits[key]);
        const int cc[3] = {cx, cy, cz};
        float d2c = 0.0f;
        for (int a = 0; a < 3; ++a) {
          const float lo_a = g.minc[a] + float(cc[a]) * g.cs[a];
          const float mg = 1e-4f * g.cs[a] + 4e-7f * (fabsf(g.minc[a]) + fabsf(lo_a) + g.cs[a]);
          const float d = fmaxf(fmaxf((lo_a - mg) - bmax[a], bmin[a] - (lo_a + g.cs[a] + mg)), 0.0f);
          d2c = fmaf(d, d, d2c);
        }
        if (d2c < mh * mh * sup2) {
          // the cell's particles come by descending h (top 16 bits): the
          // prefix whose h, rounded up to its bucket, reaches the tile
          uint32_t a = 0, b = ce - cb;
          while (a < b) {
            const uint32_t m = (a + b) >> 1;
            const float hu = __uint_as_float(__float_as_uint(g.hs[cb + m]) | 0xffffu);
            if (d2c < hu * hu * sup2)
              a = m + 1;
            else
              b = m;
          }
          cn = a;
        }
        const bool full = REF && cx >= ilo[0] && cx <= ihi[0] && cy >= ilo[1] && cy <= ihi[1] &&
                          cz >= ilo[2] && cz <= ihi[2];
        cxy = cx | (cy << 16);
        czf = (cz << 1) | (full ? 1 : 0);
      }
    }
    uint32_t tot;
    const uint32_t coff = block_scan_excl(cn, &tot, sm.scan_ws);
    sm.cell_b[t] = cb;
    sm.cell_off[t] = coff;
    sm.cell_xy[t] = cxy;
    sm.cell_zf[t] = czf;
    __syncthreads();
    tot = uint32_t(uniform(int(tot)));
    for (uint32_t v0 = 0; v0 < tot; v0 += kBlock) {
      const uint32_t v = v0 + uint32_t(t);
      bool keep = false, full = false;
      float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
      float w = 0.0f;
      int2 cinf = make_int2(0, 0);
      if (v < tot) {
        int l = 0;
#pragma unroll
        for (int step = kBlock / 2; step > 0; step >>= 1)
          if (sm.cell_off[l + step] <= v) l += step;
        const uint32_t sidx = sm.cell_b[l] + (v - sm.cell_off[l]);
        const float4 pp = g.part[sidx];  // (x, y, coef, 1/h^2)
        const float pzv = g.pz[sidx];
        if (pp.z != 0.0f) {
          const float pc[3] = {pp.x, pp.y, pzv};
          float d2 = 0.0f;
          for (int a = 0; a < 3; ++a) {
            const float d = fmaxf(fmaxf(bmin[a] - pc[a], pc[a] - bmax[a]), 0.0f);
            d2 = fmaf(d, d, d2);
          }
          keep = d2 * pp.w < sup2;
        }
        full = keep && (!REF || (sm.cell_zf[l] & 1));
        e = make_float4(pp.x, pp.y, pzv, pp.w);
        w = pp.z;
        cinf = make_int2(sm.cell_xy[l], sm.cell_zf[l] >> 1);
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
    __syncthreads();  // the cell table is rewritten by the next chunk
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

}  // namespace

void launch_smooth3d(const grid_view& g, const point_view& p, int op, double* acc, hipStream_t s,
                     unsigned long long* cnt) {
  if (!p.n) return;
  ILC_REQUIRE(g.pz != nullptr, "the grid is not a volumetric grid");
  ILC_REQUIRE(op == OP_REFERENCE || op == OP_EXACT || op == OP_COUNT, "bad op %d", op);
  ILC_REQUIRE(op == OP_COUNT ? cnt != nullptr : acc != nullptr, "null output");
  const uint32_t tiles_x = (p.lnx + kTile3 - 1) / kTile3;
  const uint64_t tiles = uint64_t(tiles_x) * ((p.lny + kTile3 - 1) / kTile3);
  ILC_REQUIRE(tiles < (uint64_t(1) << 31), "too many tiles");
  const dim3 grid{static_cast<uint32_t>(tiles)}, block{kBlock};
  if (op == OP_REFERENCE)
    hipLaunchKernelGGL((k_smooth3d<true, false>), grid, block, 0, s, g, p, tiles_x, acc, cnt);
  else if (op == OP_EXACT)
    hipLaunchKernelGGL((k_smooth3d<false, false>), grid, block, 0, s, g, p, tiles_x, acc, cnt);
  else
    hipLaunchKernelGGL((k_smooth3d<false, true>), grid, block, 0, s, g, p, tiles_x, acc, cnt);
  ILC_HIP(hipGetLastError());
}

}  // namespace ilc
