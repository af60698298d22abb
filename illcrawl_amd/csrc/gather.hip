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
//      min(max_h[cell] + cell_radius, max_h), index box of p -/+ radius, from
//      the unclamped 64-bit key of the reference);
//   2. reduces the tile's bounding box and the union / intersection of the
//      points' index boxes;
//   3. visits the candidate cells 256 at a time, skips every cell whose max h
//      cannot reach the tile, and streams the particles of the others with a
//      block-wide offset table (coalesced 16-byte loads).  A cell holds its
//      particles by descending h, so a binary search on h finds the prefix
//      that can reach the tile and the rest is never read;
//   4. culls each particle against the tile (support disk vs tile box) and
//      compacts the survivors into LDS lists with wave ballots:
//        big    - every point may use it and its support is at least twice
//                 the tile's half-diagonal: stored as the coefficients of
//                 its kernel argument expanded about the tile centre;
//        small  - every point may use it, stored by position;
//        masked - reference mode, its cell is inside some points' index box
//                 and outside others': stored by position and cell;
//   5. when a list is full, every thread evaluates the listed particles for
//      its points from LDS (broadcast reads), FP32 over 32 particles, then
//      into a per-point FP64 accumulator.
// The small and masked lists share one LDS array from its two ends (as the
// big list's two kinds do), which keeps the workgroup at 17 KB of LDS.
//
// Kernel arithmetic.  With c = coef^0.4, the quartic term coef*s^2.5,
// s = 1 - r^2/h^2, equals (c*s)^2.5, and c*s is affine in the point's
// coordinates u (relative to the tile centre t) and |u|^2:
//   c*s = c(1 - |x-t|^2/h^2) + (2c/h^2)(x-t).u - (c/h^2)|u|^2,
// three FMAs per point for big particles (their expansion stays well
// conditioned: |x-t| <= 1.5 h), then max, sqrt, mul, fma.  Small and masked
// particles use c*s = c - (c/h^2) r^2 directly.
//
// The cell-distance cull of the reference (smoothing_particle_grid.cl:95-100)
// never removes a particle whose support covers the point (a cell's centre
// is within the cell half-diagonal of every particle in it), so it is not
// repeated; particles outside their support contribute exactly 0 in the
// reference and are culled here.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"
#include "pixel.hpp"

namespace ilc {

namespace {

// a workgroup of nb threads evaluates a list when it reaches flush_of(nb)
// entries (one-wave workgroups: fewer, to keep their LDS small)
constexpr int flush_of(int nb) { return nb >= 256 ? 64 : 32; }
// list capacity (entries): a batch of nb candidates lands on a list holding
// at most flush_of(nb) - 1
constexpr int cap_of(int nb) { return nb >= 256 ? 320 : 96; }
constexpr int kCap = cap_of(256);
// a tile of a sparse grid leaves the row streaming for the cell enumeration
// above this many particles per candidate cell
constexpr uint32_t kDenseRows = 2;
static_assert(flush_of(256) - 1 + 256 <= cap_of(256) && flush_of(64) - 1 + 64 <= cap_of(64),
              "list capacity");
constexpr int kCubicTable = 4096;

__device__ inline int32_t floor_to_int(float f) {
  const float fl = floorf(f);
  if (!(fl == fl)) return 0;
  if (fl >= 2147483647.0f) return INT32_MAX;
  if (fl <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(fl);
}

// grid3d_get_cell_indices of p + r on one axis, FP32, no FMA.
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


// Exclusive scan over the NB threads of the workgroup; *total = sum.
// Contains two workgroup barriers.
template <int NB>
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
  for (int w = 0; w < NB / 64; ++w) {
    if (w < wave) off += ws[w];
    tot += ws[w];
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// (c s)^2.5 for a scaled kernel argument cs = c*s, clamped at the support.
__device__ inline float pow25(float cs) {
  cs = fmaxf(cs, 0.0f);
  return (cs * cs) * __builtin_amdgcn_sqrtf(cs);
}

// The same for an argument known to be positive.
__device__ inline float pow25_pos(float cs) { return (cs * cs) * __builtin_amdgcn_sqrtf(cs); }

// The projected quartic clipped along the line of sight.  The 3-D kernel
// along a line at projected distance r from the particle is proportional to
// (a^2 - z^2)^2 on |z| < a = h s^(1/2), s = 1 - r^2/h^2; its integral over
// the depth range [z0, z1] relative to the particle is the full column's
// times (15/16) (G(u1) - G(u0)), G(u) = u - 2u^3/3 + u^5/5, u = z / a clamped
// to [-1, 1].  With v = c s (the argument of pow25) and t = (z / h) c^(1/2),
// u = t / v^(1/2).  tc: (t0, t1) of the range's two ends.
__device__ inline float clip_pow(float v, float2 tc) {
  v = fmaxf(v, 0.0f);
  const float sv = __builtin_amdgcn_sqrtf(v);
  const float iv = v > 0.0f ? __builtin_amdgcn_rcpf(sv) : 0.0f;
  const float u0 = __builtin_amdgcn_fmed3f(tc.x * iv, -1.0f, 1.0f);
  const float u1 = __builtin_amdgcn_fmed3f(tc.y * iv, -1.0f, 1.0f);
  const float q0 = u0 * u0, q1 = u1 * u1;
  const float g0 = u0 * fmaf(q0, fmaf(q0, 0.2f, -2.0f / 3.0f), 1.0f);
  const float g1 = u1 * fmaf(q1, fmaf(q1, 0.2f, -2.0f / 3.0f), 1.0f);
  return ((v * v) * sv) * ((g1 - g0) * 0.9375f);
}

// Tabulated projected cubic spline on u = r^2/h^2 in [0, 4), linear.
__device__ inline float cubic_shape(float u, const float* tab) {
  const float x = fminf(u * (float(kCubicTable - 1) / 4.0f), float(kCubicTable - 1));
  const int i = min(int(x), kCubicTable - 2);
  const float f = x - float(i);
  const float a = tab[i], b = tab[i + 1];
  return fmaf(f, b - a, a);
}

// Adjoint inputs and outputs (OP_ADJOINT only).
struct adjoint_io {
  const float* grad;  // dL/d(point value), one per point
  float scale;        // points' value = scale * accumulator
  float norm;         // kernel normalisation K (W = K shape / h^2)
  const float* ftab;  // cubic: F(u) and dF/du on u = r^2/h^2 in [0, 4] (kCubicTable nodes)
  const float* dtab;
  float* gw;          // per sorted particle: dL/dw, dL/dh, dL/dx, dL/dy (atomic adds)
  float* gh;
  float* gx;
  float* gy;
};

// the small and masked lists share one array from its two ends: before a
// batch of nb candidates each holds at most flush_of(nb) - 1 entries
constexpr int pool_of(int nb) { return cap_of(nb) + flush_of(nb); }
static_assert(2 * (flush_of(256) - 1) + 256 <= pool_of(256) &&
                  2 * (flush_of(64) - 1) + 64 <= pool_of(64),
              "pool capacity");

template <bool REF, int KERNEL, bool ADJ = false, bool CLIP = false, int NB = 256>
struct gather_smem {
  static constexpr int CAP = cap_of(NB), POOL = pool_of(NB);
  float4 Lb[KERNEL == 0 && !ADJ ? CAP : 1];  // big: expansion coefficients
  // small (position form) from the front; reference mode: masked (position
  // form) from the back, the m-th at Ls[POOL - 1 - m]
  float4 Ls[REF ? POOL : CAP];
  alignas(16) int Lmc[REF ? CAP : 1];  // masked: packed cell (cx | cy << 16)
  // line-of-sight clip: per big / small entry, the depth range relative to
  // the particle in units of h, times c^(1/2) (see clip_pow)
  float2 Lbz[CLIP ? CAP : 1];
  float2 Lsz[CLIP ? CAP : 1];
  uint32_t wcnt[2][16];
  float redf[4][4];
  int redi[4][10];
  float tab[KERNEL == 2 ? kCubicTable : 1];
  float dtab[KERNEL == 2 && ADJ ? kCubicTable : 1];
  uint32_t cell_b[NB];
  uint32_t cell_off[NB];
  int cell_info[NB];  // cells: packed cell cx | cy << 16 (unsigned); rows: base key
  int cell_y[NB];     // cells: 1 when in every point's search box; rows: the row's cy
  uint32_t scan_ws[4];
  uint32_t Lidx[ADJ ? CAP : 1];    // adjoint: sorted particle index
  float4 Pt[ADJ ? NB * 4 : 1];  // adjoint: tile points (x, y, dL/dvalue, -), row-major
  float4 Sb[ADJ ? 16 : 1];          // adjoint: 8x8 point blocks' boxes (x0, y0, x1, y1)
};

// One tile (tx, ty) of the point image: everything between the point loads
// and the accumulator update.  All NB threads of the workgroup (NB / 64
// waves) call it with the same arguments.
template <int TW, int PPT, int OP, int KERNEL, int NB = 256>
__device__ __attribute__((always_inline)) inline void gather_tile(
    const grid_view& g, const point_view& p, uint32_t tx, uint32_t ty, double* __restrict__ acc_out,
    unsigned long long* __restrict__ counts, float support,
    gather_smem<OP == OP_REFERENCE, KERNEL, OP == OP_ADJOINT, OP == OP_CLIP, NB>& sm,
    const adjoint_io* adj = nullptr, float* __restrict__ img_out = nullptr, float img_scale = 1.0f,
    int fused = 0) {
  constexpr int NW = NB / 64;
  constexpr int kCap = cap_of(NB), kPool = pool_of(NB), kFlush = flush_of(NB);
  constexpr int ROWS = NB / TW;
  constexpr int TH = ROWS * PPT;
  constexpr bool REF = OP == OP_REFERENCE;
  constexpr bool COUNT = OP == OP_COUNT;
  constexpr bool ADJ = OP == OP_ADJOINT;
  constexpr bool CLIP = OP == OP_CLIP;
  constexpr bool QUARTIC = KERNEL == 0 && !COUNT && !ADJ;
  static_assert(!CLIP || KERNEL == 0, "the line-of-sight clip is the quartic's");
  static_assert(!ADJ || (NB == 256 && PPT <= 4), "adjoint tile table");

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  // Point ownership: the wave's 64 * PPT points, in row-major order within
  // its (64 * PPT / TW) rows of the tile; lane l owns point
  // 16 (l / 16 + 4 k) + l % 16 of them, k < PPT.
  static_assert((64 * PPT) % TW == 0, "a wave covers whole rows");
  constexpr int WROWS = (64 * PPT) / TW;
  int prow[PPT], pcol[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int pw = 16 * ((lane >> 4) + 4 * k) + (lane & 15);
    prow[k] = wave * WROWS + pw / TW;
    pcol[k] = pw % TW;
  }

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
    const uint32_t px = tx * TW + uint32_t(pcol[k]);
    const uint32_t py = ty * TH + uint32_t(prow[k]);
    pid[k] = py * p.lnx + px;
    pv[k] = px < p.lnx && py < p.lny && pid[k] < p.n;
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (pv[k]) q = p.gen ? pixel_point(p.pg, px, py) : p.pts[pid[k]];
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
      // waves without a valid point must not narrow the intersection
      if (!av) {
        sm.redi[wave][4] = INT32_MIN;
        sm.redi[wave][5] = INT32_MIN;
        sm.redi[wave][6] = INT32_MAX;
        sm.redi[wave][7] = INT32_MAX;
        sm.redi[wave][9] = 1;
      }
      sm.wcnt[0][wave] = av;
    }
  }
  __syncthreads();
  float tbx0 = sm.redf[0][0], tby0 = sm.redf[0][1], tbx1 = sm.redf[0][2], tby1 = sm.redf[0][3];
  int r_ulx = sm.redi[0][0], r_uly = sm.redi[0][1], r_uhx = sm.redi[0][2], r_uhy = sm.redi[0][3];
  int r_ilx = sm.redi[0][4], r_ily = sm.redi[0][5], r_ihx = sm.redi[0][6], r_ihy = sm.redi[0][7];
  int r_anyz = sm.redi[0][8], r_allz = sm.redi[0][9];
  int any_valid = sm.wcnt[0][0];
  for (int w = 1; w < NW; ++w) {
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
    any_valid |= sm.wcnt[0][w];
  }
  __syncthreads();
  if (!uniform(any_valid)) return;

  if constexpr (ADJ) {
    // the tile's points with their incoming gradients (0 outside the image),
    // then the bounding box of the points with a nonzero gradient in each
    // 8x8 block of the tile (one wave per block)
    static_assert(TW == 32 && TH == 32, "adjoint tile is 32 x 32");
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const float gk = pv[k] ? adj->grad[pid[k]] * adj->scale : 0.0f;
      sm.Pt[prow[k] * TW + pcol[k]] = make_float4(ex[k], ey[k], gk, 0.0f);
    }
    __syncthreads();
    for (int sb = wave; sb < 16; sb += 4) {
      const float4 q = sm.Pt[((sb >> 2) * 8 + (lane >> 3)) * TW + (sb & 3) * 8 + (lane & 7)];
      const bool on = q.z != 0.0f;
      const float x0 = wmin(on ? q.x : INFINITY), y0 = wmin(on ? q.y : INFINITY);
      const float x1 = wmax(on ? q.x : -INFINITY), y1 = wmax(on ? q.y : -INFINITY);
      if (lane == 0) sm.Sb[sb] = make_float4(x0, y0, x1, y1);
    }
    __syncthreads();
    int live = 0;
    for (int sb = 0; sb < 16; ++sb) live |= sm.Sb[sb].x <= sm.Sb[sb].z ? 1 : 0;
    if (!uniform(live)) return;
  }

  // tile centre and the points relative to it (big-particle expansion)
  const float tcx = 0.5f * (tbx0 + tbx1), tcy = 0.5f * (tby0 + tby1);
  float lx[PPT], ly[PPT], l2[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    lx[k] = ex[k] - tcx;
    ly[k] = ey[k] - tcy;
    l2[k] = fmaf(ly[k], ly[k], lx[k] * lx[k]);
  }
  // a particle is "big" when h >= 2 * tile half-diagonal
  const float hd_x = 0.5f * (tbx1 - tbx0), hd_y = 0.5f * (tby1 - tby0);
  const float big_r2 = 4.0f * fmaf(hd_y, hd_y, hd_x * hd_x) * 1.0001f + 1e-30f;

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
  constexpr int kPoolN = REF ? kPool : kCap;

  // list sizes (identical in every thread).  The big list holds the
  // particles whose support covers the whole tile (argument positive at every
  // point, no clamp) from the front of Lb, the others from the back.
  int nb = 0, nbp = 0, ns = 0, nm = 0;
  int slot = 0;

  // ---- evaluation of the lists (loop bounds are wave-uniform)
  auto eval_big = [&](int cnt, int cntp) {
    if constexpr (QUARTIC) {
      auto run = [&](int base, int n, auto clamp) {
        n = uniform(n);
        for (int j0 = 0; j0 < n; j0 += 32) {
          const int jn = min(32, n - j0);
          float a[PPT];
#pragma unroll
          for (int k = 0; k < PPT; ++k) a[k] = 0.0f;
          int j = 0;
          for (; j + 4 <= jn; j += 4) {
            float4 q[4];
            float2 z[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              q[u] = sm.Lb[base + j0 + j + u];
              if constexpr (CLIP) z[u] = sm.Lbz[base + j0 + j + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
              for (int k = 0; k < PPT; ++k) {
                const float v = fmaf(q[u].y, lx[k], fmaf(q[u].z, ly[k], fmaf(q[u].w, l2[k], q[u].x)));
                if constexpr (CLIP)
                  a[k] += clip_pow(v, z[u]);
                else
                  a[k] += decltype(clamp)::value ? pow25(v) : pow25_pos(v);
              }
          }
          for (; j < jn; ++j) {
            const float4 q = sm.Lb[base + j0 + j];
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
              const float v = fmaf(q.y, lx[k], fmaf(q.z, ly[k], fmaf(q.w, l2[k], q.x)));
              if constexpr (CLIP)
                a[k] += clip_pow(v, sm.Lbz[base + j0 + j]);
              else
                a[k] += decltype(clamp)::value ? pow25(v) : pow25_pos(v);
            }
          }
#pragma unroll
          for (int k = 0; k < PPT; ++k) a64[k] += double(a[k]);
        }
      };
      run(0, cnt, std::false_type{});
      run(kCap - cntp, cntp, std::true_type{});
    }
  };
  auto eval_small = [&](int cnt) {
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
          for (int u = 0; u < 4; ++u) q[u] = sm.Ls[j0 + j + u];
          float2 z[4];
          if constexpr (CLIP) {
#pragma unroll
            for (int u = 0; u < 4; ++u) z[u] = sm.Lsz[j0 + j + u];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
              const float dx = q[u].x - ex[k], dy = q[u].y - ey[k];
              const float r2 = fmaf(dy, dy, dx * dx);
              if constexpr (KERNEL == 2)
                a[k] = fmaf(q[u].z, cubic_shape(r2 * q[u].w, sm.tab), a[k]);
              else if constexpr (CLIP)
                a[k] += clip_pow(fmaf(-r2, q[u].z, q[u].w), z[u]);
              else
                a[k] += pow25(fmaf(-r2, q[u].z, q[u].w));
            }
          }
        }
      }
      for (; j < jn; ++j) {
        const float4 q = sm.Ls[j0 + j];
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
            if constexpr (KERNEL == 2)
              a[k] = fmaf(q.z, cubic_shape(r2 * q.w, sm.tab), a[k]);
            else if constexpr (CLIP)
              a[k] += clip_pow(fmaf(-r2, q.z, q.w), sm.Lsz[j0 + j]);
            else
              a[k] += pow25(fmaf(-r2, q.z, q.w));
          }
        }
      }
      if constexpr (!COUNT) {
#pragma unroll
        for (int k = 0; k < PPT; ++k) a64[k] += double(a[k]);
      }
    }
  };
  // masked entries: each point tests the entry's cell against its own search
  // box, as unsigned(cx - x0) <= width (and the same for y) on the box
  // clamped to the cells a list entry can have (0..65535 per axis; an empty
  // box or a point outside the grid layer fails every cell).  Four entries and
  // their cells are read at once.
  auto eval_masked = [&](int cnt) {
    if constexpr (REF) {
      cnt = uniform(cnt);
      int bx0[PPT], by0[PPT];
      uint32_t bxw[PPT], byw[PPT];
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        const int x0 = max(lox[k], 0), x1 = min(hix[k], 65535);
        const int y0 = max(loy[k], 0), y1 = min(hiy[k], 65535);
        const bool ok = zin[k] && x0 <= x1 && y0 <= y1;
        bx0[k] = ok ? x0 : INT32_MAX;
        by0[k] = y0;
        bxw[k] = ok ? uint32_t(x1 - x0) : 0u;
        byw[k] = ok ? uint32_t(y1 - y0) : 0u;
      }
      auto add = [&](float* a, const float4& q, int c) {
        const int cx = c & 0xffff, cy = int(uint32_t(c) >> 16);
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          const bool in = uint32_t(cx - bx0[k]) <= bxw[k] && uint32_t(cy - by0[k]) <= byw[k];
          const float dx = q.x - ex[k], dy = q.y - ey[k];
          const float r2 = fmaf(dy, dy, dx * dx);
          a[k] = fmaf(in ? 1.0f : 0.0f, pow25(fmaf(-r2, q.z, q.w)), a[k]);
        }
      };
      for (int j0 = 0; j0 < cnt; j0 += 32) {
        const int jn = min(32, cnt - j0);
        float a[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) a[k] = 0.0f;
        int j = 0;
        for (; j + 4 <= jn; j += 4) {
          const int4 cc = *reinterpret_cast<const int4*>(&sm.Lmc[j0 + j]);
          float4 q[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) q[u] = sm.Ls[kPoolN - 1 - (j0 + j + u)];
          add(a, q[0], uniform(cc.x));
          add(a, q[1], uniform(cc.y));
          add(a, q[2], uniform(cc.z));
          add(a, q[3], uniform(cc.w));
        }
        for (; j < jn; ++j) add(a, sm.Ls[kPoolN - 1 - (j0 + j)], uniform(sm.Lmc[j0 + j]));
#pragma unroll
        for (int k = 0; k < PPT; ++k) a64[k] += double(a[k]);
      }
    }
  };

  // adjoint: one listed particle per thread.  The thread visits the 8x8
  // point blocks its support reaches and sums, over their points, five
  // moments of g = dL/dvalue * s^1.5 (s = 1 - r^2/h^2); the weights are
  // recomputed, not stored.  One atomic add per gradient component per
  // particle and tile:
  //   dL/dw = K/h^2        sum g s
  //   dL/dh = w K/h^3      sum g (7 r^2/h^2 - 2)
  //   dL/dx = -5 w K/h^4   sum g (x - x_point)     (same for y)
  auto eval_adjoint = [&](int cnt) {
    if constexpr (ADJ) {
      cnt = uniform(cnt);
      for (int j0 = 0; j0 < cnt; j0 += NB) {
        const int j = j0 + t;
        const bool act = j < cnt;
        const float4 q = act ? sm.Ls[j] : make_float4(0.f, 0.f, 0.f, 0.f);  // (x, y, 1/h^2, coef)
        float A0 = 0.f, A1 = 0.f, A2 = 0.f, A3 = 0.f, A4 = 0.f;
        for (int sb = 0; sb < 16; ++sb) {
          const float4 bx = sm.Sb[sb];
          const float bdx = fmaxf(fmaxf(bx.x - q.x, q.x - bx.z), 0.0f);
          const float bdy = fmaxf(fmaxf(bx.y - q.y, q.y - bx.w), 0.0f);
          if (!(act && fmaf(bdy, bdy, bdx * bdx) * q.z < cull2)) continue;
          const int base = (sb >> 2) * 8 * TW + (sb & 3) * 8;
          for (int rr = 0; rr < 8; ++rr) {
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) {
              const float4 pt = sm.Pt[base + rr * TW + cc];
              const float dx = q.x - pt.x, dy = q.y - pt.y;
              const float r2 = fmaf(dy, dy, dx * dx);
              if constexpr (KERNEL == 2) {
                // F and dF/du at u = r^2/h^2, linear in the tables (0 beyond u = 4)
                const float uu = r2 * q.z;
                const float xx =
                    fminf(uu * (float(kCubicTable - 1) / 4.0f), float(kCubicTable - 1));
                const int ti = min(int(xx), kCubicTable - 2);
                const float tf = xx - float(ti);
                const float F = fmaf(tf, sm.tab[ti + 1] - sm.tab[ti], sm.tab[ti]);
                const float D = fmaf(tf, sm.dtab[ti + 1] - sm.dtab[ti], sm.dtab[ti]);
                const float gd = pt.z * D;
                A0 = fmaf(pt.z, F, A0);
                A1 = fmaf(gd, uu, A1);
                A3 = fmaf(gd, dx, A3);
                A4 = fmaf(gd, dy, A4);
              } else {
                const float sv = fmaxf(fmaf(-r2, q.z, 1.0f), 0.0f);
                const float gq = pt.z * (sv * __builtin_amdgcn_sqrtf(sv));
                A0 = fmaf(gq, sv, A0);
                A1 += gq;
                A2 = fmaf(gq, r2, A2);
                A3 = fmaf(gq, dx, A3);
                A4 = fmaf(gq, dy, A4);
              }
            }
          }
        }
        if (act) {
          const uint32_t si = sm.Lidx[j];
          if constexpr (KERNEL == 2) {
            // value = w K F(u) / h^2: d/dh = -2 w K / h^3 (F + u F'),
            // d/dx = 2 w K / h^4 F' (x - x_point)
            const float c2 = 2.0f * q.w * q.z;
            atomicAdd(adj->gw + si, adj->norm * q.z * A0);
            atomicAdd(adj->gh + si, -2.0f * q.w * __builtin_amdgcn_sqrtf(q.z) * (A0 + A1));
            atomicAdd(adj->gx + si, c2 * A3);
            atomicAdd(adj->gy + si, c2 * A4);
          } else {
            const float c5 = -5.0f * q.w * q.z;  // -5 w K / h^4
            atomicAdd(adj->gw + si, adj->norm * q.z * A0);
            atomicAdd(adj->gh + si,
                      q.w * __builtin_amdgcn_sqrtf(q.z) * fmaf(7.0f * q.z, A2, -2.0f * A1));
            atomicAdd(adj->gx + si, c5 * A3);
            atomicAdd(adj->gy + si, c5 * A4);
          }
        }
      }
    }
  };

  // ---- candidates.  The particles of the cells cx0..cx1 of one row (y, z)
  // are contiguous in the sorted order, so the box's particles form nrows *
  // nlays ranges.  Dense grids visit the box's cells 256 at a time (x
  // fastest) and skip the cells whose max h cannot reach the tile; sparse
  // grids (well under one particle per cell, e.g. the depth-cut lists of a
  // slab) stream the row ranges and cull per particle.  Both give the same
  // candidates in the same order (rows, cells by x, particles in sorted
  // order).  A sparse grid can still be dense where the tile is (a halo
  // inside a slab): when the first chunk of rows holds more than
  // kDenseRows particles per cell of the box, the tile switches to the cell
  // enumeration, whose max-h skip and prefix search read far fewer.
  const int ncols = cx1 >= cx0 ? cx1 - cx0 + 1 : 0;
  const int nrows = cy1 >= cy0 ? cy1 - cy0 + 1 : 0;
  const int nlays = cz1 >= cz0 ? cz1 - cz0 + 1 : 0;
  // the box lies inside the grid, whose cells have 32-bit keys: 32-bit
  // counts and indices (one layer, as in reference mode, skips a division)
  const uint32_t ncand = uint32_t(ncols) * uint32_t(nrows) * uint32_t(nlays);
  const uint32_t nrowz = ncols > 0 ? uint32_t(nrows) * uint32_t(nlays) : 0u;
  const uint32_t nxny = uint32_t(g.nc[0]) * uint32_t(g.nc[1]);
  const float sup_h2 = support * support * 1.00002f;
  bool rows_mode = g.sparse != 0 && ncand > uint32_t(NB);
  const bool one_layer = nlays == 1;
  uint64_t nitems = rows_mode ? nrowz : ncand;
  for (uint64_t c0 = 0; c0 < nitems; c0 += NB) {
    const bool in_box = c0 + uint64_t(t) < nitems;
    const uint32_t ci = uint32_t(c0) + uint32_t(t);
    uint32_t cb = 0, cn = 0;
    int ccx = 0, ccy = 0, cfull = 0;
    if (rows_mode) {
      if (in_box) {
        const uint32_t lay = one_layer ? 0u : ci / uint32_t(nrows);
        ccy = cy0 + int(ci - lay * uint32_t(nrows));
        const uint32_t kb = uint32_t(ccy) * uint32_t(g.nc[0]) + uint32_t(cz0 + int(lay)) * nxny;
        cb = g.start[kb + uint32_t(cx0)];
        cn = g.start[kb + uint32_t(cx1) + 1] - cb;
        ccx = int(kb);  // stored as the row's base key
      }
    } else if (in_box) {
      const uint32_t rest = ci / uint32_t(ncols);
      const int col_c = int(ci - rest * uint32_t(ncols));
      const uint32_t lay = one_layer ? 0u : rest / uint32_t(nrows);
      ccx = cx0 + col_c;
      ccy = cy0 + int(rest - lay * uint32_t(nrows));
      const int ccz = cz0 + int(lay);
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
        const float d2c = dx * dx + dy * dy;
        if (d2c < mh * mh * sup_h2) {
          // the cell's particles come by descending h (to the top 16 bits of
          // its pattern): only a prefix can reach the tile.  With h rounded
          // up to its 16-bit bucket the test is monotone along the cell.
          uint32_t lo = 0, hi = ce - cb;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            const float hu = __uint_as_float(__float_as_uint(g.hs[cb + mid]) | 0xffffu);
            if (d2c < hu * hu * sup_h2)
              lo = mid + 1;
            else
              hi = mid;
          }
          cn = lo;
        }
        if constexpr (REF)
          cfull = (allz_u && ccy >= r_ily && ccy <= r_ihy && ccx >= r_ilx && ccx <= r_ihx) ? 1 : 0;
      }
    }
    uint32_t tot;
    const uint32_t coff = block_scan_excl<NB>(cn, &tot, sm.scan_ws);
    tot = uint32_t(uniform(int(tot)));
    if (rows_mode && c0 == 0 && uint64_t(tot) > uint64_t(kDenseRows) * ncand) {
      // nothing of this chunk is in LDS yet: start over with the cells
      rows_mode = false;
      nitems = ncand;
      c0 = uint64_t(0) - uint64_t(NB);
      continue;
    }
    sm.cell_b[t] = cb;
    sm.cell_off[t] = coff;
    if (rows_mode) {
      sm.cell_info[t] = ccx;
      sm.cell_y[t] = ccy;
    } else {
      sm.cell_info[t] = int(uint32_t(ccx) | (uint32_t(ccy) << 16));
      sm.cell_y[t] = cfull;
    }
    __syncthreads();
    for (uint32_t v0 = 0; v0 < tot; v0 += NB) {
      const uint32_t v = v0 + uint32_t(t);
      float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
      bool keep = false, full = false, big = false, cover = false;
      int cinf = 0, cf = 0;
      uint32_t adj_idx = 0;
      float2 tcl = make_float2(0.0f, 0.0f);
      if (v < tot) {
        // last cell whose offset is <= v (it has a nonzero count)
        int lo = 0;
#pragma unroll
        for (int step = NB / 2; step > 0; step >>= 1)
          if (sm.cell_off[lo + step] <= v) lo += step;
        const uint32_t sidx = sm.cell_b[lo] + (v - sm.cell_off[lo]);
        cinf = sm.cell_info[lo];
        if constexpr (REF) {
          if (!rows_mode) cf = sm.cell_y[lo];
          if (rows_mode) {
            // the particle's cell from its key and the row's base key
            const int px_c = int(g.key[sidx] - uint32_t(cinf));
            const int py_c = sm.cell_y[lo];
            const int pf = (allz_u && py_c >= r_ily && py_c <= r_ihy && px_c >= r_ilx &&
                            px_c <= r_ihx) ? 1 : 0;
            cinf = int(uint32_t(px_c) | (uint32_t(py_c) << 16));
            cf = pf;
          }
        }
        // (x, y, coef, 1/h^2); the quartic's z is c = coef^0.4
        const float4 q = g.part[sidx];
        if (ADJ || q.z != 0.0f) {
          const float dx = fmaxf(fmaxf(tbx0 - q.x, q.x - tbx1), 0.0f);
          const float dy = fmaxf(fmaxf(tby0 - q.y, q.y - tby1), 0.0f);
          const float d2 = fmaf(dy, dy, dx * dx);
          keep = d2 * q.w < cull2;
        }
        if constexpr (CLIP) {
          // the depth range relative to the particle, in units of h: its
          // support must overlap it
          const float2 zc = g.clip[sidx];
          keep = keep && zc.x < 1.0f && zc.y > -1.0f;
          const float rc = __builtin_amdgcn_sqrtf(q.z);
          tcl = make_float2(zc.x * rc, zc.y * rc);
        }
        full = keep && (!REF || cf);
        if (keep) {
          if constexpr (COUNT) {
            e = make_float4(q.x, q.y, g.hs[sidx], 0.0f);
          } else if constexpr (ADJ) {
            // (x, y, 1/h^2, coef)
            e = make_float4(q.x, q.y, q.w, KERNEL == 2 ? q.z : pow25_pos(q.z));
            adj_idx = sidx;
          } else if constexpr (KERNEL == 2) {
            e = q;
          } else {
            const float c = q.z;        // coef^0.4
            const float ci2 = c * q.w;  // c / h^2
            big = full && big_r2 * q.w <= 1.0f;
            if (big) {
              const float xr = q.x - tcx, yr = q.y - tcy;
              const float d2r = fmaf(yr, yr, xr * xr);
              e = make_float4(c * fmaf(-q.w, d2r, 1.0f), 2.0f * ci2 * xr, 2.0f * ci2 * yr, -ci2);
              // s >= 0.01 at every point of the tile box: far beyond the
              // rounding of the expanded argument (~1e-6 of c)
              const float fx = fmaxf(q.x - tbx0, tbx1 - q.x), fy = fmaxf(q.y - tby0, tby1 - q.y);
              cover = fmaf(fy, fy, fx * fx) * q.w < 0.99f;
            } else {
              e = make_float4(q.x, q.y, ci2, c);
            }
          }
        }
      }
      const uint64_t bb = __ballot(big && cover);
      const uint64_t bp = __ballot(big && !cover);
      const uint64_t bs = __ballot(full && !big);
      const uint64_t bm = __ballot(keep && !full);
      if (lane == 0) {
        sm.wcnt[slot][wave] = __popcll(bb);
        sm.wcnt[slot][4 + wave] = __popcll(bs);
        sm.wcnt[slot][8 + wave] = __popcll(bm);
        sm.wcnt[slot][12 + wave] = __popcll(bp);
      }
      __syncthreads();
      int baseb = nb, basebp = nbp, bases = ns, basem = nm, totb = 0, totbp = 0, tots = 0,
          totm = 0;
      for (int w = 0; w < NW; ++w) {
        const int cb_ = sm.wcnt[slot][w], cs_ = sm.wcnt[slot][4 + w],
                  cm_ = sm.wcnt[slot][8 + w], cp_ = sm.wcnt[slot][12 + w];
        if (w < wave) {
          baseb += cb_;
          basebp += cp_;
          bases += cs_;
          basem += cm_;
        }
        totb += cb_;
        totbp += cp_;
        tots += cs_;
        totm += cm_;
      }
      if constexpr (QUARTIC) {
        // nb + nbp <= kFlush - 1 + NB <= kCap: the two ends never meet
        if (big && cover) {
          const int pb = baseb + int(prefix_count(bb));
          sm.Lb[pb] = e;
          if constexpr (CLIP) sm.Lbz[pb] = tcl;
        }
        if (big && !cover) {
          const int pc = kCap - 1 - (basebp + int(prefix_count(bp)));
          sm.Lb[pc] = e;
          if constexpr (CLIP) sm.Lbz[pc] = tcl;
        }
      }
      if (full && !big) {
        const int pos = bases + prefix_count(bs);
        sm.Ls[pos] = e;
        if constexpr (ADJ) sm.Lidx[pos] = adj_idx;
        if constexpr (CLIP) sm.Lsz[pos] = tcl;
      }
      if constexpr (REF) {
        if (keep && !full) {
          const int pos = basem + prefix_count(bm);
          sm.Ls[kPool - 1 - pos] = e;
          sm.Lmc[pos] = cinf;
        }
      }
      nb += totb;
      nbp += totbp;
      ns += tots;
      nm += totm;
      slot ^= 1;
      if (nb + nbp >= kFlush || ns >= kFlush || nm >= kFlush) {
        __syncthreads();
        if constexpr (ADJ) {
          eval_adjoint(ns);
        } else {
          eval_big(nb, nbp);
          eval_small(ns);
          eval_masked(nm);
        }
        __syncthreads();
        nb = nbp = ns = nm = 0;
      }
    }
    __syncthreads();  // the cell table is rewritten by the next chunk
  }
  if (nb > 0 || nbp > 0 || ns > 0 || nm > 0) {
    __syncthreads();
    if constexpr (ADJ) {
      eval_adjoint(ns);
    } else {
      eval_big(nb, nbp);
      eval_small(ns);
      eval_masked(nm);
    }
  }
  if constexpr (ADJ) return;

#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (!pv[k]) continue;
    if constexpr (COUNT) {
      if (counts) counts[pid[k]] += cnt_acc[k];
    } else if (fused) {
      // the only visit of this point in the launch: no read of the
      // accumulator, and the image as the retrieval computes it
      acc_out[pid[k]] = a64[k];
      img_out[pid[k]] = static_cast<float>(a64[k] * static_cast<double>(img_scale));
    } else {
      acc_out[pid[k]] += a64[k];
    }
  }
}

#ifndef ILC_WG64_WAVES
#define ILC_WG64_WAVES 5
#endif
// waves per SIMD asked of the compiler: one point per lane keeps the
// candidate pipeline's latency hidden behind many waves
constexpr int gather_waves(int ppt, int nb) {
  return ppt <= 1 ? 8 : (nb == 64 ? ILC_WG64_WAVES : (ppt <= 2 ? 5 : 4));
}

template <int TW, int PPT, int OP, int KERNEL, int NB>
__global__ __launch_bounds__(NB) __attribute__((amdgpu_waves_per_eu(gather_waves(PPT, NB), 8))) void k_gather(grid_view g, point_view p, uint32_t tiles_x,
                                                   uint32_t ntiles, double* __restrict__ acc_out,
                                                   unsigned long long* __restrict__ counts,
                                                   const float* __restrict__ cubic_table,
                                                   float support) {
  __shared__ gather_smem<OP == OP_REFERENCE, KERNEL, false, OP == OP_CLIP, NB> sm;
  if constexpr (KERNEL == 2) {
    for (int i = threadIdx.x; i < kCubicTable; i += NB) sm.tab[i] = cubic_table[i];
  }
  // XCD-contiguous tile order: workgroups are dealt round-robin to the 8
  // XCDs, so give XCD x the x-th eighth of the tiles (shared L2 reuse of the
  // neighbouring tiles' particle spans).
  const uint32_t bid = blockIdx.x;
  const uint32_t per = ntiles / 8u;
  const uint32_t tile = bid < per * 8u ? (bid % 8u) * per + bid / 8u : bid;
  const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
  gather_tile<TW, PPT, OP, KERNEL, NB>(g, p, tx, ty, acc_out, counts, support, sm);
}

// Persistent gather over (view, tile) items.  Workgroups are dealt to the 8
// XCDs round-robin by id, and each XCD has its own L2: queue q holds the q-th
// eighth of every view's tiles (contiguous rows), so the tiles one XCD works
// on at a time are neighbours that read the same particle spans.  A
// workgroup drains its XCD's queue, then helps with the others; every queue
// counter only grows, so every workgroup leaves once all queues are empty.
template <int TW, int PPT, int OP, int KERNEL, int NB>
__global__ __launch_bounds__(NB) __attribute__((amdgpu_waves_per_eu(gather_waves(PPT, NB), 8))) void k_gather_batch(const view_desc* __restrict__ views,
                                                         const uint2* __restrict__ items,
                                                         queue_segs segs,
                                                         uint32_t* __restrict__ counters,
                                                         const float* __restrict__ cubic_table,
                                                         float support) {
  __shared__ gather_smem<OP == OP_REFERENCE, KERNEL, false, OP == OP_CLIP, NB> sm;
  __shared__ uint32_t next;
  if constexpr (KERNEL == 2) {
    for (int i = threadIdx.x; i < kCubicTable; i += NB) sm.tab[i] = cubic_table[i];
  }
  const uint32_t home = blockIdx.x % uint32_t(kQueues);
  for (uint32_t k = 0; k < uint32_t(kQueues); ++k) {
    const uint32_t q = (home + k) % uint32_t(kQueues);
    const uint32_t qb = segs.seg[q], qe = segs.seg[q + 1];
    for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) next = atomicAdd(counters + q * kQueueStride, 1u);
      __syncthreads();
      const uint32_t got = uint32_t(uniform(int(next)));
      if (got >= qe - qb) break;
      const uint2 it = items[qb + got];
      const uint32_t vi = uint32_t(uniform(int(it.x))), ti = uint32_t(uniform(int(it.y)));
      const view_desc& vd = views[vi];
      const uint32_t tiles_x = vd.tiles_x;
      gather_tile<TW, PPT, OP, KERNEL, NB>(vd.g, vd.p, ti % tiles_x, ti / tiles_x, vd.acc,
                                           nullptr, support, sm, nullptr, vd.img, vd.scale,
                                           vd.fused);
    }
  }
}

// Adjoint of the exact-mode gather over one point image (quartic or the
// tabulated cubic spline).
template <int TW, int PPT, int KERNEL>
__global__ __launch_bounds__(kBlock) void k_adjoint(grid_view g, point_view p, uint32_t tiles_x,
                                                    uint32_t ntiles, adjoint_io io) {
  __shared__ gather_smem<false, KERNEL, true> sm;
  if constexpr (KERNEL == 2) {
    for (int i = threadIdx.x; i < kCubicTable; i += kBlock) {
      sm.tab[i] = io.ftab[i];
      sm.dtab[i] = io.dtab[i];
    }
  }
  const uint32_t bid = blockIdx.x;
  const uint32_t per = ntiles / 8u;
  const uint32_t tile = bid < per * 8u ? (bid % 8u) * per + bid / 8u : bid;
  gather_tile<TW, PPT, OP_ADJOINT, KERNEL>(g, p, tile % tiles_x, tile / tiles_x, nullptr, nullptr,
                                           KERNEL == 2 ? 2.0f : 1.0f, sm, &io);
}

// gradients in sorted order -> input order
__global__ __launch_bounds__(kBlock) void k_unpermute4(const uint32_t* __restrict__ perm, uint32_t n,
                                                        const float* __restrict__ a,
                                                        const float* __restrict__ b,
                                                        const float* __restrict__ c,
                                                        const float* __restrict__ d, float* oa,
                                                        float* ob, float* oc, float* od) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t i = perm[s];
  oa[i] = a[s];
  ob[i] = b[s];
  oc[i] = c[s];
  od[i] = d[s];
}

template <int TW, int PPT, int OP, int KERNEL, int NB>
struct plain_launcher {
  static int run(const grid_view& g, const point_view& p, double* acc,
                 unsigned long long* counts, const float* tab, float support, hipStream_t s) {
    constexpr int TH = (NB / TW) * PPT;
    const uint32_t tiles_x = (p.lnx + TW - 1) / TW;
    const uint32_t tiles_y = (p.lny + TH - 1) / TH;
    const uint32_t ntiles = tiles_x * tiles_y;
    if (!ntiles) return 0;
    hipLaunchKernelGGL((k_gather<TW, PPT, OP, KERNEL, NB>), dim3(ntiles), dim3(NB), 0, s, g, p,
                       tiles_x, ntiles, acc, counts, tab, support);
    ILC_HIP(hipGetLastError());
    return 0;
  }
};

template <int TW, int PPT, int OP, int KERNEL, int NB>
struct batch_launcher {
  static int run(const view_desc* v, const uint2* it, const queue_segs& segs, uint32_t* c,
                 const float* tab, float sup, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_gather_batch<TW, PPT, OP, KERNEL, NB>), dim3(grid), dim3(NB), 0, s, v,
                       it, segs, c, tab, sup);
    ILC_HIP(hipGetLastError());
    return 0;
  }
};

template <int TW, int PPT, int OP, int KERNEL, int NB>
struct batch_occ {
  static int run() {
    int per_cu = 0;
    ILC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, k_gather_batch<TW, PPT, OP, KERNEL, NB>, NB, 0));
    return per_cu;
  }
};

// dispatch over (tile shape, op, kernel) to F<TW, PPT, OP, KERNEL, NB>::run
template <template <int, int, int, int, int> class F, class... A>
int dispatch(const tile_shape& ts, int op, int kernel, A&&... a) {
  auto by_op = [&](auto tw, auto ppt, auto nb) -> int {
    constexpr int TW = decltype(tw)::value, PPT = decltype(ppt)::value;
    constexpr int NB = decltype(nb)::value;
    if (op == OP_CLIP) {
      if (kernel == ILC_KERNEL_CUBIC) fail(ILC_EINVAL, "the line-of-sight clip needs a quartic kernel");
      return F<TW, PPT, OP_CLIP, 0, NB>::run(a...);
    }
    if (op == OP_REFERENCE) {
      if (kernel == ILC_KERNEL_CUBIC) fail(ILC_EINVAL, "reference mode uses the quartic kernel");
      return F<TW, PPT, OP_REFERENCE, 0, NB>::run(a...);
    }
    if (kernel == ILC_KERNEL_CUBIC) return F<TW, PPT, OP_EXACT, 2, NB>::run(a...);
    return F<TW, PPT, OP_EXACT, 0, NB>::run(a...);
  };
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I4 = std::integral_constant<int, 4>;
  using W64 = std::integral_constant<int, 64>;
  using W256 = std::integral_constant<int, 256>;
  if (ts.nb == 256) {
    if (ts.tw == 16 && ts.ppt == 1) return by_op(I16{}, I1{}, W256{});
    if (ts.tw == 32 && ts.ppt == 2) return by_op(I32{}, I2{}, W256{});
    if (ts.tw == 32 && ts.ppt == 4) return by_op(I32{}, I4{}, W256{});
    if (ts.tw == 64 && ts.ppt == 4) return by_op(I64{}, I4{}, W256{});
  } else if (ts.nb == 64) {
    if (ts.tw == 16 && ts.ppt == 4) return by_op(I16{}, I4{}, W64{});
    if (ts.tw == 8 && ts.ppt == 1) return by_op(std::integral_constant<int, 8>{}, I1{}, W64{});
  }
  fail(ILC_EINVAL, "gather: unsupported tile shape tw=%d ppt=%d nb=%d", ts.tw, ts.ppt, ts.nb);
}

}  // namespace

void launch_adjoint(const grid_view& g, const point_view& p, int kernel, const float* grad,
                    float scale, float norm, const float* ftab, const float* dtab, float* gw,
                    float* gh, float* gx, float* gy, hipStream_t s) {
  constexpr int TW = 32, PPT = 4, TH = (kBlock / TW) * PPT;
  const uint32_t tiles_x = (p.lnx + TW - 1) / TW;
  const uint32_t ntiles = tiles_x * ((p.lny + TH - 1) / TH);
  if (!ntiles) return;
  adjoint_io io{grad, scale, norm, ftab, dtab, gw, gh, gx, gy};
  if (kernel == ILC_KERNEL_CUBIC) {
    if (!ftab || !dtab) fail(ILC_EINVAL, "cubic adjoint without its tables");
    hipLaunchKernelGGL((k_adjoint<TW, PPT, 2>), dim3(ntiles), dim3(kBlock), 0, s, g, p, tiles_x,
                       ntiles, io);
  } else {
    hipLaunchKernelGGL((k_adjoint<TW, PPT, 0>), dim3(ntiles), dim3(kBlock), 0, s, g, p, tiles_x,
                       ntiles, io);
  }
  ILC_HIP(hipGetLastError());
}

void launch_unpermute4(const uint32_t* perm, uint32_t n, const float* a, const float* b,
                       const float* c, const float* d, float* oa, float* ob, float* oc, float* od,
                       hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_unpermute4, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, perm, n, a,
                     b, c, d, oa, ob, oc, od);
  ILC_HIP(hipGetLastError());
}

void tile_grid(const tile_shape& ts, uint32_t lnx, uint32_t lny, uint32_t* tiles_x,
               uint32_t* ntiles) {
  const uint32_t th = uint32_t(ts.nb / ts.tw) * uint32_t(ts.ppt);
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

void launch_gather_batch(const view_desc* views, const uint2* items, const queue_segs& segs,
                         uint32_t* counters, int op, int kernel, const tile_shape& ts,
                         const float* cubic_table, int grid_blocks, hipStream_t s) {
  if (kernel == ILC_KERNEL_CUBIC && !cubic_table)
    fail(ILC_EINVAL, "cubic gather without its table");
  const float support = kernel == ILC_KERNEL_CUBIC ? 2.0f : 1.0f;
  // whole rounds of kQueues workgroups, so that every XCD gets the same number
  const int grid = std::max(kQueues, grid_blocks / kQueues * kQueues);
  dispatch<batch_launcher>(ts, op, kernel, views, items, segs, counters, cubic_table, support,
                           grid, s);
}

void launch_gather(const grid_view& g, const point_view& p, int op, int kernel,
                   const tile_shape& ts, double* acc, unsigned long long* counts,
                   const float* cubic_table, hipStream_t s) {
  const float support = kernel == ILC_KERNEL_CUBIC ? 2.0f : 1.0f;
  if (op == OP_COUNT) {
    // the count has its own tile shape: its cost is independent of the render's
    plain_launcher<32, 4, OP_COUNT, 0, 256>::run(g, p, nullptr, counts, nullptr, support, s);
    return;
  }
  if (kernel == ILC_KERNEL_CUBIC && !cubic_table)
    fail(ILC_EINVAL, "cubic gather without its table");
  if (op == OP_CLIP) ILC_REQUIRE(g.clip != nullptr, "clip gather without clip parameters");
  dispatch<plain_launcher>(ts, op, kernel, g, p, acc, counts, cubic_table, support, s);
}

}  // namespace ilc
