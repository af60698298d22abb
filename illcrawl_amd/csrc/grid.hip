// Particle-grid build on gfx950: bounding box, cell keys, stable LSD radix
// sort, cell offsets, per-cell max h, sorted SoA.
//
// Replaces particle_grid / smoothing_particle_grid (particle_grid.cpp:25-340,
// smoothing_particle_grid.cpp:29-54, particle_grid.cl:45-216, util.cl:24-49).
// The reference sorts ulong keys with boost::compute::sort_by_key (stable for
// integral keys) and then derives cell ranges; here a 8-bit-digit LSD radix
// sort on 32-bit keys gives the same permutation (each pass is a stable
// counting sort, ranks computed with wave64 ballots), and the cell ranges
// come from a scan of per-cell counts.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace ilc {

namespace {

__device__ inline float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// lanes below the calling lane
__device__ inline uint64_t lanes_below() {
  const uint32_t lane = threadIdx.x & 63u;
  return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Per-block min/max of x, y, z, max h, and a count of invalid particles.
// Non-finite coordinates or w, and h that is not a positive normal float
// whose square stays finite, are invalid.
__global__ __launch_bounds__(kBlock) void k_bbox_partial(const float4* __restrict__ p4,
                                                          const float* __restrict__ h,
                                                          uint64_t n, float* __restrict__ part,
                                                          uint32_t* __restrict__ bad) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  float mh = 0.0f;
  uint32_t nbad = 0;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock) {
    const float4 q = p4[i];
    const float hh = h[i];
    const bool ok = isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w) &&
                    hh >= 1.1e-19f && hh <= 1.8e19f;
    if (!ok) {
      ++nbad;
      continue;
    }
    mn[0] = fminf(mn[0], q.x); mx[0] = fmaxf(mx[0], q.x);
    mn[1] = fminf(mn[1], q.y); mx[1] = fmaxf(mx[1], q.y);
    mn[2] = fminf(mn[2], q.z); mx[2] = fmaxf(mx[2], q.z);
    mh = fmaxf(mh, hh);
  }
  __shared__ float red[4][7];
  __shared__ uint32_t redb[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float v[7] = {wave_min(mn[0]), wave_min(mn[1]), wave_min(mn[2]), wave_max(mx[0]),
                wave_max(mx[1]), wave_max(mx[2]), wave_max(mh)};
  uint32_t b = wave_sum(nbad);
  if (lane == 0) {
    for (int k = 0; k < 7; ++k) red[wave][k] = v[k];
    redb[wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float r[7];
    for (int k = 0; k < 7; ++k) r[k] = red[0][k];
    uint32_t bb = redb[0];
    for (int w = 1; w < 4; ++w) {
      for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], red[w][k]);
      for (int k = 3; k < 7; ++k) r[k] = fmaxf(r[k], red[w][k]);
      bb += redb[w];
    }
    for (int k = 0; k < 7; ++k) part[blockIdx.x * 8 + k] = r[k];
    if (bb) atomicAdd(bad, bb);
  }
}

__global__ __launch_bounds__(kBlock) void k_bbox_final(const float* __restrict__ part, int nparts,
                                                        const uint32_t* __restrict__ bad,
                                                        bbox_result* __restrict__ out) {
  __shared__ float red[kBlock][7];
  float r[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], part[i * 8 + k]);
    for (int k = 3; k < 7; ++k) r[k] = fmaxf(r[k], part[i * 8 + k]);
  }
  for (int k = 0; k < 7; ++k) red[threadIdx.x][k] = r[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kBlock; ++i) {
      for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], red[i][k]);
      for (int k = 3; k < 7; ++k) r[k] = fmaxf(r[k], red[i][k]);
    }
    for (int k = 0; k < 3; ++k) {
      out->mn[k] = r[k];
      out->mx[k] = r[3 + k];
    }
    out->max_h = r[6];
    out->bad = *bad;
  }
}

// Cell indices exactly as grid3d_get_cell_indices (particle_grid.cl:45-53):
// floor((p - min) / size) in FP32 with a correctly rounded division.
__device__ inline int32_t floor_to_int(float f) {
  const float fl = floorf(f);
  if (!(fl == fl)) return 0;
  if (fl >= 2147483647.0f) return INT32_MAX;
  if (fl <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(fl);
}

// keys[i] = cell of particle i (grid3d_generate_sort_keys,
// particle_grid.cl:96-112), vals[i] = i.  Keys outside the grid (the
// reference writes out of bounds there) are counted in *err and clamped.
__global__ __launch_bounds__(kBlock) void k_keys(const float4* __restrict__ p4, uint32_t n,
                                                  grid_params g, uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals,
                                                  uint32_t* __restrict__ err) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 q = p4[i];
  const float c[3] = {q.x, q.y, q.z};
  int32_t idx[3];
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    const float d = c[a] - g.minc[a];
    idx[a] = floor_to_int(d / g.cs[a]);
    ok = ok && idx[a] >= 0 && idx[a] < g.nc[a];
  }
  uint32_t key = 0;
  if (ok)
    key = uint32_t(idx[0]) + uint32_t(idx[1]) * uint32_t(g.nc[0]) +
          uint32_t(idx[2]) * uint32_t(g.nc[0]) * uint32_t(g.nc[1]);
  else
    atomicAdd(err, 1u);
  keys[i] = key;
  vals[i] = i;
}

// exclusive scan over the workgroup (kBlock threads)
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t ws[4];
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

constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;  // 4096 keys per workgroup

// The lanes of the wave whose digit equals this lane's (8 ballots), among
// the valid lanes.
__device__ inline uint64_t digit_peers(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Per-tile digit counts, one LDS add per distinct digit of a wave (a digit
// with few distinct values would serialise a wave's 64 adds on one word).
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                        uint32_t n, int shift, uint32_t ntiles,
                                                        uint32_t* __restrict__ hist) {
  __shared__ uint32_t hh[256];
  hh[threadIdx.x] = 0;
  const uint32_t base = blockIdx.x * kTile;
  uint32_t k[kItems];
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t i = base + r * kBlock + threadIdx.x;
    k[r] = i < n ? keys[i] : 0u;
  }
  __syncthreads();
  const uint64_t below = lanes_below();
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const bool valid = base + r * kBlock + threadIdx.x < n;
    const uint32_t d = (k[r] >> shift) & 255u;
    const uint64_t peers = digit_peers(d, valid);
    if (valid && (peers & below) == 0) atomicAdd(&hh[d], uint32_t(__popcll(peers)));
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = hh[threadIdx.x];
}

// Stable scatter of one 8-bit digit.  Items of a tile are visited in input
// order (round r, thread t -> item r*256 + t); within a wave, lanes sharing a
// digit are found with 8 ballots and ranked by popcount of lower lanes;
// across waves and rounds, per-digit running counts in LDS keep the order.
// The tile is first ordered by digit in LDS and then written out in that
// order, so each digit's items leave as one contiguous run: item i of the
// ordered tile, of digit d, goes to offs[d, tile] + i - (the tile's items
// below digit d).
// V2: a second value array moves with the first.
template <bool V2>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ kout,
                                                           uint32_t* __restrict__ vout, uint32_t n,
                                                           int shift, uint32_t ntiles,
                                                           const uint32_t* __restrict__ offs,
                                                           const uint32_t* __restrict__ vin2,
                                                           uint32_t* __restrict__ vout2) {
  __shared__ uint32_t sk[kTile];
  __shared__ uint32_t sv[kTile];
  __shared__ uint32_t sv2[V2 ? kTile : 1];
  __shared__ uint32_t gofs[256];
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[4][256];
  const uint32_t t = threadIdx.x;
  const int wave = t >> 6;
  const uint32_t tile0 = blockIdx.x * kTile;
  uint32_t k[kItems], v[kItems], v2[V2 ? kItems : 1];
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t i = tile0 + r * kBlock + t;
    k[r] = i < n ? kin[i] : 0u;
    v[r] = i < n ? vin[i] : 0u;
    if constexpr (V2) v2[r] = i < n ? vin2[i] : 0u;
  }
  // the tile's count of digit t from the scanned counts (the last entry's
  // successor is the total, n)
  const uint64_t e = uint64_t(t) * ntiles + blockIdx.x;
  const uint32_t g0 = offs[e];
  const uint32_t cnt = (e + 1 < uint64_t(256) * ntiles ? offs[e + 1] : n) - g0;
  uint32_t tot;
  const uint32_t ls = block_exclusive_scan(cnt, &tot);
  run[t] = ls;
  gofs[t] = g0 - ls;
  for (int w = 0; w < 4; ++w) wc[w][t] = 0;
  __syncthreads();
  const uint64_t below = lanes_below();
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const bool valid = tile0 + r * kBlock + t < n;
    const uint32_t d = (k[r] >> shift) & 255u;
    const uint64_t peers = digit_peers(d, valid);
    const uint32_t rank = __popcll(peers & below);
    if (valid && rank == 0) wc[wave][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wc[w][d];
      sk[pos] = k[r];
      sv[pos] = v[r];
      if constexpr (V2) sv2[pos] = v2[r];
    }
    __syncthreads();
    run[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
    for (int w = 0; w < 4; ++w) wc[w][t] = 0;
    __syncthreads();
  }
  const uint32_t nv = min(uint32_t(kTile), n - tile0);
  for (uint32_t i = t; i < nv; i += kBlock) {
    const uint32_t key = sk[i];
    const uint32_t pos = gofs[(key >> shift) & 255u] + i;
    kout[pos] = key;
    vout[pos] = sv[i];
    if constexpr (V2) vout2[pos] = sv2[i];
  }
}

// ---- exclusive scan of uint32 (reduce, scan of block sums, apply)


// A workgroup scans one tile of kTile words.  A whole tile of a 16-byte
// aligned array is read as kItems / 4 passes of one 16-byte vector per
// thread over consecutive 4 KB (coalesced); vector j * kBlock + t holds words
// 4 (j * kBlock + t) .. + 3 of the tile.  Other tiles read one word per thread
// and pass, the same way.
constexpr int kVecPasses = kItems / 4;
static_assert(kItems % 4 == 0, "whole vectors per thread");

__device__ inline bool whole_vec_tile(const void* p, uint64_t base, uint64_t n) {
  return base + kTile <= n && (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                         uint32_t* __restrict__ sums) {
  const uint64_t b0 = uint64_t(blockIdx.x) * kTile;
  uint32_t s = 0;
  if (whole_vec_tile(in, b0, n)) {
    const uint4* v = reinterpret_cast<const uint4*>(in + b0);
#pragma unroll
    for (int j = 0; j < kVecPasses; ++j) {
      const uint4 q = v[j * kBlock + threadIdx.x];
      s += (q.x + q.y) + (q.z + q.w);
    }
  } else {
    for (int k = 0; k < kItems; ++k) {
      const uint64_t i = b0 + uint64_t(k) * kBlock + threadIdx.x;
      if (i < n) s += in[i];
    }
  }
  s = wave_sum(s);
  __shared__ uint32_t ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(kBlock) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t nb) {
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < nb; c0 += kBlock * kItems) {
    const uint32_t base = c0 + threadIdx.x * kItems;
    uint32_t v[kItems], s = 0;
    for (int k = 0; k < kItems; ++k) {
      v[k] = base + k < nb ? sums[base + k] : 0u;
      s += v[k];
    }
    uint32_t tot;
    uint32_t off = block_exclusive_scan(s, &tot) + carry;
    for (int k = 0; k < kItems; ++k) {
      if (base + k < nb) sums[base + k] = off;
      off += v[k];
    }
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out, uint64_t n,
                                                        const uint32_t* __restrict__ sums) {
  const uint64_t b0 = uint64_t(blockIdx.x) * kTile;
  if (whole_vec_tile(in, b0, n) && (reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
    // the tile's vectors in order, kBlock at a time: a block scan of the
    // vector sums of each pass, carried over the passes
    const uint4* vin = reinterpret_cast<const uint4*>(in + b0);
    uint4* vout = reinterpret_cast<uint4*>(out + b0);
    uint4 q[kVecPasses];
#pragma unroll
    for (int j = 0; j < kVecPasses; ++j) q[j] = vin[j * kBlock + threadIdx.x];
    uint32_t carry = sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kVecPasses; ++j) {
      uint32_t tot;
      const uint32_t off = block_exclusive_scan((q[j].x + q[j].y) + (q[j].z + q[j].w), &tot) + carry;
      uint4 o;
      o.x = off;
      o.y = o.x + q[j].x;
      o.z = o.y + q[j].y;
      o.w = o.z + q[j].z;
      vout[j * kBlock + threadIdx.x] = o;
      carry += tot;
    }
    return;
  }
  const uint64_t base = b0 + uint64_t(threadIdx.x) * kItems;
  uint32_t v[kItems], s = 0;
  for (int k = 0; k < kItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t off = block_exclusive_scan(s, &tot) + sums[blockIdx.x];
  for (int k = 0; k < kItems; ++k) {
    if (base + k < n) out[base + k] = off;
    off += v[k];
  }
}

// Sorted SoA: part[s] = (x, y, coef, 1/h^2) of particle perm[s], hs[s] = h.
// coef = w * C / h^2 with C the kernel's normalisation, so that the kernel
// value is coef * shape(r^2 / h^2).  For the projected quartic (shape
// s^2.5, s = 1 - r^2/h^2) the z slot holds c = coef^0.4 instead: the gather
// evaluates coef s^2.5 as (c s)^2.5, and the power belongs to the particle,
// not to each tile that lists it.
// The same pass derives the per-cell particle counts and max h
// (grid3d_determine_cells_begin/_end and grid3d_determine_max_per_cell,
// particle_grid.cl:114-216): keys are sorted, so the particles of a cell form
// one run; each wave reduces its part of every run (count, max h) and one
// lane per run segment adds it with an atomic (a cell spans at most a few
// waves, so there is no contention).  count and maxh_bits start at 0;
// maxh_bits may be null (no max h wanted).
__global__ __launch_bounds__(kBlock) void k_build_soa(const float4* __restrict__ p4,
                                                       const float* __restrict__ h,
                                                       const uint32_t* __restrict__ perm,
                                                       const uint32_t* __restrict__ skeys,
                                                       uint32_t n, float norm, bool pow04,
                                                       float4* __restrict__ part,
                                                       float* __restrict__ hs,
                                                       uint32_t* __restrict__ count,
                                                       uint32_t* __restrict__ maxh_bits,
                                                       const float2* __restrict__ zc_in,
                                                       float2* __restrict__ zc_out) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = s < n;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t key = 0xffffffffu;
  float hh = 0.0f;
  if (valid) {
    const uint32_t i = perm[s];
    const float4 q = p4[i];
    hh = h ? h[i] : q.z;
    key = skeys[s];
    const float ih2 = 1.0f / (hh * hh);
    const float coef = q.w * (norm * ih2);
    part[s] = make_float4(q.x, q.y, pow04 ? powf(coef, 0.4f) : coef, ih2);
    hs[s] = hh;
    if (zc_out) zc_out[s] = zc_in[i];
  }
  // run heads within the wave
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(head);
  // end (exclusive) of this lane's run inside the wave
  const uint64_t above = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
  const uint32_t run_end = above ? uint32_t(__builtin_ctzll(above)) : 64u;
  // segmented max towards the run head
  float m = hh;
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const float other = __shfl_down(m, o, 64);
    if (lane + o < run_end) m = fmaxf(m, other);
  }
  if (valid && head) {
    atomicAdd(&count[key], run_end - lane);
    if (maxh_bits) atomicMax(&maxh_bits[key], __float_as_uint(m));
  }
}

// Within-cell order for the gather: descending h, to the top 16 bits of its
// float pattern (h > 0, so the pattern orders like the value).  A stable sort
// of positions by this key, then the stable sort by cell key, leaves every
// cell's particles biggest first (ties in input order); a tile then only
// needs the prefix of a cell whose h can reach it.
__global__ __launch_bounds__(kBlock) void k_h_order_keys(const float* __restrict__ h,
                                                         const uint32_t* __restrict__ perm,
                                                         uint32_t n, uint32_t* __restrict__ hk,
                                                         uint32_t* __restrict__ hv) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  hk[j] = 0x7fffu - (__float_as_uint(h[perm[j]]) >> 16);
  hv[j] = j;
}

// out_key[i] = key[order[i]], out_perm[i] = perm[order[i]]
__global__ __launch_bounds__(kBlock) void k_permute_pairs(const uint32_t* __restrict__ order,
                                                          const uint32_t* __restrict__ key,
                                                          const uint32_t* __restrict__ perm,
                                                          uint32_t n,
                                                          uint32_t* __restrict__ out_key,
                                                          uint32_t* __restrict__ out_perm) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = order[i];
  out_key[i] = key[j];
  out_perm[i] = perm[j];
}

// Cell key of every input index (by_index[perm[j]] = key[j]) and the
// identity, the input of a stable sort by cell key that gives the
// reference's order (stable by input index within a cell).
__global__ __launch_bounds__(kBlock) void k_keys_by_index(const uint32_t* __restrict__ key,
                                                          const uint32_t* __restrict__ perm,
                                                          uint32_t n,
                                                          uint32_t* __restrict__ by_index,
                                                          uint32_t* __restrict__ iota) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  by_index[perm[j]] = key[j];
  iota[j] = j;
}

// Reference layout of the grid for parity checks.
__global__ __launch_bounds__(kBlock) void k_export_cells(const uint32_t* __restrict__ start,
                                                          const uint32_t* __restrict__ maxh_bits,
                                                          uint32_t ncells, int2* cells2,
                                                          float* maxh) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= ncells) return;
  const uint32_t b = start[c], e = start[c + 1];
  if (cells2) cells2[c] = e > b ? make_int2(int(b), int(e)) : make_int2(0, 0);
  if (maxh) maxh[c] = __uint_as_float(maxh_bits[c]);
}

inline uint32_t blocks_for(uint64_t n, uint64_t per) {
  return static_cast<uint32_t>((n + per - 1) / per);
}

// Per-cell max h without the sort (a slab group's shared grid: its views
// read only the per-cell max h and the grid parameters): each particle's cell
// as k_keys computes it, and an atomic max that is only issued when the cell's
// current max is smaller (clustered cells hold many particles).  maxh_bits
// starts at 0; particles outside the grid count in err.
__global__ __launch_bounds__(kBlock) void k_cell_maxh(const float4* __restrict__ p4,
                                                      const float* __restrict__ h, uint32_t n,
                                                      grid_params g,
                                                      uint32_t* __restrict__ maxh_bits,
                                                      uint32_t* __restrict__ err) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 q = p4[i];
  const float c[3] = {q.x, q.y, q.z};
  int32_t idx[3];
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    const float d = c[a] - g.minc[a];
    idx[a] = floor_to_int(d / g.cs[a]);
    ok = ok && idx[a] >= 0 && idx[a] < g.nc[a];
  }
  if (!ok) {
    atomicAdd(err, 1u);
    return;
  }
  const uint32_t key = uint32_t(idx[0]) + uint32_t(idx[1]) * uint32_t(g.nc[0]) +
                       uint32_t(idx[2]) * uint32_t(g.nc[0]) * uint32_t(g.nc[1]);
  const uint32_t hb = __float_as_uint(h[i]);
  uint32_t* m = maxh_bits + key;
  if (hb > __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(m, hb);
}

}  // namespace

void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  uint32_t nb = blocks_for(n, kBlock);
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_fill_u32, dim3(nb), dim3(kBlock), 0, s, p, v, n);
  ILC_HIP(hipGetLastError());
}

void launch_bbox(const float4* p4, const float* h, uint64_t n, float* partial, int nparts,
                 uint32_t* bad, bbox_result* out, hipStream_t s) {
  launch_fill_u32(bad, 0u, 1, s);
  hipLaunchKernelGGL(k_bbox_partial, dim3(nparts), dim3(kBlock), 0, s, p4, h, n, partial, bad);
  ILC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, s, partial, nparts, bad, out);
  ILC_HIP(hipGetLastError());
}

void launch_bbox_final(const float* partial, int nparts, const uint32_t* bad, bbox_result* out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, s, partial, nparts, bad, out);
  ILC_HIP(hipGetLastError());
}

void launch_cell_maxh(const float4* p4, const float* h, uint32_t n, const grid_params& g,
                      uint32_t* maxh_bits, uint32_t* err, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_cell_maxh, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, p4, h, n, g,
                     maxh_bits, err);
  ILC_HIP(hipGetLastError());
}

void launch_keys(const float4* p4, uint32_t n, const grid_params& g, uint32_t* keys,
                 uint32_t* vals, uint32_t* err, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, p4, n, g, keys,
                     vals, err);
  ILC_HIP(hipGetLastError());
}

uint64_t scan_temp_words(uint64_t n) { return blocks_for(n, kTile) + 1; }

void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* temp,
                           hipStream_t s) {
  if (!n) return;
  const uint32_t nb = blocks_for(n, kTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, in, n, temp);
  ILC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, temp, nb);
  ILC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, s, in, out, n, temp);
  ILC_HIP(hipGetLastError());
}

uint64_t radix_temp_words(uint64_t n) {
  const uint64_t ntiles = blocks_for(n, kTile);
  return 2 * n + 256 * ntiles + scan_temp_words(256 * ntiles);
}

namespace {
void radix_sort_impl(uint32_t* keys, uint32_t* vals, uint32_t* vals2, uint32_t n, int key_bits,
                     uint32_t* temp, hipStream_t s) {
  if (n <= 1 || key_bits <= 0) return;
  const uint32_t ntiles = blocks_for(n, kTile);
  uint32_t* k2 = temp;
  uint32_t* v2 = temp + n;
  uint32_t* hist = temp + 2 * uint64_t(n);
  uint32_t* stemp = hist + 256 * uint64_t(ntiles);
  uint32_t* w2 = temp + radix_temp_words(n);  // the second values' other half (vals2 only)
  uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2, *wa = vals2, *wb = w2;
  int passes = (key_bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kBlock), 0, s, ka, n, shift, ntiles, hist);
    ILC_HIP(hipGetLastError());
    launch_exclusive_scan(hist, hist, uint64_t(256) * ntiles, stemp, s);
    if (vals2)
      hipLaunchKernelGGL(k_radix_scatter<true>, dim3(ntiles), dim3(kBlock), 0, s, ka, va, kb, vb, n,
                         shift, ntiles, hist, wa, wb);
    else
      hipLaunchKernelGGL(k_radix_scatter<false>, dim3(ntiles), dim3(kBlock), 0, s, ka, va, kb, vb,
                         n, shift, ntiles, hist, nullptr, nullptr);
    ILC_HIP(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
    std::swap(wa, wb);
  }
  if (ka != keys) {
    ILC_HIP(hipMemcpyAsync(keys, ka, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    ILC_HIP(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    if (vals2)
      ILC_HIP(hipMemcpyAsync(vals2, wa, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  }
}
}  // namespace

void launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t n, int key_bits, uint32_t* temp,
                       hipStream_t s) {
  radix_sort_impl(keys, vals, nullptr, n, key_bits, temp, s);
}

uint64_t radix2_temp_words(uint64_t n) { return radix_temp_words(n) + n; }

void launch_radix_sort2(uint32_t* keys, uint32_t* vals, uint32_t* vals2, uint32_t n, int key_bits,
                        uint32_t* temp, hipStream_t s) {
  ILC_REQUIRE(vals2 != nullptr, "second value array");
  radix_sort_impl(keys, vals, vals2, n, key_bits, temp, s);
}

void launch_build_soa(const float4* p4, const float* h, const uint32_t* perm,
                      const uint32_t* skeys, uint32_t n, float norm, bool pow04, float4* part, float* hs,
                      uint32_t* count, uint32_t* maxh_bits, hipStream_t s, const float2* zc_in,
                      float2* zc_out) {
  if (!n) return;
  ILC_REQUIRE((zc_in == nullptr) == (zc_out == nullptr), "clip parameters need both arrays");
  hipLaunchKernelGGL(k_build_soa, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, p4, h, perm,
                     skeys, n, norm, pow04, part, hs, count, maxh_bits, zc_in, zc_out);
  ILC_HIP(hipGetLastError());
}

void launch_h_order(const float* h, uint32_t* key, uint32_t* perm, uint32_t n, uint32_t* horder,
                    uint32_t* temp, hipStream_t s) {
  if (n <= 1) return;
  uint32_t* hk = horder;
  uint32_t* hv = horder + n;
  hipLaunchKernelGGL(k_h_order_keys, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, h, perm, n,
                     hk, hv);
  ILC_HIP(hipGetLastError());
  launch_radix_sort(hk, hv, n, 16, temp, s);
  hipLaunchKernelGGL(k_permute_pairs, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, hv, key,
                     perm, n, temp, temp + n);
  ILC_HIP(hipGetLastError());
  ILC_HIP(hipMemcpyAsync(key, temp, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  ILC_HIP(hipMemcpyAsync(perm, temp + n, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
}

void launch_reference_perm(const uint32_t* key, const uint32_t* perm, uint32_t n, int key_bits,
                           uint32_t* out_perm, uint32_t* by_index, uint32_t* temp, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_keys_by_index, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, key, perm,
                     n, by_index, out_perm);
  ILC_HIP(hipGetLastError());
  launch_radix_sort(by_index, out_perm, n, key_bits, temp, s);
}

void launch_export_cells(const uint32_t* start, const uint32_t* maxh_bits, uint32_t ncells,
                         int32_t* cells2, float* maxh, hipStream_t s) {
  if (!ncells) return;
  hipLaunchKernelGGL(k_export_cells, dim3(blocks_for(ncells, kBlock)), dim3(kBlock), 0, s, start,
                     maxh_bits, ncells, reinterpret_cast<int2*>(cells2), maxh);
  ILC_HIP(hipGetLastError());
}

}  // namespace ilc
