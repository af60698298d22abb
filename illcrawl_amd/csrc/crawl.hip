// Block body of the data crawler on the device: periodic correction, render
// volume cutout and order-preserving compaction of the kept rows with their
// extra columns (reconstructing_data_crawler.cpp:180-213,
// coordinate_system.hpp:24-47, reconstructing_data_crawler.cpp:59-71).
//
// The reference reads each row's coordinates as float, corrects them in
// double and stores float(x) in the particle (to_device_vector4).  Here the
// same double operations run per row on the device; the kept rows keep their
// input order (per-tile counts, a scan, then a ballot-ranked scatter), so the
// particle block handed to the grid is the reference's.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace ilc {

namespace {

constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;

// Scratch of the allocation-free block body: per-tile counts, their
// exclusive scan and the scan's block sums, for blocks of up to
// kCrawlStaticMaxRows rows (1 MiB; every GPU holds its own copy for the life
// of the process).  Host code never passes `scratch` itself to a kernel: its
// device address comes from hipGetSymbolAddress, once per device.
constexpr uint64_t kStaticTiles = kCrawlStaticMaxRows / kTile;
constexpr uint64_t kStaticWords = 2 * (kStaticTiles + 1) + 64;
static_assert(kCrawlStaticMaxRows % kTile == 0, "whole tiles");
static __device__ __align__(256) uint32_t scratch[kStaticWords];

struct crawl_params {
  double period[3];
  double gmin[3], gmax[3];  // volume grown by the border (periodic test)
  double vmin[3], vmax[3];  // volume (cutout test)
};

__device__ inline bool crawl_row(const float* xyz, uint64_t i, const crawl_params& cp,
                                 double* c) {
  bool inside = true;
  for (int j = 0; j < 3; ++j) {
    double x = double(xyz[3 * i + j]);
    const double lo = x - cp.period[j], hi = x + cp.period[j];
    if (cp.gmin[j] <= lo && lo <= cp.gmax[j])
      x = lo;
    else if (cp.gmin[j] <= hi && hi <= cp.gmax[j])
      x = hi;
    c[j] = x;
    inside = inside && !(x < cp.vmin[j]) && !(x > cp.vmax[j]);
  }
  return inside;
}

__global__ __launch_bounds__(kBlock) void k_crawl_count(const float* __restrict__ xyz, uint64_t n,
                                                         crawl_params cp,
                                                         uint32_t* __restrict__ counts) {
  __shared__ uint32_t wc[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint32_t c = 0;
  for (int r = 0; r < kItems; ++r) {
    const uint64_t i = base + uint64_t(r) * kBlock + threadIdx.x;
    double x[3];
    const bool keep = i < n && crawl_row(xyz, i, cp, x);
    c += __popcll(__ballot(keep));
  }
  if (lane == 0) wc[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(kBlock) void k_crawl_scatter(const float* __restrict__ xyz,
                                                           uint64_t n, crawl_params cp,
                                                           crawl_columns cols,
                                                           const uint32_t* __restrict__ offsets,
                                                           float4* __restrict__ out4,
                                                           uint64_t* __restrict__ index_out) {
  __shared__ uint32_t wc[2][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint64_t out = offsets[blockIdx.x];
  for (int r = 0; r < kItems; ++r) {
    const uint64_t i = base + uint64_t(r) * kBlock + threadIdx.x;
    double x[3] = {0.0, 0.0, 0.0};
    const bool keep = i < n && crawl_row(xyz, i, cp, x);
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
      const uint64_t j = out + before +
                         __builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32),
                                                   __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u));
      out4[j] = make_float4(float(x[0]), float(x[1]), float(x[2]), 0.0f);
      for (int k = 0; k < cols.n; ++k) cols.out[k][j] = cols.in[k][i];
      if (index_out) index_out[j] = i;
    }
    out += tot;
  }
}

}  // namespace

uint64_t crawl_tiles(uint64_t n) { return (n + kTile - 1) / kTile; }

uint32_t* crawl_static_scratch(uint64_t words) {
  ILC_REQUIRE(words <= kStaticWords, "crawl scratch of %llu words exceeds the static %llu",
              (unsigned long long)words, (unsigned long long)kStaticWords);
  int dev = 0;
  ILC_HIP(hipGetDevice(&dev));
  static std::mutex mu;
  static std::vector<uint32_t*> addr;  // per device
  std::lock_guard<std::mutex> lock(mu);
  if (addr.size() <= size_t(dev)) addr.resize(size_t(dev) + 1, nullptr);
  if (!addr[size_t(dev)]) {
    void* p = nullptr;
    ILC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ILC_REQUIRE(p != nullptr, "no device address for the crawl scratch on device %d", dev);
    addr[size_t(dev)] = static_cast<uint32_t*>(p);
  }
  return addr[size_t(dev)];
}

void launch_crawl(const double period[3], const double gmin[3], const double gmax[3],
                  const double vmin[3], const double vmax[3], const float* xyz, uint64_t n,
                  const crawl_columns& cols, uint32_t* counts, uint32_t* offsets,
                  uint32_t* scan_temp, float4* out4, uint64_t* index_out, hipStream_t s) {
  const uint64_t nt = crawl_tiles(n);
  if (!nt) return;
  ILC_REQUIRE(nt < (uint64_t(1) << 31), "crawl block too large");
  crawl_params cp;
  for (int j = 0; j < 3; ++j) {
    cp.period[j] = period[j];
    cp.gmin[j] = gmin[j];
    cp.gmax[j] = gmax[j];
    cp.vmin[j] = vmin[j];
    cp.vmax[j] = vmax[j];
  }
  hipLaunchKernelGGL(k_crawl_count, dim3(uint32_t(nt)), dim3(kBlock), 0, s, xyz, n, cp, counts);
  ILC_HIP(hipGetLastError());
  launch_exclusive_scan(counts, offsets, nt + 1, scan_temp, s);
  hipLaunchKernelGGL(k_crawl_scatter, dim3(uint32_t(nt)), dim3(kBlock), 0, s, xyz, n, cp, cols,
                     offsets, out4, index_out);
  ILC_HIP(hipGetLastError());
}

}  // namespace ilc
