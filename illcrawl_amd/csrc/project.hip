// Camera projection, pixel points, quantity functors and small utilities.
//
// project_particles / project_evaluation_points (projection.cl:24-122),
// camera_generate_pixel_coordinates (camera.cl:23-41), vector_scale
// (math.cl:94-102), unprocessed_quantity / constant_quantity
// (quantities.cl:701-726) and the bolometric X-ray emissivity built from the
// reference's ion-density convention and tabulated-function sampler
// (quantities.cl:97-136, tabulated_function.cl:26-44).
//
// The FP32 arithmetic of the projection follows the reference operation by
// operation without contraction, so projected coordinates (and therefore
// the cell keys computed from them) are bit-identical to a correctly
// rounded evaluation of the reference kernels.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "pixel.hpp"

namespace ilc {

namespace {

struct proj_args {
  float m[9];
  float c[3];
};

__device__ inline float dot3(const float* r, float x, float y, float z) {
#pragma clang fp contract(off)
  const float a = r[0] * x;
  const float b = r[1] * y;
  const float s = a + b;
  const float d = r[2] * z;
  return s + d;
}

__global__ __launch_bounds__(kBlock) void k_project_particles(proj_args pa, float depth,
                                                               const float4* in, uint64_t n,
                                                               float4* out) {
#pragma clang fp contract(off)
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i >= n) return;
  const float4 q = in[i];
  const float vx = q.x - pa.c[0], vy = q.y - pa.c[1], vz = q.z - pa.c[2];
  const float px = dot3(pa.m + 0, vx, vy, vz);
  const float py = dot3(pa.m + 3, vx, vy, vz);
  const float pz = dot3(pa.m + 6, vx, vy, vz);
  const float w = (pz < 0.0f || pz > depth) ? 0.0f : q.w;
  out[i] = make_float4(px, py, 0.0f, w);
}

struct pix_args {
  proj_args pa;
  float mn[3], b0[3], b1[3];
  float ps;
};

__global__ __launch_bounds__(kBlock) void k_pixel_points(pixel_gen a, uint32_t nx, uint32_t ny,
                                                          float4* out) {
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i >= uint64_t(nx) * ny) return;
  const uint32_t iy = uint32_t(i / nx), ix = uint32_t(i - uint64_t(iy) * nx);
  out[i] = pixel_point(a, ix, iy);
}

// camera_generate_pixel_coordinates (camera.cl:23-41): world-frame corners,
// min + ix*ps*b0 + iy*ps*b1 evaluated left to right in FP32, w = 0.
__global__ __launch_bounds__(kBlock) void k_world_pixels(pix_args a, uint32_t nx, uint32_t ny,
                                                          float4* out) {
#pragma clang fp contract(off)
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i >= uint64_t(nx) * ny) return;
  const uint32_t iy = uint32_t(i / nx), ix = uint32_t(i - uint64_t(iy) * nx);
  const float sx = float(ix) * a.ps;
  const float sy = float(iy) * a.ps;
  float q[3];
  for (int c = 0; c < 3; ++c) {
    const float t0 = sx * a.b0[c];
    const float t1 = a.mn[c] + t0;
    const float t2 = sy * a.b1[c];
    q[c] = t1 + t2;
  }
  out[i] = make_float4(q[0], q[1], q[2], 0.0f);
}

// project_evaluation_points (projection.cl:97-122): z and w kept.
__global__ __launch_bounds__(kBlock) void k_project_points(proj_args pa, const float4* in,
                                                            uint64_t n, float4* out) {
#pragma clang fp contract(off)
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i >= n) return;
  const float4 q = in[i];
  const float vx = q.x - pa.c[0], vy = q.y - pa.c[1], vz = q.z - pa.c[2];
  out[i] = make_float4(dot3(pa.m + 0, vx, vy, vz), dot3(pa.m + 3, vx, vy, vz),
                       dot3(pa.m + 6, vx, vy, vz), q.w);
}

// vector_scale (math.cl:94-102): x *= alpha in FP32.
// p4[4 i + 3] = w[i]
__global__ __launch_bounds__(kBlock) void k_set_w(float4* __restrict__ p4, const float* __restrict__ w,
                                                   uint64_t n) {
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i < n) p4[i].w = w[i];
}

__global__ __launch_bounds__(kBlock) void k_vector_scale(float* x, float alpha, uint64_t n) {
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i < n) x[i] *= alpha;
}

__global__ __launch_bounds__(kBlock) void k_retrieve(const double* acc, uint64_t n, float scale,
                                                      float* out) {
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i < n) out[i] = static_cast<float>(acc[i] * static_cast<double>(scale));
}

__global__ __launch_bounds__(kBlock) void k_quantity_scaled(const float* in, double scale,
                                                             uint64_t n, float* w) {
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i >= n) return;
  w[i] = in ? static_cast<float>(scale * static_cast<double>(in[i])) : static_cast<float>(scale);
}

// stellar_mass (quantities.cl:728-750): the mass of star particles, 0 for
// wind cells (formation time <= 0)
__global__ __launch_bounds__(kBlock) void k_quantity_stellar_mass(const float* sft,
                                                                   const float* mass, uint64_t n,
                                                                   float* w) {
  const uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (i >= n) return;
  w[i] = sft[i] > 0.0f ? mass[i] : 0.0f;
}

constexpr int kMaxTable = 4096;

// read_imagef with a linear, clamp-to-border, unnormalised sampler on a 1-D
// R/float image: texel centres at i + 0.5, border value 0.
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

__global__ __launch_bounds__(kBlock) void k_xray_bolometric(const float* rho, const float* temp,
                                                             const float* xe, uint64_t n,
                                                             const float* table, int ntab,
                                                             float x0, float dx, float* w) {
#pragma clang fp contract(off)
  __shared__ float tab[kMaxTable];
  for (int i = threadIdx.x; i < ntab; i += kBlock) tab[i] = table[i];
  __syncthreads();
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock) {
    const float T = temp[i];
    float ion = 4.046263f * rho[i];
    ion = ion * xe[i];
    const float g = sample_1d(tab, ntab, x0, dx, logf(T));
    float e = ion * ion;
    e = e * sqrtf(T);
    w[i] = e * g;
  }
}

__global__ void k_fill_f64(double* p, double v, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_sum_u64(const unsigned long long* p, uint64_t n,
                                                     unsigned long long* out) {
  unsigned long long s = 0;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock)
    s += p[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

inline uint32_t nblocks(uint64_t n) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  return static_cast<uint32_t>(b ? b : 1);
}

proj_args make_proj(const camera_frame& f) {
  proj_args pa;
  for (int k = 0; k < 9; ++k) pa.m[k] = f.minv[k];
  for (int k = 0; k < 3; ++k) pa.c[k] = f.center[k];
  return pa;
}

}  // namespace

void launch_project_particles(const camera_frame& f, float depth, const float4* in, uint64_t n,
                              float4* out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_project_particles, dim3(nblocks(n)), dim3(kBlock), 0, s, make_proj(f),
                     depth, in, n, out);
  ILC_HIP(hipGetLastError());
}

pixel_gen pixel_gen_of(const camera_frame& f) {
  pixel_gen a{};
  for (int k = 0; k < 9; ++k) a.m[k] = f.minv[k];
  for (int c = 0; c < 3; ++c) {
    a.c[c] = f.center[c];
    a.mn[c] = f.min_corner[c];
    a.b0[c] = f.b0f[c];
    a.b1[c] = f.b1f[c];
  }
  a.ps = f.psf;
  return a;
}

void launch_pixel_points(const camera_frame& f, uint32_t nx, uint32_t ny, float4* out,
                         hipStream_t s) {
  const pixel_gen a = pixel_gen_of(f);
  hipLaunchKernelGGL(k_pixel_points, dim3(nblocks(uint64_t(nx) * ny)), dim3(kBlock), 0, s, a, nx,
                     ny, out);
  ILC_HIP(hipGetLastError());
}

void launch_world_pixels(const camera_frame& f, uint32_t nx, uint32_t ny, float4* out,
                         hipStream_t s) {
  if (!nx || !ny) return;
  pix_args a;
  a.pa = make_proj(f);
  for (int c = 0; c < 3; ++c) {
    a.mn[c] = f.min_corner[c];
    a.b0[c] = f.b0f[c];
    a.b1[c] = f.b1f[c];
  }
  a.ps = f.psf;
  hipLaunchKernelGGL(k_world_pixels, dim3(nblocks(uint64_t(nx) * ny)), dim3(kBlock), 0, s, a, nx,
                     ny, out);
  ILC_HIP(hipGetLastError());
}

void launch_project_points(const camera_frame& f, const float4* in, uint64_t n, float4* out,
                           hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_project_points, dim3(nblocks(n)), dim3(kBlock), 0, s, make_proj(f), in, n,
                     out);
  ILC_HIP(hipGetLastError());
}

void launch_set_w(float4* p4, const float* w, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_set_w, dim3(nblocks(n)), dim3(kBlock), 0, s, p4, w, n);
  ILC_HIP(hipGetLastError());
}

void launch_vector_scale(float* x, float alpha, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_vector_scale, dim3(nblocks(n)), dim3(kBlock), 0, s, x, alpha, n);
  ILC_HIP(hipGetLastError());
}

void launch_retrieve(const double* acc, uint64_t n, float scale, float* out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_retrieve, dim3(nblocks(n)), dim3(kBlock), 0, s, acc, n, scale, out);
  ILC_HIP(hipGetLastError());
}

void launch_quantity_scaled(const float* in, double scale, uint64_t n, float* w, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_quantity_scaled, dim3(nblocks(n)), dim3(kBlock), 0, s, in, scale, n, w);
  ILC_HIP(hipGetLastError());
}

void launch_quantity_stellar_mass(const float* sft, const float* mass, uint64_t n, float* w,
                                  hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_quantity_stellar_mass, dim3(nblocks(n)), dim3(kBlock), 0, s, sft, mass, n,
                     w);
  ILC_HIP(hipGetLastError());
}

void launch_xray_bolometric(const float* rho, const float* temp, const float* xe, uint64_t n,
                            const float* table_dev, int32_t ntab, float x0, float dx, float* w,
                            hipStream_t s) {
  if (ntab < 1 || ntab > kMaxTable) fail(ILC_EINVAL, "xray: table size %d not in [1, %d]", ntab,
                                         kMaxTable);
  if (!n) return;
  uint32_t nb = nblocks(n);
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(k_xray_bolometric, dim3(nb), dim3(kBlock), 0, s, rho, temp, xe, n, table_dev,
                     ntab, x0, dx, w);
  ILC_HIP(hipGetLastError());
}

void launch_fill_f64(double* p, double v, uint64_t n, hipStream_t s) {
  if (!n) return;
  uint32_t nb = nblocks(n);
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(kBlock), 0, s, p, v, n);
  ILC_HIP(hipGetLastError());
}

unsigned long long launch_sum_u64(const unsigned long long* p, uint64_t n,
                                  unsigned long long* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, s, reinterpret_cast<double*>(scratch), 0.0,
                     uint64_t(1));
  uint32_t nb = nblocks(n);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_sum_u64, dim3(nb), dim3(kBlock), 0, s, p, n, scratch);
  ILC_HIP(hipGetLastError());
  unsigned long long v = 0;
  ILC_HIP(hipMemcpyAsync(&v, scratch, sizeof v, hipMemcpyDeviceToHost, s));
  ILC_HIP(hipStreamSynchronize(s));
  return v;
}

}  // namespace ilc
