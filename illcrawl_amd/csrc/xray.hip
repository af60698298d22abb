// The reference's X-ray emissivity: thermal bremsstrahlung of a hydrogen
// plasma with the thermally averaged free-free Gaunt factor of van Hoof
// (2014) read from a 2-D table, per energy bin (xray_spectral_flux) or
// integrated over an energy band with the reference's adaptive
// Runge-Kutta-Fehlberg scheme (xray_flux): quantities.cl:39-239, 354-430,
// integration.cl:38-131, tabulated_function.cl:46-59.
//
// One thread per particle.  FP32 as the reference, evaluated operation by
// operation without contraction; exp, log, pow and rsqrt are computed in
// FP64 and rounded to FP32 (correctly rounded FP32 results: the reference
// leaves their precision to the OpenCL runtime), so that the step-size
// control takes the same decisions as a correctly rounded CPU evaluation.
// The band integration stops after kMaxSteps steps (the reference's loop is
// unbounded; every realistic input ends in a few tens of steps).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace ilc {

namespace {

constexpr float kB = 8.61733e-5f;   // Boltzmann constant, eV/K
constexpr float kRyd = 13.605f;     // Rydberg energy, eV
constexpr int kMaxSteps = 65536;

__device__ inline float cr_exp(float x) { return static_cast<float>(exp(static_cast<double>(x))); }
__device__ inline float cr_log(float x) { return static_cast<float>(log(static_cast<double>(x))); }
__device__ inline float cr_pow(float x, float y) {
  return static_cast<float>(pow(static_cast<double>(x), static_cast<double>(y)));
}
__device__ inline float cr_rsqrt(float x) {
  return static_cast<float>(1.0 / sqrt(static_cast<double>(x)));
}

__device__ inline int32_t floor_to_int(float f) {
  const float fl = floorf(f);
  if (!(fl == fl)) return 0;
  if (fl >= 2147483647.0f) return INT32_MAX;
  if (fl <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(fl);
}

struct gaunt2d {
  const float* t;  // t[i + nx j]: i along ln(gamma^2), j along ln(u)
  int nx, ny;
};

// read_imagef with a linear, unnormalised, clamp-to-border sampler
// (tabulated_function.cl:46-59): texel centres at i + 1/2, border value 0
__device__ inline float sample_2d(const gaunt2d& g, float x, float y) {
#pragma clang fp contract(off)
  const float u = (x - -6.f) / 0.2f, v = (y - -16.f) / 0.2f;
  const float us = u - 0.5f, vs = v - 0.5f;
  const float a = us - floorf(us), b = vs - floorf(vs);
  const int32_t i0 = floor_to_int(us), j0 = floor_to_int(vs);
  float t[2][2];
  for (int dj = 0; dj < 2; ++dj)
    for (int di = 0; di < 2; ++di) {
      const int64_t i = int64_t(i0) + di, j = int64_t(j0) + dj;
      t[dj][di] = (i >= 0 && i < g.nx && j >= 0 && j < g.ny) ? g.t[i + int64_t(g.nx) * j] : 0.0f;
    }
  const float wa = 1 - a, wb = 1 - b;
  float r = wa * wb;
  r = r * t[0][0];
  float s2 = a * wb;
  s2 = s2 * t[0][1];
  r = r + s2;
  float s3 = wa * b;
  s3 = s3 * t[1][0];
  r = r + s3;
  float s4 = a * b;
  s4 = s4 * t[1][1];
  return r + s4;
}

// get_temperature (quantities.cl:39-61)
__device__ inline float temperature(float u, float xe) {
#pragma clang fp contract(off)
  const float X_H = 0.76f;
  float den = 1.f + 3 * X_H;
  den = den + (4 * X_H) * xe;
  const float mu = 4.f / den;
  const float g1 = (5.f / 3.f) - 1.f;
  float t = g1 * u;
  t = t * mu;
  return t * 121.1475502f;
}

// g_ff (quantities.cl:69-83)
__device__ inline float gff(float E, float T, const gaunt2d& g) {
#pragma clang fp contract(off)
  const float Z = 1.f;
  const float kT = kB * T;
  float gam2 = Z * Z;
  gam2 = gam2 * kRyd;
  gam2 = gam2 / kT;
  float u = 1.e3f * E;
  u = u / kT;
  return sample_2d(g, cr_log(gam2), cr_log(u));
}

// get_spectral_brems_emission (quantities.cl:97-136)
__device__ inline float brems(float E, float dE, float T, float xe, float rho, float r,
                              const gaunt2d& g) {
#pragma clang fp contract(off)
  float ion = 4.046263f * rho;
  ion = ion * xe;
  const float Z = 1.f;
  const float t_inv = 1000.f / (kB * T);
  const float r2 = r * r;
  float e = 0.002535f * dE;
  e = e * Z;
  e = e * Z;
  e = e * ion;
  e = e * ion;
  e = e * gff(E, T, g);
  e = e * cr_rsqrt(T);
  e = e * cr_exp(-E * t_inv);
  return e / r2;
}

__device__ inline void rkf_points(float pos, float step, float* p) {
#pragma clang fp contract(off)
  p[0] = pos + (3.f / 8.f) * step;
  p[1] = pos + (12.f / 13.f) * step;
  p[2] = pos + step;
  p[3] = pos + 0.5f * step;
}

// integrate_spectral_flux (XRAY_FLUX) with rkf_advance as the reference
// calls it: absolute tolerance = rel_tol, relative tolerance = 1, epsilon 0,
// step sizes in [1e-3, 100] keV
__device__ float xray_band(float e_min, float e_max, float rel_tol, float T, float xe, float rho,
                           float dl, const gaunt2d& g, int* steps) {
#pragma clang fp contract(off)
  float state = 0.0f;
  float pos = e_min;
  const float end = e_max;
  float step = (end - pos) * rel_tol;
  float rb = brems(pos, 1.0f, T, xe, rho, dl, g);
  float pts[4];
  rkf_points(pos, step, pts);
  const float abs_tol = rel_tol, rel = 1.0f, eps = 0.0f, min_step = 1.e-3f, max_step = 100.f;
  int it = 0;
  while (pos < end && it < kMaxSteps) {
    ++it;
    float v[4];
    for (int k = 0; k < 4; ++k) v[k] = brems(pts[k], 1.0f, T, xe, rho, dl, g);
    float d4 = (25.f / 216.f) * rb;
    d4 = d4 + (1408.f / 2565.f) * v[0];
    d4 = d4 + (2197.f / 4101.f) * v[1];
    d4 = d4 - (1.f / 5.f) * v[2];
    float d5 = (16.f / 135.f) * rb;
    d5 = d5 + (6656.f / 12825.f) * v[0];
    d5 = d5 + (28561.f / 56430.f) * v[1];
    d5 = d5 - (9.f / 50.f) * v[2];
    d5 = d5 + (2.f / 55.f) * v[3];
    d4 = d4 * step;
    d5 = d5 * step;
    const float est4 = state + d4;
    const float old = pos;
    pos = pos + step;
    float s = 2.0f;
    if (d4 != d5) {
      const float err = fabsf(d5 - d4);
      float srel = state / pos;
      srel = fabsf(srel * rel);
      const float tol = fmaxf(abs_tol, srel);
      float q = tol * step;
      q = q / (2.f * err);
      s = cr_pow(q, 0.25f);
    }
    float nstep = s * step;
    if (nstep < min_step)
      nstep = min_step;
    else if (nstep > max_step)
      nstep = max_step;
    s = nstep / step;
    if (s < 0.95f) {
      pos = old;
    } else {
      state = est4;
      rb = v[2];
    }
    step = nstep;
    if (pos + step >= end) {
      step = end - pos;
      step = step + eps;
    }
    rkf_points(pos, step, pts);
  }
  if (steps) *steps = it;
  return state;
}

__global__ __launch_bounds__(kBlock) void k_xray_flux(const float* __restrict__ rho,
                                                       const float* __restrict__ u,
                                                       const float* __restrict__ xe, uint64_t n,
                                                       gaunt2d g, float zf, float dl, float e_min,
                                                       float e_max, float rel_tol,
                                                       float* __restrict__ out,
                                                       int32_t* __restrict__ steps) {
#pragma clang fp contract(off)
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock) {
    const float T = temperature(u[i], xe[i]);
    int st = 0;
    out[i] = xray_band(zf * e_min, zf * e_max, rel_tol, T, xe[i], rho[i], dl, g, &st);
    if (steps) steps[i] = st;
  }
}

__global__ __launch_bounds__(kBlock) void k_xray_spectral_flux(const float* __restrict__ rho,
                                                                const float* __restrict__ u,
                                                                const float* __restrict__ xe,
                                                                uint64_t n, gaunt2d g, float zf,
                                                                float dl, float E, float dE,
                                                                float* __restrict__ out) {
#pragma clang fp contract(off)
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock) {
    const float T = temperature(u[i], xe[i]);
    out[i] = brems(zf * E, zf * dE, T, xe[i], rho[i], dl, g);
  }
}

uint32_t grid_for(uint64_t n) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  if (b > 16384) b = 16384;
  return static_cast<uint32_t>(b ? b : 1);
}

}  // namespace

void launch_xray_flux(const float* rho, const float* u, const float* xe, uint64_t n,
                      const float* gaunt_dev, int nx, int ny, float z, float dl, float e_min,
                      float e_max, int num_samples, float* out, int32_t* steps, hipStream_t s) {
  if (!n) return;
  const gaunt2d g{gaunt_dev, nx, ny};
  const float zf = 1.f + z;
  const float rel_tol = 1.f / static_cast<float>(num_samples);
  hipLaunchKernelGGL(k_xray_flux, dim3(grid_for(n)), dim3(kBlock), 0, s, rho, u, xe, n, g, zf, dl,
                     e_min, e_max, rel_tol, out, steps);
  ILC_HIP(hipGetLastError());
}

void launch_xray_spectral_flux(const float* rho, const float* u, const float* xe, uint64_t n,
                               const float* gaunt_dev, int nx, int ny, float z, float dl, float E,
                               float dE, float* out, hipStream_t s) {
  if (!n) return;
  const gaunt2d g{gaunt_dev, nx, ny};
  hipLaunchKernelGGL(k_xray_spectral_flux, dim3(grid_for(n)), dim3(kBlock), 0, s, rho, u, xe, n,
                     g, 1.f + z, dl, E, dE, out);
  ILC_HIP(hipGetLastError());
}

}  // namespace ilc
