// Pixel-corner evaluation points of a camera, computed on the device.
//
// camera_generate_pixel_coordinates (camera.cl:23-41) and the projection of
// project_evaluation_points (projection.cl:97-122): the world-frame corner
// min + ix ps b0 + iy ps b1, left to right in FP32, then its camera-frame
// coordinates (dot products as (x + y) + z), no contraction.  Shared by the
// point kernel and the gathers that generate their points instead of
// loading them, so that both give the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace ilc {

__device__ inline float pixel_dot3(const float* r, float x, float y, float z) {
#pragma clang fp contract(off)
  const float a = r[0] * x;
  const float b = r[1] * y;
  const float s = a + b;
  const float d = r[2] * z;
  return s + d;
}

__device__ inline float4 pixel_point(const pixel_gen& a, uint32_t ix, uint32_t iy) {
#pragma clang fp contract(off)
  const float sx = float(ix) * a.ps;
  const float sy = float(iy) * a.ps;
  float q[3];
  for (int c = 0; c < 3; ++c) {
    const float t0 = sx * a.b0[c];
    const float t1 = a.mn[c] + t0;
    const float t2 = sy * a.b1[c];
    q[c] = t1 + t2;
  }
  const float vx = q[0] - a.c[0], vy = q[1] - a.c[1], vz = q[2] - a.c[2];
  return make_float4(pixel_dot3(a.m + 0, vx, vy, vz), pixel_dot3(a.m + 3, vx, vy, vz),
                     pixel_dot3(a.m + 6, vx, vy, vz), 0.0f);
}

}  // namespace ilc
