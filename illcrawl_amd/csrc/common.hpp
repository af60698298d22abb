// Internal helpers shared by the HIP sources and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "illcrawl_psg.h"

namespace ilc {

// Exception carried inside the library; the C ABI turns it into a status.
struct error : std::runtime_error {
  ilc_status code;
  error(ilc_status c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] void fail(ilc_status code, const char* fmt, ...);
void set_last_error(const char* msg);

#define ILC_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      ::ilc::fail(ILC_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                             \
  } while (0)

#define ILC_REQUIRE(cond, ...)                     \
  do {                                             \
    if (!(cond)) ::ilc::fail(ILC_EINVAL, __VA_ARGS__); \
  } while (0)

// Owning device buffer with capacity reuse (hipMalloc only when growing).
template <class T>
struct dbuf {
  T* p = nullptr;
  size_t cap = 0;
  dbuf() = default;
  dbuf(const dbuf&) = delete;
  dbuf& operator=(const dbuf&) = delete;
  ~dbuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  T* ensure(size_t n) {
    if (n <= cap && p) return p;
    release();
    size_t bytes = (n ? n : 1) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      fail(ILC_ENOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    cap = n;
    return p;
  }
};

// Reference grid parameters (particle_grid.cpp:43-100).
struct grid_params {
  int32_t nc[3];
  float minc[3];
  float cs[3];
  uint64_t ncells() const { return uint64_t(nc[0]) * uint64_t(nc[1]) * uint64_t(nc[2]); }
};

// Bounding box + validity reduction result of the particle pre-pass.
struct bbox_result {
  float mn[3];
  float mx[3];
  float max_h;
  uint32_t bad;  // count of particles with non-finite coordinates or h <= 0 / NaN
};

// Host part of the grid construction, in double exactly as the reference.
grid_params compute_grid_params(const bbox_result& bb, uint64_t n, uint64_t target);

// Camera math in double (camera.cpp:124-155) and the FP32 projection
// matrix (projection.cl:24-43, math.cl:57-91).
struct camera_frame {
  double b0[3], b1[3], min_pos[3], pixel_size;
  float minv[9];       // row-major inverse of [b0 | b1 | look_at] in FP32
  float center[3];     // float(position)
  float min_corner[3]; // float(min_pos)
  float b0f[3], b1f[3];
  float psf;
};
camera_frame make_camera_frame(const ilc_camera& cam);

constexpr int kBlock = 256;  // threads per workgroup (4 waves of 64)

}  // namespace ilc
