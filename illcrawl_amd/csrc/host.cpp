// Host-side pieces of the path: error state, grid parameters, camera frame.
//
// Every FP operation here is evaluated on its own (no contraction), in the
// precision the reference uses for it: double for the camera and the grid
// parameters, float for the projection matrix.
#pragma clang fp contract(off)

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace ilc {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }

const char* last_error_cstr() { return g_last_error.c_str(); }

void fail(ilc_status code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw error(code, buf);
}

// particle_grid.cpp:43-100: the bounding box (exact float extrema widened to
// double) is padded by 0.1 on every side, the cell width is the cube root of
// target/N times the padded volume, and the cell count per axis is the
// ceiling of extent / width; corner and width are then cast to float.
grid_params compute_grid_params(const bbox_result& bb, uint64_t n, uint64_t target) {
  if (n == 0) fail(ILC_EINVAL, "grid: no particles");
  double lo[3], hi[3];
  double volume = 1.0;
  for (int a = 0; a < 3; ++a) {
    lo[a] = static_cast<double>(bb.mn[a]);
    hi[a] = static_cast<double>(bb.mx[a]);
    lo[a] -= 0.1;
    hi[a] += 0.1;
    volume *= (hi[a] - lo[a]);
  }
  const double per_cell = static_cast<double>(target) / static_cast<double>(n) * volume;
  const double width = std::cbrt(per_cell);
  grid_params g;
  for (int a = 0; a < 3; ++a) {
    double cells = std::ceil((hi[a] - lo[a]) / width);
    if (!(cells >= 1.0) || cells > 2147483647.0)
      fail(ILC_ERANGE, "grid: %g cells along axis %d", cells, a);
    g.nc[a] = static_cast<int32_t>(cells);
    g.minc[a] = static_cast<float>(lo[a]);
    g.cs[a] = static_cast<float>(width);
  }
  if (g.ncells() >= (uint64_t(1) << 32))
    fail(ILC_ERANGE, "grid: %llu cells exceed 32-bit keys", (unsigned long long)g.ncells());
  return g;
}

namespace {

double dot3(const double* a, const double* b) {
  double r = 0.0;
  for (int i = 0; i < 3; ++i) r += a[i] * b[i];
  return r;
}

void cross3(const double* a, const double* b, double* o) {
  double x = a[1] * b[2] - a[2] * b[1];
  double y = a[2] * b[0] - a[0] * b[2];
  double z = a[0] * b[1] - a[1] * b[0];
  o[0] = x;
  o[1] = y;
  o[2] = z;
}

// Rodrigues rotation about a unit axis (math.hpp:107-132 layout).
void rotation(const double* k, double angle, double m[9]) {
  const double c = std::cos(angle), s = std::sin(angle), t = 1.0 - c;
  const double xy = k[0] * k[1], xz = k[0] * k[2], yz = k[1] * k[2];
  m[0] = k[0] * k[0] * t + c;
  m[1] = xy * t - k[2] * s;
  m[2] = xz * t + k[1] * s;
  m[3] = xy * t + k[2] * s;
  m[4] = k[1] * k[1] * t + c;
  m[5] = yz * t - k[0] * s;
  m[6] = xz * t - k[1] * s;
  m[7] = yz * t + k[0] * s;
  m[8] = k[2] * k[2] * t + c;
}

void unit_of(const double* m, const double* v, double* out) {
  double r[3];
  for (int i = 0; i < 3; ++i) r[i] = dot3(m + 3 * i, v);
  const double inv = 1. / std::sqrt(dot3(r, r));
  for (int i = 0; i < 3; ++i) out[i] = inv * r[i];
}

}  // namespace

camera_frame make_camera_frame(const ilc_camera& cam) {
  if (cam.nx == 0 || cam.ny == 0) fail(ILC_EINVAL, "camera: empty screen");
  if (!(cam.width > 0.0)) fail(ILC_EINVAL, "camera: width must be positive");
  camera_frame f;
  const double* la = cam.look_at;
  // screen axes (camera.cpp:124-147): u = z x look_at (y x look_at when
  // looking along +z), v = look_at x u, both rolled about look_at, normalised
  double up[3] = {0.0, 0.0, 1.0};
  if (la[0] == 0.0 && la[1] == 0.0 && la[2] == 1.0) {
    up[1] = 1.0;
    up[2] = 0.0;
  }
  double u[3], v[3], roll[9];
  cross3(up, la, u);
  cross3(la, u, v);
  rotation(la, cam.roll, roll);
  unit_of(roll, u, f.b0);
  unit_of(roll, v, f.b1);
  f.pixel_size = cam.width / static_cast<double>(cam.nx);
  // lower-left corner (camera.cpp:149-155)
  double mp[3] = {cam.position[0], cam.position[1], cam.position[2]};
  const double sx = (static_cast<double>(cam.nx) / 2.0) * f.pixel_size;
  for (int i = 0; i < 3; ++i) mp[i] -= sx * f.b0[i];
  const double sy = (static_cast<double>(cam.ny) / 2.0) * f.pixel_size;
  for (int i = 0; i < 3; ++i) mp[i] -= sy * f.b1[i];
  for (int i = 0; i < 3; ++i) {
    f.min_pos[i] = mp[i];
    f.center[i] = static_cast<float>(cam.position[i]);
    f.min_corner[i] = static_cast<float>(mp[i]);
    f.b0f[i] = static_cast<float>(f.b0[i]);
    f.b1f[i] = static_cast<float>(f.b1[i]);
  }
  f.psf = static_cast<float>(f.pixel_size);

  // FP32 inverse of M = [b0 | b1 | look_at] by cofactors over the float
  // determinant (projection.cl:24-43, math.cl:57-91).
  const float a = f.b0f[0], b = f.b1f[0], c = static_cast<float>(la[0]);
  const float d = f.b0f[1], e = f.b1f[1], ff = static_cast<float>(la[1]);
  const float g = f.b0f[2], h = f.b1f[2], i = static_cast<float>(la[2]);
  float det = 0.0f;
  det += (a * e) * i;
  det += (b * ff) * g;
  det += (c * d) * h;
  det -= (g * e) * c;
  det -= (h * ff) * a;
  det -= (i * d) * b;
  if (!(det != 0.0f)) fail(ILC_EINVAL, "camera: degenerate basis (look_at not unit?)");
  const float inv = 1.f / det;
  const float cof[9] = {e * i - ff * h, c * h - b * i, b * ff - c * e,
                        ff * g - d * i, a * i - c * g, c * d - a * ff,
                        d * h - e * g,  b * g - a * h, a * e - b * d};
  for (int k = 0; k < 9; ++k) f.minv[k] = cof[k] * inv;
  return f;
}

}  // namespace ilc

namespace ilc {

namespace {

// M4 cubic spline in 3-D with h = 1 (support 2), including 1/pi
// (smoothing_kernel.cl:64-84).
double m4(double q) {
  if (q >= 2.0) return 0.0;
  if (q <= 1.0) return (1.0 - 1.5 * q * q + 0.75 * q * q * q) / M_PI;
  const double t = 2.0 - q;
  return 0.25 * t * t * t / M_PI;
}

// 10-point Gauss-Legendre on [a, b] of m4(sqrt(b2 + z^2))
double column_piece(double b2, double a, double b) {
  static const double x[5] = {0.1488743389816312, 0.4333953941292472, 0.6794095682990244,
                              0.8650633666889845, 0.9739065285171717};
  static const double w[5] = {0.2955242247147529, 0.2692667193099963, 0.2190863625159820,
                              0.1494513491505806, 0.0666713443086881};
  const double c = 0.5 * (a + b), hw = 0.5 * (b - a);
  double s = 0.0;
  for (int k = 0; k < 5; ++k) {
    const double z0 = c - hw * x[k], z1 = c + hw * x[k];
    s += w[k] * (m4(std::sqrt(b2 + z0 * z0)) + m4(std::sqrt(b2 + z1 * z1)));
  }
  return s * hw;
}

// d/du of m4(sqrt(u + z^2)) = m4'(rho) / (2 rho), rho = sqrt(u + z^2)
double m4_du(double rho) {
  if (rho >= 2.0) return 0.0;
  if (rho <= 1.0) return (-1.5 + 1.125 * rho) / M_PI;
  const double t = 2.0 - rho;
  return -0.375 * t * t / (M_PI * rho);
}

double column_piece_du(double b2, double a, double b) {
  static const double x[5] = {0.1488743389816312, 0.4333953941292472, 0.6794095682990244,
                              0.8650633666889845, 0.9739065285171717};
  static const double w[5] = {0.2955242247147529, 0.2692667193099963, 0.2190863625159820,
                              0.1494513491505806, 0.0666713443086881};
  const double c = 0.5 * (a + b), hw = 0.5 * (b - a);
  double s = 0.0;
  for (int k = 0; k < 5; ++k) {
    const double z0 = c - hw * x[k], z1 = c + hw * x[k];
    s += w[k] * (m4_du(std::sqrt(b2 + z0 * z0)) + m4_du(std::sqrt(b2 + z1 * z1)));
  }
  return s * hw;
}

}  // namespace

// dF/du of the projected M4 column (for the adjoint), tabulated like F:
// d/du int m4(sqrt(u + z^2)) dz = int m4'(rho) / (2 rho) dz (the column's end
// moves with u but the integrand vanishes there).
std::vector<float> cubic_projected_dtable(int n) {
  std::vector<float> t(n);
  for (int i = 0; i < n; ++i) {
    const double u = 4.0 * double(i) / double(n - 1);
    if (u >= 4.0) {
      t[i] = 0.0f;
      continue;
    }
    const double z2 = std::sqrt(4.0 - u);
    const double z1 = u < 1.0 ? std::sqrt(1.0 - u) : 0.0;
    double s = 0.0;
    const int m = 8;
    for (int k = 0; k < m; ++k) {
      if (z1 > 0.0) s += column_piece_du(u, z1 * k / m, z1 * (k + 1) / m);
      s += column_piece_du(u, z1 + (z2 - z1) * k / m, z1 + (z2 - z1) * (k + 1) / m);
    }
    t[i] = static_cast<float>(2.0 * s);
  }
  return t;
}

// Projected M4 column density F(u) = int m4(sqrt(u + z^2)) dz for h = 1,
// tabulated on n nodes u_i = 4 i / (n - 1).  The integrand is a polynomial
// in z on each side of q = 1, so each piece is integrated with composite
// Gauss-Legendre.
std::vector<float> cubic_projected_table(int n) {
  std::vector<float> t(n);
  for (int i = 0; i < n; ++i) {
    const double u = 4.0 * double(i) / double(n - 1);
    if (u >= 4.0) {
      t[i] = 0.0f;
      continue;
    }
    const double z2 = std::sqrt(4.0 - u);
    const double z1 = u < 1.0 ? std::sqrt(1.0 - u) : 0.0;
    double s = 0.0;
    const int m = 8;
    for (int k = 0; k < m; ++k) {
      if (z1 > 0.0) s += column_piece(u, z1 * k / m, z1 * (k + 1) / m);
      s += column_piece(u, z1 + (z2 - z1) * k / m, z1 + (z2 - z1) * (k + 1) / m);
    }
    t[i] = static_cast<float>(2.0 * s);
  }
  return t;
}

}  // namespace ilc
