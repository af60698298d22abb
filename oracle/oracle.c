/*
 * oracle.c -- CPU restatement of illcrawl's projective SPH smoothing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker.  It is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, and by nothing else: the product (illcrawl_amd) never links,
 * imports or calls it, and it never stands in for a missing GPU path.
 *
 * Each function names the reference file:line whose behaviour it restates
 * (paths relative to the illcrawl source tree).
 *
 * Arithmetic contract of the "reference" functions (orc_*_ref, camera,
 * projection, grid, quantities): every FP32 operation is evaluated on its
 * own in IEEE-754 binary32 round-to-nearest, with no fused multiply-add
 * (oracle/Makefile builds with -ffp-contract=off), division and sqrtf
 * correctly rounded, floorf exact.  The reference leaves three points to its
 * OpenCL runtime; this restatement fixes them as follows:
 *   - dot(a,b) of float3   = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   - distance(a,b)        = sqrtf(dot(a-b, a-b))
 *   - read_imagef, linear  = OpenCL 1.2 s8.2 formula with an unquantised
 *     weight (GPU samplers quantise it to 8 bits).
 * The FP64 functions (orc_gather_exact, orc_count_evals) compute the same
 * physical quantity in double and are the "exact" oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 2
#define ORC_ERANGE 3

/* ------------------------------------------------------------------ */
/* host double-precision vector helpers: math.hpp:68-77 (dot),         */
/* :81-89 (normalize), :91-97 (cross), :107-132 (rotation matrix),     */
/* :134-142 (matrix_vector_mult)                                      */
/* ------------------------------------------------------------------ */
static double d_dot3(const double* a, const double* b)
{
  double r = 0.0;
  for (int i = 0; i < 3; ++i)
    r += a[i] * b[i];
  return r;
}

static void d_cross3(const double* a, const double* b, double* out)
{
  double t0 = a[1] * b[2] - a[2] * b[1];
  double t1 = a[2] * b[0] - a[0] * b[2];
  double t2 = a[0] * b[1] - a[1] * b[0];
  out[0] = t0; out[1] = t1; out[2] = t2;
}

static void d_normalize3(const double* v, double* out)
{
  double linv = 1. / sqrt(d_dot3(v, v));
  for (int i = 0; i < 3; ++i)
    out[i] = linv * v[i];
}

static void d_rotation_matrix(double m[9], const double* axis, double alpha)
{
  double a1 = axis[0], a2 = axis[1], a3 = axis[2];
  double a1a2 = a1 * a2, a1a3 = a1 * a3, a2a3 = a2 * a3;
  double ca = cos(alpha), sa = sin(alpha);
  double fc = 1.0 - ca;
  m[0] = a1 * a1 * fc + ca;  m[1] = a1a2 * fc - a3 * sa; m[2] = a1a3 * fc + a2 * sa;
  m[3] = a1a2 * fc + a3 * sa; m[4] = a2 * a2 * fc + ca;  m[5] = a2a3 * fc - a1 * sa;
  m[6] = a1a3 * fc - a2 * sa; m[7] = a2a3 * fc + a1 * sa; m[8] = a3 * a3 * fc + ca;
}

static void d_matvec3(const double m[9], const double* v, double* out)
{
  double r[3];
  for (int i = 0; i < 3; ++i)
    r[i] = d_dot3(m + 3 * i, v);
  out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}

/*
 * Camera basis, pixel size and lower-left screen corner.
 * camera.cpp:27-41 (ctor: pixel_size = width/nx), :124-147
 * (update_basis_vectors), :149-155 (update_min_position).
 */
void orc_camera_setup(const double pos[3], const double look_at[3], double roll,
                      double width, uint64_t nx, uint64_t ny,
                      double b0[3], double b1[3], double min_pos[3],
                      double* pixel_size)
{
  double ps = width / (double)nx;
  double v1[3] = {0, 0, 1};
  if (look_at[0] == v1[0] && look_at[1] == v1[1] && look_at[2] == v1[2]) {
    v1[0] = 0; v1[1] = 1; v1[2] = 0;
  }
  double t[3], v2[3];
  d_cross3(v1, look_at, t);
  v1[0] = t[0]; v1[1] = t[1]; v1[2] = t[2];
  d_cross3(look_at, v1, v2);
  double roll_m[9];
  d_rotation_matrix(roll_m, look_at, roll);
  double r1[3], r2[3];
  d_matvec3(roll_m, v1, r1);
  d_matvec3(roll_m, v2, r2);
  d_normalize3(r1, b0);
  d_normalize3(r2, b1);

  double mp[3] = {pos[0], pos[1], pos[2]};
  double s0 = ((double)nx / 2.0) * ps;
  for (int i = 0; i < 3; ++i) mp[i] -= s0 * b0[i];
  double s1 = ((double)ny / 2.0) * ps;
  for (int i = 0; i < 3; ++i) mp[i] -= s1 * b1[i];
  min_pos[0] = mp[0]; min_pos[1] = mp[1]; min_pos[2] = mp[2];
  *pixel_size = ps;
}

/* ------------------------------------------------------------------ */
/* FP32 projection: math.cl:45-91, projection.cl:24-122                */
/* ------------------------------------------------------------------ */
static float f_dot3(const float* a, const float* b)
{
  float t = a[0] * b[0];
  float u = a[1] * b[1];
  float s = t + u;
  float w = a[2] * b[2];
  return s + w;
}

/*
 * Inverse of M = [b0 | b1 | n] (columns) by cofactors in FP32.
 * projection.cl:24-43 (create_projection_matrix), math.cl:57-91
 * (matrix3x3_det, matrix3x3_invert).  minv is row-major.
 */
void orc_projection_matrix(const float b0[3], const float b1[3], const float nrm[3],
                           float minv[9])
{
  float a = b0[0], b = b1[0], c = nrm[0];
  float d = b0[1], e = b1[1], f = nrm[1];
  float g = b0[2], h = b1[2], i = nrm[2];
  float det = 0.0f;
  det += (a * e) * i;
  det += (b * f) * g;
  det += (c * d) * h;
  det -= (g * e) * c;
  det -= (h * f) * a;
  det -= (i * d) * b;
  float inv = 1.f / det;
  float r[9] = {e * i - f * h, c * h - b * i, b * f - c * e,
                f * g - d * i, a * i - c * g, c * d - a * f,
                d * h - e * g, b * g - a * h, a * e - b * d};
  for (int k = 0; k < 9; ++k)
    minv[k] = r[k] * inv;
}

static void f_project(const float minv[9], const float center[3], const float* x, float* out)
{
  float v[3] = {x[0] - center[0], x[1] - center[1], x[2] - center[2]};
  out[0] = f_dot3(minv + 0, v);
  out[1] = f_dot3(minv + 3, v);
  out[2] = f_dot3(minv + 6, v);
}

/*
 * project_particles: projection.cl:56-95.  in/out are float4 (x,y,z,w).
 * A particle with camera-frame z outside [0, depth] gets w = 0; z := 0.
 */
void orc_project_particles(const float* in4, uint64_t n, const float minv[9],
                           const float center[3], float depth, float* out4)
{
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)n; ++k) {
    float p[3];
    f_project(minv, center, in4 + 4 * k, p);
    float w = in4[4 * k + 3];
    if (p[2] < 0.0f || p[2] > depth)
      w = 0.0f;
    out4[4 * k + 0] = p[0];
    out4[4 * k + 1] = p[1];
    out4[4 * k + 2] = 0.0f;
    out4[4 * k + 3] = w;
  }
}

/* The same projection keeping the camera depth z' in the z slot and w
 * unmasked (input of orc_gather_clip). */
void orc_project_particles_depth(const float* in4, uint64_t n, const float minv[9],
                                 const float center[3], float* out4)
{
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)n; ++k) {
    float p[3];
    f_project(minv, center, in4 + 4 * k, p);
    out4[4 * k + 0] = p[0];
    out4[4 * k + 1] = p[1];
    out4[4 * k + 2] = p[2];
    out4[4 * k + 3] = in4[4 * k + 3];
  }
}

/*
 * Pixel corner coordinates, then their projection:
 * camera.cl:23-41 (camera_generate_pixel_coordinates, index iy*nx+ix),
 * projection.cl:97-122 (project_evaluation_points, w kept).
 * min_corner, b0, b1: float casts of the camera's double vectors
 * (cl_types.hpp:55-62); pixel_size: float cast (camera.cpp:170).
 */
void orc_pixel_points(uint64_t nx, uint64_t ny, float pixel_size,
                      const float min_corner[3], const float b0[3], const float b1[3],
                      const float minv[9], const float center[3], float* out4)
{
#pragma omp parallel for schedule(static)
  for (int64_t iy = 0; iy < (int64_t)ny; ++iy) {
    for (uint64_t ix = 0; ix < nx; ++ix) {
      float sx = (float)ix * pixel_size;
      float sy = (float)iy * pixel_size;
      float q[3];
      for (int c = 0; c < 3; ++c) {
        float t = min_corner[c] + sx * b0[c];
        q[c] = t + sy * b1[c];
      }
      float qw = (0.0f + sx * 0.0f) + sy * 0.0f;
      float* o = out4 + 4 * (iy * nx + ix);
      f_project(minv, center, q, o);
      o[3] = qw;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Particle grid: particle_grid.cpp:43-119 (parameters, host FP64),    */
/* particle_grid.cl:45-93 (cell indices and keys, FP32)               */
/* ------------------------------------------------------------------ */

/*
 * Grid parameters from the particle bounding box.  particle_grid.cpp:43-100:
 * bbox in double over the float coordinates, +-0.1, V = prod(extent),
 * w = cbrt(target/N * V), nc = ceil(extent/w), min and w cast to float.
 */
int orc_grid_params(const float* p4, uint64_t n, uint64_t target,
                    int32_t nc[3], float minc[3], float cs[3])
{
  if (n == 0)
    return ORC_EINVAL;
  double mn[3], mx[3];
  for (int j = 0; j < 3; ++j)
    mn[j] = mx[j] = (double)p4[j];
  for (uint64_t k = 0; k < n; ++k)
    for (int i = 0; i < 3; ++i) {
      double v = (double)p4[4 * k + i];
      if (v < mn[i]) mn[i] = v;
      if (v > mx[i]) mx[i] = v;
    }
  double vol = 1.0;
  for (int i = 0; i < 3; ++i) {
    mn[i] -= 0.1;
    mx[i] += 0.1;
    vol *= (mx[i] - mn[i]);
  }
  double cell_volume = (double)target / (double)n * vol;
  double w = cbrt(cell_volume);
  for (int i = 0; i < 3; ++i) {
    nc[i] = (int32_t)ceil((mx[i] - mn[i]) / w);
    minc[i] = (float)mn[i];
    cs[i] = (float)w;
  }
  return ORC_OK;
}

/* (int)floor(x) with the out-of-range cases saturated (the reference's
 * conversion is implementation-defined there). */
static int32_t f_floor_to_int(float f)
{
  float fl = floorf(f);
  if (!(fl == fl)) return 0;
  if (fl >= 2147483647.0f) return INT32_MAX;
  if (fl <= -2147483648.0f) return INT32_MIN;
  return (int32_t)fl;
}

/* grid3d_get_cell_indices: particle_grid.cl:45-53 */
static void f_cell_indices(const float minc[3], const float cs[3], const float* pt, int32_t c[3])
{
  for (int i = 0; i < 3; ++i) {
    float d = pt[i] - minc[i];
    c[i] = f_floor_to_int(d / cs[i]);
  }
}

/* grid3d_get_cell_key_from_indices: particle_grid.cl:77-85 (ulong wrap) */
static uint64_t cell_key(const int32_t nc[3], const int32_t c[3])
{
  uint64_t nx = (uint64_t)(int64_t)nc[0], ny = (uint64_t)(int64_t)nc[1];
  return (uint64_t)(int64_t)c[0] + (uint64_t)(int64_t)c[1] * nx
       + (uint64_t)(int64_t)c[2] * nx * ny;
}

/* keys of float4 points: grid3d_generate_sort_keys, particle_grid.cl:96-112 */
void orc_cell_keys(const float* p4, uint64_t n, const int32_t nc[3],
                   const float minc[3], const float cs[3], uint64_t* keys)
{
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)n; ++k) {
    int32_t c[3];
    f_cell_indices(minc, cs, p4 + 4 * k, c);
    keys[k] = cell_key(nc, c);
  }
}

/*
 * Grid build: keys, stable sort by key (within-cell order = input order,
 * the order boost::compute::sort_by_key produces for ulong keys with less<>),
 * permuted particles and h, cell (begin, end) with (0,0) for empty cells,
 * overall max h and per-cell max h (0 for empty cells).
 * particle_grid.cpp:151-222 (build_tiles), :226-251 (sorted_to_unsorted map),
 * :253-275 (sort scalars), :315-340 (max per cell); particle_grid.cl:114-216;
 * smoothing_particle_grid.cpp:29-54.
 * Restated with a counting sort (stable by construction).
 * Returns ORC_ERANGE if any key is >= the number of cells (the reference
 * writes out of bounds there).
 */
int orc_grid_build(const float* p4, const float* h, uint64_t n,
                   const int32_t nc[3], const float minc[3], const float cs[3],
                   uint32_t* perm, int32_t* cells2, float* sorted_p4, float* sorted_h,
                   float* maxh_cell, float* max_h)
{
  uint64_t ncells = (uint64_t)nc[0] * (uint64_t)nc[1] * (uint64_t)nc[2];
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
  uint32_t* count = (uint32_t*)calloc(ncells + 1, sizeof(uint32_t));
  if (!keys || !count) { free(keys); free(count); return ORC_ENOMEM; }
  orc_cell_keys(p4, n, nc, minc, cs, keys);
  for (uint64_t k = 0; k < n; ++k) {
    if (keys[k] >= ncells) { free(keys); free(count); return ORC_ERANGE; }
    count[keys[k] + 1]++;
  }
  for (uint64_t c = 0; c < ncells; ++c)
    count[c + 1] += count[c];
  /* count[c] is now the first sorted position of cell c */
  for (uint64_t c = 0; c < ncells; ++c) {
    cells2[2 * c] = 0;
    cells2[2 * c + 1] = 0;
    if (count[c + 1] > count[c]) {
      cells2[2 * c] = (int32_t)count[c];
      cells2[2 * c + 1] = (int32_t)count[c + 1];
    }
  }
  for (uint64_t k = 0; k < n; ++k)
    perm[count[keys[k]]++] = (uint32_t)k;
  float mh = 0.0f;
  int have = 0;
  for (uint64_t s = 0; s < n; ++s) {
    uint32_t k = perm[s];
    memcpy(sorted_p4 + 4 * s, p4 + 4 * (uint64_t)k, 4 * sizeof(float));
    sorted_h[s] = h[k];
    if (!have || sorted_h[s] > mh) { mh = sorted_h[s]; have = 1; }
  }
  for (uint64_t c = 0; c < ncells; ++c) {
    float cur = 0.0f;
    for (int32_t i = cells2[2 * c]; i < cells2[2 * c + 1]; ++i)
      if (sorted_h[i] > cur) cur = sorted_h[i];
    maxh_cell[c] = cur;
  }
  *max_h = mh;
  free(keys);
  free(count);
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Projected quartic kernel: smoothing_kernel.cl:46-61                 */
/* ------------------------------------------------------------------ */
static float f_quartic_projection(float r, float h)
{
  float h_inv = 1.f / h;
  float h_inv3 = (h_inv * h_inv) * h_inv;
  float norm = 1.04445f * h_inv3;
  float hh = h * h;
  float rr = r * r;
  float s2 = fmaxf(hh - rr, 0.0f);
  float integral = 1.066667f * s2;
  integral = integral * s2;
  integral = integral * sqrtf(s2);
  integral = integral * h_inv;
  integral = integral * h_inv3;
  return (2.f * norm) * integral;
}

float orc_quartic_projection(float r, float h) { return f_quartic_projection(r, h); }

/*
 * The gather with reference semantics: spg_run_reconstruction with
 * SPG_PROJECTION (smoothing_particle_grid.cl:31-127), called through
 * projective_smoothing_grid (projective_smoothing_backend_grid.cl:30-56).
 * Per point: search radius from max h of the point's own cell plus the
 * cell half-diagonal, capped by the overall max (or the overall max when
 * the unclamped key is out of range), cell loop z->y->x over the index
 * box, 3-D cell cull, per-particle w * W2D(|dxy|, h).
 * acc_mode 0: FP32 accumulation starting from 0 (the reference's zeroed
 * result buffer, dm_reconstruction_backend_grid.cpp:95-99); the float
 * result is returned widened to double.
 * acc_mode 1: the same FP32 terms accumulated in double.
 * If pair_count is not NULL it receives, per point, the number of
 * (point, particle) terms the reference loop evaluates.
 */
void orc_gather_ref(const int32_t nc[3], const float minc[3], const float cs[3],
                    const int32_t* cells2, const float* sorted_p4, const float* sorted_h,
                    const float* maxh_cell, float max_h,
                    const float* pts4, uint64_t npts, int acc_mode,
                    double* out, uint64_t* pair_count)
{
  const uint64_t total = (uint64_t)nc[0] * (uint64_t)nc[1] * (uint64_t)nc[2];
  float dd = cs[0] * cs[0];
  dd = dd + cs[1] * cs[1];
  dd = dd + cs[2] * cs[2];
  const float cell_radius = 0.5f * sqrtf(dd);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < (int64_t)npts; ++t) {
    const float* ep = pts4 + 4 * t;
    float res32 = 0.0f;
    double res64 = 0.0;
    uint64_t pairs = 0;
    int32_t ec[3];
    f_cell_indices(minc, cs, ep, ec);
    uint64_t ekey = cell_key(nc, ec);
    float R = max_h;
    if (ekey < total)
      R = fminf(maxh_cell[ekey] + cell_radius, max_h);
    float lo_pt[3] = {ep[0] - R, ep[1] - R, ep[2] - R};
    float hi_pt[3] = {ep[0] + R, ep[1] + R, ep[2] + R};
    int32_t lo[3], hi[3];
    f_cell_indices(minc, cs, lo_pt, lo);
    f_cell_indices(minc, cs, hi_pt, hi);
    for (int64_t cz = lo[2]; cz <= hi[2]; ++cz)
      for (int64_t cy = lo[1]; cy <= hi[1]; ++cy)
        for (int64_t cx = lo[0]; cx <= hi[0]; ++cx) {
          if (!(0 <= cx && cx < nc[0] && 0 <= cy && cy < nc[1] && 0 <= cz && cz < nc[2]))
            continue;
          int32_t c[3] = {(int32_t)cx, (int32_t)cy, (int32_t)cz};
          uint64_t ck = cell_key(nc, c);
          int32_t b = cells2[2 * ck], e = cells2[2 * ck + 1];
          float cell_cutoff = maxh_cell[ck];
          float center[3];
          for (int i = 0; i < 3; ++i) {
            float m = minc[i] + (float)c[i] * cs[i];
            center[i] = m + 0.5f * cs[i];
          }
          float d[3] = {ep[0] - center[0], ep[1] - center[1], ep[2] - center[2]};
          float dist = sqrtf(f_dot3(d, d));
          if (!(dist < cell_cutoff + cell_radius))
            continue;
          for (int32_t i = b; i < e; ++i) {
            float hs = sorted_h[i];
            const float* p = sorted_p4 + 4 * (int64_t)i;
            float dx = p[0] - ep[0], dy = p[1] - ep[1];
            float r = sqrtf(dx * dx + dy * dy);
            float term = p[3] * f_quartic_projection(r, hs);
            res32 += term;
            res64 += (double)term;
            ++pairs;
          }
        }
    out[t] = acc_mode == 0 ? (double)res32 : res64;
    if (pair_count)
      pair_count[t] = pairs;
  }
}

/* ------------------------------------------------------------------ */
/* Volumetric smoothing (dm::grid): smoothing_kernel.cl:27-37 and       */
/* spg_run_reconstruction without SPG_PROJECTION                        */
/* ------------------------------------------------------------------ */
static float f_quartic3d(float r, float h)
{
  float h_inv = 1.f / h;
  float norm = 1.04445f * h_inv;
  norm = norm * h_inv;
  norm = norm * h_inv;
  float q = fminf(r * h_inv, 1.f);
  float q2 = q * q;
  float q4 = q2 * q2;
  float a = 1.f + q4;
  a = a - 2.f * q2;
  return norm * a;
}

float orc_quartic3d(float r, float h) { return f_quartic3d(r, h); }

/* The 3-D grid reconstruction with the reference's search: as
 * orc_gather_ref, with the particle term w * quartic_polynomial3d(r, h) of
 * the 3-D distance r = distance(particle.xyz, point.xyz).
 * form 0: the term as the reference evaluates it (FP32, 1 + q^4 - 2 q^2);
 * form 1: the same kernel, 1.04445 / h^3 (1 - r^2/h^2)^2, in FP64 (the
 * reference's expression cancels near r = h). */
void orc_smooth3d_ref(const int32_t nc[3], const float minc[3], const float cs[3],
                      const int32_t* cells2, const float* sorted_p4, const float* sorted_h,
                      const float* maxh_cell, float max_h,
                      const float* pts4, uint64_t npts, int acc_mode, int form, double* out)
{
  const uint64_t total = (uint64_t)nc[0] * (uint64_t)nc[1] * (uint64_t)nc[2];
  float dd = cs[0] * cs[0];
  dd = dd + cs[1] * cs[1];
  dd = dd + cs[2] * cs[2];
  const float cell_radius = 0.5f * sqrtf(dd);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < (int64_t)npts; ++t) {
    const float* ep = pts4 + 4 * t;
    float res32 = 0.0f;
    double res64 = 0.0;
    int32_t ec[3];
    f_cell_indices(minc, cs, ep, ec);
    uint64_t ekey = cell_key(nc, ec);
    float R = max_h;
    if (ekey < total)
      R = fminf(maxh_cell[ekey] + cell_radius, max_h);
    float lo_pt[3] = {ep[0] - R, ep[1] - R, ep[2] - R};
    float hi_pt[3] = {ep[0] + R, ep[1] + R, ep[2] + R};
    int32_t lo[3], hi[3];
    f_cell_indices(minc, cs, lo_pt, lo);
    f_cell_indices(minc, cs, hi_pt, hi);
    for (int64_t cz = lo[2]; cz <= hi[2]; ++cz)
      for (int64_t cy = lo[1]; cy <= hi[1]; ++cy)
        for (int64_t cx = lo[0]; cx <= hi[0]; ++cx) {
          if (!(0 <= cx && cx < nc[0] && 0 <= cy && cy < nc[1] && 0 <= cz && cz < nc[2]))
            continue;
          int32_t c[3] = {(int32_t)cx, (int32_t)cy, (int32_t)cz};
          uint64_t ck = cell_key(nc, c);
          int32_t b = cells2[2 * ck], e = cells2[2 * ck + 1];
          float cell_cutoff = maxh_cell[ck];
          float center[3];
          for (int i = 0; i < 3; ++i) {
            float m = minc[i] + (float)c[i] * cs[i];
            center[i] = m + 0.5f * cs[i];
          }
          float d[3] = {ep[0] - center[0], ep[1] - center[1], ep[2] - center[2]};
          float dist = sqrtf(f_dot3(d, d));
          if (!(dist < cell_cutoff + cell_radius))
            continue;
          for (int32_t i = b; i < e; ++i) {
            float hs = sorted_h[i];
            const float* p = sorted_p4 + 4 * (int64_t)i;
            if (form == 1) {
              double dx = (double)p[0] - ep[0], dy = (double)p[1] - ep[1];
              double dz = (double)p[2] - ep[2];
              double hh = (double)hs;
              double s = 1.0 - (dx * dx + dy * dy + dz * dz) / (hh * hh);
              double term = s > 0.0 ? (double)p[3] * (double)1.04445f / (hh * hh * hh) * s * s : 0.0;
              res32 += (float)term;
              res64 += term;
            } else {
              float dv[3] = {p[0] - ep[0], p[1] - ep[1], p[2] - ep[2]};
              float r = sqrtf(f_dot3(dv, dv));
              float term = p[3] * f_quartic3d(r, hs);
              res32 += term;
              res64 += (double)term;
            }
          }
        }
    out[t] = acc_mode == 0 ? (double)res32 : res64;
  }
}

/* Every particle whose support contains the point, in FP64:
 * w * 105 / (32 pi h^3) (1 - r^2/h^2)^2. */
void orc_smooth3d_exact(const float* p4, const float* h, uint64_t n, const float* pts4,
                        uint64_t npts, double* out)
{
  const double K = 105.0 / (32.0 * M_PI);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t t = 0; t < (int64_t)npts; ++t) {
    double acc = 0.0;
    for (uint64_t j = 0; j < n; ++j) {
      double dx = (double)p4[4 * j] - pts4[4 * t], dy = (double)p4[4 * j + 1] - pts4[4 * t + 1];
      double dz = (double)p4[4 * j + 2] - pts4[4 * t + 2];
      double hh = (double)h[j];
      double s = 1.0 - (dx * dx + dy * dy + dz * dz) / (hh * hh);
      if (s > 0.0) acc += (double)p4[4 * j + 3] * K / (hh * hh * hh) * s * s;
    }
    out[t] = acc;
  }
}

/* ------------------------------------------------------------------ */
/* FP64 exact oracle                                                   */
/* ------------------------------------------------------------------ */

/* Kernel families.  QUARTIC_REF: the reference's projected quartic with its
 * float constants and its factor 2 (smoothing_kernel.cl:46-61).
 * QUARTIC: the same shape normalised to unit integral, 7/(2 pi h^2).
 * CUBIC: projection of the M4 cubic spline (smoothing_kernel.cl:64-84,
 * support 2h), evaluated by Gauss-Legendre quadrature along the line of
 * sight on each piece. */
#define ORC_K_QUARTIC_REF 0
#define ORC_K_QUARTIC 1
#define ORC_K_CUBIC 2

static double d_cubic3d(double q)
{
  /* M4 spline shape times 1/pi (h = 1) */
  if (q >= 2.0) return 0.0;
  if (q <= 1.0) return (1.0 - 1.5 * q * q + 0.75 * q * q * q) / M_PI;
  double t = 2.0 - q;
  return 0.25 * t * t * t / M_PI;
}

/* 16-point Gauss-Legendre nodes/weights on [-1,1] (positive half) */
static const double gl_x[8] = {0.0950125098376374, 0.2816035507792589, 0.4580167776572274,
                               0.6178762444026438, 0.7554044083550030, 0.8656312023878318,
                               0.9445750230732326, 0.9894009349916499};
static const double gl_w[8] = {0.1894506104550685, 0.1826034150449236, 0.1691565193950025,
                               0.1495959888165767, 0.1246289712555339, 0.0951585116824928,
                               0.0622535239386479, 0.0271524594117541};

static double gl_integrate_cubic(double b, double z0, double z1)
{
  double c = 0.5 * (z0 + z1), hw = 0.5 * (z1 - z0), s = 0.0;
  for (int k = 0; k < 8; ++k) {
    double za = c - hw * gl_x[k], zb = c + hw * gl_x[k];
    s += gl_w[k] * (d_cubic3d(sqrt(b * b + za * za)) + d_cubic3d(sqrt(b * b + zb * zb)));
  }
  return s * hw;
}

/* projected M4 cubic spline for h = 1 at impact parameter b (0 <= b < 2) */
double orc_cubic_projected_unit(double b)
{
  if (b >= 2.0) return 0.0;
  /* the integrand is piecewise polynomial in q; split at q = 1 and
   * integrate each piece with composite Gauss-Legendre in z */
  double z2 = sqrt(4.0 - b * b);
  double z1 = b < 1.0 ? sqrt(1.0 - b * b) : 0.0;
  double s = 0.0;
  const int m = 16;
  if (z1 > 0.0)
    for (int k = 0; k < m; ++k)
      s += gl_integrate_cubic(b, z1 * k / m, z1 * (k + 1) / m);
  for (int k = 0; k < m; ++k)
    s += gl_integrate_cubic(b, z1 + (z2 - z1) * k / m, z1 + (z2 - z1) * (k + 1) / m);
  return 2.0 * s;
}

/* d/du of the projected cubic at u = b^2 (h = 1): the integral over z of
 * d/du M4(sqrt(u + z^2)) = M4'(rho) / (2 rho); the column end moves with u
 * but the integrand vanishes there. */
static double d_cubic3d_du(double rho)
{
  if (rho >= 2.0) return 0.0;
  if (rho <= 1.0) return (-1.5 + 1.125 * rho) / M_PI;
  double t = 2.0 - rho;
  return -0.375 * t * t / (M_PI * rho);
}

static double gl_integrate_cubic_du(double b, double z0, double z1)
{
  double c = 0.5 * (z0 + z1), hw = 0.5 * (z1 - z0), s = 0.0;
  for (int k = 0; k < 8; ++k) {
    double za = c - hw * gl_x[k], zb = c + hw * gl_x[k];
    s += gl_w[k] * (d_cubic3d_du(sqrt(b * b + za * za)) + d_cubic3d_du(sqrt(b * b + zb * zb)));
  }
  return s * hw;
}

double orc_cubic_projected_du(double b)
{
  if (b >= 2.0) return 0.0;
  double z2 = sqrt(4.0 - b * b);
  double z1 = b < 1.0 ? sqrt(1.0 - b * b) : 0.0;
  double s = 0.0;
  const int m = 16;
  if (z1 > 0.0)
    for (int k = 0; k < m; ++k)
      s += gl_integrate_cubic_du(b, z1 * k / m, z1 * (k + 1) / m);
  for (int k = 0; k < m; ++k)
    s += gl_integrate_cubic_du(b, z1 + (z2 - z1) * k / m, z1 + (z2 - z1) * (k + 1) / m);
  return 2.0 * s;
}

static double d_kernel2d(int kind, double r2, double h)
{
  double h2 = h * h;
  if (kind == ORC_K_CUBIC) {
    double b = sqrt(r2) / h;
    return orc_cubic_projected_unit(b) / h2;
  }
  double s = 1.0 - r2 / h2;
  if (s <= 0.0) return 0.0;
  double shape = s * s * sqrt(s);
  double k = kind == ORC_K_QUARTIC_REF
                 ? 2.0 * (double)1.04445f * (double)1.066667f
                 : 3.5 / M_PI;
  return k * shape / h2;
}

/*
 * Brute-force FP64 projection: out[t] = sum_j w_j W(|x_j - p_t|, h_j) over
 * every particle with w_j != 0 and |x_j - p_t| < support(h_j)
 * (support = h for the quartics, 2h for the cubic spline).  Points and
 * particles are float4 in the camera frame (projected).  counts, if not
 * NULL, receives the number of such pairs per point (the unit of work:
 * "particle-kernel evaluations").
 */
void orc_gather_exact(const float* p4, const float* h, uint64_t n,
                      const float* pts4, uint64_t npts, int kind,
                      double* out, uint64_t* counts)
{
  const double sup = kind == ORC_K_CUBIC ? 2.0 : 1.0;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t t = 0; t < (int64_t)npts; ++t) {
    double ex = pts4[4 * t], ey = pts4[4 * t + 1];
    double acc = 0.0;
    uint64_t cnt = 0;
    for (uint64_t j = 0; j < n; ++j) {
      float w = p4[4 * j + 3];
      if (w == 0.0f) continue;
      double dx = (double)p4[4 * j] - ex, dy = (double)p4[4 * j + 1] - ey;
      double r2 = dx * dx + dy * dy;
      double hs = sup * (double)h[j];
      if (r2 < hs * hs) {
        acc += (double)w * d_kernel2d(kind, r2, (double)h[j]);
        ++cnt;
      }
    }
    out[t] = acc;
    if (counts) counts[t] = cnt;
  }
}

/*
 * Line-of-sight clipped projection (the reference has none: it masks by
 * the particle centre, projection.cl:80-91).  Each particle contributes the
 * part of its 3-D quartic's column that lies in the view's depth range
 * [0, depth]:
 *   out[t] = sum_j w_j int_{0 - z_j}^{depth - z_j} W3(sqrt(r^2 + z^2), h_j) dz,
 * with W3(q h) = K (15/16) (1 - q^2)^2 / h^3, K the projected kernel's
 * normalisation of d_kernel2d (the reference's constants, with its factor 2,
 * for ORC_K_QUARTIC_REF), so that the whole column equals d_kernel2d.  The
 * integral runs over the part of the range inside the support (|z| < a,
 * a^2 = h^2 - r^2) by 8-point Gauss-Legendre quadrature in FP64 (the
 * integrand is a polynomial of degree 4 in z, so the rule is exact up to
 * rounding).  p4: (x', y', z', w) camera frame (orc_project_particles_depth).
 */
static const double orc_gl8_x[8] = {-0.9602898564975363, -0.7966664774136267,
                                    -0.5255324099163290, -0.1834346424956498,
                                    0.1834346424956498,  0.5255324099163290,
                                    0.7966664774136267,  0.9602898564975363};
static const double orc_gl8_w[8] = {0.1012285362903763, 0.2223810344533745,
                                    0.3137066458778873, 0.3626837833783620,
                                    0.3626837833783620, 0.3137066458778873,
                                    0.2223810344533745, 0.1012285362903763};

void orc_gather_clip(const float* p4, const float* h, uint64_t n, const float* pts4,
                     uint64_t npts, double depth, int kind, double* out)
{
  const double K = kind == ORC_K_QUARTIC_REF ? 2.0 * (double)1.04445f * (double)1.066667f
                                             : 3.5 / M_PI;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t t = 0; t < (int64_t)npts; ++t) {
    const double ex = pts4[4 * t], ey = pts4[4 * t + 1];
    double acc = 0.0;
    for (uint64_t j = 0; j < n; ++j) {
      const double w = p4[4 * j + 3];
      if (w == 0.0) continue;
      const double hj = h[j];
      const double dx = (double)p4[4 * j] - ex, dy = (double)p4[4 * j + 1] - ey;
      const double r2 = dx * dx + dy * dy;
      if (r2 >= hj * hj) continue;
      const double a = sqrt(hj * hj - r2);
      const double zj = p4[4 * j + 2];
      const double lo = fmax(0.0 - zj, -a), hi = fmin(depth - zj, a);
      if (!(hi > lo)) continue;
      double s = 0.0;
      for (int k = 0; k < 8; ++k) {
        const double z = 0.5 * (lo + hi) + 0.5 * (hi - lo) * orc_gl8_x[k];
        const double q2 = (r2 + z * z) / (hj * hj);
        const double f = q2 < 1.0 ? (1.0 - q2) * (1.0 - q2) : 0.0;
        s += orc_gl8_w[k] * f;
      }
      s *= 0.5 * (hi - lo);
      acc += w * K * (15.0 / 16.0) * s / (hj * hj * hj);
    }
    out[t] = acc;
  }
}

/*
 * Exact number of (point, particle) pairs with w != 0 and distance below
 * the support, over a regular nx x ny pixel lattice whose point (ix, iy)
 * is origin + ix*ps*e0 + iy*ps*e1 in the camera frame with e0, e1 the
 * camera-frame images of the screen axes (the identity for an exact
 * inverse; passed explicitly so that the FP32 inverse can be used).
 * Visits, per particle, only the lattice rows and columns its support can
 * reach (widened by two pixels for rounding) and tests each visited point
 * with the same double predicate as orc_gather_exact on the exact point
 * positions in pts4 (row-major): the count is exact, and its cost is the
 * number of pairs, not points x particles.
 */
uint64_t orc_count_evals_lattice(const float* p4, const float* h, uint64_t n,
                                 const float* pts4, uint64_t nx, uint64_t ny,
                                 double support_factor)
{
  uint64_t total = 0;
  /* row / column coordinates from the first column / row: the lattice is
   * axis-aligned in the camera frame up to rounding, which the two-pixel
   * widening absorbs. */
  double* rowy = (double*)malloc(sizeof(double) * ny);
  double* colx = (double*)malloc(sizeof(double) * nx);
  if (!rowy || !colx) { free(rowy); free(colx); return UINT64_MAX; }
  for (uint64_t iy = 0; iy < ny; ++iy) rowy[iy] = pts4[4 * (iy * nx) + 1];
  for (uint64_t ix = 0; ix < nx; ++ix) colx[ix] = pts4[4 * ix];
  double ps_x = nx > 1 ? (colx[nx - 1] - colx[0]) / (double)(nx - 1) : 1.0;
  double ps_y = ny > 1 ? (rowy[ny - 1] - rowy[0]) / (double)(ny - 1) : 1.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int64_t j = 0; j < (int64_t)n; ++j) {
    float w = p4[4 * j + 3];
    if (w == 0.0f) continue;
    double px = p4[4 * j], py = p4[4 * j + 1];
    double hs = support_factor * (double)h[j];
    double hs2 = hs * hs;
    double fy0 = (py - hs - rowy[0]) / ps_y, fy1 = (py + hs - rowy[0]) / ps_y;
    int64_t iy0 = (int64_t)floor(fy0) - 2, iy1 = (int64_t)ceil(fy1) + 2;
    if (iy0 < 0) iy0 = 0;
    if (iy1 > (int64_t)ny - 1) iy1 = (int64_t)ny - 1;
    for (int64_t iy = iy0; iy <= iy1; ++iy) {
      double dy0 = (double)pts4[4 * (iy * nx) + 1] - py;
      double rem = hs2 - dy0 * dy0;
      if (rem < -4.0 * ps_y * hs) continue;
      double half = rem > 0 ? sqrt(rem) : 0.0;
      int64_t ix0 = (int64_t)floor((px - half - colx[0]) / ps_x) - 2;
      int64_t ix1 = (int64_t)ceil((px + half - colx[0]) / ps_x) + 2;
      if (ix0 < 0) ix0 = 0;
      if (ix1 > (int64_t)nx - 1) ix1 = (int64_t)nx - 1;
      if (ix1 < ix0) continue;
      const float* row = pts4 + 4 * (iy * nx);
      uint64_t c = 0;
      for (int64_t ix = ix0; ix <= ix1; ++ix) {
        double dx = (double)row[4 * ix] - px, dy = (double)row[4 * ix + 1] - py;
        c += (dx * dx + dy * dy < hs2);
      }
      total += c;
    }
  }
  free(rowy);
  free(colx);
  return total;
}

/* ------------------------------------------------------------------ */
/* Quantities: quantities.cl:29-61 (temperature), tabulated_function.cl */
/* :26-59 (linear sampler), quantity.cpp:84-93 (input scaling)          */
/* ------------------------------------------------------------------ */

/* mean_molecular_weight / get_temperature: quantities.cl:39-61 */
float orc_temperature(float internal_energy, float electron_abundance)
{
  const float X_H = 0.76f;
  float den = 1.f + 3 * X_H;
  den = den + (4 * X_H) * electron_abundance;
  float mu = 4.f / den;
  float g1 = (5.f / 3.f) - 1.f;
  float t = g1 * internal_energy;
  t = t * mu;
  return t * 121.1475502f;
}

/*
 * read_imagef on a 1-D CL_R/CL_FLOAT image with CLK_FILTER_LINEAR |
 * CLK_ADDRESS_CLAMP | unnormalised coordinates (OpenCL 1.2 s8.2):
 * u' = u - 0.5, i0 = floor(u'), a = frac(u'), texels outside [0, n) read
 * the border colour 0.  evaluate_tabulated_function: tabulated_function.cl
 * :31-44 computes u = (x - x_min) / dx.
 */
float orc_tabulated_1d(const float* y, int32_t n, float x_min, float dx, float x)
{
  float u = (x - x_min) / dx;
  float us = u - 0.5f;
  float fl = floorf(us);
  float a = us - fl;
  int32_t i0 = f_floor_to_int(us);
  float t0 = (i0 >= 0 && i0 < n) ? y[i0] : 0.0f;
  float t1 = (i0 >= -1 && i0 + 1 < n) ? y[i0 + 1] : 0.0f;
  return (1.0f - a) * t0 + a * t1;
}

/* Same sampler on a 2-D image (row size nx, data[ix + nx*iy]):
 * tabulated_function.cl:46-59; layout data_table_loader.hpp:88-96,
 * multi_array.hpp:503-536 (first index fastest). */
float orc_tabulated_2d(const float* data, int32_t nx, int32_t ny,
                       float x_min, float y_min, float dx, float dy, float x, float y)
{
  float u = (x - x_min) / dx, v = (y - y_min) / dy;
  float us = u - 0.5f, vs = v - 0.5f;
  float a = us - floorf(us), b = vs - floorf(vs);
  int32_t i0 = f_floor_to_int(us), j0 = f_floor_to_int(vs);
  float t[2][2];
  for (int dj = 0; dj < 2; ++dj)
    for (int di = 0; di < 2; ++di) {
      int32_t i = i0 + di, j = j0 + dj;
      t[dj][di] = (i >= 0 && i < nx && j >= 0 && j < ny) ? data[i + nx * j] : 0.0f;
    }
  return (1 - a) * (1 - b) * t[0][0] + a * (1 - b) * t[0][1]
       + (1 - a) * b * t[1][0] + a * b * t[1][1];
}

/*
 * Bolometric free-free emissivity of the synthetic X-ray workload,
 * composed from the reference's pieces: ion density 4.046263*rho*x_e
 * (quantities.cl:112), a 1-D tabulated Gaunt factor on ln T read with the
 * sampler above, and w = n_e n_i sqrt(T) g(T) with n_e = n_i = ion density
 * (quantities.cl:128-130 uses ion_density^2).
 */
void orc_xray_bolometric(const float* rho, const float* temp, const float* xe, uint64_t n,
                         const float* gtab, int32_t gn, float gx0, float gdx, float* w)
{
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)n; ++k) {
    float ion = 4.046263f * rho[k];
    ion = ion * xe[k];
    float g = orc_tabulated_1d(gtab, gn, gx0, gdx, logf(temp[k]));
    float e = ion * ion;
    e = e * sqrtf(temp[k]);
    w[k] = e * g;
  }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* FP64 adjoint of orc_gather_exact (quartic kernels)                  */
/* ------------------------------------------------------------------ */

/*
 * Gradients of L with dL/d out[t] = G[t] (out as in orc_gather_exact, with
 * kind QUARTIC_REF or QUARTIC) w.r.t. each particle's w, h and camera-frame
 * x, y.  With s = 1 - r^2/h^2, W = K s^2.5 / h^2:
 *   dW/dh = K s^1.5 (7 r^2/h^2 - 2) / h^3,
 *   dW/dx = -5 K s^1.5 (x - p_x) / h^4   (and y alike).
 * No reference counterpart (the reference has no adjoint); the formulas are
 * pinned by central differences of orc_gather_exact in tests/.
 */
void orc_backward_exact(const float* p4, const float* h, uint64_t n, const float* pts4,
                        uint64_t npts, const double* G, int kind, double* gw, double* gh,
                        double* gx, double* gy)
{
  const double K = kind == ORC_K_QUARTIC_REF ? 2.0 * (double)1.04445f * (double)1.066667f
                                             : 3.5 / M_PI;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t j = 0; j < (int64_t)n; ++j) {
    const double w = p4[4 * j + 3];
    const double px = p4[4 * j], py = p4[4 * j + 1], hh = h[j];
    const double h2 = hh * hh;
    double sw = 0, sh = 0, sx = 0, sy = 0;
    if (kind == ORC_K_CUBIC) {
      /* value = w F(u) / h^2, u = r^2 / h^2, support 2h */
      for (uint64_t t = 0; t < npts; ++t) {
        const double dx = px - (double)pts4[4 * t], dy = py - (double)pts4[4 * t + 1];
        const double r2 = dx * dx + dy * dy;
        if (!(r2 < 4.0 * h2)) continue;
        const double b = sqrt(r2) / hh, u = r2 / h2;
        const double F = orc_cubic_projected_unit(b), D = orc_cubic_projected_du(b);
        const double g = G[t];
        sw += g * F / h2;
        sh += g * (-2.0) * (F + u * D) / (h2 * hh);
        sx += g * 2.0 * D * dx / (h2 * h2);
        sy += g * 2.0 * D * dy / (h2 * h2);
      }
      gw[j] = sw;
      gh[j] = w * sh;
      gx[j] = w * sx;
      gy[j] = w * sy;
      continue;
    }
    for (uint64_t t = 0; t < npts; ++t) {
      const double dx = px - (double)pts4[4 * t], dy = py - (double)pts4[4 * t + 1];
      const double r2 = dx * dx + dy * dy;
      if (!(r2 < h2)) continue;
      const double s = 1.0 - r2 / h2;
      const double s15 = s * sqrt(s);
      const double g = G[t];
      sw += g * K * s15 * s / h2;
      sh += g * K * s15 * (7.0 * r2 / h2 - 2.0) / (h2 * hh);
      sx += g * (-5.0) * K * s15 * dx / (h2 * h2);
      sy += g * (-5.0) * K * s15 * dy / (h2 * h2);
    }
    gw[j] = sw;
    gh[j] = w * sh;
    gx[j] = w * sx;
    gy[j] = w * sy;
  }
}

/* ------------------------------------------------------------------ */
/* Block body of reconstructing_data_crawler::run                      */
/* (reconstructing_data_crawler.cpp:180-213): each row's coordinates,  */
/* read as float (:183-185), are moved by one period towards the       */
/* render volume grown by `border` (coordinate_system.hpp:24-47), and   */
/* kept if the volume contains them (reconstructing_data_crawler.cpp:  */
/* 59-71, tolerance 0); kept rows go out in input order as float4      */
/* (x, y, z, 0) (to_device_vector4) with their extra columns.           */
/* Returns the number kept.                                            */
/* ------------------------------------------------------------------ */
uint64_t orc_crawl_block(const double center[3], const double extent[3],
                         const double period[3], double border, const float* xyz,
                         uint64_t n, const float* const* cols, int32_t ncols, float* out4,
                         float* const* cols_out, uint64_t* index_out)
{
  double bmin[3], bmax[3];
  for (int j = 0; j < 3; ++j) {
    bmin[j] = center[j] - 0.5 * extent[j];
    bmax[j] = center[j] + 0.5 * extent[j];
  }
  uint64_t m = 0;
  for (uint64_t i = 0; i < n; ++i) {
    double c[3];
    for (int j = 0; j < 3; ++j) c[j] = (double)xyz[3 * i + j];
    for (int j = 0; j < 3; ++j) {
      const double gmin = bmin[j] - border, gmax = bmax[j] + border;
      const double lo = c[j] - period[j], hi = c[j] + period[j];
      if (gmin <= lo && lo <= gmax)
        c[j] = lo;
      else if (gmin <= hi && hi <= gmax)
        c[j] = hi;
    }
    int inside = 1;
    for (int j = 0; j < 3; ++j) {
      if (c[j] < center[j] - 0.5 * extent[j] - 0.0) inside = 0;
      if (c[j] > center[j] + 0.5 * extent[j] + 0.0) inside = 0;
    }
    if (!inside) continue;
    for (int j = 0; j < 3; ++j) out4[4 * m + j] = (float)c[j];
    out4[4 * m + 3] = 0.0f;
    for (int k = 0; k < ncols; ++k) cols_out[k][m] = cols[k][i];
    if (index_out) index_out[m] = i;
    ++m;
  }
  return m;
}

/* ------------------------------------------------------------------ */
/* Spectral bremsstrahlung X-ray emissivity (quantities.cl:69-239,     */
/* 354-430) with the Runge-Kutta-Fehlberg energy integration of        */
/* integration.cl:38-131.                                               */
/*                                                                      */
/* The reference leaves exp, log, pow and rsqrt to the OpenCL runtime  */
/* (unspecified precision).  This restatement uses correctly rounded   */
/* FP32 versions (the FP64 function rounded to FP32), the convention    */
/* the HIP path shares, so that both take the same accept/reject        */
/* decisions in the step-size control.  The step loop stops after       */
/* ORC_RKF_MAX_STEPS steps (the reference has no bound).                 */
/* ------------------------------------------------------------------ */
#define ORC_K_B 8.61733e-5f
#define ORC_RYD 13.605f
#define ORC_RKF_MAX_STEPS 65536

static float crf_expf(float x) { return (float)exp((double)x); }
static float crf_logf(float x) { return (float)log((double)x); }
static float crf_powf(float x, float y) { return (float)pow((double)x, (double)y); }
static float crf_rsqrtf(float x) { return (float)(1.0 / sqrt((double)x)); }

/* g_ff (quantities.cl:69-83): van Hoof table on (ln gamma^2, ln u), nodes
 * from (-6, -16) in steps of 0.2, sampled as read_imagef linear */
static float orc_gff(float E, float T, float Z, const float* tab, int32_t nx, int32_t ny)
{
  float kT = ORC_K_B * T;
  float gam2 = Z * Z;
  gam2 = gam2 * ORC_RYD;
  gam2 = gam2 / kT;
  float u = 1.e3f * E;
  u = u / kT;
  return orc_tabulated_2d(tab, nx, ny, -6.f, -16.f, 0.2f, 0.2f, crf_logf(gam2), crf_logf(u));
}

/* get_spectral_brems_emission (quantities.cl:97-136) */
static float orc_brems(float E, float dE, float T, float xe, float rho, float r,
                       const float* tab, int32_t nx, int32_t ny)
{
  float ion = 4.046263f * rho;
  ion = ion * xe;
  const float Z = 1.f;
  float t_inv = 1000.f / (ORC_K_B * T);
  float r2 = r * r;
  float e = 0.002535f * dE;
  e = e * Z;
  e = e * Z;
  e = e * ion;
  e = e * ion;
  e = e * orc_gff(E, T, Z, tab, nx, ny);
  e = e * crf_rsqrtf(T);
  e = e * crf_expf(-E * t_inv);
  return e / r2;
}

static void orc_rkf_points(float pos, float step, float* p)
{
  p[0] = pos + (3.f / 8.f) * step;
  p[1] = pos + (12.f / 13.f) * step;
  p[2] = pos + step;
  p[3] = pos + 0.5f * step;
}

/* integrate_spectral_flux (quantities.cl:156-239) for XRAY_FLUX, with
 * rkf_advance (integration.cl:49-131) called as the reference calls it:
 * absolute tolerance = rel_tol, relative tolerance = 1, epsilon = 0,
 * step sizes in [1e-3, 100]. */
static float orc_xray_integrate(float e_min, float e_max, float rel_tol, float T, float xe,
                                float rho, float dl, const float* tab, int32_t nx, int32_t ny,
                                int32_t* steps_out)
{
  float state = 0.0f;
  float pos = e_min;
  const float end = e_max;
  float step = (end - pos) * rel_tol;
  float rb = orc_brems(pos, 1.0f, T, xe, rho, dl, tab, nx, ny);
  float pts[4];
  orc_rkf_points(pos, step, pts);
  const float abs_tol = rel_tol, rel = 1.0f, eps = 0.0f, min_step = 1.e-3f, max_step = 100.f;
  int32_t it = 0;
  while (pos < end && it < ORC_RKF_MAX_STEPS) {
    ++it;
    float v[4];
    for (int k = 0; k < 4; ++k) v[k] = orc_brems(pts[k], 1.0f, T, xe, rho, dl, tab, nx, ny);
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
      s = crf_powf(q, 0.25f);
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
    orc_rkf_points(pos, step, pts);
  }
  if (steps_out) *steps_out = it;
  return state;
}

/* xray_flux (quantities.cl:354-386): out[i] in keV/s/kpc^3/m^2 for the
 * band [E_min, E_max] keV at redshift z and luminosity distance dl kpc;
 * T from the internal energy u and electron abundance xe. */
void orc_xray_flux(const float* rho, const float* u, const float* xe, uint64_t n, float z,
                   float dl, const float* tab, int32_t nx, int32_t ny, float e_min, float e_max,
                   int32_t num_samples, float* out, int32_t* steps)
{
  const float zf = 1.f + z;
  const float rel_tol = 1.f / (float)num_samples;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const float T = orc_temperature(u[i], xe[i]);
    out[i] = orc_xray_integrate(zf * e_min, zf * e_max, rel_tol, T, xe[i], rho[i], dl, tab,
                                nx, ny, steps ? steps + i : 0);
  }
}

/* xray_spectral_flux (quantities.cl:400-430): one energy bin E +- dE/2 */
void orc_xray_spectral_flux(const float* rho, const float* u, const float* xe, uint64_t n,
                            float z, float dl, const float* tab, int32_t nx, int32_t ny, float E,
                            float dE, float* out)
{
  const float zf = 1.f + z;
  for (uint64_t i = 0; i < n; ++i) {
    const float T = orc_temperature(u[i], xe[i]);
    out[i] = orc_brems(zf * E, zf * dE, T, xe[i], rho[i], dl, tab, nx, ny);
  }
}
