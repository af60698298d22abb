// illcrawl_volumetric.hpp -- volumetric slices and tomographic cubes over the
// data crawler with a volumetric backend (reconstruction_backends::dm::grid).
//
//   illcrawl::volumetric_slice       volumetric_slice.hpp, volumetric_slice.cpp:31-108
//   illcrawl::volumetric_tomography  volumetric_tomography.hpp, volumetric_tomography.cpp:29-102
//
// A slice evaluates the reconstruction at the camera's pixel positions
// (camera::get_pixel_coordinate, computed in double and stored as float) and
// scales each value by the quantity's effective dV of a pixel volume.
// Additional samples per pixel are jittered by up to half a pixel and
// averaged; the reference draws the jitter from std::random_device, here it
// comes from a seeded std::mt19937 so that a slice can be reproduced.
//
// A tomographic cube is a stack of slices, the camera moved by one pixel
// along its look-at direction per slice.  The reference runs the crawler once
// per slice; here the points of many slices go to one crawler pass (at most
// max_points_per_pass points at a time), so every block of the snapshot is
// read and binned once for all of them.
#ifndef ILLCRAWL_VOLUMETRIC_HPP
#define ILLCRAWL_VOLUMETRIC_HPP

#include <algorithm>
#include <cstdint>
#include <random>
#include <vector>

#include "illcrawl_crawler.hpp"

namespace illcrawl {

namespace detail {

// camera pixel coordinates as the reference's device_vector4 (w = 0)
inline void pixel_points(const camera& cam, std::vector<device_vector4>& out, std::size_t offset) {
  const std::size_t nx = cam.get_num_pixels(0), ny = cam.get_num_pixels(1);
  double c[3];
  for (std::size_t y = 0; y < ny; ++y)
    for (std::size_t x = 0; x < nx; ++x) {
      cam.get_pixel_coordinate(static_cast<double>(x), static_cast<double>(y), c);
      out[offset + y * nx + x] = device_vector4{static_cast<float>(c[0]), static_cast<float>(c[1]),
                                                 static_cast<float>(c[2]), 0.0f};
    }
}

inline double pixel_dV(const camera& cam, const reconstruction_quantity::quantity& q,
                       double volume_conversion_factor) {
  const double ps = cam.get_pixel_size();
  const double v = ps * ps * ps * volume_conversion_factor;
  return q.effective_volume_integration_dV(v, v);
}

}  // namespace detail

class volumetric_slice {
 public:
  explicit volumetric_slice(const camera& cam) : _cam(cam) {}

  // output: nx * ny floats, x fastest (the reference's multi_array{nx, ny}).
  // volume_conversion_factor: the quantity's unit converter's volume factor.
  void create_slice(reconstructing_data_crawler& reconstruction,
                    const reconstruction_quantity::quantity& q, std::vector<float>& output,
                    std::size_t num_additional_samples = 0, uint64_t seed = 0,
                    double volume_conversion_factor = 1.0) const {
    const std::size_t nx = _cam.get_num_pixels(0), ny = _cam.get_num_pixels(1);
    const std::size_t npix = nx * ny, samples = num_additional_samples + 1;
    std::vector<device_vector4> pts(samples * npix);
    detail::pixel_points(_cam, pts, 0);
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::uniform_real_distribution<double> jitter(-0.5, 0.5);
    double c[3];
    for (std::size_t i = 1; i < samples; ++i)
      for (std::size_t y = 0; y < ny; ++y)
        for (std::size_t x = 0; x < nx; ++x) {
          const double sx = static_cast<double>(x) + jitter(rng);
          const double sy = static_cast<double>(y) + jitter(rng);
          _cam.get_pixel_coordinate(sx, sy, c);
          pts[i * npix + y * nx + x] = device_vector4{static_cast<float>(c[0]),
                                                       static_cast<float>(c[1]),
                                                       static_cast<float>(c[2]), 0.0f};
        }
    auto dev = device_buffer::create<device_vector4>(pts.size());
    dev.upload(pts.data(), pts.size() * sizeof(device_vector4));
    reconstruction.get_backend().set_evaluation_layout(nx);
    reconstruction.run(dev, pts.size(), q);
    std::vector<float> res(pts.size());
    reconstruction.get_reconstruction().download(res.data(), res.size() * sizeof(float));
    const double dV = detail::pixel_dV(_cam, q, volume_conversion_factor);
    output.assign(npix, 0.0f);
    for (std::size_t i = 0; i < samples; ++i)
      for (std::size_t p = 0; p < npix; ++p)
        // float += double, as output[idx] += dV * contribution / samples
        output[p] = static_cast<float>(static_cast<double>(output[p]) +
                                       dV * static_cast<double>(res[i * npix + p]) /
                                           static_cast<double>(samples));
  }

  const camera& get_camera() const { return _cam; }

 private:
  camera _cam;
};

class volumetric_tomography {
 public:
  explicit volumetric_tomography(const camera& cam) : _cam(cam) {}
  void set_camera(const camera& cam) { _cam = cam; }
  const camera& get_camera() const { return _cam; }

  // cube of size_t(z_range / pixel size) slices (at least one)
  void create_tomographic_cube(reconstructing_data_crawler& reconstruction,
                               const reconstruction_quantity::quantity& q, double z_range,
                               std::vector<float>& output, double volume_conversion_factor = 1.0,
                               std::size_t max_points_per_pass = std::size_t(1) << 27) const {
    std::size_t nz = static_cast<std::size_t>(z_range / _cam.get_pixel_size());
    if (nz == 0) nz = 1;
    create_tomographic_cube(reconstruction, q, 0, nz, output, volume_conversion_factor,
                            max_points_per_pass);
  }

  // slices initial_z_step ... initial_z_step + num_steps - 1; output:
  // nx * ny * num_steps floats, x fastest, then y, then the slice
  void create_tomographic_cube(reconstructing_data_crawler& reconstruction,
                               const reconstruction_quantity::quantity& q,
                               std::size_t initial_z_step, std::size_t num_steps,
                               std::vector<float>& output, double volume_conversion_factor = 1.0,
                               std::size_t max_points_per_pass = std::size_t(1) << 27) const {
    output.clear();
    if (num_steps == 0) return;
    const std::size_t nx = _cam.get_num_pixels(0), ny = _cam.get_num_pixels(1);
    const std::size_t npix = nx * ny;
    output.assign(npix * num_steps, 0.0f);
    const std::size_t per_pass = std::max<std::size_t>(1, max_points_per_pass / npix);
    const double dV = detail::pixel_dV(_cam, q, volume_conversion_factor);
    const double ps = _cam.get_pixel_size();
    const double* pos = _cam.get_position();
    const double* la = _cam.get_look_at();
    for (std::size_t z0 = 0; z0 < num_steps; z0 += per_pass) {
      const std::size_t nz = std::min(per_pass, num_steps - z0);
      std::vector<device_vector4> pts(nz * npix);
      for (std::size_t k = 0; k < nz; ++k) {
        // volumetric_tomography.cpp:84-86: position + step * pixel size * look_at
        const double step = static_cast<double>(initial_z_step + z0 + k) * ps;
        double p[3];
        for (int i = 0; i < 3; ++i) p[i] = pos[i] + step * la[i];
        camera moving = _cam;
        moving.set_position(p);
        detail::pixel_points(moving, pts, k * npix);
      }
      auto dev = device_buffer::create<device_vector4>(pts.size());
      dev.upload(pts.data(), pts.size() * sizeof(device_vector4));
      reconstruction.get_backend().set_evaluation_layout(nx);
      reconstruction.run(dev, pts.size(), q);
      std::vector<float> res(pts.size());
      reconstruction.get_reconstruction().download(res.data(), res.size() * sizeof(float));
      for (std::size_t j = 0; j < res.size(); ++j)
        output[z0 * npix + j] = static_cast<float>(dV * static_cast<double>(res[j]));
    }
  }

 private:
  camera _cam;
};

}  // namespace illcrawl

#endif
