// illcrawl_backend.hpp -- C++ host interface of the MI355X projective smoothing
// path, over the C ABI of illcrawl_psg.h.
//
// The classes keep the names and method sets of illcrawl's reconstruction
// backend stack so that callers written against it take this path as a
// drop-in:
//   illcrawl::reconstruction_backend        reconstruction_backend.hpp:35-58
//   illcrawl::projective_smoothing_backend  projective_smoothing_backend.hpp:35-42
//   reconstruction_backends::dm::projective_smoothing_grid
//                                            projective_smoothing_backend_grid.hpp:28-44
//   reconstruction_backends::dm::projective_smoothing
//                                            projective_smoothing_reconstruction.hpp:33-89
//   reconstruction_backends::dm::grid       dm_reconstruction_backend_grid.hpp:32-79
//                                            (volumetric smoothing at 3-D points)
// with two substitutions of types the path does not need:
//   cl::Buffer   -> illcrawl::device_buffer (HIP device memory, shared ownership
//                   like cl::Buffer's reference counting)
//   H5::DataSet  -> illcrawl::dataset_ref (the dataset's name; the data
//                   crawler of illcrawl_crawler.hpp opens it in the snapshot)
// Errors throw illcrawl::error, a std::runtime_error carrying the ABI status
// (the reference throws qcl::qcl_error / std::runtime_error, qcl.hpp:79-118).
// All device work goes to one HIP stream per object (default: the null
// stream), as the reference uses one in-order queue per context.
#ifndef ILLCRAWL_BACKEND_HPP
#define ILLCRAWL_BACKEND_HPP

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstddef>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "illcrawl_psg.h"

namespace illcrawl {

class error : public std::runtime_error {
 public:
  error(ilc_status status, const std::string& what) : std::runtime_error(what), _status(status) {}
  ilc_status status() const noexcept { return _status; }

 private:
  ilc_status _status;
};

inline void check(ilc_status s, const char* context) {
  if (s != ILC_OK) throw error(s, std::string(context) + ": " + ilc_last_error());
}

inline void check_hip(hipError_t e, const char* context) {
  if (e != hipSuccess)
    throw error(e == hipErrorOutOfMemory ? ILC_ENOMEM : ILC_EHIP,
                std::string(context) + ": " + hipGetErrorString(e));
}

// particle layout of the reference (device_vector4, particle_grid.hpp:43-48)
struct device_vector4 {
  float x, y, z, w;
};
using device_scalar = float;

// Device memory with shared ownership (stands where cl::Buffer stands).
class device_buffer {
 public:
  device_buffer() = default;

  template <class T>
  static device_buffer create(std::size_t count) {
    device_buffer b;
    void* p = nullptr;
    if (count) check_hip(hipMalloc(&p, count * sizeof(T)), "device_buffer::create");
    b._p = std::shared_ptr<void>(p, [](void* q) {
      if (q) (void)hipFree(q);
    });
    b._bytes = count * sizeof(T);
    b._count = count;
    return b;
  }

  // wraps memory the caller owns (no free)
  static device_buffer borrow(void* p, std::size_t bytes, std::size_t count) {
    device_buffer b;
    b._p = std::shared_ptr<void>(p, [](void*) {});
    b._bytes = bytes;
    b._count = count;
    return b;
  }

  void* data() const noexcept { return _p.get(); }
  template <class T>
  T* as() const noexcept {
    return static_cast<T*>(_p.get());
  }
  std::size_t size_bytes() const noexcept { return _bytes; }
  std::size_t count() const noexcept { return _count; }
  explicit operator bool() const noexcept { return _p != nullptr; }

  void upload(const void* host, std::size_t bytes, hipStream_t s = nullptr) const {
    if (bytes > _bytes) throw error(ILC_EINVAL, "device_buffer::upload: too many bytes");
    check_hip(hipMemcpyAsync(_p.get(), host, bytes, hipMemcpyHostToDevice, s), "upload");
    check_hip(hipStreamSynchronize(s), "upload");
  }
  void download(void* host, std::size_t bytes, hipStream_t s = nullptr) const {
    if (bytes > _bytes) throw error(ILC_EINVAL, "device_buffer::download: too many bytes");
    check_hip(hipMemcpyAsync(host, _p.get(), bytes, hipMemcpyDeviceToHost, s), "download");
    check_hip(hipStreamSynchronize(s), "download");
  }

 private:
  std::shared_ptr<void> _p;
  std::size_t _bytes = 0, _count = 0;
};

// Names an additional per-particle dataset a backend needs (stands where
// H5::DataSet stands in get_required_additional_datasets).
struct dataset_ref {
  std::string name;
};

class reconstruction_backend {
 public:
  using particle = device_vector4;

  virtual std::vector<dataset_ref> get_required_additional_datasets() const = 0;
  virtual const device_buffer& retrieve_results() = 0;
  virtual std::string get_backend_name() const = 0;
  virtual void init_backend(std::size_t blocksize) = 0;
  virtual void setup_particles(const std::vector<particle>& particles,
                               const std::vector<device_buffer>& additional_dataset) = 0;
  // the same for particles already on the device (float4 x, y, z, w, world
  // frame), e.g. a block the data crawler compacted there
  virtual void setup_particles(const device_buffer& particles, std::size_t num_particles,
                               const std::vector<device_buffer>& additional_dataset) = 0;
  virtual void setup_evaluation_points(const device_buffer& evaluation_points,
                                       std::size_t num_points) = 0;
  virtual void run() = 0;
  // the stream all of the backend's device work goes to
  virtual hipStream_t stream() const = 0;
  // Row length of the evaluation points when they form an image, which the
  // gather tiles; 0 = one row.  Backends that lay out their own points
  // ignore it.
  virtual void set_evaluation_layout(std::size_t /*row_length*/) {}
  virtual ~reconstruction_backend() {}
};

class projective_smoothing_backend : virtual public reconstruction_backend {
 public:
  virtual void setup_projected_particles(const device_buffer& particles,
                                         std::size_t num_particles,
                                         const std::vector<device_buffer>& additional_datasets) = 0;
};

// Camera of camera.hpp:35-56 (screen centre, unit look-at, roll, screen
// width, pixel counts).
class camera {
 public:
  camera(const double position[3], const double look_at[3], double roll, double screen_width,
         std::size_t num_pix_x, std::size_t num_pix_y) {
    for (int i = 0; i < 3; ++i) {
      _c.position[i] = position[i];
      _c.look_at[i] = look_at[i];
    }
    _c.roll = roll;
    _c.width = screen_width;
    _c.nx = static_cast<uint32_t>(num_pix_x);
    _c.ny = static_cast<uint32_t>(num_pix_y);
    _c.depth = 0.0;
  }

  std::size_t get_num_pixels(int dim) const { return dim == 0 ? _c.nx : _c.ny; }
  double get_pixel_size() const { return _c.width / _c.nx; }
  const double* get_position() const { return _c.position; }
  const double* get_look_at() const { return _c.look_at; }
  // camera::set_position (camera.cpp:88-92); the screen moves with it
  void set_position(const double pos[3]) {
    for (int i = 0; i < 3; ++i) _c.position[i] = pos[i];
  }
  // camera::get_pixel_coordinate (camera.cpp:49-61), in double:
  // min_position + x ps b0 + y ps b1, x and y in pixel units
  void get_pixel_coordinate(double x, double y, double out[3]) const {
    double b0[3], b1[3], mn[3], ps = 0.0;
    check(ilc_camera_basis(&_c, b0, b1, mn, &ps), "camera::get_pixel_coordinate");
    const double sx = x * ps, sy = y * ps;
    for (int i = 0; i < 3; ++i) out[i] = (mn[i] + sx * b0[i]) + sy * b1[i];
  }
  // camera::get_pixel_area (camera.cpp:191-195)
  double get_pixel_area() const { return get_pixel_size() * get_pixel_size(); }
  double get_roll() const { return _c.roll; }
  double get_screen_width() const { return _c.width; }
  // the screen basis and its lower-left corner (camera::get_screen_basis_vector0,
  // get_screen_basis_vector1, get_screen_min_position), in double
  void get_screen_basis(double b0[3], double b1[3], double screen_min[3]) const {
    double ps = 0.0;
    check(ilc_camera_basis(&_c, b0, b1, screen_min, &ps), "camera basis");
  }

  // camera::generate_pixel_coordinates (camera.cpp:157-189): world frame
  device_buffer generate_pixel_coordinates(hipStream_t s = nullptr) const {
    auto out = device_buffer::create<device_vector4>(std::size_t(_c.nx) * _c.ny);
    check(ilc_camera_generate_pixel_coordinates(&_c, out.as<float>(), s),
          "camera::generate_pixel_coordinates");
    check_hip(hipStreamSynchronize(s), "camera::generate_pixel_coordinates");
    return out;
  }

  ilc_camera c_struct(double depth) const {
    ilc_camera c = _c;
    c.depth = depth;
    return c;
  }

 private:
  ilc_camera _c{};
};

namespace reconstruction_backends {
namespace dm {

// dm::projective_smoothing_grid on the MI355X path.  One additional dataset:
// the smoothing lengths (float per particle).
class projective_smoothing_grid : public projective_smoothing_backend {
 public:
  struct options {
    int device = 0;
    std::size_t target_particles_per_cell = 1000;  // projective_smoothing_backend_grid.cpp:29
    ilc_mode mode = ILC_MODE_REFERENCE;
    ilc_kernel kernel = ILC_KERNEL_QUARTIC_REF;
    int tile_w = 0, tile_h = 0;  // 0: default tile
    hipStream_t stream = nullptr;
  };

  explicit projective_smoothing_grid(const dataset_ref& smoothing_lengths)
      : projective_smoothing_grid(smoothing_lengths, options()) {}

  projective_smoothing_grid(const dataset_ref& smoothing_lengths, const options& o)
      : _hsml(smoothing_lengths), _o(o) {
    ilc_psg_config cfg{};
    cfg.device = o.device;
    cfg.target_particles_per_cell = o.target_particles_per_cell;
    cfg.mode = o.mode;
    cfg.kernel = o.kernel;
    cfg.tile_w = o.tile_w;
    cfg.tile_h = o.tile_h;
    ilc_psg* h = nullptr;
    check(ilc_psg_create(&cfg, &h), "projective_smoothing_grid");
    _h = std::shared_ptr<ilc_psg>(h, [](ilc_psg* q) { (void)ilc_psg_destroy(q); });
  }

  std::vector<dataset_ref> get_required_additional_datasets() const override { return {_hsml}; }

  // FP32 image (one value per evaluation point); valid until the next
  // setup_evaluation_points.  Synchronises the stream.
  const device_buffer& retrieve_results() override {
    if (_results.count() < _num_points) _results = device_buffer::create<float>(_num_points);
    check(ilc_psg_retrieve_results(_h.get(), 1.0f, _results.as<float>(), _o.stream),
          "retrieve_results");
    check_hip(hipStreamSynchronize(_o.stream), "retrieve_results");
    return _results;
  }

  std::string get_backend_name() const override { return ilc_psg_backend_name(_h.get()); }

  void init_backend(std::size_t blocksize) override {
    check(ilc_psg_init_backend(_h.get(), blocksize), "init_backend");
  }

  // host particles (camera frame): uploaded, then as setup_projected_particles
  void setup_particles(const std::vector<particle>& particles,
                       const std::vector<device_buffer>& additional_dataset) override {
    if (_upload.count() < particles.size())
      _upload = device_buffer::create<device_vector4>(particles.size());
    if (!particles.empty())
      _upload.upload(particles.data(), particles.size() * sizeof(particle), _o.stream);
    setup_projected_particles(_upload, particles.size(), additional_dataset);
  }

  // device particles, already in the camera frame
  void setup_particles(const device_buffer& particles, std::size_t num_particles,
                       const std::vector<device_buffer>& additional_dataset) override {
    setup_projected_particles(particles, num_particles, additional_dataset);
  }

  void setup_projected_particles(const device_buffer& particles, std::size_t num_particles,
                                 const std::vector<device_buffer>& additional_datasets) override {
    if (additional_datasets.size() != 1)
      throw error(ILC_EINVAL, "projective_smoothing_grid needs one additional dataset (h)");
    if (particles.count() < num_particles || additional_datasets[0].count() < num_particles)
      throw error(ILC_EINVAL, "setup_projected_particles: buffers shorter than num_particles");
    check(ilc_psg_setup_projected_particles(_h.get(), particles.as<float>(), num_particles,
                                            additional_datasets[0].as<float>(), _o.stream),
          "setup_projected_particles");
  }

  void setup_evaluation_points(const device_buffer& evaluation_points,
                               std::size_t num_points) override {
    if (evaluation_points.count() < num_points)
      throw error(ILC_EINVAL, "setup_evaluation_points: buffer shorter than num_points");
    _points = evaluation_points;  // borrowed by the library until the next call
    _num_points = num_points;
    check(ilc_psg_setup_evaluation_points(_h.get(), evaluation_points.as<float>(), num_points,
                                          _layout_nx, _o.stream),
          "setup_evaluation_points");
  }

  void run() override { check(ilc_psg_run(_h.get(), _o.stream), "run"); }

  // Row length of the evaluation points when they form an image (the gather
  // tiles the image); 0 = one row.  The camera wrapper sets it.
  void set_evaluation_layout(std::size_t row_length) override {
    _layout_nx = static_cast<uint32_t>(row_length);
  }

  // Adjoint of run + retrieve_results scaled by `scale` (exact mode, quartic):
  // per-particle dL/dw, dL/dh, dL/dx, dL/dy (camera frame), input order.
  void backward(const device_buffer& dl_dvalue, float scale, device_buffer& dl_dw,
                device_buffer& dl_dh, device_buffer& dl_dx, device_buffer& dl_dy) {
    ilc_grid_info gi{};
    check(ilc_psg_grid_info(_h.get(), &gi), "backward");
    const std::size_t n = gi.num_particles;
    for (device_buffer* b : {&dl_dw, &dl_dh, &dl_dx, &dl_dy})
      if (b->count() < n) *b = device_buffer::create<float>(n);
    check(ilc_psg_backward(_h.get(), dl_dvalue.as<float>(), scale, dl_dw.as<float>(),
                           dl_dh.as<float>(), dl_dx.as<float>(), dl_dy.as<float>(), _o.stream),
          "backward");
  }

  ilc_psg* handle() const noexcept { return _h.get(); }
  hipStream_t stream() const noexcept override { return _o.stream; }

 private:
  std::shared_ptr<ilc_psg> _h;
  dataset_ref _hsml;
  options _o;
  device_buffer _results, _upload, _points;
  std::size_t _num_points = 0;
  uint32_t _layout_nx = 0;
};

using projective_smoothing_grid_hip = projective_smoothing_grid;

// dm::projective_smoothing: projects particles and evaluation points into the
// camera frame, forwards to a projective smoothing backend and applies the
// quantity's effective dA at retrieval (projective_smoothing_reconstruction.cpp:27-205).
// The quantity is represented by its effective_line_of_sight_integration_dA
// (pixel area, integration depth) -> scale (quantity.hpp:198-202).
class projective_smoothing : public reconstruction_backend {
 public:
  using effective_dA_fn = std::function<double(double pixel_area, double depth)>;

  projective_smoothing(const camera& cam, double max_integration_depth, effective_dA_fn dA,
                       std::unique_ptr<projective_smoothing_backend> smoothing_backend,
                       hipStream_t stream = nullptr)
      : _cam(cam),
        _backend(std::move(smoothing_backend)),
        _depth(max_integration_depth),
        _dA(std::move(dA)),
        _stream(stream) {
    if (!_backend) throw error(ILC_EINVAL, "projective_smoothing: null backend");
  }

  std::vector<dataset_ref> get_required_additional_datasets() const override {
    return _backend->get_required_additional_datasets();
  }

  // The backend's image scaled in place by the effective dA (FP32, as
  // vector_scale); synchronises.
  const device_buffer& retrieve_results() override {
    const device_scalar pixel_area = static_cast<device_scalar>(_cam.get_pixel_area());
    const device_scalar dA = static_cast<device_scalar>(_dA ? _dA(pixel_area, _depth) : 1.0);
    const device_buffer& r = _backend->retrieve_results();
    check(ilc_vector_scale(r.as<float>(), _num_points, dA, _stream), "retrieve_results");
    check_hip(hipStreamSynchronize(_stream), "retrieve_results");
    return r;
  }

  std::string get_backend_name() const override {
    return "projective_smoothing via " + _backend->get_backend_name();
  }

  void init_backend(std::size_t blocksize) override { _backend->init_backend(blocksize); }

  void setup_particles(const std::vector<particle>& particles,
                       const std::vector<device_buffer>& additional_dataset) override {
    if (_projected.count() < particles.size())
      _projected = device_buffer::create<device_vector4>(particles.size());
    if (!particles.empty())
      _projected.upload(particles.data(), particles.size() * sizeof(particle), _stream);
    const ilc_camera c = _cam.c_struct(_depth);
    check(ilc_project_particles(&c, _projected.as<float>(), particles.size(),
                                _projected.as<float>(), _stream),
          "project_particles");
    _backend->setup_projected_particles(_projected, particles.size(), additional_dataset);
  }

  // The same for particles already on the device (float4 x, y, z, w, world
  // frame), e.g. a block the data crawler compacted on the device; the
  // caller's buffer is not modified.
  void setup_particles(const device_buffer& particles, std::size_t num_particles,
                       const std::vector<device_buffer>& additional_dataset) override {
    if (particles.count() < num_particles)
      throw error(ILC_EINVAL, "setup_particles: buffer shorter than num_particles");
    if (_projected.count() < num_particles)
      _projected = device_buffer::create<device_vector4>(num_particles);
    const ilc_camera c = _cam.c_struct(_depth);
    check(ilc_project_particles(&c, particles.as<float>(), num_particles, _projected.as<float>(),
                                _stream),
          "project_particles");
    _backend->setup_projected_particles(_projected, num_particles, additional_dataset);
  }

  hipStream_t stream() const override { return _stream; }

  // world-frame points (e.g. camera::generate_pixel_coordinates); the
  // caller's buffer is not modified
  void setup_evaluation_points(const device_buffer& evaluation_points,
                               std::size_t num_points) override {
    if (evaluation_points.count() < num_points)
      throw error(ILC_EINVAL, "setup_evaluation_points: buffer shorter than num_points");
    if (_points.count() < num_points) _points = device_buffer::create<device_vector4>(num_points);
    const ilc_camera c = _cam.c_struct(_depth);
    check(ilc_project_evaluation_points(&c, evaluation_points.as<float>(), num_points,
                                        _points.as<float>(), _stream),
          "project_evaluation_points");
    if (auto* g = dynamic_cast<projective_smoothing_grid*>(_backend.get()))
      g->set_evaluation_layout(num_points == _cam.get_num_pixels(0) * _cam.get_num_pixels(1)
                                   ? _cam.get_num_pixels(0)
                                   : 0);
    _backend->setup_evaluation_points(_points, num_points);
    _num_points = num_points;
  }

  void run() override { _backend->run(); }

  void set_camera(const camera& cam) { _cam = cam; }
  void set_integration_depth(double depth) { _depth = depth; }

  projective_smoothing_backend& backend() { return *_backend; }

 private:
  camera _cam;
  std::unique_ptr<projective_smoothing_backend> _backend;
  double _depth;
  effective_dA_fn _dA;
  hipStream_t _stream;
  device_buffer _projected, _points;
  std::size_t _num_points = 0;
};

// dm::grid: volumetric smoothing of 3-D particles at 3-D evaluation points
// with the 3-D quartic (dm_reconstruction_backend_grid.cpp:57-134,
// smoothing_kernel.cl:27-37).  One additional dataset: the smoothing
// lengths.  Results accumulate over run() calls until the next
// setup_evaluation_points.
class grid : public reconstruction_backend {
 public:
  struct options {
    int device = 0;
    std::size_t target_particles_per_cell = 64;  // dm_reconstruction_backend_grid.hpp:39
    ilc_mode mode = ILC_MODE_REFERENCE;
    ilc_kernel kernel = ILC_KERNEL_QUARTIC_REF;
    hipStream_t stream = nullptr;
  };

  explicit grid(const dataset_ref& smoothing_lengths) : grid(smoothing_lengths, options()) {}

  grid(const dataset_ref& smoothing_lengths, const options& o) : _hsml(smoothing_lengths), _o(o) {
    ilc_psg_config cfg{};
    cfg.device = o.device;
    cfg.target_particles_per_cell = o.target_particles_per_cell;
    cfg.mode = o.mode;
    cfg.kernel = o.kernel;
    ilc_grid3d* h = nullptr;
    check(ilc_grid3d_create(&cfg, &h), "dm::grid");
    _h = std::shared_ptr<ilc_grid3d>(h, [](ilc_grid3d* q) { (void)ilc_grid3d_destroy(q); });
  }

  std::vector<dataset_ref> get_required_additional_datasets() const override { return {_hsml}; }

  // FP32 values, one per evaluation point; valid until the next
  // setup_evaluation_points.  Synchronises the stream.
  const device_buffer& retrieve_results() override {
    if (_results.count() < _num_points) _results = device_buffer::create<float>(_num_points);
    check(ilc_grid3d_retrieve_results(_h.get(), 1.0f, _results.as<float>(), _o.stream),
          "retrieve_results");
    check_hip(hipStreamSynchronize(_o.stream), "retrieve_results");
    return _results;
  }

  std::string get_backend_name() const override { return "dm_smoothing/grid3d (hip gfx950)"; }

  void init_backend(std::size_t blocksize) override {
    check(ilc_grid3d_init_backend(_h.get(), blocksize), "init_backend");
  }

  void setup_particles(const std::vector<particle>& particles,
                       const std::vector<device_buffer>& additional_dataset) override {
    if (_upload.count() < particles.size())
      _upload = device_buffer::create<device_vector4>(particles.size());
    if (!particles.empty())
      _upload.upload(particles.data(), particles.size() * sizeof(particle), _o.stream);
    setup_particles(_upload, particles.size(), additional_dataset);
  }

  void setup_particles(const device_buffer& particles, std::size_t num_particles,
                       const std::vector<device_buffer>& additional_dataset) override {
    if (additional_dataset.size() != 1)
      throw error(ILC_EINVAL, "dm::grid needs one additional dataset (h)");
    if (particles.count() < num_particles || additional_dataset[0].count() < num_particles)
      throw error(ILC_EINVAL, "setup_particles: buffers shorter than num_particles");
    check(ilc_grid3d_setup_particles(_h.get(), particles.as<float>(), num_particles,
                                     additional_dataset[0].as<float>(), _o.stream),
          "setup_particles");
  }

  void setup_evaluation_points(const device_buffer& evaluation_points,
                               std::size_t num_points) override {
    if (evaluation_points.count() < num_points)
      throw error(ILC_EINVAL, "setup_evaluation_points: buffer shorter than num_points");
    _points = evaluation_points;  // borrowed by the library until the next call
    _num_points = num_points;
    check(ilc_grid3d_setup_evaluation_points(_h.get(), evaluation_points.as<float>(), num_points,
                                             _layout_nx, _o.stream),
          "setup_evaluation_points");
  }

  void run() override { check(ilc_grid3d_run(_h.get(), _o.stream), "run"); }

  void set_evaluation_layout(std::size_t row_length) override {
    _layout_nx = static_cast<uint32_t>(row_length);
  }

  hipStream_t stream() const noexcept override { return _o.stream; }
  ilc_grid3d* handle() const noexcept { return _h.get(); }

 private:
  std::shared_ptr<ilc_grid3d> _h;
  dataset_ref _hsml;
  options _o;
  device_buffer _results, _upload, _points;
  std::size_t _num_points = 0;
  uint32_t _layout_nx = 0;
};

}  // namespace dm
}  // namespace reconstruction_backends
}  // namespace illcrawl

#endif
