// illcrawl_render on the MI355X path: the reference's render tool
// (illcrawl_render.cpp:116-343) for the quantities this build reconstructs by
// projective smoothing (dark matter, stars, and gas projected by smoothing
// its cells).
//
//   -t image       one projection of the snapshot: the data crawler streams
//                  the quantity's particle group in row blocks through
//                  dm::projective_smoothing over the grid backend
//                  (illcrawl_render.cpp:124-146, projection.cpp:55-110);
//   -t animation   frames of the camera rotating about the gas distribution's
//                  centre, phi about y at animation.phi_speed turns, theta
//                  about x at one turn (illcrawl_render.cpp:189-228,
//                  animation.cpp:41-178).  The crawler collects the
//                  quantity's particles once and every frame is a view of one
//                  batched render (ilc_batch_*): one persistent gather launch
//                  instead of one crawl per frame;
//   -t tomography  the cube of volumetric slices one pixel apart along the
//                  camera's look-at direction over the recommended depth, the
//                  3-D smoothing dm::grid through the crawler
//                  (illcrawl_render.cpp:230-247, volumetric_tomography.cpp:29-102).
//
// The output is a FITS image (nx x ny) or cube (nx x ny x frames or slices)
// with the reference's header keys (illcrawl_render.cpp:252-270).  The
// spectra targets need the Voronoi reconstruction and the spectral
// quantities, which this build does not have; they are refused by name.
// Unlike the reference, a failed run exits with status 1.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "illcrawl_app.hpp"
#include "illcrawl_crawler.hpp"
#include "illcrawl_volumetric.hpp"

namespace {

using namespace illcrawl;
namespace dm = reconstruction_backends::dm;

// A backend that keeps what the crawler hands it: the kept particles of every
// block (world frame, w = the quantity) and their smoothing lengths, appended
// on the device.  Used to render all animation frames from one crawl.
class particle_collector : public reconstruction_backend {
 public:
  std::vector<dataset_ref> get_required_additional_datasets() const override {
    return {dataset_ref{io::illustris_data_loader::get_dm_smoothing_length_identifier()}};
  }
  const device_buffer& retrieve_results() override {
    throw error(ILC_EINVAL, "particle_collector has no results");
  }
  std::string get_backend_name() const override { return "particle collector"; }
  void init_backend(std::size_t) override {}
  void setup_particles(const std::vector<particle>& particles,
                       const std::vector<device_buffer>& additional_dataset) override {
    auto d = device_buffer::create<device_vector4>(particles.size());
    if (!particles.empty()) d.upload(particles.data(), particles.size() * sizeof(particle));
    setup_particles(d, particles.size(), additional_dataset);
  }
  void setup_particles(const device_buffer& particles, std::size_t n,
                       const std::vector<device_buffer>& additional_dataset) override {
    if (additional_dataset.size() != 1) throw error(ILC_EINVAL, "collector: one extra dataset");
    if (!n) return;
    reserve(_n + n);
    check_hip(hipMemcpyAsync(_p4.as<device_vector4>() + _n, particles.data(),
                             n * sizeof(device_vector4), hipMemcpyDeviceToDevice, nullptr),
              "collector");
    check_hip(hipMemcpyAsync(_h.as<float>() + _n, additional_dataset[0].data(), n * sizeof(float),
                             hipMemcpyDeviceToDevice, nullptr),
              "collector");
    _n += n;
  }
  void setup_evaluation_points(const device_buffer&, std::size_t) override {}
  void run() override {}
  hipStream_t stream() const override { return nullptr; }

  std::size_t size() const { return _n; }
  // positions (x, y, z interleaved) and w as separate arrays, as the batch
  // takes them
  void split(device_buffer& pos3, device_buffer& w) const {
    pos3 = device_buffer::create<float>(3 * std::max<std::size_t>(_n, 1));
    w = device_buffer::create<float>(std::max<std::size_t>(_n, 1));
    if (!_n) return;
    check_hip(hipMemcpy2DAsync(pos3.data(), 3 * sizeof(float), _p4.data(), sizeof(device_vector4),
                               3 * sizeof(float), _n, hipMemcpyDeviceToDevice, nullptr),
              "collector split");
    check_hip(hipMemcpy2DAsync(w.data(), sizeof(float), _p4.as<float>() + 3,
                               sizeof(device_vector4), sizeof(float), _n, hipMemcpyDeviceToDevice,
                               nullptr),
              "collector split");
    check_hip(hipStreamSynchronize(nullptr), "collector split");
  }
  const device_buffer& hsml() const { return _h; }

 private:
  void reserve(std::size_t n) {
    if (n <= _p4.count()) return;
    const std::size_t cap = std::max(n, 2 * _p4.count());
    auto p4 = device_buffer::create<device_vector4>(cap);
    auto h = device_buffer::create<float>(cap);
    if (_n) {
      check_hip(hipMemcpyAsync(p4.data(), _p4.data(), _n * sizeof(device_vector4),
                               hipMemcpyDeviceToDevice, nullptr),
                "collector");
      check_hip(hipMemcpyAsync(h.data(), _h.data(), _n * sizeof(float), hipMemcpyDeviceToDevice,
                               nullptr),
                "collector");
      check_hip(hipStreamSynchronize(nullptr), "collector");
    }
    _p4 = p4;
    _h = h;
  }
  device_buffer _p4, _h;
  std::size_t _n = 0;
};

// math::matrix_create_rotation_matrix (math.hpp:109-140): rotation by alpha
// about the unit axis, right-handed
void rotation_matrix(const double axis[3], double alpha, double R[9]) {
  const double c = std::cos(alpha), s = std::sin(alpha), k = 1.0 - c;
  const double x = axis[0], y = axis[1], z = axis[2];
  const double m[9] = {x * x * k + c,     x * y * k - z * s, x * z * k + y * s,
                       x * y * k + z * s, y * y * k + c,     y * z * k - x * s,
                       x * z * k - y * s, y * z * k + x * s, z * z * k + c};
  for (int i = 0; i < 9; ++i) R[i] = m[i];
}

void mat_mul(const double A[9], const double B[9], double C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

void mat_vec(const double R[9], const double v[3], double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = R[3 * i] * v[0] + R[3 * i + 1] * v[1] + R[3 * i + 2] * v[2];
}

void cross(const double a[3], const double b[3], double out[3]) {
  out[0] = a[1] * b[2] - a[2] * b[1];
  out[1] = a[2] * b[0] - a[0] * b[2];
  out[2] = a[0] * b[1] - a[1] * b[0];
}

// camera::rotate (camera.cpp:107-122) rotates the position about the centre,
// the look-at vector and the screen basis with one matrix.  The camera here
// derives its basis from look_at and roll (camera::update_basis_vectors), so
// the rotated camera gets the roll whose basis is the rotated one: the angle
// of the rotated b0 from the unrolled x axis of the new look_at.
camera rotated_camera(const camera& cam, const double R[9], const double center[3]) {
  double b0[3], b1[3], mn[3];
  cam.get_screen_basis(b0, b1, mn);
  double rel[3], pos[3], la[3], t[3];
  for (int i = 0; i < 3; ++i) rel[i] = cam.get_position()[i] - center[i];
  mat_vec(R, rel, pos);
  for (int i = 0; i < 3; ++i) pos[i] += center[i];
  mat_vec(R, cam.get_look_at(), la);
  mat_vec(R, b0, t);
  double v[3] = {0.0, 0.0, 1.0};
  if (la[0] == 0.0 && la[1] == 0.0 && la[2] == 1.0) {
    v[1] = 1.0;
    v[2] = 0.0;
  }
  double u[3], w[3];
  cross(v, la, u);
  const double nu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (double& x : u) x /= nu;
  cross(la, u, w);
  const double roll = std::atan2(t[0] * w[0] + t[1] * w[1] + t[2] * w[2],
                                 t[0] * u[0] + t[1] * u[1] + t[2] * u[2]);
  return camera(pos, la, roll, cam.get_screen_width(), cam.get_num_pixels(0),
                cam.get_num_pixels(1));
}

// camera_movement::dual_axis_rotation_around_point (animation.cpp:142-178):
// frame i of n rotates the initial camera by M_phi(a) M_theta(a * theta_range
// / phi_range), a = i * phi_range / n (ranges in degrees)
camera animation_frame(const camera& initial, const double center[3], double phi_speed,
                       std::size_t frame, std::size_t num_frames) {
  const double phi_range = (2.0 * M_PI / 360.0) * (phi_speed * 360.0);
  const double per_frame = phi_range / static_cast<double>(num_frames);
  const double alpha_phi = static_cast<double>(frame) * per_frame;
  const double alpha_theta = alpha_phi * (360.0 / (phi_speed * 360.0));
  const double phi_axis[3] = {0.0, 1.0, 0.0}, theta_axis[3] = {1.0, 0.0, 0.0};
  double Mp[9], Mt[9], R[9];
  rotation_matrix(phi_axis, alpha_phi, Mp);
  rotation_matrix(theta_axis, alpha_theta, Mt);
  mat_mul(Mp, Mt, R);
  return rotated_camera(initial, R, center);
}

ilc_view view_of(const camera& cam, double depth, float dA) {
  ilc_view v{};
  v.cam = cam.c_struct(depth);
  v.quantity = ILC_QUANTITY_MASS;
  v.scale = dA;
  return v;
}

float projected_dA_f32(const reconstruction_quantity::quantity& q, const camera& cam,
                       double depth) {
  const float pixel_area = static_cast<float>(cam.get_pixel_area());
  return static_cast<float>(reconstruction_quantity::projected_dA(q, pixel_area, depth));
}

std::vector<float> render_image(const illcrawl_app& app, const reconstruction_quantity::quantity& q,
                                const camera& cam) {
  const double depth = app.get_recommended_integration_depth();
  auto grid = std::make_unique<dm::projective_smoothing_grid>(
      dataset_ref{io::illustris_data_loader::get_dm_smoothing_length_identifier()});
  const reconstruction_quantity::quantity* qp = &q;
  dm::projective_smoothing ps(
      cam, depth,
      [qp](double pixel_area, double range) {
        return reconstruction_quantity::projected_dA(*qp, pixel_area, range);
      },
      std::move(grid));
  reconstructing_data_crawler crawler(app.get_gas_distribution_volume_cutout(),
                                      &app.get_data_loader(), app.get_data_crawling_blocksize(),
                                      &ps);
  const std::size_t P = cam.get_num_pixels(0) * cam.get_num_pixels(1);
  device_buffer pts = cam.generate_pixel_coordinates();
  crawler.run(pts, P, q);
  const device_buffer& img = ps.retrieve_results();
  std::vector<float> out(P);
  img.download(out.data(), P * sizeof(float));
  return out;
}

std::vector<float> render_animation(const illcrawl_app& app,
                                    const reconstruction_quantity::quantity& q,
                                    const camera& cam, std::size_t num_frames, double phi_speed) {
  if (phi_speed <= 0) throw std::invalid_argument("animation.phi_speed must be > 0.");
  const double depth = app.get_recommended_integration_depth();
  particle_collector collect;
  reconstructing_data_crawler crawler(app.get_gas_distribution_volume_cutout(),
                                      &app.get_data_loader(), app.get_data_crawling_blocksize(),
                                      &collect);
  device_buffer none = device_buffer::create<device_vector4>(1);
  crawler.run(none, 0, q);
  const std::size_t nx = cam.get_num_pixels(0), ny = cam.get_num_pixels(1);
  std::vector<float> cube(nx * ny * num_frames, 0.0f);
  if (num_frames == 0) return cube;
  const float dA = projected_dA_f32(q, cam, depth);
  std::vector<ilc_view> views;
  for (std::size_t f = 0; f < num_frames; ++f)
    views.push_back(view_of(
        animation_frame(cam, app.get_gas_distribution_center(), phi_speed, f, num_frames), depth,
        dA));
  device_buffer pos3, w;
  collect.split(pos3, w);
  ilc_psg_config cfg{};
  cfg.target_particles_per_cell = 1000;  // projective_smoothing_backend_grid.cpp:29
  cfg.mode = ILC_MODE_REFERENCE;
  cfg.kernel = ILC_KERNEL_QUARTIC_REF;
  ilc_batch* b = nullptr;
  check(ilc_batch_create(&cfg, &b), "animation");
  std::unique_ptr<ilc_batch, void (*)(ilc_batch*)> guard(b, [](ilc_batch* p) {
    (void)ilc_batch_destroy(p);
  });
  check(ilc_batch_set_particles(b, pos3.as<float>(), w.as<float>(), collect.hsml().as<float>(),
                                nullptr, nullptr, nullptr, collect.size()),
        "animation particles");
  check(ilc_batch_set_views(b, views.data(), int32_t(views.size()), nullptr), "animation views");
  check(ilc_batch_render(b, 0, nullptr), "animation render");
  auto img = device_buffer::create<float>(nx * ny);
  for (std::size_t f = 0; f < num_frames; ++f) {
    check(ilc_batch_image(b, int32_t(f), img.as<float>(), nullptr), "animation frame");
    img.download(cube.data() + f * nx * ny, nx * ny * sizeof(float));
  }
  return cube;
}

std::vector<float> render_tomography(const illcrawl_app& app,
                                     const reconstruction_quantity::quantity& q,
                                     const camera& cam, std::size_t& num_slices) {
  dm::grid grid(dataset_ref{io::illustris_data_loader::get_dm_smoothing_length_identifier()});
  reconstructing_data_crawler crawler(app.get_gas_distribution_volume_cutout(),
                                      &app.get_data_loader(), app.get_data_crawling_blocksize(),
                                      &grid);
  std::vector<float> cube;
  volumetric_tomography tomography(cam);
  tomography.create_tomographic_cube(crawler, q, app.get_recommended_integration_depth(), cube,
                                     app.get_unit_converter().volume_conversion_factor());
  num_slices = cube.size() / (cam.get_num_pixels(0) * cam.get_num_pixels(1));
  return cube;
}

}  // namespace

int main(int argc, char** argv) {
  illcrawl_app app;
  option_set& o = app.options();
  o.add("output", 'o', "illcrawl_render.fits", "the output fits file");
  o.add_required("render_target", 't',
                 "the target type of the render: image, animation, tomography (xray_spectrum "
                 "and chandra_spectrum are not in this build)");
  o.add("xray_spectrum.min_energy", 0, "0.1", "minimum energy for X-Ray spectra [keV]");
  o.add("xray_spectrum.max_energy", 0, "10", "maximum energy for X-Ray spectra [keV]");
  o.add("xray_spectrum.num_energies", 0, "100", "Number of sampled energies for X-ray spectra");
  o.add("chandra_spectrum.num_energies", 0, "100", "Number of sampled energies for chandra spectra");
  o.add("animation.num_frames", 0, "100", "Number of frames for animations");
  o.add("animation.phi_speed", 0, "3",
        "Speed of the rotation around the phi axis relative to the rotation speed around the "
        "theta axis");
  o.add("tomography.num_slices", 0, "100",
        "Number of slices for tomographies (recorded; the cube spans the recommended depth)");
  o.add("integration.absolute_tolerance", 0, "0",
        "Absolute tolerance for the line-of-sight integration for voronoi (gas) quantities");
  o.add("integration.relative_tolerance", 0, "0.01",
        "Relative tolerance for line-of-sight integration for voronoi (gas) quantities");
  o.add("output.resolution.x", 0, "1024", "Number of pixels in x direction");
  o.add("output.resolution.y", 0, "1024", "Number of pixels in y direction");
  quantity_command_line_parser quantity_parser;
  quantity_parser.register_options(o);
  try {
    app.parse_command_line(argc, argv);
    std::unique_ptr<reconstruction_quantity::quantity> quantity =
        quantity_parser.create_quantity(app, o);
    const double abs_tol = o.get_double("integration.absolute_tolerance");
    const double rel_tol = o.get_double("integration.relative_tolerance");
    if (abs_tol <= 0.0 && rel_tol <= 0.0)
      throw std::invalid_argument(
          "Either absolute integration tolerance or relative tolerance has to be > 0");
    const std::string target = o.get("render_target");
    const std::string output = o.get("output");
    const std::size_t nx = static_cast<std::size_t>(o.get_ull("output.resolution.x"));
    const std::size_t ny = static_cast<std::size_t>(o.get_ull("output.resolution.y"));
    if (nx == 0 || ny == 0) throw std::invalid_argument("output resolution must be > 0");
    camera cam = app.create_distribution_centered_camera(nx, ny);
    std::string camera_name = "camera";
    std::cout << "Gas distribution center: " << app.get_gas_distribution_center()[0] << ", "
              << app.get_gas_distribution_center()[1] << ", "
              << app.get_gas_distribution_center()[2] << std::endl;
    if (target == "image") {
      std::cout << "Starting projection..." << std::endl;
      std::vector<float> img = render_image(app, *quantity, cam);
      io::fits<float>(output).save(img.data(), {long(nx), long(ny)});
    } else if (target == "animation") {
      camera_name = "camera_frame0";
      std::cout << "Starting animation..." << std::endl;
      const std::size_t frames = static_cast<std::size_t>(o.get_ull("animation.num_frames"));
      std::vector<float> cube =
          render_animation(app, *quantity, cam, frames, o.get_double("animation.phi_speed"));
      io::fits<float>(output).save(cube.data(), {long(nx), long(ny), long(frames)});
    } else if (target == "tomography") {
      camera_name = "camera_slice0";
      std::cout << "Starting tomography..." << std::endl;
      std::size_t nz = 0;
      std::vector<float> cube = render_tomography(app, *quantity, cam, nz);
      io::fits<float>(output).save(cube.data(), {long(nx), long(ny), long(nz)});
    } else if (target == "xray_spectrum" || target == "chandra_spectrum") {
      throw std::invalid_argument("render target " + target +
                                  " needs the Voronoi reconstruction, which this build does "
                                  "not have");
    } else {
      throw std::runtime_error("Invalid render target: " + target);
    }
    io::fits_header header(output);
    app.save_settings_to_fits(header);
    app.save_camera_settings_to_fits_header(header, camera_name, cam);
    quantity_parser.save_configuration_to_fits_header(header, o);
    header.add_entry("integration.abs_tol", abs_tol, "Absolute integration tolerance");
    header.add_entry("integration.rel_tol", rel_tol, "Relative integration tolerance");
    header.add_entry("original_filename", output, "Original output filename");
    header.add_entry("type", target, "Type of this file");
  } catch (const std::invalid_argument& e) {
    std::cout << "Error: " << e.what() << "\n\n" << o.usage() << std::endl;
    return 1;
  } catch (const std::exception& e) {
    std::cout << "Error: " << e.what() << std::endl;
    return 1;
  } catch (const H5::Exception& e) {
    std::cout << "Error: HDF5: " << e.getFuncName() << ": " << e.getDetailMsg() << std::endl;
    return 1;
  }
  return 0;
}
