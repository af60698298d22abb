// Renders one image through the C++ backend classes of illcrawl_backend.hpp,
// the way projection::create_projection drives them (projection.cpp:68-110):
// dm::projective_smoothing over dm::projective_smoothing_grid,
// init_backend -> setup_particles -> setup_evaluation_points (camera pixel
// coordinates) -> run -> retrieve_results.
//
//   backend_demo DIR NX NY WIDTH DEPTH PX PY PZ LX LY LZ ROLL MODE KERNEL
//
// reads DIR/particles.f32 (n x float4, world frame) and DIR/hsml.f32 (n
// floats), writes DIR/image.f32 (NX*NY floats) and, in exact mode with a
// quartic kernel, DIR/grad.f32 (4 x n floats: dL/dw, dL/dh, dL/dx, dL/dy for
// L = sum of the image).  Also checks that a call out of order throws
// illcrawl::error with ILC_ENOTREADY.  Exit status 0 on success.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "illcrawl_backend.hpp"

namespace {

std::vector<float> read_f32(const std::string& path) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<float> v(static_cast<std::size_t>(bytes) / sizeof(float));
  if (std::fread(v.data(), sizeof(float), v.size(), f) != v.size())
    throw std::runtime_error("short read " + path);
  std::fclose(f);
  return v;
}

void write_f32(const std::string& path, const std::vector<float>& v) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(v.data(), sizeof(float), v.size(), f);
  std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
  namespace dm = illcrawl::reconstruction_backends::dm;
  if (argc != 15) {
    std::fprintf(stderr, "usage: %s DIR NX NY WIDTH DEPTH PX PY PZ LX LY LZ ROLL MODE KERNEL\n",
                 argv[0]);
    return 2;
  }
  try {
    const std::string dir = argv[1];
    const std::size_t nx = std::strtoul(argv[2], nullptr, 10);
    const std::size_t ny = std::strtoul(argv[3], nullptr, 10);
    const double width = std::atof(argv[4]), depth = std::atof(argv[5]);
    const double pos[3] = {std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8])};
    const double look[3] = {std::atof(argv[9]), std::atof(argv[10]), std::atof(argv[11])};
    const double roll = std::atof(argv[12]);
    const int mode = std::atoi(argv[13]), kernel = std::atoi(argv[14]);

    // call order is checked and reported as an exception
    {
      dm::projective_smoothing_grid g(illcrawl::dataset_ref{"SubfindHsml"});
      bool thrown = false;
      try {
        g.run();
      } catch (const illcrawl::error& e) {
        thrown = e.status() == ILC_ENOTREADY;
      }
      if (!thrown) {
        std::fprintf(stderr, "run before setup did not throw ILC_ENOTREADY\n");
        return 1;
      }
      std::printf("error-path ok\n");
    }

    const std::vector<float> p = read_f32(dir + "/particles.f32");
    const std::vector<float> hs = read_f32(dir + "/hsml.f32");
    const std::size_t n = hs.size();
    if (p.size() != 4 * n) throw std::runtime_error("particles.f32 / hsml.f32 size mismatch");
    std::vector<illcrawl::device_vector4> particles(n);
    for (std::size_t i = 0; i < n; ++i)
      particles[i] = {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]};
    auto h_dev = illcrawl::device_buffer::create<float>(n);
    if (n) h_dev.upload(hs.data(), n * sizeof(float));

    dm::projective_smoothing_grid::options o;
    o.mode = static_cast<ilc_mode>(mode);
    o.kernel = static_cast<ilc_kernel>(kernel);
    auto grid = std::make_unique<dm::projective_smoothing_grid>(
        illcrawl::dataset_ref{"SubfindHsml"}, o);
    dm::projective_smoothing_grid* grid_ptr = grid.get();
    illcrawl::camera cam(pos, look, roll, width, nx, ny);
    dm::projective_smoothing ps(cam, depth, [](double, double) { return 1.0; }, std::move(grid));
    std::printf("backend: %s\n", ps.get_backend_name().c_str());

    ps.init_backend(n);
    ps.setup_particles(particles, {h_dev});
    const illcrawl::device_buffer pix = cam.generate_pixel_coordinates();
    ps.setup_evaluation_points(pix, nx * ny);
    ps.run();
    const illcrawl::device_buffer& img = ps.retrieve_results();
    std::vector<float> out(nx * ny);
    img.download(out.data(), out.size() * sizeof(float));
    write_f32(dir + "/image.f32", out);
    std::printf("image written (%zu x %zu)\n", nx, ny);

    if (mode == ILC_MODE_EXACT && kernel != ILC_KERNEL_CUBIC) {
      std::vector<float> ones(nx * ny, 1.0f);
      auto g = illcrawl::device_buffer::create<float>(nx * ny);
      g.upload(ones.data(), ones.size() * sizeof(float));
      illcrawl::device_buffer gw, gh, gx, gy;
      grid_ptr->backward(g, 1.0f, gw, gh, gx, gy);
      illcrawl::check_hip(hipDeviceSynchronize(), "backward");
      std::vector<float> grads(4 * n);
      gw.download(grads.data(), n * sizeof(float));
      gh.download(grads.data() + n, n * sizeof(float));
      gx.download(grads.data() + 2 * n, n * sizeof(float));
      gy.download(grads.data() + 3 * n, n * sizeof(float));
      write_f32(dir + "/grad.f32", grads);
      std::printf("gradients written\n");
    }
  } catch (const illcrawl::error& e) {
    std::fprintf(stderr, "illcrawl::error (status %d): %s\n", static_cast<int>(e.status()),
                 e.what());
    return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
