// The data crawler end to end through include/illcrawl_crawler.hpp:
//
//   crawl_demo write-snapshot <dir> <n>
//       <dir>/{pos.f64 (n*3), rho.f32, u.f32, xe.f32, mass.f32, hsml.f32}
//       -> <dir>/snap.hdf5, group PartType0 (Coordinates float64 n x 3,
//       Density, InternalEnergy, ElectronAbundance, Masses, SmoothingLength
//       float32), the Illustris layout; with <dir>/subfind_hsml.f32 also
//       PartType0/SubfindHsml; with <dir>/dm_pos.f64 (n1*3) and dm_hsml.f32
//       group PartType1 (Coordinates, SubfindHsml); with <dir>/star_pos.f64
//       (n4*3), star_hsml.f32, star_sft.f32 and star_mass.f32 group PartType4
//       (Coordinates, SubfindHsml, GFM_StellarFormationTime, Masses)
//   crawl_demo write-fits <raw.f32> <nx> <ny> <out.fits>   (cfitsio; checks its own read back)
//   crawl_demo fits-to-raw <in.fits> <out.f32>              (cfitsio read; prints the extents)
//   crawl_demo render <dir> <blocksize> <quantity: density|mass|xray> <nx> <width> <box>
//       renders <dir>/snap.hdf5 with the crawler (volume: the box, periodic,
//       no border; camera along z through the box centre) into
//       <dir>/image_<quantity>.fits; xray reads <dir>/gaunt.f32 (gnx*gny,
//       dims in <dir>/gaunt.dims)
//   crawl_demo volume <dir> <blocksize> <quantity: density|mass> <nx> <width> <box> <nz> <samples>
//       volumetric smoothing (dm::grid) of <dir>/snap.hdf5 through the
//       crawler: camera along z through the box centre; nz = 0 writes the
//       slice with <samples> additional jittered samples per pixel (seed 7)
//       to <dir>/slice_<quantity>.fits, nz > 0 the tomographic cube of nz
//       slices to <dir>/cube_<quantity>.fits
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "illcrawl_crawler.hpp"
#include "illcrawl_volumetric.hpp"

namespace {

template <class T>
std::vector<T> read_raw(const std::string& path, std::size_t n) {
  std::vector<T> v(n);
  std::ifstream f(path, std::ios::binary);
  if (!f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(T))))
    throw std::runtime_error("cannot read " + path);
  return v;
}

bool exists(const std::string& path) { return std::ifstream(path).good(); }

std::size_t file_size(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + path);
  return static_cast<std::size_t>(f.tellg());
}

template <class T>
void write_dataset(H5::Group& g, const std::string& name, const std::vector<T>& v, hsize_t rows,
                   hsize_t cols, const H5::PredType& type) {
  hsize_t dims[2] = {rows, cols};
  H5::DataSpace sp(cols > 1 ? 2 : 1, dims);
  H5::DataSet d = g.createDataSet(name, type, sp);
  d.write(v.data(), type);
}

int write_snapshot(const std::string& dir, std::size_t n) {
  H5::H5File f(dir + "/snap.hdf5", H5F_ACC_TRUNC);
  H5::Group g = f.createGroup("PartType0");
  write_dataset(g, "Coordinates", read_raw<double>(dir + "/pos.f64", 3 * n), n, 3,
                H5::PredType::NATIVE_DOUBLE);
  const char* names[5] = {"Density", "InternalEnergy", "ElectronAbundance", "Masses",
                          "SmoothingLength"};
  const char* files[5] = {"rho", "u", "xe", "mass", "hsml"};
  for (int k = 0; k < 5; ++k)
    write_dataset(g, names[k], read_raw<float>(dir + "/" + files[k] + ".f32", n), n, 1,
                  H5::PredType::NATIVE_FLOAT);
  if (exists(dir + "/subfind_hsml.f32"))
    write_dataset(g, "SubfindHsml", read_raw<float>(dir + "/subfind_hsml.f32", n), n, 1,
                  H5::PredType::NATIVE_FLOAT);
  // other particle groups: rows from the size of the coordinate file
  auto group = [&](const std::string& name, const std::string& prefix,
                   const std::vector<std::pair<std::string, std::string>>& cols) {
    const std::string pos = dir + "/" + prefix + "_pos.f64";
    if (!exists(pos)) return;
    const std::size_t m = file_size(pos) / (3 * sizeof(double));
    H5::Group gg = f.createGroup(name);
    write_dataset(gg, "Coordinates", read_raw<double>(pos, 3 * m), m, 3,
                  H5::PredType::NATIVE_DOUBLE);
    for (const auto& c : cols)
      write_dataset(gg, c.first, read_raw<float>(dir + "/" + prefix + "_" + c.second + ".f32", m),
                    m, 1, H5::PredType::NATIVE_FLOAT);
  };
  group("PartType1", "dm", {{"SubfindHsml", "hsml"}});
  group("PartType4", "star",
        {{"SubfindHsml", "hsml"}, {"GFM_StellarFormationTime", "sft"}, {"Masses", "mass"}});
  return 0;
}

int render(const std::string& dir, std::size_t blocksize, const std::string& qname, int nx,
           double width, double box) {
  namespace dm = illcrawl::reconstruction_backends::dm;
  illcrawl::io::illustris_data_loader data(dir + "/snap.hdf5", 0u);
  const double c = 0.5 * box;
  const double position[3] = {c, c, -1.0}, look_at[3] = {0.0, 0.0, 1.0};
  illcrawl::camera cam(position, look_at, 0.0, width, nx, nx);
  dm::projective_smoothing_grid::options o;  // reference mode, quartic
  auto grid = std::make_unique<dm::projective_smoothing_grid>(
      illcrawl::dataset_ref{illcrawl::io::illustris_data_loader::get_smoothing_length_identifier()},
      o);
  std::unique_ptr<illcrawl::reconstruction_quantity::quantity> q;
  if (qname == "density") {
    q.reset(new illcrawl::reconstruction_quantity::projected_density());
  } else if (qname == "mass") {
    q.reset(new illcrawl::reconstruction_quantity::mass());
  } else if (qname == "xray") {
    std::ifstream dims(dir + "/gaunt.dims");
    int gnx = 0, gny = 0;
    dims >> gnx >> gny;
    q.reset(new illcrawl::reconstruction_quantity::xray_flux(
        read_raw<float>(dir + "/gaunt.f32", std::size_t(gnx) * gny), gnx, gny, 0.0, 1.0e5, 0.5,
        5.0, 100));
  } else {
    throw std::runtime_error("unknown quantity " + qname);
  }
  const double depth = box + 2.0;
  const illcrawl::reconstruction_quantity::quantity* qp = q.get();
  dm::projective_smoothing ps(cam, depth,
                              [qp](double dA, double range) {
                                return qp->effective_line_of_sight_integration_dA(dA, range);
                              },
                              std::move(grid));
  ilc_volume vol{};
  for (int a = 0; a < 3; ++a) {
    vol.center[a] = c;
    vol.extent[a] = box;
    vol.periodic_wraparound[a] = box;
  }
  vol.border = 0.0;
  illcrawl::reconstructing_data_crawler crawler(vol, &data, blocksize, &ps);
  const std::size_t P = std::size_t(nx) * nx;
  illcrawl::device_buffer pts = cam.generate_pixel_coordinates();
  crawler.run(pts, P, *q);
  const illcrawl::device_buffer& img = ps.retrieve_results();
  std::vector<float> host(P);
  img.download(host.data(), P * sizeof(float));
  illcrawl::io::fits<float>(dir + "/image_" + qname + ".fits").save(host.data(), {nx, nx});
  std::printf("crawl_demo: %s %dx%d, blocksize %zu, sum %.9g\n", qname.c_str(), nx, nx,
              blocksize, [&] {
                double s = 0;
                for (float v : host) s += v;
                return s;
              }());
  return 0;
}

int volume(const std::string& dir, std::size_t blocksize, const std::string& qname, int nx,
           double width, double box, int nz, int samples) {
  namespace dm = illcrawl::reconstruction_backends::dm;
  illcrawl::io::illustris_data_loader data(dir + "/snap.hdf5", 0u);
  const double c = 0.5 * box;
  const double position[3] = {c, c, c}, look_at[3] = {0.0, 0.0, 1.0};
  illcrawl::camera cam(position, look_at, 0.0, width, nx, nx);
  dm::grid grid(
      illcrawl::dataset_ref{illcrawl::io::illustris_data_loader::get_smoothing_length_identifier()});
  std::unique_ptr<illcrawl::reconstruction_quantity::quantity> q;
  if (qname == "density")
    q.reset(new illcrawl::reconstruction_quantity::projected_density());
  else if (qname == "mass")
    q.reset(new illcrawl::reconstruction_quantity::mass());
  else
    throw std::runtime_error("unknown quantity " + qname);
  ilc_volume vol{};
  for (int a = 0; a < 3; ++a) {
    vol.center[a] = c;
    vol.extent[a] = box;
    vol.periodic_wraparound[a] = box;
  }
  vol.border = 0.0;
  illcrawl::reconstructing_data_crawler crawler(vol, &data, blocksize, &grid);
  std::vector<float> out;
  std::string path;
  if (nz == 0) {
    illcrawl::volumetric_slice(cam).create_slice(crawler, *q, out, std::size_t(samples), 7);
    path = dir + "/slice_" + qname + ".fits";
    illcrawl::io::fits<float>(path).save(out.data(), {nx, nx});
  } else {
    illcrawl::volumetric_tomography(cam).create_tomographic_cube(crawler, *q, 0, std::size_t(nz),
                                                                 out);
    path = dir + "/cube_" + qname + ".fits";
    illcrawl::io::fits<float>(path).save(out.data(), {nx, nx, nz});
  }
  double s = 0;
  for (float v : out) s += v;
  std::printf("crawl_demo: volume %s %dx%dx%d, samples %d, blocksize %zu, sum %.9g\n",
              qname.c_str(), nx, nx, nz, samples, blocksize, s);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "write-snapshot" && argc == 4) return write_snapshot(argv[2], std::stoul(argv[3]));
    if (mode == "write-fits" && argc == 6) {
      const long nx = std::stol(argv[3]), ny = std::stol(argv[4]);
      std::vector<float> v = read_raw<float>(argv[2], std::size_t(nx * ny));
      illcrawl::io::fits<float>(argv[5]).save(v.data(), {nx, ny});
      std::vector<long> ext;
      std::vector<float> back = illcrawl::io::fits<float>(argv[5]).load(ext);
      if (ext.size() != 2 || ext[0] != nx || ext[1] != ny || back != v) {
        std::fprintf(stderr, "crawl_demo: fits round trip differs\n");
        return 1;
      }
      return 0;
    }
    if (mode == "fits-to-raw" && argc == 4) {
      std::vector<long> ext;
      std::vector<float> v = illcrawl::io::fits<float>(argv[2]).load(ext);
      std::ofstream f(argv[3], std::ios::binary);
      f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(float)));
      std::printf("extents");
      for (long e : ext) std::printf(" %ld", e);
      std::printf("\n");
      return f ? 0 : 1;
    }
    if (mode == "volume" && argc == 10)
      return volume(argv[2], std::stoul(argv[3]), argv[4], std::stoi(argv[5]), std::stod(argv[6]),
                    std::stod(argv[7]), std::stoi(argv[8]), std::stoi(argv[9]));
    if (mode == "render" && argc == 8)
      return render(argv[2], std::stoul(argv[3]), argv[4], std::stoi(argv[5]), std::stod(argv[6]),
                    std::stod(argv[7]));
    std::fprintf(stderr, "usage: see the header of crawl_demo.cpp\n");
    return 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "crawl_demo: %s\n", e.what());
    return 1;
  }
}
