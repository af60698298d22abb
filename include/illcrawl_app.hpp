// illcrawl_app.hpp -- the application layer of the illcrawl command-line
// tools on the MI355X path: options, the gas distribution that centres the
// camera and the render volume, the quantity selection and the FITS header.
//
//   illcrawl::option_set                    the options of boost::program_options
//                                           the reference declares (illcrawl_app.cpp:45-83,
//                                           illcrawl_app.cpp:513-592, illcrawl_render.cpp:291-334)
//   illcrawl::coordinate_system::correct_periodicity
//                                           coordinate_system.hpp:84-103 (pivot form)
//   illcrawl::particle_distribution         particle_distribution.cpp:27-90
//   illcrawl::illcrawl_app                  illcrawl_app.cpp:36-300, illcrawl_app.hpp:45-160
//   illcrawl::quantity_command_line_parser  illcrawl_app.cpp:510-890
//
// Link as illcrawl_io.hpp says (HDF5 C++, cfitsio) plus libilcpsg.
#ifndef ILLCRAWL_APP_HPP
#define ILLCRAWL_APP_HPP

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdlib>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "illcrawl_crawler.hpp"
#include "illcrawl_io.hpp"

namespace illcrawl {

// Named options: "--name value", "--name=value" or "-c value" for a short
// letter.  Every option has a default (as text); required options have none
// and must be given.  Unknown options are an error.
class option_set {
 public:
  void add(const std::string& name, char short_name, const std::string& default_value,
           const std::string& help) {
    _opts.push_back(opt{name, short_name, default_value, help, false, false});
  }
  void add_required(const std::string& name, char short_name, const std::string& help) {
    _opts.push_back(opt{name, short_name, "", help, true, false});
  }

  void parse(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      opt* o = nullptr;
      std::string value;
      bool have_value = false;
      if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
        std::string name = a.substr(2);
        const std::size_t eq = name.find('=');
        if (eq != std::string::npos) {
          value = name.substr(eq + 1);
          name = name.substr(0, eq);
          have_value = true;
        }
        o = find(name);
        if (!o) throw std::invalid_argument("unrecognised option '--" + name + "'");
      } else if (a.size() == 2 && a[0] == '-') {
        for (opt& c : _opts)
          if (c.short_name == a[1]) o = &c;
        if (!o) throw std::invalid_argument("unrecognised option '" + a + "'");
      } else {
        throw std::invalid_argument("unexpected argument '" + a + "'");
      }
      if (!have_value) {
        if (i + 1 >= argc) throw std::invalid_argument("option '--" + o->name + "' needs a value");
        value = argv[++i];
      }
      o->value = value;
      o->given = true;
    }
    for (const opt& o : _opts)
      if (o.required && !o.given)
        throw std::invalid_argument("the option '--" + o.name + "' is required but missing");
  }

  const std::string& get(const std::string& name) const { return at(name).value; }
  double get_double(const std::string& name) const {
    return convert<double>(name, [](const char* s, char** e) { return std::strtod(s, e); });
  }
  unsigned long long get_ull(const std::string& name) const {
    const std::string& v = get(name);
    if (!v.empty() && v[0] == '-') throw std::invalid_argument("option '--" + name + "': negative");
    return convert<unsigned long long>(
        name, [](const char* s, char** e) { return std::strtoull(s, e, 10); });
  }
  bool given(const std::string& name) const { return at(name).given; }

  std::string usage() const {
    std::ostringstream o;
    for (const opt& c : _opts) {
      o << "  --" << c.name;
      if (c.short_name) o << " [-" << c.short_name << "]";
      if (c.required)
        o << " (required)";
      else
        o << " (=" << c.value << ")";
      o << "\n      " << c.help << "\n";
    }
    return o.str();
  }

 private:
  struct opt {
    std::string name;
    char short_name;
    std::string value;
    std::string help;
    bool required;
    bool given;
  };
  opt* find(const std::string& name) {
    for (opt& o : _opts)
      if (o.name == name) return &o;
    return nullptr;
  }
  const opt& at(const std::string& name) const {
    for (const opt& o : _opts)
      if (o.name == name) return o;
    throw std::logic_error("option not registered: " + name);
  }
  template <class T, class F>
  T convert(const std::string& name, F f) const {
    const std::string& v = get(name);
    char* end = nullptr;
    const T r = static_cast<T>(f(v.c_str(), &end));
    if (v.empty() || end == v.c_str() || *end != '\0')
      throw std::invalid_argument("option '--" + name + "': cannot convert '" + v + "'");
    return r;
  }
  std::vector<opt> _opts;
};

namespace coordinate_system {

// Moves a point by whole periods, axis by axis, where that brings it closer
// to the pivot (coordinate_system.hpp:84-103): the candidate carries the
// axes already moved.
inline void correct_periodicity(const double periodic[3], const double pivot[3],
                                double coordinates[3]) {
  auto d2 = [&](const double* c) {
    double r[3];
    for (int k = 0; k < 3; ++k) r[k] = pivot[k] - c[k];
    return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  };
  for (int i = 0; i < 3; ++i) {
    double cand[3] = {coordinates[0], coordinates[1], coordinates[2]};
    cand[i] = coordinates[i] + periodic[i];
    if (d2(cand) < d2(coordinates)) {
      coordinates[i] += periodic[i];
    } else {
      cand[i] = coordinates[i] - periodic[i];
      if (d2(cand) < d2(coordinates)) coordinates[i] -= periodic[i];
    }
  }
}

}  // namespace coordinate_system

// Extent and centre of a particle group's coordinates, each point first
// moved to the periodic image nearest the running mean of the points before
// it (particle_distribution.cpp:27-90; rows read as double in blocks of
// 100000).  The maxima start at the smallest positive double, as the
// reference's do.
class particle_distribution {
 public:
  particle_distribution(const H5::DataSet& coordinates, const double periodic[3]) {
    io::dataset_streamer streamer({coordinates});
    if (streamer.get_num_columns(0) != 3) throw std::runtime_error("Coordinates: 3 columns");
    double mn[3], mx[3], est[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 3; ++j) {
      mn[j] = std::numeric_limits<double>::max();
      mx[j] = std::numeric_limits<double>::min();
    }
    std::size_t count = 0;
    const std::size_t rows = streamer.get_num_rows(), block = 100000;
    std::vector<double> buf;
    for (std::size_t first = 0; first < rows; first += block) {
      const std::size_t n = std::min(block, rows - first);
      streamer.read_f64(0, first, n, buf);
      for (std::size_t i = 0; i < n; ++i) {
        double c[3] = {buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]};
        // the first point is taken as it is: the running mean is still 0
        if (first != 0 || i != 0) coordinate_system::correct_periodicity(periodic, est, c);
        const double f = 1.0 / static_cast<double>(count + 1);
        for (int j = 0; j < 3; ++j) est[j] = f * (static_cast<double>(count) * est[j] + c[j]);
        for (int j = 0; j < 3; ++j) {
          if (c[j] > mx[j]) mx[j] = c[j];
          if (c[j] < mn[j]) mn[j] = c[j];
        }
        ++count;
      }
    }
    for (int j = 0; j < 3; ++j) {
      _center[j] = 0.5 * (mn[j] + mx[j]);
      _size[j] = mx[j] - mn[j];
      _mean[j] = est[j];
    }
  }
  const double* get_extent_center() const { return _center; }
  const double* get_distribution_size() const { return _size; }
  const double* get_mean_particle_position() const { return _mean; }

 private:
  double _center[3], _size[3], _mean[3];
};

// The settings every illcrawl tool shares (illcrawl_app.cpp:36-300): the
// snapshot, cosmology, reconstructor names and data crawling block size;
// after parsing, the gas (PartType0) distribution that centres the camera,
// the render volume and the integration depth.
class illcrawl_app {
 public:
  explicit illcrawl_app(const double periodic_wraparound[3] = nullptr) {
    for (int i = 0; i < 3; ++i)
      _periodic[i] = periodic_wraparound ? periodic_wraparound[i] : 75000.0;
    _options.add_required("input", 'i', "the hdf5 file containing the data");
    _options.add("hubble_constant", 'h', "0.704", "Hubble constant in units of 100 km/s/Mpc");
    _options.add("redshift", 'z', "0.2", "Redshift");
    _options.add("luminosity_distance", 'd', "978500",
                 "Luminosity distance of the dataset in kpc");
    _options.add("reconstructor.voronoi", 0, "projective_smoothing",
                 "The reconstructor of gas (Voronoi cell) quantities.  This build has no "
                 "Voronoi reconstructor (the reference's nn8 and tree): gas quantities are "
                 "projected by smoothing their cells with SubfindHsml");
    _options.add("reconstructor.voronoi.tree.opening_angle", 0, "0.4",
                 "The opening angle of the tree reconstructor (recorded, not used)");
    _options.add("reconstructor.smoothing", 0, "grid",
                 "The reconstructor of smoothed quantities for volumes: grid");
    _options.add("reconstructor.smoothing.projective", 0, "grid",
                 "The projective reconstructor of smoothed quantities: grid");
    _options.add("data_crawler.blocksize", 0, "40000000",
                 "The number of particles (or cells) that are read and processed in one block.");
  }
  illcrawl_app(const illcrawl_app&) = delete;
  illcrawl_app& operator=(const illcrawl_app&) = delete;

  option_set& options() { return _options; }

  // parses argv, opens the snapshot and scans the gas distribution
  void parse_command_line(int argc, char** argv) {
    _options.parse(argc, argv);
    _hdf5_filename = _options.get("input");
    _h = _options.get_double("hubble_constant");
    _z = _options.get_double("redshift");
    _luminosity_distance = _options.get_double("luminosity_distance");
    _blocksize = static_cast<std::size_t>(_options.get_ull("data_crawler.blocksize"));
    if (_blocksize == 0) throw std::invalid_argument("data_crawler.blocksize must be > 0");
    if (_options.get("reconstructor.voronoi") != "projective_smoothing")
      throw std::invalid_argument("Unknown voronoi reconstructor: " +
                                  _options.get("reconstructor.voronoi") +
                                  " (this build projects gas by smoothing: projective_smoothing)");
    if (_options.get("reconstructor.smoothing") != "grid")
      throw std::invalid_argument("Unknown dark matter reconstructor: " +
                                  _options.get("reconstructor.smoothing"));
    if (_options.get("reconstructor.smoothing.projective") != "grid")
      throw std::invalid_argument("Unknown projective smoothing reconstruction backend: " +
                                  _options.get("reconstructor.smoothing.projective"));
    _a = 1.0 / (1. + _z);
    _units = unit_converter(_a, _h);
    _data = std::make_unique<io::illustris_data_loader>(_hdf5_filename, 0u);
    _distribution = std::make_unique<particle_distribution>(
        _data->get_dataset(io::illustris_data_loader::get_coordinate_identifier()), _periodic);
    const double* s = _distribution->get_distribution_size();
    _radius = 0.5 * std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  }

  const double* get_gas_distribution_center() const { return _distribution->get_extent_center(); }
  const double* get_gas_distribution_size() const {
    return _distribution->get_distribution_size();
  }
  double get_distribution_radius() const { return _radius; }

  // illcrawl_app.cpp:179-189: from the distribution's centre, one radius back
  // along look_at; screen width = the distribution's x extent
  camera create_distribution_centered_camera(std::size_t nx, std::size_t ny,
                                             const double look_at[3] = nullptr) const {
    const double la[3] = {look_at ? look_at[0] : 0.0, look_at ? look_at[1] : 0.0,
                          look_at ? look_at[2] : 1.0};
    const double* c = get_gas_distribution_center();
    double pos[3];
    for (int i = 0; i < 3; ++i) pos[i] = c[i] - _radius * la[i];
    return camera(pos, la, 0.0, get_gas_distribution_size()[0], nx, ny);
  }

  // illcrawl_app.cpp:191-198 and the crawler's 100 ckpc/h border for the
  // periodic correction (reconstructing_data_crawler.hpp:151)
  ilc_volume get_gas_distribution_volume_cutout() const {
    ilc_volume v{};
    for (int i = 0; i < 3; ++i) {
      v.center[i] = get_gas_distribution_center()[i];
      v.extent[i] = get_gas_distribution_size()[i];
      v.periodic_wraparound[i] = _periodic[i];
    }
    v.border = 100.0;
    return v;
  }

  double get_recommended_integration_depth() const { return 2.0 * _radius; }
  io::illustris_data_loader& get_data_loader() const { return *_data; }
  const unit_converter& get_unit_converter() const { return _units; }
  double get_redshift() const { return _z; }
  double get_luminosity_distance() const { return _luminosity_distance; }
  std::size_t get_data_crawling_blocksize() const { return _blocksize; }

  // illcrawl_app.cpp:239-300
  void save_settings_to_fits(io::fits_header& header) const {
    header.add_entry("renderer", "Illcrawl suite");
    header.add_entry("renderer.voronoi_reconstructor", _options.get("reconstructor.voronoi"));
    header.add_entry("renderer.dm_reconstructor", _options.get("reconstructor.smoothing"));
    header.add_entry("renderer.dm_projective_reconstructor",
                     _options.get("reconstructor.smoothing.projective"));
    header.add_entry("renderer.partitioner", "uniform", "The parallel work partitioning strategy");
    header.add_entry("data_crawler.blocksize", static_cast<unsigned long long>(_blocksize),
                     "The number of particles/cells processed in one block");
    header.add_entry("environment.num_mpi_processes", 1, "Number of parallel MPI processes");
    std::string name = "unknown", driver = "unknown", runtime = "HIP";
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    int v = 0;
    if (hipDriverGetVersion(&v) == hipSuccess) driver = std::to_string(v);
    if (hipRuntimeGetVersion(&v) == hipSuccess) runtime = "HIP " + std::to_string(v);
    header.add_entry("environment.device_name", name, "Name of Compute device of master process");
    header.add_entry("environment.device_vendor", "Advanced Micro Devices, Inc.",
                     "Name of compute device vendor on the master process");
    header.add_entry("environment.device_opencl_version", runtime,
                     "The runtime version of the compute device of the master process");
    header.add_entry("environment.device_driver_version", driver,
                     "Version of the driver of the compute device of the master process");
    header.add_entry("data_source", _hdf5_filename);
    header.add_entry("a", _a, "Scale factor");
    header.add_entry("z", _z, "Redshift");
    header.add_entry("h", _h, "Hubble constant [100 km/s/Mpc]");
    header.add_entry("dL", _luminosity_distance, "Luminosity distance [kpc]");
    save_length_scalar(header, "particle_distribution", "radius", _radius,
                       "Radius within which all particles lie");
    save_comoving_vector3(header, "particle_distribution", "center", get_gas_distribution_center(),
                          "Particle distribution's box center");
    save_comoving_vector3(header, "particle_distribution", "size", get_gas_distribution_size(),
                          "The size of the distribution");
  }

  // illcrawl_app.cpp:302-332
  void save_camera_settings_to_fits_header(io::fits_header& header, const std::string& name,
                                           const camera& cam) const {
    double b0[3], b1[3], mn[3];
    cam.get_screen_basis(b0, b1, mn);
    save_comoving_vector3(header, name, "position", cam.get_position(),
                          "Center position of the camera plane");
    save_comoving_vector3(header, name, "screen_min", mn,
                          "Lower left corner position of the screen");
    save_vector3(header, name + ".look_at", cam.get_look_at(), "Camera look-at vector");
    save_vector3(header, name + ".screen_basis0", b0,
                 "The basis vector of the x-direction of the camera screen");
    save_vector3(header, name + ".screen_basis1", b1,
                 "The basis vector of the y-direction of the camera screen");
    save_length_scalar(header, name, "pixel_size", cam.get_pixel_size(),
                       "The physical width of one pixel");
  }

 private:
  void save_length_scalar(io::fits_header& header, const std::string& parent,
                          const std::string& key, double value, const std::string& comment) const {
    header.add_entry(parent + "." + key, value, comment + " (Comoving length [ckpc/h])");
    header.add_entry(parent + ".proper_" + key, _units.length_conversion_factor() * value,
                     comment + " (Proper length [kpc])");
  }
  void save_vector3(io::fits_header& header, const std::string& key, const double* v,
                    const std::string& comment) const {
    const char* axes[3] = {"x", "y", "z"};
    for (int i = 0; i < 3; ++i) header.add_entry(key + "." + axes[i], v[i], comment);
  }
  void save_comoving_vector3(io::fits_header& header, const std::string& parent,
                             const std::string& key, const double* v,
                             const std::string& comment) const {
    save_vector3(header, parent + "." + key, v, comment + " (Comoving length [ckpc/h])");
    double p[3];
    for (int i = 0; i < 3; ++i) p[i] = _units.length_conversion_factor() * v[i];
    save_vector3(header, parent + ".proper_" + key, p, comment + " (Proper length [kpc])");
  }

  option_set _options;
  double _periodic[3];
  std::string _hdf5_filename;
  double _a = 1 / 1.2, _z = 0.2, _h = 0.704, _luminosity_distance = 978500.0;
  double _radius = 0.0;
  std::size_t _blocksize = 40000000;
  unit_converter _units{1 / 1.2, 0.704};
  std::unique_ptr<io::illustris_data_loader> _data;
  std::unique_ptr<particle_distribution> _distribution;
};

// The --quantity option and its parameters (illcrawl_app.cpp:510-890).  The
// quantities this build reconstructs by projective smoothing are made; the
// others of the reference (Chandra count rates, photon and spectral fluxes,
// temperatures, metallicities, potential) are refused by name.
class quantity_command_line_parser {
 public:
  void register_options(option_set& o) const {
    o.add("quantity", 'q', "chandra_count_rate",
          "The quantity which shall be used for reconstruction: xray_flux, projected_density, "
          "mass, dm_mean_density, dm_mass, stellar_mean_density, stellar_mass");
    o.add("quantity.xray_spectral_flux.energy", 0, "1", "xray_spectral_flux energy [keV]");
    o.add("quantity.xray_spectral_flux.energy_bin_width", 0, "0.1",
          "xray_spectral_flux energy bin width [keV]");
    o.add("quantity.chandra_spectral_count_rate.energy", 0, "1",
          "chandra_spectral_count_rate energy [keV]");
    o.add("quantity.chandra_spectral_count_rate.energy_bin_width", 0, "0.1",
          "chandra_spectral_count_rate energy bin width [keV]");
    o.add("quantity.xray_flux.min_energy", 0, "0.1", "xray_flux energy range start [keV]");
    o.add("quantity.xray_flux.max_energy", 0, "10", "xray_flux energy range end [keV]");
    o.add("quantity.xray_flux.num_samples", 0, "100", "xray_flux energy integration samples");
    o.add("quantity.xray_flux.gaunt_table", 0, "gaunt_factor.dat",
          "the van Hoof free-free Gaunt factor table the reference reads as gaunt_factor.dat");
    o.add("quantity.xray_photon_flux.min_energy", 0, "0.1", "xray_photon_flux range start [keV]");
    o.add("quantity.xray_photon_flux.max_energy", 0, "10", "xray_photon_flux range end [keV]");
    o.add("quantity.xray_photon_flux.num_samples", 0, "100", "xray_photon_flux samples");
    o.add("quantity.luminosity_weighted_temperature.min_energy", 0, "0.1",
          "luminosity_weighted_temperature range start [keV]");
    o.add("quantity.luminosity_weighted_temperature.max_energy", 0, "10",
          "luminosity_weighted_temperature range end [keV]");
    o.add("quantity.luminosity_weighted_metallicity.min_energy", 0, "0.1",
          "luminosity_weighted_metallicity range start [keV]");
    o.add("quantity.luminosity_weighted_metallicity.max_energy", 0, "10",
          "luminosity_weighted_metallicity range end [keV]");
    o.add("quantity.dm.particle_mass", 0, "0.000440896524361",
          "The mass of a dark matter particle [10^10 M_sun/h]");
  }

  std::unique_ptr<reconstruction_quantity::quantity> create_quantity(
      const illcrawl_app& app, const option_set& o) const {
    namespace rq = reconstruction_quantity;
    const std::string q = o.get("quantity");
    const unit_converter& u = app.get_unit_converter();
    if (q == "xray_flux") {
      const double e0 = o.get_double("quantity.xray_flux.min_energy");
      const double e1 = o.get_double("quantity.xray_flux.max_energy");
      const unsigned long long ns = o.get_ull("quantity.xray_flux.num_samples");
      if (ns == 0) throw std::invalid_argument("Number of integration samples must be > 0");
      valid_range(e0, e1);
      std::size_t gnx = 0, gny = 0;
      std::vector<float> tab =
          io::load_data_table(o.get("quantity.xray_flux.gaunt_table"), gnx, gny);
      // density, internal energy and electron abundance in proper units
      // (density_temperature_electron_abundance_based_quantity, quantity.hpp:137-175)
      auto x = std::make_unique<rq::xray_flux>(
          std::move(tab), int(gnx), int(gny), app.get_redshift(), app.get_luminosity_distance(),
          e0, e1, unsigned(ns),
          std::vector<double>{u.density_conversion_factor(), 1.0, 1.0});
      x->set_unit_converter(u);
      return x;
    }
    if (q == "projected_density" || q == "mass") {
      // the density in proper units (density_based_quantity, quantity.hpp:123-146)
      std::unique_ptr<rq::quantity> d;
      if (q == "mass")
        d = std::make_unique<rq::mass>(u.density_conversion_factor());
      else
        d = std::make_unique<rq::projected_density>(u.density_conversion_factor());
      d->set_unit_converter(u);
      return d;
    }
    if (q == "dm_mean_density" || q == "dm_mass") {
      const double m = o.get_double("quantity.dm.particle_mass");
      if (m <= 0) throw std::invalid_argument("Dark matter particle mass must be > 0");
      if (q == "dm_mass") return std::make_unique<rq::dm_mass>(u, m);
      return std::make_unique<rq::dm_density>(u, m);
    }
    if (q == "stellar_mean_density") return std::make_unique<rq::stellar_density>(u);
    if (q == "stellar_mass") return std::make_unique<rq::stellar_mass>(u);
    static const char* absent[] = {"chandra_count_rate", "chandra_spectral_count_rate",
                                   "xray_photon_flux", "xray_spectral_flux", "mean_temperature",
                                   "luminosity_weighted_temperature", "mean_density",
                                   "potential", "metallicity", "density_weighted_metallicity",
                                   "luminosity_weighted_metallicity"};
    for (const char* a : absent)
      if (q == a)
        throw std::invalid_argument("quantity " + q +
                                    " is not reconstructed by this build (projective "
                                    "smoothing quantities only)");
    throw std::invalid_argument("Unknwon quantity: " + q);
  }

  // illcrawl_app.cpp:837-880
  void save_configuration_to_fits_header(io::fits_header& h, const option_set& o) const {
    h.add_entry("quantity", o.get("quantity"), "The rendered quantity");
    h.add_entry("quantity.xray_flux.min_energy", o.get_double("quantity.xray_flux.min_energy"),
                "xray_flux quantity: Minimum energy for the energy integration [keV]");
    h.add_entry("quantity.xray_flux.max_energy", o.get_double("quantity.xray_flux.max_energy"),
                "xray_flux quantity: Maximum energy for the energy integration [keV]");
    h.add_entry("quantity.xray_flux.num_samples",
                unsigned(o.get_ull("quantity.xray_flux.num_samples")),
                "xray_flux quantity: Number of samples for the energy integration");
    h.add_entry("quantity.xray_photon_flux.min_energy",
                o.get_double("quantity.xray_photon_flux.min_energy"),
                "xray_photon_flux quantity: Minimum energy for the energy integration [keV]");
    h.add_entry("quantity.xray_photon_flux.max_energy",
                o.get_double("quantity.xray_photon_flux.max_energy"),
                "xray_photon_flux quantity: Maximum energy for the energy integration [keV]");
    h.add_entry("quantity.xray_photon_flux.num_samples",
                unsigned(o.get_ull("quantity.xray_photon_flux.num_samples")),
                "xray_photon_flux quantity: Number of samples for the energy integration");
    h.add_entry("quantity.xray_spectral_flux.energy",
                o.get_double("quantity.xray_spectral_flux.energy"),
                "xray_spectral_flux quantity: Central energy of the evaluated energy channel [keV]");
    h.add_entry("quantity.xray_spectral_flux.energy_bin_width",
                o.get_double("quantity.xray_spectral_flux.energy_bin_width"),
                "xray_spectral_flux quantity: Width of the evaluated energy channel [keV]");
    h.add_entry("quantity.chandra_spectral_count_rate.energy",
                o.get_double("quantity.chandra_spectral_count_rate.energy"),
                "chandra_spectral_count_rate quantity: Central energy of the evaluated energy "
                "channel [keV]");
    // the reference writes the energy under the bin width's key as well
    h.add_entry("quantity.chandra_spectral_count_rate.energy_bin_width",
                o.get_double("quantity.chandra_spectral_count_rate.energy"),
                "chandra_spectral_count_rate quantity: Width of the evaluated energy channel "
                "[keV]");
    h.add_entry("quantity.luminosity_weighted_temp.min_energy",
                o.get_double("quantity.luminosity_weighted_temperature.min_energy"),
                "luminosity_weighted_temperature quantity: Minimum energy for the energy "
                "integration [keV]");
    h.add_entry("quantity.luminosity_weighted_temp.max_energy",
                o.get_double("quantity.luminosity_weighted_temperature.max_energy"),
                "luminosity_weighted_temperature quantity: Maximum energy for the energy "
                "integration [keV]");
    h.add_entry("quantity.dm.particle_mass", o.get_double("quantity.dm.particle_mass"),
                "The dark matter particle mass [10^10 Msun/h]");
  }

 private:
  static void valid_range(double mn, double mx) {
    if (mn < 0) throw std::invalid_argument("Start of integration range must be > 0");
    if (mx < 0) throw std::invalid_argument("End of integration range must be > 0");
    if (mx <= mn)
      throw std::invalid_argument(
          "End of integration range must be > than start of integration range");
  }
};

}  // namespace illcrawl

#endif
