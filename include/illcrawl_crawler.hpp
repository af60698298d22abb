// illcrawl_crawler.hpp -- the data crawler over the MI355X path.
//
//   illcrawl::unit_converter               units.hpp:30-90
//   illcrawl::reconstruction_quantity::quantity and its projective-smoothing
//   kinds                                  quantity.hpp:41-66, 177-228, 323-376,
//                                          547-605, 805-1017; quantity.cpp:84-131
//   illcrawl::reconstructing_data_crawler  reconstructing_data_crawler.cpp:112-260
//
// The crawler streams row blocks of a snapshot group (io::dataset_streamer,
// the next block read while the current one renders), uploads them, and runs
// the reference's per-row body on the device (ilc_crawl_block: periodic
// correction, render-volume cutout, order-preserving compaction of the kept
// rows and their columns), then the quantity transformation, and hands the
// block to a reconstruction backend (setup_particles + run): the projective
// smoothing of dm::projective_smoothing or the volumetric smoothing of
// dm::grid.
// Blocks accumulate in the backend until retrieve_results, as in the
// reference.
#ifndef ILLCRAWL_CRAWLER_HPP
#define ILLCRAWL_CRAWLER_HPP

#include <string>
#include <vector>

#include "illcrawl_backend.hpp"
#include "illcrawl_io.hpp"

namespace illcrawl {

// Illustris units (comoving ckpc/h, 1e10 Msun/h) to proper kpc and Msun at
// scale factor a and Hubble parameter h (units.hpp:30-90).
class unit_converter {
 public:
  unit_converter(double a = 1.0, double h = 1.0) : _a(a), _h(h) {}
  double length_conversion_factor() const { return _a / _h; }
  double mass_conversion_factor() const { return 1.e10 / _h; }
  double volume_conversion_factor() const { return _a * _a * _a / (_h * _h * _h); }
  double density_conversion_factor() const {
    return mass_conversion_factor() / volume_conversion_factor();
  }
  double area_conversion_factor() const {
    return length_conversion_factor() * length_conversion_factor();
  }
  double potential_conversion_factor() const { return 1. / _a; }
  double get_scale_factor() const { return _a; }
  double get_hubble_parameter() const { return _h; }

 private:
  double _a, _h;
};

namespace reconstruction_quantity {

// A reconstructed quantity: the particle group and datasets it reads, their
// scaling factors (the unit conversion: input = float(factor * value),
// quantity.cpp:84-93), the transformation of the scaled inputs into the
// particle weight w, and the effective dA of the line-of-sight integration
// and dV of a volume integration.
class quantity {
 public:
  virtual ~quantity() = default;
  virtual unsigned get_particle_type_id() const = 0;
  virtual std::vector<std::string> get_required_datasets() const = 0;
  virtual std::vector<double> get_quantity_scaling_factors() const {
    return std::vector<double>(get_required_datasets().size(), 1.0);
  }
  // w_dev[i] from the scaled inputs in_dev[k][i], k over the datasets
  virtual void transform(const float* const* in_dev, std::size_t n, float* w_dev,
                         hipStream_t s) const = 0;
  virtual double effective_line_of_sight_integration_dA(double dA, double range) const = 0;
  virtual double effective_volume_integration_dV(double dV, double volume) const = 0;
  // the units of the pixel area and range passed to the dA above
  // (projective_smoothing::retrieve_results); identity unless set
  const unit_converter& get_unit_converter() const { return _units; }
  void set_unit_converter(const unit_converter& u) { _units = u; }

 private:
  unit_converter _units;
};

// The effective dA of a projection as dm::projective_smoothing applies it
// (projective_smoothing_reconstruction.cpp:49-84): the FP32 pixel area times
// the area conversion, the depth times the FP32 length conversion.
inline double projected_dA(const quantity& q, double pixel_area_f32, double depth) {
  const unit_converter& u = q.get_unit_converter();
  const double lc = static_cast<double>(static_cast<float>(u.length_conversion_factor()));
  return q.effective_line_of_sight_integration_dA(pixel_area_f32 * u.area_conversion_factor(),
                                                  lc * depth);
}

// the gas density as is (unprocessed_quantity, quantities.cl:701-711)
class density_based_quantity : public quantity {
 public:
  explicit density_based_quantity(double density_scale = 1.0) : _scale(density_scale) {}
  unsigned get_particle_type_id() const override { return 0; }
  std::vector<std::string> get_required_datasets() const override {
    return {io::illustris_data_loader::get_density_identifier()};
  }
  std::vector<double> get_quantity_scaling_factors() const override { return {_scale}; }
  void transform(const float* const* in_dev, std::size_t n, float* w_dev,
                 hipStream_t s) const override {
    check(ilc_quantity_scaled(in_dev[0], 1.0, n, w_dev, s), "quantity transformation");
  }

 private:
  double _scale;
};

// projected_density: the density integrated along the line of sight, per
// unit area (quantity.hpp:547-575)
class projected_density : public density_based_quantity {
 public:
  using density_based_quantity::density_based_quantity;
  double effective_line_of_sight_integration_dA(double, double) const override { return 1.0; }
  // the mean density over the volume
  double effective_volume_integration_dV(double dV, double volume) const override {
    return dV / volume;
  }
};

// mass: the same integral times the pixel area (quantity.hpp:577-604)
class mass : public density_based_quantity {
 public:
  using density_based_quantity::density_based_quantity;
  double effective_line_of_sight_integration_dA(double dA, double) const override { return dA; }
  double effective_volume_integration_dV(double dV, double) const override { return dV; }
};

// xray_flux: bremsstrahlung in the band [e_min, e_max] keV from density,
// internal energy and electron abundance (quantity.hpp:177-228, 323-376,
// quantities.cl:354-385), with the van Hoof Gaunt table the caller loaded
// (gny rows of ln u, gnx columns of ln gamma^2, row-major)
class xray_flux : public quantity {
 public:
  xray_flux(std::vector<float> gaunt_table, int gnx, int gny, double redshift,
            double luminosity_distance, double e_min, double e_max, unsigned num_samples,
            std::vector<double> scaling_factors = {1.0, 1.0, 1.0})
      : _tab(std::move(gaunt_table)),
        _gnx(gnx),
        _gny(gny),
        _z(redshift),
        _dl(luminosity_distance),
        _emin(e_min),
        _emax(e_max),
        _ns(num_samples),
        _scale(std::move(scaling_factors)) {
    if (_tab.size() != std::size_t(gnx) * std::size_t(gny))
      throw error(ILC_EINVAL, "xray_flux: Gaunt table size is not gnx * gny");
    if (_scale.size() != 3) throw error(ILC_EINVAL, "xray_flux: three scaling factors");
  }
  unsigned get_particle_type_id() const override { return 0; }
  std::vector<std::string> get_required_datasets() const override {
    return {io::illustris_data_loader::get_density_identifier(),
            io::illustris_data_loader::get_internal_energy_identifier(),
            io::illustris_data_loader::get_electron_abundance_identifier()};
  }
  std::vector<double> get_quantity_scaling_factors() const override { return _scale; }
  void transform(const float* const* in_dev, std::size_t n, float* w_dev,
                 hipStream_t s) const override {
    // the kernel arguments are cast to device_scalar (quantity.hpp:212-216, 358-364)
    check(ilc_quantity_xray_flux(in_dev[0], in_dev[1], in_dev[2], n, _tab.data(), _gnx, _gny,
                                 static_cast<float>(_z), static_cast<float>(_dl),
                                 static_cast<float>(_emin), static_cast<float>(_emax),
                                 static_cast<int32_t>(_ns), w_dev, nullptr, s),
          "quantity transformation");
  }
  double effective_line_of_sight_integration_dA(double dA, double) const override { return dA; }
  // quantity.hpp:193-197
  double effective_volume_integration_dV(double dV, double) const override { return dV; }

 private:
  std::vector<float> _tab;
  int _gnx, _gny;
  double _z, _dl, _emin, _emax;
  unsigned _ns;
  std::vector<double> _scale;
};

// Dark matter (PartType1): every particle carries the same mass, given in
// 1e10 Msun/h and converted to Msun per comoving volume unit (the smoothing
// sees ckpc/h coordinates), as constant_quantity (quantity.hpp:805-857,
// quantities.cl:713-726).
class dm_quantity : public quantity {
 public:
  dm_quantity(const unit_converter& converter, double dm_particle_mass)
      : _m(converter.mass_conversion_factor() * dm_particle_mass) {
    _m /= converter.volume_conversion_factor();
    set_unit_converter(converter);
  }
  unsigned get_particle_type_id() const override { return 1; }
  std::vector<std::string> get_required_datasets() const override { return {}; }
  std::vector<double> get_quantity_scaling_factors() const override { return {}; }
  void transform(const float* const*, std::size_t n, float* w_dev, hipStream_t s) const override {
    // the mass is a kernel argument of type device_scalar
    check(ilc_quantity_scaled(nullptr, _m, n, w_dev, s), "quantity transformation");
  }
  double get_dm_particle_mass() const { return _m; }

 private:
  double _m;
};

// the mean dark matter density along the line of sight (quantity.hpp:859-883)
class dm_density : public dm_quantity {
 public:
  using dm_quantity::dm_quantity;
  double effective_volume_integration_dV(double dV, double volume) const override {
    return dV / volume;
  }
  double effective_line_of_sight_integration_dA(double, double range) const override {
    return 1.0 / range;
  }
};

// the dark matter mass per pixel (quantity.hpp:885-912)
class dm_mass : public dm_quantity {
 public:
  using dm_quantity::dm_quantity;
  double effective_volume_integration_dV(double dV, double) const override { return dV; }
  double effective_line_of_sight_integration_dA(double dA, double) const override { return dA; }
};

// Stars (PartType4): the particle mass where GFM_StellarFormationTime > 0,
// 0 for wind cells; the masses scaled to Msun per comoving volume unit
// (quantity.hpp:918-1017, quantities.cl:728-750).
class stellar_density : public quantity {
 public:
  explicit stellar_density(const unit_converter& converter)
      : _mass_conversion(converter.mass_conversion_factor() /
                         converter.volume_conversion_factor()) {
    set_unit_converter(converter);
  }
  unsigned get_particle_type_id() const override { return 4; }
  std::vector<std::string> get_required_datasets() const override {
    return {"GFM_StellarFormationTime", io::illustris_data_loader::get_masses_identifier()};
  }
  std::vector<double> get_quantity_scaling_factors() const override {
    return {1.0, _mass_conversion};
  }
  void transform(const float* const* in_dev, std::size_t n, float* w_dev,
                 hipStream_t s) const override {
    check(ilc_quantity_stellar_mass(in_dev[0], in_dev[1], n, w_dev, s),
          "quantity transformation");
  }
  double get_mass_conversion_factor() const { return _mass_conversion; }
  double effective_volume_integration_dV(double dV, double volume) const override {
    return dV / volume;
  }
  double effective_line_of_sight_integration_dA(double, double range) const override {
    return 1.0 / range;
  }

 private:
  double _mass_conversion;
};

// the stellar mass per pixel (quantity.hpp:993-1017)
class stellar_mass : public stellar_density {
 public:
  using stellar_density::stellar_density;
  double effective_volume_integration_dV(double dV, double) const override { return dV; }
  double effective_line_of_sight_integration_dA(double dA, double) const override { return dA; }
};

}  // namespace reconstruction_quantity

// reconstructing_data_crawler: streams the quantity's particle group through
// a reconstruction backend in blocks of `blocksize` rows.
class reconstructing_data_crawler {
 public:
  // volume: the render volume, the periodic box and the additional border
  // region (reconstructing_data_crawler.cpp:188-193)
  reconstructing_data_crawler(const ilc_volume& volume, io::illustris_data_loader* data,
                              std::size_t blocksize, reconstruction_backend* backend)
      : _vol(volume), _data(data), _blocksize(blocksize), _backend(backend) {
    if (!_data || !_backend) throw error(ILC_EINVAL, "reconstructing_data_crawler: null argument");
    if (!_blocksize) throw error(ILC_EINVAL, "reconstructing_data_crawler: zero blocksize");
    _backend->init_backend(_blocksize);
  }

  // reconstructing_data_crawler::run (reconstructing_data_crawler.cpp:112-260):
  // evaluation points in the world frame; the image is the backend's
  // retrieve_results afterwards
  void run(const device_buffer& evaluation_points, std::size_t num_points,
           const reconstruction_quantity::quantity& q) {
    const hipStream_t s = _backend->stream();
    _backend->setup_evaluation_points(evaluation_points, num_points);
    _data->select_group(q.get_particle_type_id());
    std::vector<H5::DataSet> ds{_data->get_dataset(io::illustris_data_loader::get_coordinate_identifier())};
    const std::vector<dataset_ref> extra = _backend->get_required_additional_datasets();
    for (const dataset_ref& r : extra) ds.push_back(_data->get_dataset(r.name));
    const std::vector<std::string> qn = q.get_required_datasets();
    for (const std::string& name : qn) ds.push_back(_data->get_dataset(name));
    const std::vector<double> factors = q.get_quantity_scaling_factors();
    if (factors.size() != qn.size())
      throw error(ILC_EINVAL, "quantity: one scaling factor per dataset");
    io::dataset_streamer streamer(ds);
    if (streamer.get_num_columns(0) != 3) throw error(ILC_EINVAL, "Coordinates: three columns");
    for (std::size_t d = 1; d < ds.size(); ++d)
      if (streamer.get_num_columns(d) != 1)
        throw error(ILC_EINVAL, "crawler: extra and quantity datasets have one column");
    const std::size_t ncol = ds.size() - 1;
    if (ncol > 8) throw error(ILC_EINVAL, "crawler: at most 8 extra and quantity datasets");
    ensure(ncol);
    streamer.for_each_block(_blocksize, [&](std::size_t, const io::dataset_streamer::block& blk) {
      const std::size_t n = blk[0].size() / 3;
      if (!n) return;
      _xyz.upload(blk[0].data(), n * 3 * sizeof(float), s);
      std::vector<const float*> in(ncol);
      std::vector<float*> out(ncol);
      for (std::size_t k = 0; k < ncol; ++k) {
        _in[k].upload(blk[1 + k].data(), n * sizeof(float), s);
        in[k] = _in[k].as<float>();
        out[k] = _out[k].as<float>();
      }
      uint64_t kept = 0;
      check(ilc_crawl_block(&_vol, _xyz.as<float>(), n, in.data(), int32_t(ncol),
                            _particles.as<float>(), out.data(), nullptr, &kept, s),
            "crawl block");
      if (!kept) return;
      // quantity inputs scaled (quantity_transformation::queue_input_quantities)
      std::vector<const float*> qin(qn.size());
      for (std::size_t k = 0; k < qn.size(); ++k) {
        float* col = out[extra.size() + k];
        if (factors[k] != 1.0)
          check(ilc_quantity_scaled(col, factors[k], kept, col, s), "quantity scaling");
        qin[k] = col;
      }
      q.transform(qin.data(), kept, _w.as<float>(), s);
      check(ilc_particles_set_w(_particles.as<float>(), _w.as<float>(), kept, s), "set w");
      std::vector<device_buffer> extra_buffers;
      for (std::size_t k = 0; k < extra.size(); ++k) extra_buffers.push_back(_out[k]);
      _backend->setup_particles(_particles, kept, extra_buffers);
      _backend->run();
    });
    check_hip(hipStreamSynchronize(s), "crawler");
  }

  // the backend's results of the last run (reconstructing_data_crawler.hpp:127);
  // synchronises
  const device_buffer& get_reconstruction() { return _backend->retrieve_results(); }
  reconstruction_backend& get_backend() { return *_backend; }

 private:
  void ensure(std::size_t ncol) {
    if (_xyz.count() < 3 * _blocksize) {
      _xyz = device_buffer::create<float>(3 * _blocksize);
      _particles = device_buffer::create<device_vector4>(_blocksize);
      _w = device_buffer::create<float>(_blocksize);
      _in.clear();
      _out.clear();
    }
    while (_in.size() < ncol) {
      _in.push_back(device_buffer::create<float>(_blocksize));
      _out.push_back(device_buffer::create<float>(_blocksize));
    }
  }

  ilc_volume _vol;
  io::illustris_data_loader* _data;
  std::size_t _blocksize;
  reconstruction_backend* _backend;
  device_buffer _xyz, _particles, _w;
  std::vector<device_buffer> _in, _out;
};

}  // namespace illcrawl

#endif
