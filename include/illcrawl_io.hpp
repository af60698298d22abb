// illcrawl_io.hpp -- snapshot input and image output of the data crawler.
//
//   illcrawl::io::illustris_data_loader   hdf5_io.hpp:39-107 (HDF5 C++ API)
//   illcrawl::io::dataset_streamer        the row-block reader that
//                                          io::async_dataset_streamer and
//                                          async_for_each_block provide to
//                                          reconstructing_data_crawler::run
//                                          (reconstructing_data_crawler.cpp:146-178)
//   illcrawl::io::fits<T>                 fits.hpp:210-300 (cfitsio)
//   illcrawl::io::fits_header             fits.hpp:142-201
//   illcrawl::io::load_data_table         util::data_table_loader
//                                          (data_table_loader.hpp:35-97)
//
// Link with -lhdf5_cpp -lhdf5 -lcfitsio (the image has them under /opt/conda).
#ifndef ILLCRAWL_IO_HPP
#define ILLCRAWL_IO_HPP

#include <H5Cpp.h>
#include <fitsio.h>

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <fstream>
#include <future>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace illcrawl {
namespace io {

// One particle group (PartType<id>) of an Illustris/TNG snapshot file.
class illustris_data_loader {
 public:
  illustris_data_loader(const std::string& hdf5_filename, unsigned particle_type_id)
      : _filename(hdf5_filename), _file(hdf5_filename, H5F_ACC_RDONLY) {
    select_group(particle_type_id);
  }
  illustris_data_loader(const std::string& hdf5_filename, const std::string& group)
      : _filename(hdf5_filename), _file(hdf5_filename, H5F_ACC_RDONLY) {
    select_group(group);
  }
  illustris_data_loader(const illustris_data_loader&) = delete;
  illustris_data_loader& operator=(const illustris_data_loader&) = delete;
  ~illustris_data_loader() {
    try {
      _file.close();
    } catch (...) {
    }
  }

  H5::DataSet get_dataset(const std::string& name) const { return _group.openDataSet(name); }

  // dataset names of the snapshot format (hdf5_io.hpp:70-77)
  static std::string get_coordinate_identifier() { return "Coordinates"; }
  static std::string get_density_identifier() { return "Density"; }
  static std::string get_internal_energy_identifier() { return "InternalEnergy"; }
  static std::string get_smoothing_length_identifier() { return "SmoothingLength"; }
  static std::string get_volume_identifier() { return "Volume"; }
  static std::string get_masses_identifier() { return "Masses"; }
  static std::string get_electron_abundance_identifier() { return "ElectronAbundance"; }
  static std::string get_dm_smoothing_length_identifier() { return "SubfindHsml"; }

  void select_group(unsigned particle_type_id) {
    select_group("PartType" + std::to_string(particle_type_id));
  }
  void select_group(const std::string& group_name) {
    try {
      _group.close();
    } catch (...) {
    }
    _group = _file.openGroup(group_name);
    _current_group_name = group_name;
  }
  std::string get_current_group_name() const { return _current_group_name; }

 private:
  std::string _filename;
  H5::H5File _file;
  H5::Group _group;
  std::string _current_group_name;
};

// Row blocks of datasets with the same number of rows (Coordinates: 3
// columns, the others 1).  A block holds, per dataset, its rows converted to
// double and then to float, row-major: the reference reads every value as
// double and casts it to device_scalar (reconstructing_data_crawler.cpp:180-222).
// The next block is read on a background thread while the caller works on
// the current one, as the reference's async worker does (async_worker.hpp:36-80).
class dataset_streamer {
 public:
  using block = std::vector<std::vector<float>>;

  explicit dataset_streamer(std::vector<H5::DataSet> datasets) : _ds(std::move(datasets)) {
    if (_ds.empty()) throw std::runtime_error("dataset_streamer: no datasets");
    for (std::size_t d = 0; d < _ds.size(); ++d) {
      H5::DataSpace sp = _ds[d].getSpace();
      const int rank = sp.getSimpleExtentNdims();
      if (rank < 1 || rank > 2) throw std::runtime_error("dataset_streamer: rank not 1 or 2");
      hsize_t dims[2] = {0, 1};
      sp.getSimpleExtentDims(dims);
      if (d == 0) _rows = dims[0];
      if (dims[0] != _rows) throw std::runtime_error("dataset_streamer: row counts differ");
      _cols.push_back(rank == 2 ? dims[1] : 1);
    }
  }

  std::size_t get_num_rows() const { return _rows; }
  std::size_t get_num_columns(std::size_t dataset) const { return _cols.at(dataset); }
  std::size_t get_num_datasets() const { return _ds.size(); }

  // rows [first, first + count) of every dataset
  block read(std::size_t first, std::size_t count) const {
    block out(_ds.size());
    std::vector<double> tmp;
    for (std::size_t d = 0; d < _ds.size(); ++d) {
      read_f64(d, first, count, tmp);
      out[d].resize(tmp.size());
      for (std::size_t k = 0; k < tmp.size(); ++k) out[d][k] = static_cast<float>(tmp[k]);
    }
    return out;
  }

  // rows [first, first + count) of one dataset as double, row-major
  void read_f64(std::size_t d, std::size_t first, std::size_t count,
                std::vector<double>& out) const {
    if (first + count > _rows) throw std::runtime_error("dataset_streamer: rows out of range");
    const hsize_t cols = _cols.at(d);
    out.resize(count * cols);
    if (!count) return;
    H5::DataSpace fs = _ds[d].getSpace();
    const int rank = fs.getSimpleExtentNdims();
    hsize_t off[2] = {first, 0}, cnt[2] = {count, cols};
    fs.selectHyperslab(H5S_SELECT_SET, cnt, off);
    H5::DataSpace ms(rank, cnt);
    _ds[d].read(out.data(), H5::PredType::NATIVE_DOUBLE, ms, fs);
  }

  // f(first_row, block) for the blocks of blocksize rows, in order; block
  // k + 1 is read while f runs on block k
  template <class F>
  void for_each_block(std::size_t blocksize, F&& f) const {
    if (!blocksize) throw std::runtime_error("dataset_streamer: zero blocksize");
    if (!_rows) return;
    auto load = [this, blocksize](std::size_t first) {
      return read(first, std::min<std::size_t>(blocksize, std::size_t(_rows) - first));
    };
    std::future<block> next = std::async(std::launch::async, load, std::size_t(0));
    for (std::size_t first = 0; first < _rows; first += blocksize) {
      block cur = next.get();
      if (first + blocksize < _rows) next = std::async(std::launch::async, load, first + blocksize);
      f(first, cur);
    }
  }

 private:
  std::vector<H5::DataSet> _ds;
  std::vector<hsize_t> _cols;
  hsize_t _rows = 0;
};

template <class T>
struct fits_datatype;
template <>
struct fits_datatype<float> {
  static int image_type() { return FLOAT_IMG; }
  static int datatype() { return TFLOAT; }
};
template <>
struct fits_datatype<double> {
  static int image_type() { return DOUBLE_IMG; }
  static int datatype() { return TDOUBLE; }
};

// FITS images and cubes through cfitsio, as the reference writes them: the
// first extent is NAXIS1 (the fastest index), so an image of nx * ny pixels
// stored as ix + nx * iy has extents {nx, ny}.
template <class T>
class fits {
 public:
  explicit fits(const std::string& filename) : _filename(filename) {}

  void save(const T* data, const std::vector<long>& extents) const {
    fitsfile* file = nullptr;
    int status = 0;
    const std::string name = "!" + _filename;  // overwrite
    if (fits_create_file(&file, name.c_str(), &status))
      throw std::runtime_error("Could not create fits file: " + _filename);
    std::vector<long> naxes(extents);
    long n = 1;
    for (long e : naxes) n *= e;
    std::vector<long> fpixel(naxes.size(), 1);
    if (fits_create_img(file, fits_datatype<T>::image_type(), int(naxes.size()), naxes.data(),
                        &status)) {
      fits_close_file(file, &status);
      throw std::runtime_error("Could not create fits image: " + _filename);
    }
    fits_write_pix(file, fits_datatype<T>::datatype(), fpixel.data(), n, const_cast<T*>(data),
                   &status);
    int close_status = 0;
    fits_close_file(file, &close_status);
    if (status || close_status) throw std::runtime_error("Could not write fits image: " + _filename);
  }

  // the image's data and extents
  std::vector<T> load(std::vector<long>& extents) const {
    fitsfile* file = nullptr;
    int status = 0;
    if (fits_open_file(&file, _filename.c_str(), READONLY, &status))
      throw std::runtime_error("Could not open fits file: " + _filename);
    int dim = 0, bitpix = 0;
    fits_get_img_dim(file, &dim, &status);
    extents.assign(std::size_t(dim > 0 ? dim : 0), 0);
    if (dim > 0) fits_get_img_param(file, dim, &bitpix, &dim, extents.data(), &status);
    long n = 1;
    for (long e : extents) n *= e;
    std::vector<T> out(status ? 0 : std::size_t(n));
    std::vector<long> fpixel(extents.size(), 1);
    if (!status && n)
      fits_read_pix(file, fits_datatype<T>::datatype(), fpixel.data(), n, nullptr, out.data(),
                    nullptr, &status);
    int close_status = 0;
    fits_close_file(file, &close_status);
    if (status) throw std::runtime_error("Could not read fits image: " + _filename);
    return out;
  }

 private:
  std::string _filename;
};

// Keys of a FITS file's primary header, written as the reference writes
// them: a COMMENT card with the comment, then the key (cfitsio uses the
// HIERARCH convention for the long dotted names).
class fits_header {
 public:
  explicit fits_header(const std::string& filename) : _filename(filename) {
    int status = 0;
    if (fits_open_file(&_file, _filename.c_str(), READWRITE, &status))
      throw std::runtime_error("fits_header: Could not open fits file " + filename);
  }
  fits_header(const fits_header&) = delete;
  fits_header& operator=(const fits_header&) = delete;
  ~fits_header() {
    int status = 0;
    if (_file) fits_close_file(_file, &status);
  }

  void add_entry(const std::string& name, double value, const std::string& comment = "") {
    write(name, TDOUBLE, &value, comment, false);
  }
  void add_entry(const std::string& name, int value, const std::string& comment = "") {
    write(name, TINT, &value, comment, false);
  }
  void add_entry(const std::string& name, unsigned value, const std::string& comment = "") {
    write(name, TUINT, &value, comment, false);
  }
  void add_entry(const std::string& name, unsigned long long value,
                 const std::string& comment = "") {
    write(name, TULONGLONG, &value, comment, false);
  }
  // strings carry the comment on the key card as well
  void add_entry(const std::string& name, const std::string& value,
                 const std::string& comment = "") {
    write(name, TSTRING, const_cast<char*>(value.c_str()), comment, true);
  }
  void add_entry(const std::string& name, const char* value, const std::string& comment = "") {
    add_entry(name, std::string(value), comment);
  }

 private:
  void write(const std::string& name, int type, void* value, const std::string& comment,
             bool key_comment) {
    int status = 0;
    fits_write_comment(_file, comment.c_str(), &status);
    status = 0;
    if (fits_write_key(_file, type, name.c_str(), value,
                       key_comment ? comment.c_str() : nullptr, &status)) {
      char text[FLEN_STATUS] = {};
      fits_get_errstatus(status, text);
      throw std::runtime_error("Error while updating fits key: " + std::string(text));
    }
  }
  std::string _filename;
  fitsfile* _file = nullptr;
};

// A whitespace-separated numeric table in a text file (lines starting with
// '#' and empty lines skipped, every row the same length), as
// util::data_table_loader reads gaunt_factor.dat: row-major values, cols
// entries per row.  Values parse to float with correct rounding.
inline std::vector<float> load_data_table(const std::string& filename, std::size_t& cols,
                                          std::size_t& rows) {
  std::ifstream file(filename);
  if (!file.is_open()) throw std::runtime_error("Could not load data table from file: " + filename);
  std::vector<float> data;
  std::string line;
  cols = 0;
  rows = 0;
  while (std::getline(file, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream in(line);
    std::string tok;
    std::size_t n = 0;
    while (in >> tok) {
      char* end = nullptr;
      const float v = std::strtof(tok.c_str(), &end);
      if (end == tok.c_str() || *end != '\0')
        throw std::runtime_error("Could not convert entry in data table " + filename + ": " + tok);
      data.push_back(v);
      ++n;
    }
    if (n == 0) continue;
    if (cols != 0 && n != cols)
      throw std::runtime_error("Invalid row in file " + filename + ": expected " +
                               std::to_string(cols) + " entries, got " + std::to_string(n));
    cols = n;
    ++rows;
  }
  return data;
}

}  // namespace io
}  // namespace illcrawl

#endif
