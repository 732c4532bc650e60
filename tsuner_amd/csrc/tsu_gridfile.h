// tsu_gridfile.h -- read-only access to the gridded input files of the
// real-data path: NetCDF classic (CDF-1, 64-bit-offset CDF-2 and 64-bit-data
// CDF-5, read with the library's own code from the published format
// specification) and NetCDF-4 (an HDF5 file: read through the HDF5 C
// library, loaded at run time so that nothing else in the library depends
// on it).
//
// The reference opens every input with nc_open and asks for dimensions and
// variables by name (Problem_loader.cxx:57-175, interpolations.cxx:368-460);
// this class answers the same questions: the length of a named dimension and
// the whole of a named variable converted to double (nc_get_var_double
// semantics: plain type conversion, no scale/offset or fill handling).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace tsu {

class GridFile {
public:
    GridFile();
    ~GridFile();
    GridFile(const GridFile&) = delete;
    GridFile& operator=(const GridFile&) = delete;

    // Opens `path`; false (with error()) when the file is missing, is not a
    // NetCDF classic or HDF5 file, or HDF5 is needed and not loadable.
    bool open(const std::string& path);
    bool is_hdf5() const { return hdf5_; }
    // First name of `names` that exists as a dimension (classic) or dataset
    // (NetCDF-4 stores every dimension as a dataset of that name); "" if none.
    std::string find_dim(const std::vector<std::string>& names) const;
    std::string find_var(const std::vector<std::string>& names) const;
    bool dim_len(const std::string& name, size_t* len) const;
    // Shape of a variable, slowest dimension first.
    bool var_shape(const std::string& name, std::vector<size_t>* shape) const;
    // The whole variable, row-major, converted to double.
    bool read_double(const std::string& name, std::vector<double>* out) const;
    const std::string& error() const { return err_; }

private:
    struct Impl;
    std::unique_ptr<Impl> im_;
    bool hdf5_ = false;
    mutable std::string err_;
};

}  // namespace tsu
