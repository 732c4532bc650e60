// tsu_netcdf.h -- minimal NetCDF classic writer (64-bit offset format,
// "CDF\002"), enough for the solver output file: fixed dimensions plus one
// unlimited record dimension, double and text attributes, double variables.
//
// The library does not depend on libnetcdf: the format is written from
// the published NetCDF classic file format specification (big-endian
// header: magic, numrecs, dim_list, gatt_list, var_list; non-record
// variables stored contiguously after the header; record variables
// interleaved per record).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace tsu {

class NcWriter {
public:
    ~NcWriter();
    int add_dim(const std::string& name, size_t len);   // len 0 = unlimited
    int add_var(const std::string& name, const std::vector<int>& dims);
    void att_text(int var, const std::string& name, const std::string& value);   // var -1 = global
    void att_double(int var, const std::string& name, double value);
    bool create(const std::string& path);   // writes the header; returns false on I/O error
    bool put_var(int var, const double* data);               // whole non-record variable
    bool put_record(int var, size_t rec, const double* data); // one record of a record variable
    bool sync();
    bool close();
    bool ok() const { return ok_; }

private:
    struct Att {
        std::string name;
        bool text;
        std::string s;
        double d;
    };
    struct Dim {
        std::string name;
        size_t len;
    };
    struct Var {
        std::string name;
        std::vector<int> dims;
        std::vector<Att> atts;
        bool rec = false;
        uint64_t vsize = 0;
        uint64_t begin = 0;
        uint64_t slice = 0;   // elements per record (or total for non-record)
    };
    void put_header(std::vector<uint8_t>& h) const;
    std::vector<Dim> dims_;
    std::vector<Att> gatts_;
    std::vector<Var> vars_;
    std::FILE* f_ = nullptr;
    uint64_t recsize_ = 0;
    uint64_t numrecs_ = 0;
    bool ok_ = true;
};

}  // namespace tsu
