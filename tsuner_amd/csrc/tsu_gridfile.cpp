// tsu_gridfile.cpp -- NetCDF classic reader and the run-time HDF5 binding
// behind tsu::GridFile (declared in tsu_gridfile.h).
//
// Classic format (big-endian): magic "CDF" + version byte (1: 32-bit
// offsets, 2: 64-bit offsets, 5: 64-bit offsets and 64-bit counts, the
// "64-bit data" format), numrecs, then the dimension, global attribute and
// variable lists; a list is ABSENT (a zero tag and a zero count) or a tag
// (0x0A dimensions, 0x0C attributes, 0x0B variables) and a count.  Counts,
// name lengths, dimension lengths and ids, vsize and numrecs are 32-bit
// (versions 1 and 2) or 64-bit (version 5); type codes and tags stay
// 32-bit.  Version 5 adds the unsigned and 64-bit integer types.  Names and
// attribute values are padded to 4 bytes.  A variable whose first dimension
// has length 0 is a record variable: record r of it starts at
// begin + r*recsize, recsize being the sum of the record variables' vsize
// (or the unpadded size when there is a single record variable).
#include "tsu_gridfile.h"

#include <dlfcn.h>

#include <algorithm>

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace tsu {

namespace {

enum {
    NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6,
    // version 5 only
    NC_UBYTE = 7, NC_USHORT = 8, NC_UINT = 9, NC_INT64 = 10, NC_UINT64 = 11
};

int type_size(int t)
{
    switch (t) {
    case NC_BYTE:
    case NC_CHAR:
    case NC_UBYTE: return 1;
    case NC_SHORT:
    case NC_USHORT: return 2;
    case NC_INT:
    case NC_FLOAT:
    case NC_UINT: return 4;
    case NC_DOUBLE:
    case NC_INT64:
    case NC_UINT64: return 8;
    }
    return 0;
}

uint64_t be(const uint8_t* p, int n)
{
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v = (v << 8) | p[i];
    return v;
}

// n big-endian values of type t from raw to out, converted to double.
void decode_all(const uint8_t* raw, size_t n, int t, double* out)
{
    switch (t) {
    case NC_BYTE:
        for (size_t i = 0; i < n; i++) out[i] = (double)(int8_t)raw[i];
        break;
    case NC_CHAR:
    case NC_UBYTE:
        for (size_t i = 0; i < n; i++) out[i] = (double)raw[i];
        break;
    case NC_USHORT:
        for (size_t i = 0; i < n; i++) {
            uint16_t u;
            std::memcpy(&u, raw + 2 * i, 2);
            out[i] = (double)__builtin_bswap16(u);
        }
        break;
    case NC_UINT:
        for (size_t i = 0; i < n; i++) {
            uint32_t u;
            std::memcpy(&u, raw + 4 * i, 4);
            out[i] = (double)__builtin_bswap32(u);
        }
        break;
    case NC_INT64:
        for (size_t i = 0; i < n; i++) {
            uint64_t u;
            std::memcpy(&u, raw + 8 * i, 8);
            out[i] = (double)(int64_t)__builtin_bswap64(u);
        }
        break;
    case NC_UINT64:
        for (size_t i = 0; i < n; i++) {
            uint64_t u;
            std::memcpy(&u, raw + 8 * i, 8);
            out[i] = (double)__builtin_bswap64(u);
        }
        break;
    case NC_SHORT:
        for (size_t i = 0; i < n; i++) {
            uint16_t u;
            std::memcpy(&u, raw + 2 * i, 2);
            out[i] = (double)(int16_t)__builtin_bswap16(u);
        }
        break;
    case NC_INT:
        for (size_t i = 0; i < n; i++) {
            uint32_t u;
            std::memcpy(&u, raw + 4 * i, 4);
            out[i] = (double)(int32_t)__builtin_bswap32(u);
        }
        break;
    case NC_FLOAT:
        for (size_t i = 0; i < n; i++) {
            uint32_t u;
            std::memcpy(&u, raw + 4 * i, 4);
            u = __builtin_bswap32(u);
            float f;
            std::memcpy(&f, &u, 4);
            out[i] = (double)f;
        }
        break;
    case NC_DOUBLE:
        for (size_t i = 0; i < n; i++) {
            uint64_t u;
            std::memcpy(&u, raw + 8 * i, 8);
            u = __builtin_bswap64(u);
            std::memcpy(out + i, &u, 8);
        }
        break;
    }
}

// Forward-only cursor over the file, reading more as the header needs it.
struct Cursor {
    std::FILE* f;
    int nn = 4;   // bytes of a count, length, id or size (8 in version 5)
    int max_type = 6;
    std::vector<uint8_t> buf;
    size_t pos = 0;
    bool ok = true;
    bool need(size_t n)
    {
        while (ok && buf.size() < pos + n) {
            const size_t old = buf.size();
            const size_t want = std::max<size_t>(n, 1 << 16);
            buf.resize(old + want);
            const size_t got = std::fread(buf.data() + old, 1, want, f);
            buf.resize(old + got);
            if (got == 0) ok = false;
        }
        return ok;
    }
    uint64_t u(int n)
    {
        if (!need((size_t)n)) return 0;
        const uint64_t v = be(buf.data() + pos, n);
        pos += (size_t)n;
        return v;
    }
    std::string name()
    {
        const size_t n = (size_t)u(nn);
        const size_t padded = (n + 3) & ~(size_t)3;
        if (n > (1u << 20) || !need(padded)) {
            ok = false;
            return std::string();
        }
        std::string s((const char*)buf.data() + pos, n);
        pos += padded;
        return s;
    }
    void skip_atts()
    {
        const uint64_t tag = u(4), count = u(nn);
        if (tag == 0 && count == 0) return;
        if (tag != 0x0C || count > (1u << 20)) {
            ok = false;
            return;
        }
        for (uint64_t a = 0; a < count && ok; a++) {
            name();
            const int t = (int)u(4);
            const uint64_t n = u(nn);
            const uint64_t bytes = ((uint64_t)type_size(t) * n + 3) & ~(uint64_t)3;
            if (t > max_type || type_size(t) == 0 || n > (1u << 26) || bytes > (1u << 26) || !need((size_t)bytes)) {
                ok = false;
                return;
            }
            pos += (size_t)bytes;
        }
    }
};

// ---------------------------------------------------------------- HDF5 (run-time binding)
// HDF5 1.10 C API: hid_t is int64_t, herr_t and htri_t are int, hsize_t is
// unsigned long long; H5P_DEFAULT, H5S_ALL and H5E_DEFAULT are 0 and
// H5F_ACC_RDONLY is 0u.  H5T_NATIVE_DOUBLE is the exported variable
// H5T_NATIVE_DOUBLE_g, valid after H5open().
typedef int64_t hid_t;
typedef unsigned long long hsize_t;
struct H5 {
    void* so = nullptr;
    int (*open)(void);
    int (*eset_auto)(hid_t, void*, void*);
    hid_t (*fopen)(const char*, unsigned, hid_t);
    int (*fclose)(hid_t);
    int (*lexists)(hid_t, const char*, hid_t);
    hid_t (*dopen)(hid_t, const char*, hid_t);
    int (*dclose)(hid_t);
    hid_t (*dget_space)(hid_t);
    int (*sndims)(hid_t);
    int (*sdims)(hid_t, hsize_t*, hsize_t*);
    int (*sclose)(hid_t);
    int (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    hid_t native_double = -1;
    std::string error;
};

template <class F>
bool sym(void* so, const char* name, F* f)
{
    *f = (F)dlsym(so, name);
    return *f != nullptr;
}

H5& h5_state()
{
    static H5 h;
    return h;
}

bool load_hdf5(H5& h)
{
    std::vector<std::string> names;
    if (const char* e = std::getenv("TSU_HDF5_LIB")) names.push_back(e);
    for (const char* n : {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"}) names.push_back(n);
    for (const auto& n : names)
        if ((h.so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!h.so) {
        h.error = "the HDF5 C library (libhdf5.so) is not loadable; set TSU_HDF5_LIB to its path";
        return false;
    }
    bool ok = sym(h.so, "H5open", &h.open) && sym(h.so, "H5Eset_auto2", &h.eset_auto) &&
              sym(h.so, "H5Fopen", &h.fopen) && sym(h.so, "H5Fclose", &h.fclose) &&
              sym(h.so, "H5Lexists", &h.lexists) && sym(h.so, "H5Dopen2", &h.dopen) &&
              sym(h.so, "H5Dclose", &h.dclose) && sym(h.so, "H5Dget_space", &h.dget_space) &&
              sym(h.so, "H5Sget_simple_extent_ndims", &h.sndims) &&
              sym(h.so, "H5Sget_simple_extent_dims", &h.sdims) && sym(h.so, "H5Sclose", &h.sclose) &&
              sym(h.so, "H5Dread", &h.dread);
    hid_t* nd = ok ? (hid_t*)dlsym(h.so, "H5T_NATIVE_DOUBLE_g") : nullptr;
    if (!ok || !nd || h.open() < 0) {
        h.error = "the loaded HDF5 library lacks the 1.10 C API";
        dlclose(h.so);
        h.so = nullptr;
        return false;
    }
    h.native_double = *nd;
    h.eset_auto(0, nullptr, nullptr);   // no error stack printing; failures are returned
    return true;
}

// The library, loaded once per process (thread-safe static initialisation);
// nullptr when unavailable.
H5* hdf5()
{
    static const bool ok = load_hdf5(h5_state());
    return ok ? &h5_state() : nullptr;
}

}  // namespace

struct GridFile::Impl {
    // classic
    struct Dim {
        std::string name;
        uint64_t len;
    };
    struct Var {
        std::string name;
        std::vector<int> dims;
        int type = 0;
        uint64_t vsize = 0, begin = 0;
        bool rec = false;
    };
    std::FILE* f = nullptr;
    std::vector<Dim> dims;
    std::vector<Var> vars;
    uint64_t numrecs = 0, recsize = 0;
    // HDF5
    hid_t h5file = -1;

    const Var* var(const std::string& n) const
    {
        for (const auto& v : vars)
            if (v.name == n) return &v;
        return nullptr;
    }
    bool h5_shape(const std::string& n, std::vector<size_t>* shape) const
    {
        H5* h = hdf5();
        if (!h || h->lexists(h5file, n.c_str(), 0) <= 0) return false;
        const hid_t d = h->dopen(h5file, n.c_str(), 0);
        if (d < 0) return false;
        const hid_t s = h->dget_space(d);
        bool ok = s >= 0;
        if (ok) {
            const int nd = h->sndims(s);
            std::vector<hsize_t> e((size_t)std::max(nd, 1));
            ok = nd >= 0 && h->sdims(s, e.data(), nullptr) >= 0;
            if (ok) shape->assign(e.begin(), e.begin() + nd);
            h->sclose(s);
        }
        h->dclose(d);
        return ok;
    }
};

GridFile::GridFile() : im_(new Impl) {}

GridFile::~GridFile()
{
    if (im_->f) std::fclose(im_->f);
    if (im_->h5file >= 0) hdf5()->fclose(im_->h5file);
}

bool GridFile::open(const std::string& path)
{
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err_ = "cannot open " + path;
        return false;
    }
    uint8_t magic[8] = {0};
    const size_t got = std::fread(magic, 1, 8, f);
    if (got >= 8 && std::memcmp(magic, "\x89HDF\r\n\x1a\n", 8) == 0) {
        std::fclose(f);
        H5* h = hdf5();
        if (!h) {
            err_ = path + " is a NetCDF-4/HDF5 file: " + h5_state().error;
            return false;
        }
        im_->h5file = h->fopen(path.c_str(), 0u, 0);
        if (im_->h5file < 0) {
            err_ = "HDF5 cannot open " + path;
            return false;
        }
        hdf5_ = true;
        return true;
    }
    const int version = got >= 4 ? magic[3] : 0;
    if (got < 4 || std::memcmp(magic, "CDF", 3) != 0 || (version != 1 && version != 2 && version != 5)) {
        std::fclose(f);
        err_ = path + " is neither a NetCDF classic (CDF-1/CDF-2/CDF-5) nor a NetCDF-4/HDF5 file";
        return false;
    }
    const int off_bytes = version == 1 ? 4 : 8;
    std::fseek(f, 0, SEEK_SET);
    Cursor c{f};
    c.nn = version == 5 ? 8 : 4;
    c.max_type = version == 5 ? NC_UINT64 : NC_DOUBLE;
    c.u(4);
    im_->numrecs = c.u(c.nn);
    const uint64_t streaming = version == 5 ? ~0ull : 0xFFFFFFFFull;
    // dimensions
    uint64_t tag = c.u(4), count = c.u(c.nn);
    if (!(tag == 0 && count == 0)) {
        if (tag != 0x0A || count > (1u << 20)) c.ok = false;
        for (uint64_t i = 0; i < count && c.ok; i++) {
            Impl::Dim d;
            d.name = c.name();
            d.len = c.u(c.nn);
            im_->dims.push_back(d);
        }
    }
    c.skip_atts();
    tag = c.u(4);
    count = c.u(c.nn);
    if (!(tag == 0 && count == 0)) {
        if (tag != 0x0B || count > (1u << 20)) c.ok = false;
        for (uint64_t i = 0; i < count && c.ok; i++) {
            Impl::Var v;
            v.name = c.name();
            const uint64_t nd = c.u(c.nn);
            if (nd > 1024) c.ok = false;
            for (uint64_t k = 0; k < nd && c.ok; k++) {
                const uint64_t id = c.u(c.nn);
                if (id >= im_->dims.size()) c.ok = false;
                v.dims.push_back((int)id);
            }
            c.skip_atts();
            v.type = (int)c.u(4);
            v.vsize = c.u(c.nn);
            v.begin = c.u(off_bytes);
            if (v.type > c.max_type || type_size(v.type) == 0) c.ok = false;
            v.rec = c.ok && !v.dims.empty() && im_->dims[(size_t)v.dims[0]].len == 0;
            im_->vars.push_back(v);
        }
    }
    if (!c.ok) {
        std::fclose(f);
        err_ = "malformed NetCDF classic header in " + path;
        return false;
    }
    int nrec = 0;
    for (const auto& v : im_->vars)
        if (v.rec) {
            nrec++;
            im_->recsize += v.vsize;
        }
    if (nrec == 1)
        for (const auto& v : im_->vars)
            if (v.rec) {
                uint64_t n = (uint64_t)type_size(v.type);
                for (size_t k = 1; k < v.dims.size(); k++) n *= im_->dims[(size_t)v.dims[k]].len;
                im_->recsize = n;
            }
    if (im_->numrecs == streaming) {   // streaming: count from the file size
        std::fseek(f, 0, SEEK_END);
        const uint64_t size = (uint64_t)std::ftell(f);
        uint64_t first = UINT64_MAX;
        for (const auto& v : im_->vars)
            if (v.rec) first = std::min(first, v.begin);
        im_->numrecs = (im_->recsize && first != UINT64_MAX && size > first) ? (size - first) / im_->recsize : 0;
    }
    im_->f = f;
    return true;
}

std::string GridFile::find_dim(const std::vector<std::string>& names) const
{
    for (const auto& n : names) {
        size_t len;
        if (dim_len(n, &len)) return n;
    }
    return std::string();
}

std::string GridFile::find_var(const std::vector<std::string>& names) const
{
    for (const auto& n : names) {
        std::vector<size_t> s;
        if (var_shape(n, &s)) return n;
    }
    return std::string();
}

bool GridFile::dim_len(const std::string& name, size_t* len) const
{
    if (hdf5_) {
        std::vector<size_t> s;
        if (!im_->h5_shape(name, &s) || s.size() != 1) return false;
        *len = s[0];
        return true;
    }
    for (const auto& d : im_->dims)
        if (d.name == name) {
            *len = (size_t)(d.len == 0 ? im_->numrecs : d.len);
            return true;
        }
    return false;
}

bool GridFile::var_shape(const std::string& name, std::vector<size_t>* shape) const
{
    if (hdf5_) return im_->h5_shape(name, shape);
    const Impl::Var* v = im_->var(name);
    if (!v) return false;
    shape->clear();
    for (int d : v->dims) {
        const uint64_t l = im_->dims[(size_t)d].len;
        shape->push_back((size_t)(l == 0 ? im_->numrecs : l));
    }
    return true;
}

bool GridFile::read_double(const std::string& name, std::vector<double>* out) const
{
    std::vector<size_t> shape;
    if (!var_shape(name, &shape)) {
        err_ = "no variable " + name;
        return false;
    }
    size_t n = 1;
    for (size_t s : shape) n *= s;
    out->assign(n, 0.0);
    if (hdf5_) {
        H5* h = hdf5();
        const hid_t d = h->dopen(im_->h5file, name.c_str(), 0);
        const bool ok = d >= 0 && (n == 0 || h->dread(d, h->native_double, 0, 0, 0, out->data()) >= 0);
        if (d >= 0) h->dclose(d);
        if (!ok) err_ = "HDF5 read of " + name + " failed";
        return ok;
    }
    const Impl::Var* v = im_->var(name);
    const int ts = type_size(v->type);
    // elements per record (record variable) or in total
    size_t per = 1;
    for (size_t k = v->rec ? 1 : 0; k < shape.size(); k++) per *= shape[k];
    const size_t nrec = v->rec ? (size_t)im_->numrecs : 1;
    std::vector<uint8_t> raw(per * (size_t)ts);
    for (size_t r = 0; r < nrec; r++) {
        const uint64_t off = v->begin + (v->rec ? r * im_->recsize : 0);
        if (std::fseek(im_->f, (long)off, SEEK_SET) != 0 || std::fread(raw.data(), 1, raw.size(), im_->f) != raw.size()) {
            err_ = "short read of " + name;
            return false;
        }
        decode_all(raw.data(), per, v->type, out->data() + r * per);
    }
    return true;
}

}  // namespace tsu
