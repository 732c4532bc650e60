// tsu_netcdf.cpp -- NetCDF classic (64-bit offset) writer.
#include "tsu_netcdf.h"

#include <cstring>
#include <sys/types.h>

namespace tsu {

namespace {

constexpr uint32_t kDimTag = 0x0A, kVarTag = 0x0B, kAttTag = 0x0C;
constexpr uint32_t kChar = 2, kDouble = 6;

void be32(std::vector<uint8_t>& b, uint32_t v)
{
    for (int i = 3; i >= 0; i--) b.push_back((uint8_t)(v >> (8 * i)));
}

void be64(std::vector<uint8_t>& b, uint64_t v)
{
    for (int i = 7; i >= 0; i--) b.push_back((uint8_t)(v >> (8 * i)));
}

void pad4(std::vector<uint8_t>& b)
{
    while (b.size() % 4) b.push_back(0);
}

void name(std::vector<uint8_t>& b, const std::string& s)
{
    be32(b, (uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad4(b);
}

uint64_t round4(uint64_t v) { return (v + 3) & ~uint64_t(3); }

void to_be_doubles(const double* src, size_t n, std::vector<uint8_t>& out)
{
    out.resize(n * 8);
    for (size_t i = 0; i < n; i++) {
        uint64_t u;
        std::memcpy(&u, &src[i], 8);
        for (int k = 0; k < 8; k++) out[i * 8 + k] = (uint8_t)(u >> (8 * (7 - k)));
    }
}

}  // namespace

NcWriter::~NcWriter()
{
    if (f_) std::fclose(f_);
}

int NcWriter::add_dim(const std::string& n, size_t len)
{
    dims_.push_back({n, len});
    return (int)dims_.size() - 1;
}

int NcWriter::add_var(const std::string& n, const std::vector<int>& dims)
{
    Var v;
    v.name = n;
    v.dims = dims;
    vars_.push_back(v);
    return (int)vars_.size() - 1;
}

void NcWriter::att_text(int var, const std::string& n, const std::string& value)
{
    Att a{n, true, value, 0.0};
    if (var < 0)
        gatts_.push_back(a);
    else
        vars_[var].atts.push_back(a);
}

void NcWriter::att_double(int var, const std::string& n, double value)
{
    Att a{n, false, std::string(), value};
    if (var < 0)
        gatts_.push_back(a);
    else
        vars_[var].atts.push_back(a);
}

void NcWriter::put_header(std::vector<uint8_t>& b) const
{
    auto atts = [&](const std::vector<Att>& list) {
        if (list.empty()) {
            be32(b, 0);
            be32(b, 0);
            return;
        }
        be32(b, kAttTag);
        be32(b, (uint32_t)list.size());
        for (const Att& a : list) {
            name(b, a.name);
            if (a.text) {
                be32(b, kChar);
                be32(b, (uint32_t)a.s.size());
                b.insert(b.end(), a.s.begin(), a.s.end());
                pad4(b);
            } else {
                be32(b, kDouble);
                be32(b, 1);
                uint64_t u;
                std::memcpy(&u, &a.d, 8);
                be64(b, u);
            }
        }
    };
    b.clear();
    b.push_back('C');
    b.push_back('D');
    b.push_back('F');
    b.push_back(2);
    be32(b, (uint32_t)numrecs_);
    if (dims_.empty()) {
        be32(b, 0);
        be32(b, 0);
    } else {
        be32(b, kDimTag);
        be32(b, (uint32_t)dims_.size());
        for (const Dim& d : dims_) {
            name(b, d.name);
            be32(b, (uint32_t)d.len);
        }
    }
    atts(gatts_);
    if (vars_.empty()) {
        be32(b, 0);
        be32(b, 0);
    } else {
        be32(b, kVarTag);
        be32(b, (uint32_t)vars_.size());
        for (const Var& v : vars_) {
            name(b, v.name);
            be32(b, (uint32_t)v.dims.size());
            for (int d : v.dims) be32(b, (uint32_t)d);
            atts(v.atts);
            be32(b, kDouble);
            be32(b, v.vsize > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)v.vsize);
            be64(b, v.begin);
        }
    }
}

bool NcWriter::create(const std::string& path)
{
    // sizes
    for (Var& v : vars_) {
        uint64_t n = 1;
        v.rec = false;
        for (size_t k = 0; k < v.dims.size(); k++) {
            const Dim& d = dims_[v.dims[k]];
            if (d.len == 0 && k == 0)
                v.rec = true;
            else
                n *= d.len;
        }
        v.slice = n;
        v.vsize = round4(n * 8);
    }
    std::vector<uint8_t> hdr;
    put_header(hdr);   // offsets are fixed width, so the size is final
    uint64_t off = hdr.size();
    for (Var& v : vars_)
        if (!v.rec) {
            v.begin = off;
            off += v.vsize;
        }
    int nrec = 0;
    for (Var& v : vars_)
        if (v.rec) nrec++;
    recsize_ = 0;
    for (Var& v : vars_)
        if (v.rec) {
            v.begin = off + recsize_;
            // a single record variable is not padded (classic format rule)
            recsize_ += (nrec == 1) ? v.slice * 8 : v.vsize;
        }
    put_header(hdr);
    f_ = std::fopen(path.c_str(), "wb+");
    if (!f_) return ok_ = false;
    ok_ = std::fwrite(hdr.data(), 1, hdr.size(), f_) == hdr.size();
    return ok_;
}

bool NcWriter::put_var(int var, const double* data)
{
    if (!f_ || var < 0 || var >= (int)vars_.size() || vars_[var].rec) return ok_ = false;
    const Var& v = vars_[var];
    std::vector<uint8_t> buf;
    to_be_doubles(data, v.slice, buf);
    buf.resize(v.vsize, 0);
    if (fseeko(f_, (off_t)v.begin, SEEK_SET) != 0) return ok_ = false;
    if (std::fwrite(buf.data(), 1, buf.size(), f_) != buf.size()) return ok_ = false;
    return true;
}

bool NcWriter::put_record(int var, size_t rec, const double* data)
{
    if (!f_ || var < 0 || var >= (int)vars_.size() || !vars_[var].rec) return ok_ = false;
    const Var& v = vars_[var];
    std::vector<uint8_t> buf;
    to_be_doubles(data, v.slice, buf);
    uint64_t pos = v.begin + (uint64_t)rec * recsize_;
    if (fseeko(f_, (off_t)pos, SEEK_SET) != 0) return ok_ = false;
    if (std::fwrite(buf.data(), 1, buf.size(), f_) != buf.size()) return ok_ = false;
    if (rec + 1 > numrecs_) {
        numrecs_ = rec + 1;
        std::vector<uint8_t> n4;
        be32(n4, (uint32_t)numrecs_);
        if (fseeko(f_, 4, SEEK_SET) != 0) return ok_ = false;
        if (std::fwrite(n4.data(), 1, 4, f_) != 4) return ok_ = false;
    }
    return true;
}

bool NcWriter::sync()
{
    if (!f_) return false;
    return std::fflush(f_) == 0;
}

bool NcWriter::close()
{
    if (!f_) return false;
    // make the file long enough for every record slot
    uint64_t first = UINT64_MAX;
    for (const Var& v : vars_)
        if (v.rec && v.begin < first) first = v.begin;
    if (first != UINT64_MAX && numrecs_ > 0) {
        uint64_t want = first + numrecs_ * recsize_;
        fseeko(f_, 0, SEEK_END);
        off_t have = ftello(f_);
        if (have >= 0 && (uint64_t)have < want) {
            std::vector<uint8_t> z((size_t)(want - (uint64_t)have), 0);
            std::fwrite(z.data(), 1, z.size(), f_);
        }
    }
    bool good = std::fclose(f_) == 0;
    f_ = nullptr;
    return good && ok_;
}

}  // namespace tsu
