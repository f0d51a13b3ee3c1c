// NetCDF classic-format (64-bit offset) writer; see lhysea_netcdf.hpp.
#include "lhysea_netcdf.hpp"

#include <cstring>

namespace lhysea {

namespace {

void be32(std::vector<uint8_t>& o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}
void be64(std::vector<uint8_t>& o, uint64_t v)
{
    for (int s = 56; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}
void pad4(std::vector<uint8_t>& o)
{
    while (o.size() % 4) o.push_back(0);
}
void name(std::vector<uint8_t>& o, const std::string& s)
{
    be32(o, (uint32_t)s.size());
    o.insert(o.end(), s.begin(), s.end());
    pad4(o);
}
int type_size(int t) { return t == NcWriter::NC_CHAR ? 1 : (t == NcWriter::NC_DOUBLE ? 8 : 4); }

void atts(std::vector<uint8_t>& o, const std::vector<NcWriter::Att>& a)
{
    if (a.empty()) {
        be32(o, 0);
        be32(o, 0);
        return;
    }
    be32(o, 0x0C);
    be32(o, (uint32_t)a.size());
    for (const auto& x : a) {
        name(o, x.name);
        be32(o, (uint32_t)x.type);
        if (x.type == NcWriter::NC_CHAR) {
            be32(o, (uint32_t)x.text.size());
            o.insert(o.end(), x.text.begin(), x.text.end());
        } else {
            be32(o, (uint32_t)x.values.size());
            for (double v : x.values) {
                float f = (float)v;
                uint32_t u;
                std::memcpy(&u, &f, 4);
                be32(o, u);
            }
        }
        pad4(o);
    }
}

}  // namespace

NcWriter::NcWriter(const std::string& path) { f_ = std::fopen(path.c_str(), "wb+"); }

NcWriter::~NcWriter()
{
    if (f_) close();
}

int NcWriter::def_dim(const std::string& n, int64_t len)
{
    dims_.push_back({n, len});
    return (int)dims_.size() - 1;
}

int NcWriter::def_var(const std::string& n, int type, const std::vector<int>& dims)
{
    Var v;
    v.name = n;
    v.type = type;
    v.dims = dims;
    vars_.push_back(v);
    return (int)vars_.size() - 1;
}

void NcWriter::put_att_text(int var, const std::string& n, const std::string& text)
{
    Att a{n, NC_CHAR, text, {}};
    (var < 0 ? gatts_ : vars_[var].atts).push_back(a);
}

void NcWriter::put_att_float(int var, const std::string& n, float v)
{
    Att a{n, NC_FLOAT, "", {(double)v}};
    (var < 0 ? gatts_ : vars_[var].atts).push_back(a);
}

void NcWriter::header(std::vector<uint8_t>& o) const
{
    o.clear();
    o.push_back('C');
    o.push_back('D');
    o.push_back('F');
    o.push_back(2);
    be32(o, (uint32_t)numrecs_);
    if (dims_.empty()) {
        be32(o, 0);
        be32(o, 0);
    } else {
        be32(o, 0x0A);
        be32(o, (uint32_t)dims_.size());
        for (const auto& d : dims_) {
            name(o, d.name);
            be32(o, (uint32_t)d.len);
        }
    }
    atts(o, gatts_);
    if (vars_.empty()) {
        be32(o, 0);
        be32(o, 0);
        return;
    }
    be32(o, 0x0B);
    be32(o, (uint32_t)vars_.size());
    for (const auto& v : vars_) {
        name(o, v.name);
        be32(o, (uint32_t)v.dims.size());
        for (int d : v.dims) be32(o, (uint32_t)d);
        atts(o, v.atts);
        be32(o, (uint32_t)v.type);
        be32(o, (uint32_t)v.vsize);
        be64(o, (uint64_t)v.begin);
    }
}

bool NcWriter::enddef()
{
    if (!f_ || defined_) return false;
    int nrec = 0;
    for (auto& v : vars_) {
        int64_t n = 1;
        v.record = !v.dims.empty() && dims_[v.dims[0]].len == 0;
        for (size_t i = v.record ? 1 : 0; i < v.dims.size(); i++) n *= dims_[v.dims[i]].len;
        v.nelems = n;
        int64_t bytes = n * type_size(v.type);
        v.vsize = (bytes + 3) / 4 * 4;
        if (v.record) nrec++;
    }
    // the header size does not depend on the offsets' values
    std::vector<uint8_t> h;
    header(h);
    int64_t off = (int64_t)h.size();
    for (auto& v : vars_)
        if (!v.record) {
            v.begin = off;
            off += v.vsize;
        }
    recsize_ = 0;
    for (auto& v : vars_)
        if (v.record) {
            v.begin = off + recsize_;
            // a file with a single record variable does not pad its records
            recsize_ += (nrec == 1) ? v.nelems * type_size(v.type) : v.vsize;
        }
    header(h);
    defined_ = true;
    if (!write_at(0, h.data(), h.size())) return false;
    // pre-size the fixed-variable section with zeros
    for (auto& v : vars_)
        if (!v.record) {
            std::vector<uint8_t> z((size_t)v.vsize, 0);
            if (!write_at(v.begin, z.data(), z.size())) return false;
        }
    return true;
}

bool NcWriter::write_at(int64_t off, const void* p, size_t n)
{
    if (fseeko(f_, (off_t)off, SEEK_SET) != 0) return false;
    return std::fwrite(p, 1, n, f_) == n;
}

static void to_be_floats(const float* d, int64_t n, std::vector<uint8_t>& o)
{
    o.resize((size_t)((n * 4 + 3) / 4 * 4), 0);
    for (int64_t i = 0; i < n; i++) {
        uint32_t u;
        std::memcpy(&u, &d[i], 4);
        o[4 * i] = (uint8_t)(u >> 24);
        o[4 * i + 1] = (uint8_t)(u >> 16);
        o[4 * i + 2] = (uint8_t)(u >> 8);
        o[4 * i + 3] = (uint8_t)u;
    }
}

bool NcWriter::put_var(int var, const float* data)
{
    const Var& v = vars_[var];
    if (!defined_ || v.record || v.type != NC_FLOAT) return false;
    std::vector<uint8_t> b;
    to_be_floats(data, v.nelems, b);
    return write_at(v.begin, b.data(), b.size());
}

bool NcWriter::write_numrecs()
{
    uint8_t b[4] = {(uint8_t)(numrecs_ >> 24), (uint8_t)(numrecs_ >> 16), (uint8_t)(numrecs_ >> 8),
                    (uint8_t)numrecs_};
    return write_at(4, b, 4);
}

bool NcWriter::put_record(int var, int64_t rec, const float* data)
{
    const Var& v = vars_[var];
    if (!defined_ || !v.record || v.type != NC_FLOAT) return false;
    std::vector<uint8_t> b;
    to_be_floats(data, v.nelems, b);
    if (!write_at(v.begin + rec * recsize_, b.data(), b.size())) return false;
    if (rec + 1 > numrecs_) {
        numrecs_ = rec + 1;
        return write_numrecs();
    }
    return true;
}

bool NcWriter::close()
{
    if (!f_) return false;
    bool ok = std::fflush(f_) == 0;
    ok = (std::fclose(f_) == 0) && ok;
    f_ = nullptr;
    return ok;
}

}  // namespace lhysea
