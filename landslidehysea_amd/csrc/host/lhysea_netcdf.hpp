// Minimal NetCDF classic-format writer (64-bit offset variant, "CDF\x02").
//
// Writes the layout the reference produces through PnetCDF
// (src/GPU/netcdf.cu:29-411): fixed-size variables, one unlimited "time"
// dimension and record variables, text and float attributes.  Big-endian,
// 4-byte padding, record variables interleaved per record as the format
// specification requires.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace lhysea {

class NcWriter {
public:
    enum Type { NC_CHAR = 2, NC_FLOAT = 5, NC_DOUBLE = 6 };
    struct Att {
        std::string name;
        int type;
        std::string text;
        std::vector<double> values;
    };

    explicit NcWriter(const std::string& path);
    ~NcWriter();
    bool ok() const { return f_ != nullptr; }

    int def_dim(const std::string& name, int64_t len);    // len 0 = unlimited
    int def_var(const std::string& name, int type, const std::vector<int>& dims);
    void put_att_text(int var, const std::string& name, const std::string& text);   // var -1: global
    void put_att_float(int var, const std::string& name, float v);
    bool enddef();
    bool put_var(int var, const float* data);            // a fixed-size variable
    bool put_record(int var, int64_t rec, const float* data);
    bool close();

private:
    struct Dim {
        std::string name;
        int64_t len;
    };
    struct Var {
        std::string name;
        int type;
        std::vector<int> dims;
        std::vector<Att> atts;
        int64_t vsize = 0;
        int64_t begin = 0;
        bool record = false;
        int64_t nelems = 0;       // elements per record (record vars) or total
    };
    void header(std::vector<uint8_t>& out) const;
    bool write_at(int64_t off, const void* p, size_t n);
    bool write_numrecs();

    FILE* f_ = nullptr;
    std::vector<Dim> dims_;
    std::vector<Var> vars_;
    std::vector<Att> gatts_;
    int64_t recsize_ = 0;
    int64_t numrecs_ = 0;
    bool defined_ = false;
};

}  // namespace lhysea
