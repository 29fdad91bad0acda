// netcdf_classic.cpp -- NetCDF classic (64-bit offset) writer and reader.
#include "netcdf_classic.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace axg {
namespace netcdf {

namespace {
const uint32_t kDimTag = 0x0A, kVarTag = 0x0B, kAttTag = 0x0C;

int typeSize(Type t) { return t == CHAR ? 1 : (t == DOUBLE ? 8 : 4); }
int64_t pad4(int64_t n) { return (n + 3) / 4 * 4; }

struct Out {
    std::vector<unsigned char> b;
    void u32(uint32_t v) {
        for (int i = 3; i >= 0; i--) b.push_back((unsigned char)(v >> (8 * i)));
    }
    void u64(uint64_t v) {
        for (int i = 7; i >= 0; i--) b.push_back((unsigned char)(v >> (8 * i)));
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s) {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
    void value(Type t, double v) {
        if (t == DOUBLE) {
            uint64_t u;
            std::memcpy(&u, &v, 8);
            u64(u);
        } else if (t == FLOAT) {
            const float f = (float)v;
            uint32_t u;
            std::memcpy(&u, &f, 4);
            u32(u);
        } else {
            u32((uint32_t)(int32_t)v);
        }
    }
    void attrs(const std::vector<Attr> &a) {
        if (a.empty()) {
            u32(0);
            u32(0);
            return;
        }
        u32(kAttTag);
        u32((uint32_t)a.size());
        for (const auto &x : a) {
            name(x.name);
            u32((uint32_t)x.type);
            if (x.type == CHAR) {
                u32((uint32_t)x.text.size());
                b.insert(b.end(), x.text.begin(), x.text.end());
            } else {
                u32((uint32_t)x.values.size());
                for (double v : x.values) value(x.type, v);
            }
            pad();
        }
    }
};

struct In {
    const std::vector<unsigned char> &b;
    size_t p = 0;
    explicit In(const std::vector<unsigned char> &v) : b(v) {}
    void need(size_t n) const {
        if (p + n > b.size()) throw std::runtime_error("NetCDF_Reader || truncated file.");
    }
    uint32_t u32() {
        need(4);
        uint32_t v = 0;
        for (int i = 0; i < 4; i++) v = (v << 8) | b[p++];
        return v;
    }
    uint64_t u64() {
        need(8);
        uint64_t v = 0;
        for (int i = 0; i < 8; i++) v = (v << 8) | b[p++];
        return v;
    }
    void skip(size_t n) {
        need(n);
        p += n;
    }
    std::string name() {
        const uint32_t n = u32();
        need(n);
        std::string s(b.begin() + p, b.begin() + p + n);
        p += pad4(n);
        return s;
    }
    void attrs() {
        const uint32_t tag = u32(), n = u32();
        if (tag == 0) return;
        if (tag != kAttTag) throw std::runtime_error("NetCDF_Reader || bad attribute list.");
        for (uint32_t i = 0; i < n; i++) {
            name();
            const Type t = (Type)u32();
            const uint32_t m = u32();
            skip(pad4((int64_t)m * typeSize(t)));
        }
    }
};
}  // namespace

int File::dim(int64_t n) {
    const std::string nm = "ncdim_" + std::to_string(n);
    for (size_t i = 0; i < dims.size(); i++)
        if (dims[i].first == nm) return (int)i;
    dims.push_back({nm, n});
    return (int)dims.size() - 1;
}

void write(const std::string &path, const File &f) {
    // a zero-length dimension is the record (unlimited) dimension, as
    // nc_def_dim(..., 0) makes it; with no records written, a variable on it
    // holds no data and its size entry is that of one record
    auto slab = [&](const Var &v) {
        int64_t n = 1;
        for (size_t k = 0; k < v.dims.size(); k++) {
            const int64_t len = f.dims[v.dims[k]].second;
            if (len == 0 && k != 0) throw std::runtime_error("NetCDF_Writer || the record dimension must come first.");
            if (len != 0) n *= len;
        }
        return n;
    };
    // the header with begin offsets of 0, to learn its size; then again
    auto header = [&](const std::vector<uint64_t> &begin) {
        Out o;
        o.b = {'C', 'D', 'F', 2};
        o.u32(0);  // no record dimension
        if (f.dims.empty()) {
            o.u32(0);
            o.u32(0);
        } else {
            o.u32(kDimTag);
            o.u32((uint32_t)f.dims.size());
            for (const auto &d : f.dims) {
                o.name(d.first);
                o.u32((uint32_t)d.second);
            }
        }
        o.attrs(f.attrs);
        if (f.vars.empty()) {
            o.u32(0);
            o.u32(0);
        } else {
            o.u32(kVarTag);
            o.u32((uint32_t)f.vars.size());
            for (size_t i = 0; i < f.vars.size(); i++) {
                const Var &v = f.vars[i];
                o.name(v.name);
                o.u32((uint32_t)v.dims.size());
                for (int d : v.dims) o.u32((uint32_t)d);
                o.attrs(v.attrs);
                o.u32((uint32_t)v.type);
                o.u32((uint32_t)pad4(slab(v) * typeSize(v.type)));
                o.u64(begin[i]);
            }
        }
        return o;
    };
    std::vector<uint64_t> begin(f.vars.size(), 0);
    const size_t hsize = header(begin).b.size();
    uint64_t at = hsize;
    for (size_t i = 0; i < f.vars.size(); i++) {
        begin[i] = at;
        int64_t n = 1;
        for (int d : f.vars[i].dims) n *= f.dims[d].second;
        if ((int64_t)f.vars[i].data.size() != n) throw std::runtime_error("NetCDF_Writer || data size mismatch.");
        slab(f.vars[i]);
        at += pad4(n * typeSize(f.vars[i].type));
    }
    Out o = header(begin);
    for (const Var &v : f.vars) {
        for (double x : v.data) o.value(v.type, x);
        o.pad();
    }
    std::FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error("NetCDF_Writer::open || Error creating NetCDF file: || " + path);
    const bool ok = std::fwrite(o.b.data(), 1, o.b.size(), fp) == o.b.size();
    if (std::fclose(fp) != 0 || !ok) throw std::runtime_error("NetCDF_Writer || Error writing file: || " + path);
}

bool readVar(const std::string &path, const std::string &want, std::vector<int64_t> &shape,
             std::vector<double> &values) {
    std::ifstream fs(path, std::ios::binary);
    if (!fs) throw std::runtime_error("NetCDF_Reader::open || Error opening NetCDF file: || " + path);
    std::vector<unsigned char> b((std::istreambuf_iterator<char>(fs)), std::istreambuf_iterator<char>());
    if (b.size() < 4 || b[0] != 'C' || b[1] != 'D' || b[2] != 'F' || (b[3] != 1 && b[3] != 2))
        throw std::runtime_error("NetCDF_Reader::open || not a NetCDF classic file (NetCDF-4 files need the "
                                 "NetCDF library): || " + path);
    const bool off64 = b[3] == 2;
    In in(b);
    in.skip(4);
    const int64_t numrecs = in.u32();
    std::vector<int64_t> dlen;
    {
        const uint32_t tag = in.u32(), n = in.u32();
        if (tag != 0 && tag != kDimTag) throw std::runtime_error("NetCDF_Reader || bad dimension list.");
        for (uint32_t i = 0; i < n && tag; i++) {
            in.name();
            dlen.push_back(in.u32());
        }
    }
    in.attrs();
    const uint32_t tag = in.u32(), nv = in.u32();
    if (tag != 0 && tag != kVarTag) throw std::runtime_error("NetCDF_Reader || bad variable list.");
    struct Info {
        std::string name;
        std::vector<int64_t> shape;
        bool rec = false;  // first dimension is the record dimension
        Type type = DOUBLE;
        int64_t vsize = 0;
        uint64_t begin = 0;
    };
    std::vector<Info> vars;
    for (uint32_t i = 0; i < nv && tag; i++) {
        Info v;
        v.name = in.name();
        const uint32_t nd = in.u32();
        for (uint32_t d = 0; d < nd; d++) {
            const uint32_t id = in.u32();
            if (id >= dlen.size()) throw std::runtime_error("NetCDF_Reader || bad dimension id.");
            if (dlen[id] == 0 && d == 0) v.rec = true;
            v.shape.push_back(dlen[id] == 0 ? numrecs : dlen[id]);
        }
        in.attrs();
        v.type = (Type)in.u32();
        v.vsize = in.u32();
        v.begin = off64 ? in.u64() : in.u32();
        vars.push_back(v);
    }
    // record variables are interleaved record by record; one record variable
    // alone is not padded
    int64_t recsize = 0, nrecvar = 0;
    for (const Info &v : vars)
        if (v.rec) {
            recsize += v.vsize;
            nrecvar++;
        }
    for (const Info &v : vars) {
        if (v.name != want) continue;
        int64_t n = 1;
        for (int64_t d : v.shape) n *= d;
        const int64_t per = v.rec ? (v.shape[0] ? n / v.shape[0] : 0) : n;
        if (v.rec && nrecvar == 1) recsize = per * typeSize(v.type);
        values.resize(n);
        for (int64_t k = 0; k < n; k++) {
            In data(b);
            const int64_t r = v.rec ? k / per : 0, j = v.rec ? k % per : k;
            data.p = v.begin + r * recsize + j * typeSize(v.type);
            if (v.type == DOUBLE) {
                const uint64_t u = data.u64();
                std::memcpy(&values[k], &u, 8);
            } else if (v.type == FLOAT) {
                const uint32_t u = data.u32();
                float f;
                std::memcpy(&f, &u, 4);
                values[k] = f;
            } else if (v.type == INT) {
                values[k] = (double)(int32_t)data.u32();
            } else {
                throw std::runtime_error("NetCDF_Reader || unsupported variable type.");
            }
        }
        shape = v.shape;
        return true;
    }
    return false;
}

}  // namespace netcdf
}  // namespace axg
