// netcdf_classic.hpp -- the NetCDF classic file format (64-bit offset variant,
// "CDF\2"), written and read without the NetCDF library: the reference writes
// its station and wisdom files through NetCDF_Writer (NetCDF_Writer.h) and
// reads them through nc_open, which takes this format as well as NetCDF-4.
// Layout (Unidata's published classic-format specification): magic, record
// count, dimension list, global attributes, variable list (name, dimension
// ids, attributes, type, size, begin offset), then every variable's data,
// big-endian, each padded to 4 bytes.  Only fixed-size variables are used.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace axg {
namespace netcdf {

enum Type { CHAR = 2, INT = 4, FLOAT = 5, DOUBLE = 6 };

struct Attr {
    std::string name;
    Type type = DOUBLE;
    std::vector<double> values;  // DOUBLE / FLOAT / INT
    std::string text;            // CHAR
};

struct Var {
    std::string name;
    std::vector<int> dims;  // dimension ids
    std::vector<Attr> attrs;
    Type type = DOUBLE;
    std::vector<double> data;  // row-major, converted to `type` on writing
};

struct File {
    std::vector<std::pair<std::string, int64_t>> dims;
    std::vector<Attr> attrs;
    std::vector<Var> vars;
    // dimension named "ncdim_<n>" of length n, as NetCDF_Writer::defineVariable
    // names them (NetCDF_Writer.h:33-47), created on first use
    int dim(int64_t n);
};

void write(const std::string &path, const File &f);
// a variable's dimensions and values (any numeric type, as double)
bool readVar(const std::string &path, const std::string &name, std::vector<int64_t> &shape,
             std::vector<double> &values);

}  // namespace netcdf
}  // namespace axg
