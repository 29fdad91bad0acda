// nu_wisdom.hpp -- the Nu wisdom of the reference's NU_TYPE wisdom path:
// rows (s, z, learned Nu, original Nu) of every GLL point (Domain::dumpWisdom,
// SOLVER/src/core/domain/Domain.cpp:404-438), their file (NuWisdom::writeToFile
// / readFromFile, SOLVER/src/preloop/nrfield/wisdom/NuWisdom.cpp:17-79) in the
// reference's "NetCDF ascii format" alternative (NetCDF_ReaderAscii.cpp), and
// the Nu of a new point from them (NuWisdom::getNu, NuWisdom.cpp:82-98).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace axg {

struct NuWisdomIndex {
    // rows [n][4]: s, z, learned Nu, original Nu
    explicit NuWisdomIndex(std::vector<double> rows);
    int64_t size() const { return (int64_t)nu_.size(); }
    // inverse-distance mean of the learned Nu of the k nearest rows, rounded; a
    // row closer than 1e-10 m gives its own Nu (NuWisdom::getNu)
    int getNu(double s, double z, int k) const;

   private:
    std::vector<double> s_, z_;
    std::vector<int> nu_;
    // uniform grid of buckets over the bounding box of the rows
    double lo_s_ = 0., lo_z_ = 0., h_ = 1.;
    int ns_ = 1, nz_ = 1;
    std::vector<int64_t> start_;
    std::vector<int64_t> item_;
};

// the ascii wisdom file: variable axisem3d_wisdom [n][4]
void writeWisdomAscii(const std::string &path, const double *rows, int64_t n);
// the same variable in a NetCDF classic file (NuWisdom.cpp:19-44)
void writeWisdomNetCDF(const std::string &path, const double *rows, int64_t n);
// rows [n][4] of a wisdom file, ascii or NetCDF classic; a 3-column variable
// takes its third column as the original Nu too (NuWisdom.cpp:60-71)
std::vector<double> readWisdomAscii(const std::string &path);

}  // namespace axg
