// elem3d.hpp -- stiffness of elements whose material varies in azimuth (the
// reference's "3D" path with FieldFFT).  Placeholder until the FFT kernels land.
#pragma once
#include "common.hpp"

namespace axg {
struct Bucket3D {
    int colour = 0, phase = 0, material = 0, att = 0;
    bool tiso = false;
    int32_t nelem = 0;
};
}  // namespace axg
