// exodus.hpp -- the reference's Exodus mesh model (ExodusModel,
// SOLVER/src/preloop/exodus/ExodusModel.cpp) and the per-quad quantities the
// reference's Quad derives from it (Quad::Quad, Quad.cpp:37-110, 527-547;
// Material::Material, Material.cpp:30-90; the spherical, linear and
// semi-spherical mappings of SOLVER/src/preloop/spectral/mapping).
//
// The mesh files the reference ships are NetCDF-4, i.e. HDF5 files; they are
// read through the HDF5 C library, loaded at run time (dlopen), so nothing
// else in the library depends on it.
#pragma once
#include <stdint.h>

#include <array>
#include <map>
#include <string>
#include <vector>

namespace axg {
namespace exodus {

struct Model {
    // raw data (ExodusModel::readRawData, ExodusModel.cpp:52-268)
    std::vector<std::string> globalNames;
    std::vector<double> globalValues;
    std::vector<std::string> recordsRaw;
    std::vector<std::array<int, 4>> conn;  // 0-based node tags per quad
    std::vector<double> ns, nz;            // nodal s and z
    double distTol = 0.;
    std::vector<std::string> ssNames;
    std::vector<std::vector<int>> ssValues;  // [set][quad]: side, -1 if none
    std::vector<double> ellipKnots, ellipCoeffs;
    std::vector<std::string> elemNamesAll;
    std::vector<std::string> axisNames;            // depth-dependent variables
    std::vector<double> axisCoords;                // z of the axis profile rows
    std::vector<std::vector<double>> axisValues;   // [variable][row]
    std::vector<double> elemType, elemDt;          // element-wise variables
    // structured (formStructured, ExodusModel.cpp:296-355)
    std::map<std::string, double> globals;
    std::map<std::string, std::string> records;
    std::string ssAxis, ssSurface;
    // auxiliary (formAuxiliary, ExodusModel.cpp:357-430)
    std::vector<double> gllSpacing;             // per node
    std::vector<std::array<int, 4>> vicinal;    // per quad

    int numQuads() const { return (int)conn.size(); }
    int numNodes() const { return (int)ns.size(); }
    bool cartesian() const;
    bool hasAttenuation() const { return globals.count("nr_lin_solids") > 0; }
    double rOuter() const { return globals.at("radius"); }
    int sideSet(const std::string &name, int quad) const;  // -1 if absent or off the set
    double elemental(const std::string &name, int quad) const;
};

// ExodusModel::initialize: read, structure, auxiliary
Model load(const std::string &path);

// per-quad geometry of the reference's Quad
struct QuadGeometry {
    int type;          // 0 spherical, 1 linear, 2 semi-spherical (element_type)
    int curvedOuter;   // -1 for linear
    int axial;         // axis side (3 after the rotation of formAuxiliary) or -1
    int fluid;
    double crd[25][2];    // (s, z) of the GLL points, ipol * 5 + jpol
    double ifact[25];     // Quad::formIntegralFactor
    double xieta[25][2];  // reference coordinates (GLJ in xi on axial quads)
};
QuadGeometry quadGeometry(const Model &m, int quad);
// Connectivity::formElemToGLL (SOLVER/src/preloop/graph/Connectivity.cpp:42-93):
// global GLL point numbers [quad][ipol * 5 + jpol], new points numbered in the
// order quads and then (ipol, jpol) are scanned, a point of a lower-numbered
// quad sharing a node or an edge keeping its number
std::vector<std::array<int, 25>> elemToGLL(const Model &m, int &ngll);
// Mapping::interpolate of NAME_0 .. NAME_3 at the GLL points (Material.cpp:30-90)
void nodalMaterial(const Model &m, int quad, const std::string &name, double out[25]);

}  // namespace exodus
}  // namespace axg

struct axg_exodus {
    axg::exodus::Model m;
};
