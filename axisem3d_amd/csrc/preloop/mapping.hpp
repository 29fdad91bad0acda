// mapping.hpp -- the reference's geometric mappings of a quad from the
// reference square (xi, eta) to the meridian (s, z), with their Jacobians:
// spherical (SphericalMapping.cpp:9-78), linear (LinearMapping.cpp:8-43) and
// semi-spherical (SemiSphericalMapping.cpp:9-85), each curved side turned to
// side 2 by the orthogonal Q2 of Mapping.cpp:55-65.
#pragma once
#include <cmath>

namespace axg {
namespace mapping {

constexpr double kPi = 3.141592653589793238463;

inline int period0123(int p) { return ((p % 4) + 4) % 4; }

// Mapping::sOrthogQ2 (Mapping.cpp:60-65)
inline const double kQ2[4][2][2] = {{{-1., 0.}, {0., -1.}},
                             {{0., -1.}, {1., 0.}},
                             {{1., 0.}, {0., 1.}},
                             {{0., 1.}, {-1., 0.}}};

// XMath::makeClose (XMath.cpp:9-16)
inline void makeClose(double &a, double &b) {
    if (a - b > kPi) b += 2. * kPi;
    if (b - a > kPi) a += 2. * kPi;
}

struct QuadGeom {
    double ns[4], nz[4];  // nodal (s, z)
    int curvedOuter;
    int type = 0;         // 0 spherical, 1 linear, 2 semi-spherical
};

struct Rot {
    double r0, r3, t0, t1, t2, t3, xi, eta;
    const double (*Q)[2];
};

inline Rot rotate(const QuadGeom &q, double xi, double eta) {
    const double(*Q)[2] = kQ2[q.curvedOuter];
    double rr[4], tt[4];
    for (int i = 0; i < 4; i++) {
        double s2 = Q[0][0] * q.ns[i] + Q[0][1] * q.nz[i];
        double z2 = Q[1][0] * q.ns[i] + Q[1][1] * q.nz[i];
        rr[i] = std::sqrt(s2 * s2 + z2 * z2);
        tt[i] = std::atan2(s2, z2);
    }
    int co = q.curvedOuter;
    Rot r;
    r.Q = Q;
    r.r0 = rr[period0123(co - 2)];
    r.r3 = rr[period0123(co + 1)];
    r.t0 = tt[period0123(co - 2)];
    r.t1 = tt[period0123(co - 1)];
    r.t2 = tt[period0123(co - 0)];
    r.t3 = tt[period0123(co + 1)];
    r.xi = Q[0][0] * xi + Q[0][1] * eta;
    r.eta = Q[1][0] * xi + Q[1][1] * eta;
    makeClose(r.t2, r.t3);
    makeClose(r.t0, r.t1);
    return r;
}

// SphericalMapping::mapping (SphericalMapping.cpp:9-38)
inline void sphMapping(const QuadGeom &q, double xi, double eta, double &s, double &z) {
    Rot R = rotate(q, xi, eta);
    double a3 = ((1. - R.xi) * R.t3 + (1. + R.xi) * R.t2) / 2.;
    double a0 = ((1. - R.xi) * R.t0 + (1. + R.xi) * R.t1) / 2.;
    double s2 = (1. + R.eta) * R.r3 / 2. * std::sin(a3) + (1. - R.eta) * R.r0 / 2. * std::sin(a0);
    double z2 = (1. + R.eta) * R.r3 / 2. * std::cos(a3) + (1. - R.eta) * R.r0 / 2. * std::cos(a0);
    // Q2^T * sz2
    s = R.Q[0][0] * s2 + R.Q[1][0] * z2;
    z = R.Q[0][1] * s2 + R.Q[1][1] * z2;
}

// SphericalMapping::jacobian (SphericalMapping.cpp:40-78): J = Q2^T J2 Q2
inline void sphJacobian(const QuadGeom &q, double xi, double eta, double J[2][2]) {
    Rot R = rotate(q, xi, eta);
    double a3 = ((1. - R.xi) * R.t3 + (1. + R.xi) * R.t2) / 2.;
    double a0 = ((1. - R.xi) * R.t0 + (1. + R.xi) * R.t1) / 2.;
    double J2[2][2];
    J2[0][0] = (1. + R.eta) * R.r3 * (R.t2 - R.t3) / 4. * std::cos(a3) +
               (1. - R.eta) * R.r0 * (R.t1 - R.t0) / 4. * std::cos(a0);
    J2[0][1] = .5 * (R.r3 * std::sin(a3) - R.r0 * std::sin(a0));
    J2[1][0] = -(1. + R.eta) * R.r3 * (R.t2 - R.t3) / 4. * std::sin(a3) -
               (1. - R.eta) * R.r0 * (R.t1 - R.t0) / 4. * std::sin(a0);
    J2[1][1] = .5 * (R.r3 * std::cos(a3) - R.r0 * std::cos(a0));
    double T[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) T[i][j] = J2[i][0] * R.Q[0][j] + J2[i][1] * R.Q[1][j];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) J[i][j] = R.Q[0][i] * T[0][j] + R.Q[1][i] * T[1][j];
}

inline double det2(const double J[2][2]) { return J[0][0] * J[1][1] - J[0][1] * J[1][0]; }

// LinearMapping (LinearMapping.cpp:8-43): bilinear in the four nodes
inline void linShape(double xi, double eta, double N[4], double dN[2][4]) {
    const double xp = 1. + xi, xm = 1. - xi, ep = 1. + eta, em = 1. - eta;
    N[0] = xm * em / 4.;
    N[1] = xp * em / 4.;
    N[2] = xp * ep / 4.;
    N[3] = xm * ep / 4.;
    dN[0][0] = -em / 4.;
    dN[0][1] = em / 4.;
    dN[0][2] = ep / 4.;
    dN[0][3] = -ep / 4.;
    dN[1][0] = -xm / 4.;
    dN[1][1] = -xp / 4.;
    dN[1][2] = xp / 4.;
    dN[1][3] = xm / 4.;
}

inline void linMapping(const QuadGeom &q, double xi, double eta, double &s, double &z) {
    double N[4], dN[2][4];
    linShape(xi, eta, N, dN);
    s = z = 0.;
    for (int i = 0; i < 4; i++) {
        s += q.ns[i] * N[i];
        z += q.nz[i] * N[i];
    }
}

inline void linJacobian(const QuadGeom &q, double xi, double eta, double J[2][2]) {
    double N[4], dN[2][4];
    linShape(xi, eta, N, dN);
    for (int c = 0; c < 2; c++) {
        J[0][c] = J[1][c] = 0.;
        for (int i = 0; i < 4; i++) {
            J[0][c] += q.ns[i] * dN[c][i];
            J[1][c] += q.nz[i] * dN[c][i];
        }
    }
}

// SemiSphericalMapping (SemiSphericalMapping.cpp:9-85): the curved side (2
// after the rotation) an arc of radius r3, the opposite side straight
struct SemiRot {
    double s0, z0, s1, z1, t2, r3, t3, xi, eta;
    const double (*Q)[2];
};

inline SemiRot semiRotate(const QuadGeom &q, double xi, double eta) {
    const double(*Q)[2] = kQ2[q.curvedOuter];
    double s2[4], z2[4];
    for (int i = 0; i < 4; i++) {
        s2[i] = Q[0][0] * q.ns[i] + Q[0][1] * q.nz[i];
        z2[i] = Q[1][0] * q.ns[i] + Q[1][1] * q.nz[i];
    }
    const int co = q.curvedOuter;
    SemiRot r;
    r.Q = Q;
    r.s0 = s2[period0123(co - 2)];
    r.z0 = z2[period0123(co - 2)];
    r.s1 = s2[period0123(co - 1)];
    r.z1 = z2[period0123(co - 1)];
    const int i2 = period0123(co), i3 = period0123(co + 1);
    r.t2 = std::atan2(s2[i2], z2[i2]);
    r.r3 = std::sqrt(s2[i3] * s2[i3] + z2[i3] * z2[i3]);
    r.t3 = std::atan2(s2[i3], z2[i3]);
    r.xi = Q[0][0] * xi + Q[0][1] * eta;
    r.eta = Q[1][0] * xi + Q[1][1] * eta;
    makeClose(r.t2, r.t3);
    return r;
}

inline void semiMapping(const QuadGeom &q, double xi, double eta, double &s, double &z) {
    const SemiRot R = semiRotate(q, xi, eta);
    const double a = ((1. - R.xi) * R.t3 + (1. + R.xi) * R.t2) / 2.;
    const double s2 = (1. + R.eta) * R.r3 / 2. * std::sin(a) + (1. - R.eta) / 2. * ((1. - R.xi) * R.s0 + (1. + R.xi) * R.s1) / 2.;
    const double z2 = (1. + R.eta) * R.r3 / 2. * std::cos(a) + (1. - R.eta) / 2. * ((1. - R.xi) * R.z0 + (1. + R.xi) * R.z1) / 2.;
    s = R.Q[0][0] * s2 + R.Q[1][0] * z2;
    z = R.Q[0][1] * s2 + R.Q[1][1] * z2;
}

inline void semiJacobian(const QuadGeom &q, double xi, double eta, double J[2][2]) {
    const SemiRot R = semiRotate(q, xi, eta);
    const double a = ((1. - R.xi) * R.t3 + (1. + R.xi) * R.t2) / 2.;
    double J2[2][2];
    J2[0][0] = (1. - R.eta) * (R.s1 - R.s0) / 4. + (1. + R.eta) * R.r3 * (R.t2 - R.t3) * std::cos(a) / 4.;
    J2[0][1] = -((1. - R.xi) * R.s0 + (1. + R.xi) * R.s1) / 4. + R.r3 / 2. * std::sin(a);
    J2[1][0] = (1. - R.eta) * (R.z1 - R.z0) / 4. - (1. + R.eta) * R.r3 * (R.t2 - R.t3) * std::sin(a) / 4.;
    J2[1][1] = -((1. - R.xi) * R.z0 + (1. + R.xi) * R.z1) / 4. + R.r3 / 2. * std::cos(a);
    double T[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) T[i][j] = J2[i][0] * R.Q[0][j] + J2[i][1] * R.Q[1][j];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) J[i][j] = R.Q[0][i] * T[0][j] + R.Q[1][i] * T[1][j];
}

// by element_type: 0 spherical, 1 linear, 2 semi-spherical
inline void mapAny(int type, const QuadGeom &q, double xi, double eta, double &s, double &z) {
    if (type == 0) sphMapping(q, xi, eta, s, z);
    else if (type == 1) linMapping(q, xi, eta, s, z);
    else semiMapping(q, xi, eta, s, z);
}

inline void jacobianAny(int type, const QuadGeom &q, double xi, double eta, double J[2][2]) {
    if (type == 0) sphJacobian(q, xi, eta, J);
    else if (type == 1) linJacobian(q, xi, eta, J);
    else semiJacobian(q, xi, eta, J);
}

}  // namespace mapping
}  // namespace axg
