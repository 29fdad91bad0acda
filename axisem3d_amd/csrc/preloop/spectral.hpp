// spectral.hpp -- GLL / GLJ(0,1) nodes, weights and derivative matrices for
// nPol = 4, generated from the polynomials (not from a table).
//
// The reference tabulates these to 12 digits in
// SOLVER/src/preloop/spectral/SpectralConstants.cpp:43-49 and stores the
// derivative matrix row-major (eigenp.h:29 RDMatPP), i.e. G[i][j] = l_i'(xi_j).
// The same convention is used here.  GLL interior nodes are the roots of
// P_{N-1}^{(1,1)}, GLJ(0,1) interior nodes the roots of P_{N-1}^{(1,2)}; weights
// are the Lobatto-type quadrature weights of the weight functions 1 and (1+xi)
// (the axial integral factor divides by (1+xi), Quad.cpp:535-541).
#pragma once
#include <array>
#include <cmath>
#include <vector>

namespace axg {
namespace spectral {

constexpr int nPol = 4;
constexpr int nP = nPol + 1;

struct Basis {
    std::array<double, nP> x;      // nodes
    std::array<double, nP> w;      // weights
    std::array<std::array<double, nP>, nP> G;  // G[i][j] = l_i'(x_j)
};

// Jacobi polynomial P_n^{(a,b)}(x) by the three-term recurrence
inline double jacobiP(int n, double a, double b, double x) {
    if (n == 0) return 1.;
    double p0 = 1.;
    double p1 = 0.5 * (a - b + (a + b + 2.) * x);
    for (int k = 2; k <= n; k++) {
        double c = 2. * k + a + b;
        double a1 = 2. * k * (k + a + b) * (c - 2.);
        double a2 = (c - 1.) * (a * a - b * b);
        double a3 = (c - 2.) * (c - 1.) * c;
        double a4 = 2. * (k + a - 1.) * (k + b - 1.) * c;
        double p2 = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// roots of P_n^{(a,b)} on (-1,1): bracket on a fine grid, then bisection
inline std::vector<double> jacobiRoots(int n, double a, double b) {
    std::vector<double> roots;
    const int ngrid = 20000;
    double xl = -1., fl = jacobiP(n, a, b, xl);
    for (int i = 1; i <= ngrid; i++) {
        double xr = -1. + 2. * i / ngrid;
        double fr = jacobiP(n, a, b, xr);
        if (fl == 0.) {
            roots.push_back(xl);
        } else if (fl * fr < 0.) {
            double lo = xl, hi = xr, flo = fl;
            for (int it = 0; it < 200; it++) {
                double mid = 0.5 * (lo + hi);
                double fm = jacobiP(n, a, b, mid);
                if ((fm < 0.) == (flo < 0.)) {
                    lo = mid;
                    flo = fm;
                } else {
                    hi = mid;
                }
            }
            roots.push_back(0.5 * (lo + hi));
        }
        xl = xr;
        fl = fr;
    }
    return roots;
}

// Lagrange basis derivative matrix on the nodes: G[i][j] = l_i'(x_j)
inline void derivativeMatrix(const std::array<double, nP> &x,
                             std::array<std::array<double, nP>, nP> &G) {
    std::array<double, nP> lam;
    for (int k = 0; k < nP; k++) {
        double p = 1.;
        for (int m = 0; m < nP; m++) {
            if (m != k) p *= (x[k] - x[m]);
        }
        lam[k] = 1. / p;
    }
    for (int i = 0; i < nP; i++) {
        for (int j = 0; j < nP; j++) {
            if (i != j) {
                G[i][j] = (lam[i] / lam[j]) / (x[j] - x[i]);
            } else {
                double s = 0.;
                for (int m = 0; m < nP; m++) {
                    if (m != i) s += 1. / (x[i] - x[m]);
                }
                G[i][i] = s;
            }
        }
    }
}

// quadrature weights from the moment equations sum_i w_i x_i^k = int x^k wf(x)
// for k = 0..nPol; wf = 1 (jacobiB = 0) or 1 + x (jacobiB = 1)
inline void momentWeights(const std::array<double, nP> &x, int jacobiB,
                          std::array<double, nP> &w) {
    double A[nP][nP + 1];
    for (int k = 0; k < nP; k++) {
        for (int i = 0; i < nP; i++) A[k][i] = std::pow(x[i], k);
        // int_{-1}^{1} x^k dx and int_{-1}^{1} x^{k+1} dx
        double m0 = (k % 2 == 0) ? 2. / (k + 1) : 0.;
        double m1 = ((k + 1) % 2 == 0) ? 2. / (k + 2) : 0.;
        A[k][nP] = (jacobiB == 1) ? m0 + m1 : m0;
    }
    // Gaussian elimination with partial pivoting
    for (int c = 0; c < nP; c++) {
        int piv = c;
        for (int r = c + 1; r < nP; r++) {
            if (std::abs(A[r][c]) > std::abs(A[piv][c])) piv = r;
        }
        for (int k = 0; k <= nP; k++) std::swap(A[c][k], A[piv][k]);
        for (int r = 0; r < nP; r++) {
            if (r == c) continue;
            double f = A[r][c] / A[c][c];
            for (int k = c; k <= nP; k++) A[r][k] -= f * A[c][k];
        }
    }
    for (int i = 0; i < nP; i++) w[i] = A[i][nP] / A[i][i];
}

inline Basis makeBasis(bool glj) {
    Basis b;
    std::vector<double> r = glj ? jacobiRoots(nPol - 1, 1., 2.) : jacobiRoots(nPol - 1, 1., 1.);
    b.x[0] = -1.;
    for (int i = 0; i < nPol - 1; i++) b.x[i + 1] = r[i];
    b.x[nPol] = 1.;
    if (!glj) b.x[nPol / 2] = 0.;  // symmetric middle node exactly zero
    momentWeights(b.x, glj ? 1 : 0, b.w);
    derivativeMatrix(b.x, b.G);
    return b;
}

inline const Basis &gll() {
    static const Basis b = makeBasis(false);
    return b;
}
inline const Basis &glj() {
    static const Basis b = makeBasis(true);
    return b;
}

// XMath::interpLagrange (SOLVER/src/preloop/utilities/XMath.cpp:32-47)
inline void interpLagrange(double target, const std::array<double, nP> &bases,
                           std::array<double, nP> &out) {
    for (int d = 0; d < nP; d++) {
        double p1 = 1., p2 = 1., x0 = bases[d];
        for (int i = 0; i < nP; i++) {
            if (i != d) {
                p1 *= target - bases[i];
                p2 *= x0 - bases[i];
            }
        }
        out[d] = p1 / p2;
    }
}

}  // namespace spectral
}  // namespace axg
