// svl_oracle.cpp -- FP64 CPU restatement of SeismoVLAB's dynamic-FEM hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: the tests, the smoke check and
// the CPU-baseline leg of bench.py call it to compare against the HIP library. The
// product (libsvlg.so) never links, loads or calls it.
//
// It restates, in plain C++17 without Eigen, the per-step algorithm of SVL
// (citations relative to the SVL source tree, 02-Run_Process = RP):
//   * lin2DQuad4 / lin3DHexa8: Jacobian, shape functions, strain-displacement matrix,
//     consistent mass with row-sum lumping, stiffness, Rayleigh damping, internal
//     force, strain update (RP/04-Elements/06-Quadrilateral/lin2DQuad4.cpp:87-107,
//     243-412, 618-740; RP/04-Elements/10-Hexahedron/lin3DHexa8.cpp:87-107, 243-412,
//     722-883);
//   * ZeroLength1D + Viscous1DLinear dashpots (RP/04-Elements/01-Zero/ZeroLength1D.cpp
//     :190-253, 321-382; RP/02-Materials/01-Linear/Viscous1DLinear.cpp:91-154);
//   * Elastic2DPlaneStrain / Elastic3DLinear (Elastic2DPlaneStrain.cpp:5-27, 96-157;
//     Elastic3DLinear.cpp:5-29, 97-158);
//   * Gauss 2x2 / 2x2x2 tables with SVL's 15-digit abscissa
//     (RP/04-Elements/11-Integration/GaussQuadrature.cpp:303-316, 721-742);
//   * Assembler: mass/damping/stiffness triplets with the tolerance filters and
//     duplicate summation in element-tag order, internal force with the ftol filter,
//     point loads by assignment then factor sum (RP/07-Assembler/Assembler.cpp:48-158,
//     240-269, 291-350, 485, 623-697);
//   * Total-to-free operator T (RP/06-Mesh/Mesh.cpp:328-383);
//   * CentralDifference (RP/10-Integrators/02-CentralDifference/CentralDifference.cpp
//     :36-71, 124-152, 206-231);
//   * NewmarkBeta / ExtendedNewmarkBeta with G = 0 (RP/10-Integrators/03-Newmark/
//     ExtendedNewmarkBeta.cpp:123-154, 208-245; NewmarkBeta.cpp:115-144, 198-235);
//   * Linear algorithm and state scatter (RP/09-Algorithms/01-Linear/Linear.cpp:22-56,
//     RP/09-Algorithms/Algorithm.cpp:18-56);
//   * Domain reduction load: masked element M, C, K times the node motion rows, exterior
//     signals negated on input (RP/04-Elements/06-Quadrilateral/lin2DQuad4.cpp:563-614,
//     RP/04-Elements/10-Hexahedron/lin3DHexa8.cpp:660-718, RP/07-Assembler/Assembler.cpp
//     :460-478, RP/12-Utilities/Driver.hpp:1669-1720); rows past the signal read as zero;
//   * DynamicAnalysis loop and index convention (RP/08-Analysis/02-Dynamic/
//     DynamicAnalysis.cpp:25-64): step k uses load sample k, rows written at k-1.
//   * Support motion (POINTLOAD_SUPPORT_MOTION): Node::GetSupportMotion (RP/01-Node/Node.cpp
//     :228-249), Assembler::ComputeSupportMotionIncrement (RP/07-Assembler/Assembler.cpp
//     :493-532), the integrators' ComputeSupportMotionVector (Feff -= T^T K SM) and the SM
//     term of ComputeNewStep (CentralDifference.cpp:137, 190-203; ExtendedNewmarkBeta.cpp
//     :134-135, 191-204; NewmarkBeta.cpp:132, 181-195 with sample k-1), with the strains
//     updated from U + T dU before SM is added (Linear.cpp:22-56, Algorithm.cpp:18-56).
//   * EigenSolver with update OFF (RP/11-Solvers/01-Direct/EigenSolver.cpp:19-59):
//     restated as a symmetric skyline LDL^T factorised once; with a diagonal matrix
//     the solve is x_i = b_i / D_ii.
//
// Parity status of each part is recorded in the repository's design notes and in the
// tests that pin it (closed forms, invariants and the OpenSees fixtures SVL ships).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../include/svlg.h"

namespace {

thread_local std::string g_err;

// GaussQuadrature.cpp:312-315 / 734-741: SVL's 15-digit abscissa, not 1/sqrt(3).
constexpr double GP = 0.577350269189626;

// Quad: (-,-),(+,-),(-,+),(+,+); Hexa: x fastest, then y, then z.
void gauss_points(int dim, std::vector<double> &w, std::vector<double> &xi) {
    if (dim == 2) {
        w.assign(4, 1.0);
        xi = {-GP, -GP, GP, -GP, -GP, GP, GP, GP};
    } else {
        w.assign(8, 1.0);
        xi.clear();
        for (int k = 0; k < 2; ++k)
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i) {
                    xi.push_back(i ? GP : -GP);
                    xi.push_back(j ? GP : -GP);
                    xi.push_back(k ? GP : -GP);
                }
    }
}

// Small dense helpers (row-major).
struct Mat {
    int r = 0, c = 0;
    std::vector<double> a;
    Mat() = default;
    Mat(int r_, int c_) : r(r_), c(c_), a(size_t(r_) * c_, 0.0) {}
    double &operator()(int i, int j) { return a[size_t(i) * c + j]; }
    double operator()(int i, int j) const { return a[size_t(i) * c + j]; }
};

// Inverse and determinant by LU with partial pivoting (the path Eigen takes for a
// dynamic-size MatrixXd::inverse()/determinant()).
void lu_inverse(const Mat &A, Mat &Ainv, double &det) {
    int n = A.r;
    Mat LU = A;
    std::vector<int> piv(n);
    det = 1.0;
    for (int i = 0; i < n; ++i) piv[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double mx = std::fabs(LU(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(LU(i, k)) > mx) { mx = std::fabs(LU(i, k)); p = i; }
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(LU(k, j), LU(p, j));
            std::swap(piv[k], piv[p]);
            det = -det;
        }
        det *= LU(k, k);
        for (int i = k + 1; i < n; ++i) {
            LU(i, k) /= LU(k, k);
            for (int j = k + 1; j < n; ++j) LU(i, j) -= LU(i, k) * LU(k, j);
        }
    }
    Ainv = Mat(n, n);
    for (int col = 0; col < n; ++col) {
        std::vector<double> b(n, 0.0);
        for (int i = 0; i < n; ++i) b[i] = (piv[i] == col) ? 1.0 : 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) b[i] -= LU(i, j) * b[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) b[i] -= LU(i, j) * b[j];
            b[i] /= LU(i, i);
        }
        for (int i = 0; i < n; ++i) Ainv(i, col) = b[i];
    }
}

// ----------------------------------------------------------------- elements
struct ElemGeom {
    int dim;        // 2 or 3
    int nen;        // 4 or 8
    const double *X; // nen*dim coordinates
};

// lin2DQuad4::ComputeJacobianMatrix (lin2DQuad4.cpp:648-667).
Mat jac_quad(const double *X, double ri, double si) {
    auto x = [&](int a, int d) { return X[a * 2 + d]; };
    Mat J(2, 2);
    for (int d = 0; d < 2; ++d) {
        J(0, d) = -1.0 / 4.0 * (1.0 - si) * x(0, d) + 1.0 / 4.0 * (1.0 - si) * x(1, d) +
                  1.0 / 4.0 * (1.0 + si) * x(2, d) - 1.0 / 4.0 * (1.0 + si) * x(3, d);
        J(1, d) = -1.0 / 4.0 * (1.0 - ri) * x(0, d) - 1.0 / 4.0 * (1.0 + ri) * x(1, d) +
                  1.0 / 4.0 * (1.0 + ri) * x(2, d) + 1.0 / 4.0 * (1.0 - ri) * x(3, d);
    }
    return J;
}

// lin2DQuad4::ComputeStrainDisplacementMatrix (lin2DQuad4.cpp:688-710): 3x8.
Mat bmat_quad(double ri, double si, const Mat &Jij) {
    Mat J; double det;
    lu_inverse(Jij, J, det);
    double dn[4][2];
    // dN/dx (row 0 of J^-1) and dN/dy (row 1).
    const double dr[4] = {-(1.0 - si), (1.0 - si), (1.0 + si), -(1.0 + si)};
    const double ds[4] = {-(1.0 - ri), -(1.0 + ri), (1.0 + ri), (1.0 - ri)};
    for (int a = 0; a < 4; ++a) {
        dn[a][0] = 1.0 / 4.0 * J(0, 1) * ds[a] + 1.0 / 4.0 * J(0, 0) * dr[a];
        dn[a][1] = 1.0 / 4.0 * J(1, 1) * ds[a] + 1.0 / 4.0 * J(1, 0) * dr[a];
    }
    Mat B(3, 8);
    for (int a = 0; a < 4; ++a) {
        B(0, 2 * a) = dn[a][0];
        B(1, 2 * a + 1) = dn[a][1];
        B(2, 2 * a) = dn[a][1];
        B(2, 2 * a + 1) = dn[a][0];
    }
    return B;
}

// Shape function matrix (lin2DQuad4.cpp:670-684): 2x8.
Mat hmat_quad(double ri, double si) {
    double h[4] = {1.0 / 4.0 * (1.0 - ri) * (1.0 - si), 1.0 / 4.0 * (1.0 + ri) * (1.0 - si),
                   1.0 / 4.0 * (1.0 + ri) * (1.0 + si), 1.0 / 4.0 * (1.0 - ri) * (1.0 + si)};
    Mat H(2, 8);
    for (int a = 0; a < 4; ++a) { H(0, 2 * a) = h[a]; H(1, 2 * a + 1) = h[a]; }
    return H;
}

// Reference derivatives of the trilinear shape functions (1/8 factor included).
void hexa_dref(double ri, double si, double ti, double d[8][3]) {
    const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const int sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const int sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int a = 0; a < 8; ++a) {
        d[a][0] = 1.0 / 8.0 * sx[a] * (1.0 + sy[a] * si) * (1.0 + sz[a] * ti);
        d[a][1] = 1.0 / 8.0 * sy[a] * (1.0 + sx[a] * ri) * (1.0 + sz[a] * ti);
        d[a][2] = 1.0 / 8.0 * sz[a] * (1.0 + sx[a] * ri) * (1.0 + sy[a] * si);
    }
}

// lin3DHexa8::ComputeJacobianMatrix (lin3DHexa8.cpp:756-785): J(i,j) = d x_j / d xi_i.
Mat jac_hexa(const double *X, double ri, double si, double ti) {
    double d[8][3];
    hexa_dref(ri, si, ti, d);
    Mat J(3, 3);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int a = 0; a < 8; ++a) s += d[a][i] * X[a * 3 + j];
            J(i, j) = s;
        }
    return J;
}

// lin3DHexa8::ComputeStrainDisplacementMatrix (lin3DHexa8.cpp:811-853): 6x24 with
// Voigt order [xx, yy, zz, xy, yz, zx] and engineering shears.
Mat bmat_hexa(double ri, double si, double ti, const Mat &Jij) {
    Mat J; double det;
    lu_inverse(Jij, J, det);
    double d[8][3];
    hexa_dref(ri, si, ti, d);
    double dn[8][3];
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i)
            dn[a][i] = J(i, 2) * d[a][2] + J(i, 1) * d[a][1] + J(i, 0) * d[a][0];
    Mat B(6, 24);
    for (int a = 0; a < 8; ++a) {
        B(0, 3 * a) = dn[a][0];
        B(1, 3 * a + 1) = dn[a][1];
        B(2, 3 * a + 2) = dn[a][2];
        B(3, 3 * a) = dn[a][1]; B(3, 3 * a + 1) = dn[a][0];
        B(4, 3 * a + 1) = dn[a][2]; B(4, 3 * a + 2) = dn[a][1];
        B(5, 3 * a) = dn[a][2]; B(5, 3 * a + 2) = dn[a][0];
    }
    return B;
}

// lin3DHexa8::ComputeShapeFunctionMatrix (lin3DHexa8.cpp:788-807): 3x24.
Mat hmat_hexa(double ri, double si, double ti) {
    const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const int sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const int sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    Mat H(3, 24);
    for (int a = 0; a < 8; ++a) {
        double h = 1.0 / 8.0 * (1.0 + sx[a] * ri) * (1.0 + sy[a] * si) * (1.0 + sz[a] * ti);
        for (int c = 0; c < 3; ++c) H(c, 3 * a + c) = h;
    }
    return H;
}

// Isotropic D (Elastic2DPlaneStrain.cpp:18-26 / Elastic3DLinear.cpp:19-28).
Mat elastic_D(int dim, double E, double nu) {
    double c1 = E * (1.0 - nu) / (1.0 - 2.0 * nu) / (1.0 + nu);
    double c2 = E * nu / (1.0 - 2.0 * nu) / (1.0 + nu);
    double c3 = E / (2.0 * (1.0 + nu));
    if (dim == 2) {
        Mat D(3, 3);
        D(0, 0) = c1; D(0, 1) = c2; D(1, 0) = c2; D(1, 1) = c1; D(2, 2) = c3;
        return D;
    }
    Mat D(6, 6);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D(i, j) = (i == j) ? c1 : c2;
    for (int i = 3; i < 6; ++i) D(i, i) = c3;
    return D;
}

struct GaussCache {
    std::vector<double> w, xi;
};

// ---------------------------------------------------------------- J2 plasticity
// Plastic3DJ2 / PlasticPlaneStrainJ2 state of one Gauss point (Plastic3DJ2.cpp:5-52). The
// 6-vectors are tensor components [11, 22, 33, 12, 23, 31] (shears not doubled).
struct J2State {
    double eps[6] = {0, 0, 0, 0, 0, 0};
    double sig[6] = {0, 0, 0, 0, 0, 0};
    double ep[6] = {0, 0, 0, 0, 0, 0};   // PlasticStrain
    double q[6] = {0, 0, 0, 0, 0, 0};    // BackStress
    double alpha = 0.0;
    double C[36];                         // TangentStiffness, row-major
};

struct J2Params {
    double K, G, sy, H, beta;
};

// K D + 2G (I - D/3): D has ones in the normal 3x3 block, I = diag(1,1,1,1/2,1/2,1/2)
// (ComputeIdentityTensor / ComputeIdentityOperator, Plastic3DJ2.cpp:282-312).
void j2_elastic_tangent(const J2Params &p, double *C) {
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            const double Dij = (i < 3 && j < 3) ? 1.0 : 0.0;
            const double Iij = (i == j) ? (i < 3 ? 1.0 : 0.5) : 0.0;
            C[i * 6 + j] = p.K * Dij + 2.0 * p.G * (Iij - Dij / 3.0);
        }
}

// ComputeTensorNorm (Plastic3DJ2.cpp:321-329): shears counted twice.
double j2_norm(const double *t) {
    return std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + 2.0 * (t[3] * t[3] + t[4] * t[4] + t[5] * t[5]));
}

// Plastic3DJ2::UpdateState with cond = 1 (Plastic3DJ2.cpp:206-258); PlasticPlaneStrainJ2 is
// the same on its embedded strain (PlasticPlaneStrainJ2.cpp:227-279). `et` is the tensor
// strain. alpha, BackStress and PlasticStrain are advanced from their current values, not
// from the committed ones, so repeated calls within one step accumulate (as in SVL).
void j2_update(J2State &s, const J2Params &p, const double *et) {
    for (int i = 0; i < 6; ++i) s.eps[i] = et[i];
    const double tr = et[0] + et[1] + et[2];
    const double one[6] = {1, 1, 1, 0, 0, 0};
    double strial[6], xi[6];
    for (int i = 0; i < 6; ++i) {
        const double dev = et[i] - 1.0 / 3.0 * tr * one[i];
        strial[i] = 2.0 * p.G * (dev - s.ep[i]);
        xi[i] = strial[i] - s.q[i];
    }
    const double nx = j2_norm(xi);
    const double F = nx - std::sqrt(2.0 / 3.0) * (p.sy + s.alpha * p.beta * p.H);
    double Id[36];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            const double Dij = (i < 3 && j < 3) ? 1.0 : 0.0;
            const double Iij = (i == j) ? (i < 3 ? 1.0 : 0.5) : 0.0;
            Id[i * 6 + j] = Iij - 1.0 / 3.0 * Dij;
        }
    if (F <= 0) {
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                const double Dij = (i < 3 && j < 3) ? 1.0 : 0.0;
                s.C[i * 6 + j] = p.K * Dij + 2.0 * p.G * Id[i * 6 + j];
            }
        for (int i = 0; i < 6; ++i) s.sig[i] = p.K * tr * one[i] + strial[i];
        return;
    }
    double n[6];
    for (int i = 0; i < 6; ++i) n[i] = xi[i] / nx;
    const double dg = F / (2.0 * p.G + 2.0 / 3.0 * p.H);
    s.alpha += std::sqrt(2.0 / 3.0) * dg;
    for (int i = 0; i < 6; ++i) {
        s.q[i] += 2.0 / 3.0 * (1.0 - p.beta) * p.H * dg * n[i];
        s.ep[i] += dg * n[i];
        s.sig[i] = p.K * tr * one[i] + strial[i] - 2.0 * p.G * dg * n[i];
    }
    const double c1 = 1.0 / (1.0 + p.H / p.G / 3.0);
    const double c2 = 4.0 * p.G * p.G * dg / nx;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            const double Dij = (i < 3 && j < 3) ? 1.0 : 0.0;
            const double nn = n[i] * n[j];
            s.C[i * 6 + j] = p.K * Dij + 2.0 * p.G * (Id[i * 6 + j] - c1 * nn) - c2 * (Id[i * 6 + j] - nn);
        }
}

// Engineering strain of the element (B u) -> tensor strain of the material:
// 3D [e_xx, e_yy, e_zz, g_xy/2, g_yz/2, g_zx/2] (Plastic3DJ2.cpp:217); plane strain
// embeds [0, e_xx, e_yy, 0, g_xy/2, 0] (PlasticPlaneStrainJ2.cpp:236) and reads back the
// components (1, 2, 4) (:115-158).
static const int kPsIdx[3] = {1, 2, 4};

// One continuum element of the restated mesh.
struct Continuum {
    int nodes[8];
    double E, nu, rho, am, ak;
    int mat = 0;                  // SVLG_MAT_ELASTIC or SVLG_MAT_J2
    J2Params j2p{0, 0, 0, 0, 0};
    std::vector<J2State> j2;      // [ngp]
    std::vector<double> strain;  // [ngp * nvoigt] material Strain (trial)
    std::vector<double> strain_n; // committed newStrain
};

struct Dashpot {
    int node, partner, dir;
    double eta;
};

struct Model {
    int dim = 2, nen = 4, nvoigt = 3, ngp = 4;
    double th = 1.0;
    bool lumped = true;
    int nmesh = 0, ntotnodes = 0;
    std::vector<double> coords; // [ntotnodes*dim]
    std::vector<int> freedof;   // [ntotnodes*dim]
    int ntotal = 0, nfree = 0;
    std::vector<Continuum> elems;
    std::vector<Dashpot> dash;
    Mat D0; // placeholder
    // loads
    int nload = 0, load_nt = 0;
    std::vector<int> load_node;
    std::vector<double> load_dir, load_factor, load_series;
    // Domain reduction load: element list, per mesh node its slot in the node signals
    // (-1: no DRM node) and exterior flag, signals [slot][nt][3*dim] with the exterior
    // ones already negated (Driver.hpp:1669-1720).
    std::vector<int> drm_elem, drm_slot, drm_ext;
    int drm_nt = 0;
    std::vector<double> drm_sig;
    double drm_factor = 1.0;
    // Support motion entries (load order): total DOF, load factor, displacement series.
    std::vector<int> sm_dof;
    std::vector<double> sm_fac;
    std::vector<std::vector<double>> sm_vals;
    GaussCache gc;

    void element_X(const Continuum &e, double *X) const {
        for (int a = 0; a < nen; ++a)
            for (int d = 0; d < dim; ++d) X[a * dim + d] = coords[size_t(e.nodes[a]) * dim + d];
    }
    Mat jac(const double *X, int g) const {
        const double *p = &gc.xi[size_t(g) * dim];
        return dim == 2 ? jac_quad(X, p[0], p[1]) : jac_hexa(X, p[0], p[1], p[2]);
    }
    Mat bmat(int g, const Mat &J) const {
        const double *p = &gc.xi[size_t(g) * dim];
        return dim == 2 ? bmat_quad(p[0], p[1], J) : bmat_hexa(p[0], p[1], p[2], J);
    }
    Mat hmat(int g) const {
        const double *p = &gc.xi[size_t(g) * dim];
        return dim == 2 ? hmat_quad(p[0], p[1]) : hmat_hexa(p[0], p[1], p[2]);
    }
    double thick() const { return dim == 2 ? th : 1.0; }
};

// lin*::ComputeMassMatrix (lin2DQuad4.cpp:243-285 / lin3DHexa8.cpp:243-285).
Mat element_mass(const Model &m, const Continuum &e) {
    int nd = m.nen * m.dim;
    Mat M(nd, nd);
    double X[24];
    m.element_X(e, X);
    for (int g = 0; g < m.ngp; ++g) {
        Mat J = m.jac(X, g);
        Mat Ji; double det;
        lu_inverse(J, Ji, det);
        Mat H = m.hmat(g);
        double s = m.gc.w[g] * e.rho * m.thick() * std::fabs(det);
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                double hh = 0.0;
                for (int k = 0; k < m.dim; ++k) hh += H(k, i) * H(k, j);
                M(i, j) += s * hh;
            }
    }
    if (m.lumped) {
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j)
                if (i != j) { M(i, i) += M(i, j); M(i, j) = 0.0; }
    }
    return M;
}

// Material matrix of Gauss point g in the element's Voigt order: the elastic D, or for J2
// the initial tangent K D + 2G (I - D/3) (GetInitialTangentStiffness, Plastic3DJ2.cpp:146-159)
// or the current consistent tangent (GetTangentStiffness), restricted to (1, 2, 4) in 2D.
Mat material_D(const Model &m, const Continuum &e, int g, bool tangent) {
    if (e.mat != SVLG_MAT_J2) return elastic_D(m.dim, e.E, e.nu);
    double C6[36];
    if (tangent)
        std::memcpy(C6, e.j2[g].C, sizeof(C6));
    else
        j2_elastic_tangent(e.j2p, C6);
    if (m.dim == 3) {
        Mat D(6, 6);
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) D(i, j) = C6[i * 6 + j];
        return D;
    }
    Mat D(3, 3);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D(i, j) = C6[kPsIdx[i] * 6 + kPsIdx[j]];
    return D;
}

// lin*::ComputeInitialStiffnessMatrix (tangent = false) / ComputeStiffnessMatrix (tangent =
// true; lin3DHexa8.cpp:289-318): sum w t |J| B^T C B.
Mat element_stiffness(const Model &m, const Continuum &e, bool tangent = false) {
    int nd = m.nen * m.dim, nv = m.nvoigt;
    Mat K(nd, nd);
    double X[24];
    m.element_X(e, X);
    for (int g = 0; g < m.ngp; ++g) {
        const Mat C = material_D(m, e, g, tangent);
        Mat J = m.jac(X, g);
        Mat Ji; double det;
        lu_inverse(J, Ji, det);
        Mat B = m.bmat(g, J);
        double s = m.gc.w[g] * m.thick() * std::fabs(det);
        Mat CB(nv, nd);
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < nd; ++j) {
                double t = 0.0;
                for (int k = 0; k < nv; ++k) t += C(i, k) * B(k, j);
                CB(i, j) = t;
            }
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                double t = 0.0;
                for (int k = 0; k < nv; ++k) t += B(k, i) * CB(k, j);
                K(i, j) += s * t;
            }
    }
    return K;
}

// lin*::ComputeDampingMatrix (lin2DQuad4.cpp:322-372): Rayleigh am*M + ak*K_initial.
Mat element_damping(const Model &m, const Continuum &e) {
    int nd = m.nen * m.dim;
    Mat Cm(nd, nd);
    if (e.am != 0.0 || e.ak != 0.0) {
        Mat M = element_mass(m, e);
        Mat K = element_stiffness(m, e);
        for (size_t i = 0; i < Cm.a.size(); ++i) Cm.a[i] += e.am * M.a[i] + e.ak * K.a[i];
    }
    return Cm;
}

// lin*::ComputeInternalForces (lin2DQuad4.cpp:383-412): sum w t |J| B^T sigma, sigma
// from the stored material strain (Elastic*::GetStress = D * Strain).
void element_fint(const Model &m, const Continuum &e, double *f) {
    int nd = m.nen * m.dim, nv = m.nvoigt;
    for (int i = 0; i < nd; ++i) f[i] = 0.0;
    double X[24];
    m.element_X(e, X);
    Mat C = elastic_D(m.dim, e.E, e.nu);
    for (int g = 0; g < m.ngp; ++g) {
        Mat J = m.jac(X, g);
        Mat Ji; double det;
        lu_inverse(J, Ji, det);
        Mat B = m.bmat(g, J);
        double sig[6];
        if (e.mat == SVLG_MAT_J2) {
            // Plastic*J2::GetStress: the stress of the last UpdateState.
            for (int i = 0; i < nv; ++i) sig[i] = m.dim == 3 ? e.j2[g].sig[i] : e.j2[g].sig[kPsIdx[i]];
        } else {
            const double *eps = &e.strain[size_t(g) * nv];
            for (int i = 0; i < nv; ++i) {
                double t = 0.0;
                for (int k = 0; k < nv; ++k) t += C(i, k) * eps[k];
                sig[i] = t;
            }
        }
        double s = m.gc.w[g] * m.thick() * std::fabs(det);
        for (int i = 0; i < nd; ++i) {
            double t = 0.0;
            for (int k = 0; k < nv; ++k) t += B(k, i) * sig[k];
            f[i] += s * t;
        }
    }
}

// lin*::UpdateState + ComputeStrain (lin2DQuad4.cpp:87-107, 618-633): strain = B (U+dU).
void element_update(const Model &m, Continuum &e, const std::vector<double> &Ut) {
    int nd = m.nen * m.dim, nv = m.nvoigt;
    double X[24], u[24];
    m.element_X(e, X);
    for (int a = 0; a < m.nen; ++a)
        for (int d = 0; d < m.dim; ++d) u[a * m.dim + d] = Ut[size_t(e.nodes[a]) * m.dim + d];
    for (int g = 0; g < m.ngp; ++g) {
        Mat J = m.jac(X, g);
        Mat B = m.bmat(g, J);
        for (int i = 0; i < nv; ++i) {
            double t = 0.0;
            for (int k = 0; k < nd; ++k) t += B(i, k) * u[k];
            e.strain[size_t(g) * nv + i] = t;
        }
        if (e.mat == SVLG_MAT_J2) {
            const double *es = &e.strain[size_t(g) * nv];
            double et[6] = {0, 0, 0, 0, 0, 0};
            if (m.dim == 3) {
                et[0] = es[0]; et[1] = es[1]; et[2] = es[2];
                et[3] = 0.5 * es[3]; et[4] = 0.5 * es[4]; et[5] = 0.5 * es[5];
            } else {
                et[1] = es[0]; et[2] = es[1]; et[4] = 0.5 * es[2];
            }
            j2_update(e.j2[g], e.j2p, et);
        }
    }
}

// ------------------------------------------------------------- sparse (total DOFs)
// Triplets summed in insertion order (what Eigen's setFromTriplets does), kept per
// column so that matrix-vector products visit entries column by column.
struct Sparse {
    int n = 0;
    std::unordered_map<int64_t, double> acc;
    std::vector<int64_t> order;
    void add(int i, int j, double v) {
        int64_t key = int64_t(j) * n + i;
        auto it = acc.find(key);
        if (it == acc.end()) { acc.emplace(key, v); order.push_back(key); }
        else it->second += v;
    }
    // Column-sorted (j, i) entries.
    std::vector<std::pair<int64_t, double>> sorted() const {
        std::vector<std::pair<int64_t, double>> v(acc.begin(), acc.end());
        std::sort(v.begin(), v.end(), [](auto &a, auto &b) { return a.first < b.first; });
        return v;
    }
};

struct CSC {
    int n = 0;
    std::vector<int> colptr, row;
    std::vector<double> val;
    static CSC from(const Sparse &s) {
        CSC c;
        c.n = s.n;
        auto e = s.sorted();
        c.colptr.assign(s.n + 1, 0);
        for (auto &p : e) c.colptr[p.first / s.n + 1]++;
        for (int j = 0; j < s.n; ++j) c.colptr[j + 1] += c.colptr[j];
        c.row.resize(e.size());
        c.val.resize(e.size());
        for (size_t k = 0; k < e.size(); ++k) {
            c.row[k] = int(e[k].first % s.n);
            c.val[k] = e[k].second;
        }
        return c;
    }
    // y = A x
    void mul(const std::vector<double> &x, std::vector<double> &y) const {
        y.assign(n, 0.0);
        for (int j = 0; j < n; ++j) {
            double xj = x[j];
            for (int k = colptr[j]; k < colptr[j + 1]; ++k) y[row[k]] += val[k] * xj;
        }
    }
};

// Combination a*A + b*B entrywise (same sparsity union).
CSC combine(double a, const CSC &A, double b, const CSC &B) {
    Sparse s;
    s.n = A.n;
    for (int j = 0; j < A.n; ++j) {
        for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k) s.add(A.row[k], j, a * A.val[k]);
    }
    // Entries present in both: a*A + b*B evaluated per entry.
    for (int j = 0; j < B.n; ++j)
        for (int k = B.colptr[j]; k < B.colptr[j + 1]; ++k) s.add(B.row[k], j, b * B.val[k]);
    return CSC::from(s);
}

// Symmetric skyline LDL^T over the free DOFs (stand-in for Eigen SimplicialLDLT,
// which is exact to roundoff; the parity of this boundary is pinned by residuals).
struct Skyline {
    int n = 0;
    bool diagonal = true;
    std::vector<int> first;          // first row index of column j
    std::vector<size_t> start;       // offset of column j in a
    std::vector<double> a;           // columns: rows first[j]..j
    bool factor(const CSC &Kf) {
        n = Kf.n;
        first.assign(n, 0);
        for (int j = 0; j < n; ++j) first[j] = j;
        diagonal = true;
        for (int j = 0; j < n; ++j)
            for (int k = Kf.colptr[j]; k < Kf.colptr[j + 1]; ++k) {
                int i = Kf.row[k];
                if (i != j && Kf.val[k] != 0.0) diagonal = false;
                int lo = std::min(i, j), hi = std::max(i, j);
                first[hi] = std::min(first[hi], lo);
            }
        start.assign(n + 1, 0);
        for (int j = 0; j < n; ++j) start[j + 1] = start[j] + size_t(j - first[j] + 1);
        a.assign(start[n], 0.0);
        for (int j = 0; j < n; ++j)
            for (int k = Kf.colptr[j]; k < Kf.colptr[j + 1]; ++k) {
                int i = Kf.row[k];
                if (i <= j) a[start[j] + (i - first[j])] = Kf.val[k];
            }
        if (diagonal) {
            for (int j = 0; j < n; ++j)
                if (a[start[j] + (j - first[j])] == 0.0) return false;
            return true;
        }
        // Column-oriented LDL^T (Crout) on the envelope.
        for (int j = 0; j < n; ++j) {
            double *cj = &a[start[j]];
            int fj = first[j];
            for (int i = fj; i < j; ++i) {
                const double *ci = &a[start[i]];
                int fi = first[i];
                int lo = std::max(fi, fj);
                double s = cj[i - fj];
                for (int k = lo; k < i; ++k) s -= ci[k - fi] * cj[k - fj];
                cj[i - fj] = s;
            }
            double d = cj[j - fj];
            for (int i = fj; i < j; ++i) {
                double dii = a[start[i] + (i - first[i])];
                double lij = cj[i - fj] / dii;
                d -= lij * cj[i - fj];
                cj[i - fj] = lij;
            }
            if (d == 0.0 || !std::isfinite(d)) return false;
            cj[j - fj] = d;
        }
        return true;
    }
    void solve(const std::vector<double> &b, std::vector<double> &x) const {
        x = b;
        if (diagonal) {
            for (int j = 0; j < n; ++j) x[j] = b[j] / a[start[j] + (j - first[j])];
            return;
        }
        for (int j = 0; j < n; ++j) {
            const double *cj = &a[start[j]];
            double s = x[j];
            for (int k = first[j]; k < j; ++k) s -= cj[k - first[j]] * x[k];
            x[j] = s;
        }
        for (int j = 0; j < n; ++j) x[j] /= a[start[j] + (j - first[j])];
        for (int j = n - 1; j >= 0; --j) {
            const double *cj = &a[start[j]];
            double xj = x[j];
            for (int k = first[j]; k < j; ++k) x[k] -= cj[k - first[j]] * xj;
        }
    }
};

// ------------------------------------------------------------------- build
Model build(const svlg_mesh_desc *d) {
    Model m;
    m.dim = d->ndim;
    if (m.dim != 2 && m.dim != 3) throw std::runtime_error("ndim must be 2 or 3");
    if ((m.dim == 2) != (d->elem_type == SVLG_ELEM_QUAD4))
        throw std::runtime_error("element type does not match ndim");
    m.nen = m.dim == 2 ? 4 : 8;
    m.nvoigt = m.dim == 2 ? 3 : 6;
    m.ngp = m.dim == 2 ? 4 : 8;
    m.th = d->thickness;
    m.lumped = d->lumped != 0;
    gauss_points(m.dim, m.gc.w, m.gc.xi);
    m.nmesh = d->nnodes;
    // Lysmer copies: one restrained node per distinct dashpot node, first appearance.
    std::vector<int> partner(d->nnodes, -1);
    int extra = 0;
    for (int i = 0; i < d->ndash; ++i) {
        int n = d->dash_node[i];
        if (n < 0 || n >= d->nnodes) throw std::runtime_error("dashpot node out of range");
        if (partner[n] < 0) partner[n] = d->nnodes + extra++;
    }
    m.ntotnodes = d->nnodes + extra;
    m.coords.assign(size_t(m.ntotnodes) * m.dim, 0.0);
    std::memcpy(m.coords.data(), d->coords, sizeof(double) * size_t(d->nnodes) * m.dim);
    for (int n = 0; n < d->nnodes; ++n)
        if (partner[n] >= 0)
            for (int c = 0; c < m.dim; ++c)
                m.coords[size_t(partner[n]) * m.dim + c] = d->coords[size_t(n) * m.dim + c];
    m.ntotal = m.ntotnodes * m.dim;
    m.freedof.assign(m.ntotal, -1);
    for (int i = 0; i < d->nnodes * m.dim; ++i) m.freedof[i] = d->free_dof[i];
    m.nfree = d->nfree;
    m.elems.resize(d->nelem);
    for (int e = 0; e < d->nelem; ++e) {
        Continuum &c = m.elems[e];
        for (int a = 0; a < m.nen; ++a) c.nodes[a] = d->conn[size_t(e) * m.nen + a];
        c.mat = d->mat_type ? d->mat_type[e] : SVLG_MAT_ELASTIC;
        if (c.mat == SVLG_MAT_J2) {
            if (!d->j2) throw std::runtime_error("oracle: J2 element without J2 parameters");
            const double *q = d->j2 + size_t(e) * 5;
            c.j2p = {q[0], q[1], q[2], q[3], q[4]};
            J2State s0;
            j2_elastic_tangent(c.j2p, s0.C);
            c.j2.assign(m.ngp, s0);
        } else if (c.mat != SVLG_MAT_ELASTIC) {
            throw std::runtime_error("oracle: unknown material type");
        }
        c.E = d->E[e]; c.nu = d->nu[e]; c.rho = d->rho[e];
        c.am = d->ray_am ? d->ray_am[e] : 0.0;
        c.ak = d->ray_ak ? d->ray_ak[e] : 0.0;
        c.strain.assign(size_t(m.ngp) * m.nvoigt, 0.0);
        c.strain_n = c.strain;
    }
    for (int i = 0; i < d->ndash; ++i)
        m.dash.push_back({d->dash_node[i], partner[d->dash_node[i]], d->dash_dir[i], d->dash_eta[i]});
    m.nload = d->nload;
    m.load_nt = d->load_nt;
    m.load_node.assign(d->load_node, d->load_node + d->nload);
    m.load_dir.assign(d->load_dir, d->load_dir + size_t(d->nload) * m.dim);
    m.load_factor.assign(d->load_factor, d->load_factor + d->nload);
    m.load_series.assign(d->load_series, d->load_series + size_t(d->nload) * d->load_nt);
    if (d->ndrm_elem > 0) {
        const int nf = 3 * m.dim;
        m.drm_factor = d->drm_factor;
        m.drm_nt = d->drm_nt;
        m.drm_elem.assign(d->drm_elem, d->drm_elem + d->ndrm_elem);
        m.drm_slot.assign(d->nnodes, -1);
        m.drm_ext.assign(d->nnodes, 0);
        m.drm_sig.assign(size_t(d->ndrm_node) * d->drm_nt * nf, 0.0);
        for (int i = 0; i < d->ndrm_node; ++i) {
            const int n = d->drm_node[i];
            if (n < 0 || n >= d->nnodes || m.drm_slot[n] >= 0) throw std::runtime_error("bad DRM node list");
            m.drm_slot[n] = i;
            m.drm_ext[n] = d->drm_cond[i] != 0;
            const double sg = d->drm_cond[i] ? -1.0 : 1.0;
            for (size_t q = 0; q < size_t(d->drm_nt) * nf; ++q)
                m.drm_sig[size_t(i) * d->drm_nt * nf + q] = sg * d->drm_signal[size_t(i) * d->drm_nt * nf + q];
        }
        for (int e : m.drm_elem) {
            if (e < 0 || e >= d->nelem) throw std::runtime_error("DRM element out of range");
            for (int a = 0; a < m.nen; ++a)
                if (m.drm_slot[m.elems[e].nodes[a]] < 0) throw std::runtime_error("DRM element node without a signal");
        }
    }
    for (int j = 0; j < std::max(d->nsm, 0); ++j) {
        const int n = d->sm_node[j], c = d->sm_dir[j];
        if (n < 0 || n >= d->nnodes || c < 0 || c >= m.dim) throw std::runtime_error("bad support motion entry");
        if (d->sm_ptr[j + 1] <= d->sm_ptr[j]) throw std::runtime_error("support motion without values");
        m.sm_dof.push_back(n * m.dim + c);
        m.sm_fac.push_back(d->sm_factor[j]);
        m.sm_vals.emplace_back(d->sm_values + d->sm_ptr[j], d->sm_values + d->sm_ptr[j + 1]);
    }
    return m;
}

// ------------------------------------------------------------------- assembler
struct Assembler {
    const Model *m;
    double mtol = 1e-12, ktol = 1e-12, ftol = 1e-12;

    void elem_dofs(const Continuum &e, int *dofs) const {
        for (int a = 0; a < m->nen; ++a)
            for (int c = 0; c < m->dim; ++c) dofs[a * m->dim + c] = e.nodes[a] * m->dim + c;
    }
    // Assembler::ComputeMassMatrix (Assembler.cpp:48-67, 661-697); no nodal masses.
    CSC mass() const {
        Sparse s; s.n = m->ntotal;
        int dofs[24];
        for (auto &e : m->elems) {
            elem_dofs(e, dofs);
            Mat Me = element_mass(*m, e);
            int nd = m->nen * m->dim;
            for (int j = 0; j < nd; ++j)
                for (int i = 0; i < nd; ++i)
                    if (std::fabs(Me(i, j)) > mtol) s.add(dofs[i], dofs[j], Me(i, j));
        }
        // ZeroLength1D::ComputeMassMatrix is zero (ZeroLength1D.cpp:180-188).
        return CSC::from(s);
    }
    // Assembler::ComputeDampingMatrix (Assembler.cpp:117-158); element order: continuum
    // elements, then dashpots (their tags follow the mesh elements).
    CSC damping() const {
        Sparse s; s.n = m->ntotal;
        int dofs[24];
        int nd = m->nen * m->dim;
        for (auto &e : m->elems) {
            if (e.am == 0.0 && e.ak == 0.0) continue;
            elem_dofs(e, dofs);
            Mat Ce = element_damping(*m, e);
            for (int j = 0; j < nd; ++j)
                for (int i = 0; i < nd; ++i)
                    if (std::fabs(Ce(i, j)) > ktol) s.add(dofs[i], dofs[j], Ce(i, j));
        }
        for (auto &dp : m->dash) {
            // C_e = eta a a^T with a = [-e_dir, +e_dir] (ZeroLength1D.cpp:211-234, 321-352).
            int di = dp.node * m->dim + dp.dir, dj = dp.partner * m->dim + dp.dir;
            int dd[2] = {di, dj};
            double av[2] = {-1.0, 1.0};
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i) {
                    double v = dp.eta * av[i] * av[j];
                    if (std::fabs(v) > ktol) s.add(dd[i], dd[j], v);
                }
        }
        return CSC::from(s);
    }
    // Assembler::ComputeStiffnessMatrix (Assembler.cpp:71-113); dashpots add k = 0. With
    // tangent = true the elements use the current material tangent (Newton), else the
    // initial one.
    CSC stiffness(bool tangent = false) const {
        Sparse s; s.n = m->ntotal;
        int dofs[24];
        int nd = m->nen * m->dim;
        for (auto &e : m->elems) {
            elem_dofs(e, dofs);
            Mat Ke = element_stiffness(*m, e, tangent);
            for (int j = 0; j < nd; ++j)
                for (int i = 0; i < nd; ++i)
                    if (std::fabs(Ke(i, j)) > ktol) s.add(dofs[i], dofs[j], Ke(i, j));
        }
        return CSC::from(s);
    }
    // Assembler::ComputeInternalForceVector (Assembler.cpp:240-269).
    void fint(std::vector<double> &F) const {
        F.assign(m->ntotal, 0.0);
        int dofs[24];
        double f[24];
        int nd = m->nen * m->dim;
        for (auto &e : m->elems) {
            elem_dofs(e, dofs);
            element_fint(*m, e, f);
            for (int j = 0; j < nd; ++j)
                if (std::fabs(f[j]) > ftol) F[dofs[j]] += f[j];
        }
        // Dashpots: Viscous1DLinear::GetStress() is 0, so every entry is filtered out.
    }
    // Assembler::ComputeDynamicInternalForceVector (Assembler.cpp:273-287, 596-619) for the
    // reactions: every element with a fixed node (Element::HasFixedNode, Element.cpp:53-64)
    // adds Fint_e + C_e V_e + M_e A_e (lin3DHexa8.cpp:416-437, lin2DQuad4 alike), unfiltered;
    // a dashpot adds C_e V_e (ZeroLength1D.cpp:257-277, its Fint and mass are zero).
    void fdyn(const std::vector<double> &V, const std::vector<double> &A, std::vector<double> &F) const {
        F.assign(m->ntotal, 0.0);
        auto fixed = [&](int node) {
            for (int c = 0; c < m->dim; ++c)
                if (m->freedof[size_t(node) * m->dim + c] < 0) return true;
            return false;
        };
        int dofs[24];
        double f[24];
        const int nd = m->nen * m->dim;
        for (auto &e : m->elems) {
            bool any = false;
            for (int a = 0; a < m->nen; ++a) any = any || fixed(e.nodes[a]);
            if (!any) continue;
            elem_dofs(e, dofs);
            element_fint(*m, e, f);
            Mat Ce = element_damping(*m, e), Me = element_mass(*m, e);
            for (int i = 0; i < nd; ++i) {
                double cv = 0.0, ma = 0.0;
                for (int j = 0; j < nd; ++j) {
                    cv += Ce(i, j) * V[dofs[j]];
                    ma += Me(i, j) * A[dofs[j]];
                }
                F[dofs[i]] += (f[i] + cv) + ma;
            }
        }
        for (auto &dp : m->dash) {
            const int di = dp.node * m->dim + dp.dir, dj = dp.partner * m->dim + dp.dir;
            const double w = -V[di] + V[dj];  // a^T V with a = [-e_dir, +e_dir]
            F[di] += dp.eta * (-1.0) * w;
            F[dj] += dp.eta * w;
        }
    }
    // Assembler::ComputeSupportMotionIncrement (Assembler.cpp:493-532): per entry (load order)
    // factor * (g(j) - g(j-1)), g(0) alone at j = 0; g(j) = Node::GetSupportMotion (Node.cpp
    // :228-249): values[j] below the series length, values[0] past it.
    void support_increment(int j, std::vector<double> &SM) const {
        SM.assign(m->ntotal, 0.0);
        for (size_t e = 0; e < m->sm_dof.size(); ++e) {
            const std::vector<double> &g = m->sm_vals[e];
            auto val = [&](int i) { return i < int(g.size()) ? g[i] : g[0]; };
            double dg = val(j);
            if (j > 0) dg -= val(j - 1);
            SM[m->sm_dof[e]] += m->sm_fac[e] * dg;
        }
    }
    // Assembler::ComputeExternalForceVector (Assembler.cpp:291-350, 485).
    void fext(int k, std::vector<double> &F) const {
        F.assign(m->ntotal, 0.0);
        std::vector<double> tmp(m->ntotal);
        for (int l = 0; l < m->nload; ++l) {
            std::fill(tmp.begin(), tmp.end(), 0.0);
            double amp = (k < m->load_nt) ? m->load_series[size_t(l) * m->load_nt + k] : 0.0;
            int n = m->load_node[l];
            for (int c = 0; c < m->dim; ++c) tmp[size_t(n) * m->dim + c] = amp * m->load_dir[size_t(l) * m->dim + c];
            for (int i = 0; i < m->ntotal; ++i) F[i] += m->load_factor[l] * tmp[i];
        }
        if (!m->drm_elem.empty()) {
            // ELEMENTLOAD_DOMAIN_REDUCTION (Assembler.cpp:460-478): per listed element
            // lin*::ComputeDomainReductionForces (lin2DQuad4.cpp:564-614, lin3DHexa8.cpp:
            // 660-718): M, C, K with every entry (i, j) of equal exterior flags zeroed, times
            // the node motion rows k (U, V, A); summed into the load's vector, then factored.
            std::fill(tmp.begin(), tmp.end(), 0.0);
            const int nd = m->nen * m->dim, nf = 3 * m->dim;
            for (int eid : m->drm_elem) {
                const Continuum &e = m->elems[eid];
                Mat Me = element_mass(*m, e), Ce = element_damping(*m, e), Ke = element_stiffness(*m, e, true);
                double Uo[24], Vo[24], Ao[24];
                int cond[24], dofs[24];
                elem_dofs(e, dofs);
                for (int a = 0; a < m->nen; ++a) {
                    const int n = e.nodes[a], sl = m->drm_slot[n];
                    for (int c = 0; c < m->dim; ++c) {
                        double u = 0.0, v = 0.0, ac = 0.0;
                        if (k < m->drm_nt) {
                            const double *row = &m->drm_sig[(size_t(sl) * m->drm_nt + k) * nf];
                            u = row[c]; v = row[m->dim + c]; ac = row[2 * m->dim + c];
                        }
                        Uo[a * m->dim + c] = u; Vo[a * m->dim + c] = v; Ao[a * m->dim + c] = ac;
                        cond[a * m->dim + c] = m->drm_ext[n];
                    }
                }
                for (int i = 0; i < nd; ++i) {
                    double ma = 0.0, cv = 0.0, ku = 0.0;
                    for (int j = 0; j < nd; ++j) {
                        if (cond[i] == cond[j]) continue;
                        ma += Me(i, j) * Ao[j];
                        cv += Ce(i, j) * Vo[j];
                        ku += Ke(i, j) * Uo[j];
                    }
                    tmp[dofs[i]] += (ma + cv) + ku;
                }
            }
            for (int i = 0; i < m->ntotal; ++i) F[i] += m->drm_factor * tmp[i];
        }
    }
};

// T^T A T restricted to free DOFs (no constraints: T is a 0/1 selection).
CSC free_block(const Model &m, const CSC &A) {
    Sparse s; s.n = m.nfree;
    for (int j = 0; j < A.n; ++j) {
        int fj = m.freedof[j];
        if (fj < 0) continue;
        for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k) {
            int fi = m.freedof[A.row[k]];
            if (fi >= 0) s.add(fi, fj, A.val[k]);
        }
    }
    return CSC::from(s);
}

void to_free(const Model &m, const std::vector<double> &t, std::vector<double> &f) {
    f.assign(m.nfree, 0.0);
    for (int i = 0; i < m.ntotal; ++i)
        if (m.freedof[i] >= 0) f[m.freedof[i]] = t[i];
}
void to_total(const Model &m, const std::vector<double> &f, std::vector<double> &t) {
    t.assign(m.ntotal, 0.0);
    for (int i = 0; i < m.ntotal; ++i)
        if (m.freedof[i] >= 0) t[i] = f[m.freedof[i]];
}

// Algorithm::UpdateStatesIncrements (Algorithm.cpp:18-56): nodes get U + dU, every
// element recomputes its Gauss-point strains.
void update_states(Model &m, const std::vector<double> &Utot_plus_dU) {
    for (auto &e : m.elems) element_update(m, e, Utot_plus_dU);
}
void commit_states(Model &m) {
    for (auto &e : m.elems) e.strain_n = e.strain;
}

}  // namespace

// ===================================================================== C API
extern "C" {

typedef struct oracle_opts {
    int32_t integrator;  // 0 CentralDifference, 1 ExtendedNewmarkBeta (G = 0), 2 NewmarkBeta
    int32_t nt;          // Analysis 'nt': steps k = 1..nt-1
    double dt;
    double mtol, ktol, ftol;
    int32_t algorithm;   // 0 Linear, 1 NewtonRaphson (Newmark integrators only)
    int32_t nmax;        // Newton 'nstep'
    int32_t cnvgtest;    // Algorithm::ComputeConvergence flag 1..8
    int32_t pad;
    double cnvgtol;      // Newton 'cnvgtol'
} oracle_opts;

// Algorithm::ComputeConvergence (Algorithm.cpp:122-188) on one process: Force = Feff,
// Delta = accumulated increment, delta = this iteration's solution, k = iteration index.
double convergence(int flag, const std::vector<double> &F, const std::vector<double> &Dl,
                   const std::vector<double> &dl, int k, double &NormFactor) {
    auto nrm = [](const std::vector<double> &v) {
        double s2 = 0.0;
        for (double x : v) s2 += x * x;
        return std::sqrt(s2);
    };
    auto energy = [&]() {
        double s2 = 0.0;
        for (size_t i = 0; i < F.size(); ++i) s2 += (dl[i] * F[i]) * (dl[i] * F[i]);
        return std::sqrt(s2);
    };
    switch (flag) {
        case 1: return nrm(F);
        case 2: return nrm(dl);
        case 3: return energy();
        case 4:
            if (k != 0) return nrm(F) / NormFactor;
            NormFactor = nrm(F);
            return NormFactor;
        case 5:
            if (k != 0) return nrm(dl) / NormFactor;
            NormFactor = nrm(Dl);
            return nrm(dl);
        case 6:
            if (k != 0) return energy() / NormFactor;
            NormFactor = energy();
            return NormFactor;
        case 7: return nrm(dl) / nrm(Dl);
        case 8: return -1.0;
        default: return 0.0;
    }
}

const char *oracle_last_error(void) { return g_err.c_str(); }

// Runs a dynamic analysis; hist_* are [(nt-1) * nrec] (any may be NULL). rec_dofs are
// free DOF indices. `seconds` (optional) receives the wall time of the time loop only.
int oracle_run(const svlg_mesh_desc *d, const oracle_opts *o, int32_t nrec, const int32_t *rec_dofs,
               double *hist_U, double *hist_V, double *hist_A, double *seconds, double *final_U,
               int32_t *newton_iters, double *newton_res, int32_t nreac, const int32_t *reac_tot,
               double *hist_R) {
    try {
        Model m = build(d);
        Assembler as{&m, o->mtol, o->ktol, o->ftol};
        const double dt = o->dt;
        std::vector<double> U(m.ntotal, 0.0), V(m.ntotal, 0.0), A(m.ntotal, 0.0), Up(m.ntotal, 0.0);
        CSC M = as.mass();
        CSC C = as.damping();
        Skyline solver;
        CSC Kt;     // total effective stiffness (CD: M/dt^2 + C/2dt)
        CSC Kforce; // CD: M/dt^2 - C/2dt
        if (o->integrator == 0) {
            // CentralDifference::Initialize (:58-71).
            for (int i = 0; i < m.ntotal; ++i) Up[i] = U[i] - dt * V[i] + dt * dt / 2.0 * A[i];
            Kt = combine(1.0 / dt / dt, M, 1.0 / 2.0 / dt, C);
            Kforce = combine(1.0 / dt / dt, M, -(1.0 / 2.0 / dt), C);
        } else {
            CSC K = as.stiffness();
            // K + 4/dt^2 M + 2/dt C (ExtendedNewmarkBeta.cpp:241; NewmarkBeta.cpp:229).
            CSC KM = combine(1.0, K, 4.0 / dt / dt, M);
            Kt = combine(1.0, KM, 2.0 / dt, C);
        }
        const bool newton = o->integrator != 0 && o->algorithm == 1;
        if (!newton) {
            CSC Kf = free_block(m, Kt);
            if (!solver.factor(Kf)) {
                g_err = "EigenSolver::SolveSystem() : Solution for the linear system failed.";
                return 1;
            }
        }
        double NormFactor = 0.0;  // Algorithm member, kept across steps
        std::vector<double> Fint, Fext, Ft(m.ntotal), Feff, x, dUt, tmp, tmp2;
        std::vector<int> rec_tot(nrec, -1);
        for (int r = 0; r < nrec; ++r) {
            if (rec_dofs[r] < 0) {  // -1 - t: total DOF t (a restrained DOF, e.g. one with support motion)
                rec_tot[r] = -1 - rec_dofs[r];
                if (rec_tot[r] >= m.ntotal) { g_err = "recorded total dof out of range"; return 2; }
                continue;
            }
            for (int i = 0; i < m.ntotal; ++i)
                if (m.freedof[i] == rec_dofs[r]) { rec_tot[r] = i; break; }
            if (rec_tot[r] < 0) { g_err = "recorded dof is not free"; return 2; }
        }
        const bool has_sm = !m.sm_dof.empty();
        if (has_sm && newton) {
            g_err = "support motion with the Newton-Raphson algorithm is not restated";
            return 2;
        }
        std::vector<double> SM(m.ntotal, 0.0);
        auto t0 = std::chrono::steady_clock::now();
        for (int k = 1; k < o->nt; ++k) {
            if (newton) {
                // NewtonRaphson::ComputeNewIncrement (NewtonRaphson.cpp:24-75): Feff and the
                // tangent Keff from the current iterate, solve, dU += du, convergence test,
                // then UpdateStatesIncrements(U + dU).
                std::vector<double> dUf(m.nfree, 0.0), du, w(m.ntotal), v2(m.ntotal), dUtot;
                double Residual = 0.0;
                int it = 0;
                do {
                    to_total(m, dUf, dUtot);
                    as.fint(Fint);
                    as.fext(k, Fext);
                    if (o->integrator == 2) {
                        // NewmarkBeta::ComputeEffectiveForce (NewmarkBeta.cpp:212-228):
                        // Fext + Fbar - Fint + M (4/dt V + A - 4/dt^2 dU) + C (V - 2/dt dU).
                        for (int i = 0; i < m.ntotal; ++i) {
                            w[i] = 4.0 / dt * V[i] + A[i] - 4.0 / dt / dt * dUtot[i];
                            v2[i] = V[i] - 2.0 / dt * dUtot[i];
                        }
                        M.mul(w, tmp2);
                        C.mul(v2, tmp);
                        for (int i = 0; i < m.ntotal; ++i) Ft[i] = (((Fext[i] + 0.0) - Fint[i]) + tmp2[i]) + tmp[i];
                    } else {
                        // ExtendedNewmarkBeta::ComputeEffectiveForce (:219-225): no dU terms.
                        for (int i = 0; i < m.ntotal; ++i) w[i] = 4.0 / dt * V[i] + A[i];
                        C.mul(V, tmp);
                        M.mul(w, tmp2);
                        for (int i = 0; i < m.ntotal; ++i) Ft[i] = (((Fext[i] + 0.0) - Fint[i]) + tmp[i]) + tmp2[i];
                    }
                    to_free(m, Ft, Feff);
                    // ComputeEffectiveStiffness: tangent K + 4/dt^2 M + 2/dt C, refactorised.
                    CSC Ktan = as.stiffness(true);
                    CSC KM = combine(1.0, Ktan, 4.0 / dt / dt, M);
                    CSC Kn = combine(1.0, KM, 2.0 / dt, C);
                    Skyline sk;
                    if (!sk.factor(free_block(m, Kn))) {
                        g_err = "EigenSolver::SolveSystem() : Solution for the linear system failed.";
                        return 1;
                    }
                    sk.solve(Feff, du);
                    for (int i = 0; i < m.nfree; ++i) dUf[i] += du[i];
                    Residual = convergence(o->cnvgtest, Feff, dUf, du, it, NormFactor);
                    if (newton_res) newton_res[size_t(k - 1) * o->nmax + it] = Residual;
                    to_total(m, dUf, dUtot);
                    std::vector<double> Un(m.ntotal);
                    for (int i = 0; i < m.ntotal; ++i) Un[i] = U[i] + dUtot[i];
                    update_states(m, Un);
                    ++it;
                } while ((Residual > o->cnvgtol) && (it < o->nmax));
                if (newton_iters) newton_iters[k - 1] = it;
                dUt = dUtot;
            } else {
            as.fint(Fint);
            as.fext(k, Fext);
            if (o->integrator == 0) {
                // CentralDifference::ComputeEffectiveForce (:206-221).
                std::vector<double> dUU(m.ntotal);
                for (int i = 0; i < m.ntotal; ++i) dUU[i] = U[i] - Up[i];
                Kforce.mul(dUU, tmp);
                for (int i = 0; i < m.ntotal; ++i) Ft[i] = ((Fext[i] + 0.0) - Fint[i]) + tmp[i];
            } else {
                // Linear: dU = 0 when the effective force is formed, so NewmarkBeta's
                // extra -M 4/dt^2 dU - C 2/dt dU terms vanish and both forms agree:
                // Fext + Fbar - Fint - G(...) + C V + M (4/dt V + A).
                C.mul(V, tmp);
                std::vector<double> w(m.ntotal);
                for (int i = 0; i < m.ntotal; ++i) w[i] = 4.0 / dt * V[i] + A[i];
                M.mul(w, tmp2);
                for (int i = 0; i < m.ntotal; ++i) Ft[i] = (((Fext[i] + 0.0) - Fint[i]) + tmp[i]) + tmp2[i];
            }
            to_free(m, Ft, Feff);
            if (has_sm) {
                // ComputeSupportMotionVector: SM of sample k (NewmarkBeta: k - 1), Feff -= T^T K SM.
                as.support_increment(o->integrator == 2 ? k - 1 : k, SM);
                std::vector<double> lg, lgf;
                Kt.mul(SM, lg);
                to_free(m, lg, lgf);
                for (int i = 0; i < m.nfree; ++i) Feff[i] -= lgf[i];
            }
            solver.solve(Feff, x);
            to_total(m, x, dUt);
            // UpdateStatesIncrements: strains from U + dU (before SM enters U).
            std::vector<double> Un(m.ntotal);
            for (int i = 0; i < m.ntotal; ++i) Un[i] = U[i] + dUt[i];
            update_states(m, Un);
            }
            if (o->integrator == 0) {
                // CentralDifference::ComputeNewStep (:137-148): dU = T x + SupportMotion.
                for (int i = 0; i < m.ntotal; ++i) {
                    const double du = has_sm ? dUt[i] + SM[i] : dUt[i];
                    double v = 1.0 / 2.0 / dt * ((U[i] + du) - Up[i]);
                    double a = 1.0 / dt / dt * ((du - U[i]) + Up[i]);
                    V[i] = v; A[i] = a;
                    Up[i] = U[i];
                    U[i] += du;
                }
            } else {
                // ExtendedNewmarkBeta::ComputeNewStep (:123-154) with G = 0 (NewmarkBeta.cpp
                // :115-144 alike): U += T x + SupportMotion, V and A from T x alone.
                for (int i = 0; i < m.ntotal; ++i) {
                    U[i] += has_sm ? dUt[i] + SM[i] : dUt[i];
                    double a = 4.0 / dt / dt * dUt[i] - 4.0 / dt * V[i] - A[i];
                    double v = 2.0 / dt * dUt[i] - V[i];
                    A[i] = a; V[i] = v;
                }
            }
            // DynamicAnalysis::UpdateDomain: commit (elastic: newStrain = Strain).
            commit_states(m);
            int row = k - 1;
            if (nreac > 0 && hist_R) {
                // Integrator::ComputeReactionForce (CentralDifference.cpp:156-171, Newmark
                // alike): R = Fdyn - Fext - Fbar at the fixed nodes (DynamicAnalysis.cpp:130-150).
                std::vector<double> Fd, Fe;
                as.fdyn(V, A, Fd);
                as.fext(k, Fe);
                for (int r = 0; r < nreac; ++r) {
                    const int td = reac_tot[r];
                    if (td < 0 || td >= m.ntotal) { g_err = "reaction dof out of range"; return 2; }
                    hist_R[size_t(row) * nreac + r] = (Fd[td] - Fe[td]) - 0.0;
                }
            }
            for (int r = 0; r < nrec; ++r) {
                int td = rec_tot[r];
                if (hist_U) hist_U[size_t(row) * nrec + r] = U[td];
                if (hist_V) hist_V[size_t(row) * nrec + r] = V[td];
                if (hist_A) hist_A[size_t(row) * nrec + r] = A[td];
            }
        }
        auto t1 = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
        if (final_U) {
            std::vector<double> f;
            to_free(m, U, f);
            std::memcpy(final_U, f.data(), sizeof(double) * f.size());
        }
        return 0;
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return 3;
    }
}

// Element matrices of element `e` of the model: K (initial), M (per massform), C
// (Rayleigh) as dense [nd*nd] row-major arrays; any may be NULL.
int oracle_element_matrices(const svlg_mesh_desc *d, int32_t e, double *K, double *M, double *C) {
    try {
        Model m = build(d);
        if (e < 0 || e >= (int)m.elems.size()) { g_err = "element index out of range"; return 1; }
        if (K) { Mat k = element_stiffness(m, m.elems[e]); std::memcpy(K, k.a.data(), k.a.size() * 8); }
        if (M) { Mat k = element_mass(m, m.elems[e]); std::memcpy(M, k.a.data(), k.a.size() * 8); }
        if (C) { Mat k = element_damping(m, m.elems[e]); std::memcpy(C, k.a.data(), k.a.size() * 8); }
        return 0;
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return 3;
    }
}

// Fint(U) in free DOFs through the SVL sequence UpdateState(U) -> ComputeInternalForces
// -> Assembler filter; also the lumped mass and damping diagonals in free DOFs.
int oracle_internal_force(const svlg_mesh_desc *d, const double *Ufree, double ftol, double *Ffree) {
    try {
        Model m = build(d);
        Assembler as{&m, 1e-12, 1e-12, ftol};
        std::vector<double> uf(Ufree, Ufree + m.nfree), ut, F, ff;
        to_total(m, uf, ut);
        update_states(m, ut);
        as.fint(F);
        to_free(m, F, ff);
        std::memcpy(Ffree, ff.data(), sizeof(double) * ff.size());
        return 0;
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return 3;
    }
}

// y = Keff x over the free DOFs with Keff = T^T (K + 4/dt^2 M + 2/dt C) T, the
// effective stiffness of ExtendedNewmarkBeta::ComputeEffectiveStiffness
// (ExtendedNewmarkBeta.cpp:233-245) with G = 0.
int oracle_newmark_keff_apply(const svlg_mesh_desc *d, double dt, const double *xfree, double *yfree) {
    try {
        Model m = build(d);
        Assembler as{&m};
        CSC K = as.stiffness();
        CSC M = as.mass();
        CSC C = as.damping();
        CSC KM = combine(1.0, K, 4.0 / dt / dt, M);
        CSC Kt = combine(1.0, KM, 2.0 / dt, C);
        CSC Kf = free_block(m, Kt);
        std::vector<double> x(xfree, xfree + m.nfree), y;
        Kf.mul(x, y);
        std::memcpy(yfree, y.data(), sizeof(double) * y.size());
        return 0;
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return 3;
    }
}

// Dense T^T (K + 4/dt^2 M + 2/dt C) T ([nfree][nfree], row-major) for small models.
int oracle_newmark_keff_dense(const svlg_mesh_desc *d, double dt, double *Kd) {
    try {
        Model m = build(d);
        Assembler as{&m};
        CSC K = as.stiffness();
        CSC M = as.mass();
        CSC C = as.damping();
        CSC KM = combine(1.0, K, 4.0 / dt / dt, M);
        CSC Kt = combine(1.0, KM, 2.0 / dt, C);
        CSC Kf = free_block(m, Kt);
        const size_t n = size_t(m.nfree);
        std::fill(Kd, Kd + n * n, 0.0);
        for (int j = 0; j < Kf.n; ++j)
            for (int k = Kf.colptr[j]; k < Kf.colptr[j + 1]; ++k) Kd[size_t(Kf.row[k]) * n + j] += Kf.val[k];
        return 0;
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return 3;
    }
}

// Drives one Plastic3DJ2 (plane = 0) or PlasticPlaneStrainJ2 (plane = 1) Gauss point
// through n engineering strains (6 or 3 per row), one UpdateState(strain, 1) per row, and
// returns per row the stress and the consistent tangent in the element's Voigt order and
// alpha. params5 = K, G, sigma_y, H, beta.
int oracle_j2_path(const double *params5, int32_t plane, int32_t n, const double *strain, double *stress,
                   double *tangent, double *alpha) {
    const J2Params p{params5[0], params5[1], params5[2], params5[3], params5[4]};
    J2State s;
    j2_elastic_tangent(p, s.C);
    const int nv = plane ? 3 : 6;
    for (int r = 0; r < n; ++r) {
        const double *es = strain + size_t(r) * nv;
        double et[6] = {0, 0, 0, 0, 0, 0};
        if (!plane) {
            et[0] = es[0]; et[1] = es[1]; et[2] = es[2];
            et[3] = 0.5 * es[3]; et[4] = 0.5 * es[4]; et[5] = 0.5 * es[5];
        } else {
            et[1] = es[0]; et[2] = es[1]; et[4] = 0.5 * es[2];
        }
        j2_update(s, p, et);
        for (int i = 0; i < nv; ++i) {
            const int ii = plane ? kPsIdx[i] : i;
            if (stress) stress[size_t(r) * nv + i] = s.sig[ii];
            for (int j = 0; j < nv; ++j) {
                const int jj = plane ? kPsIdx[j] : j;
                if (tangent) tangent[(size_t(r) * nv + i) * nv + j] = s.C[ii * 6 + jj];
            }
        }
        if (alpha) alpha[r] = s.alpha;
    }
    return 0;
}

int oracle_diagonals(const svlg_mesh_desc *d, double *Mfree, double *Cfree) {
    try {
        Model m = build(d);
        Assembler as{&m};
        CSC M = free_block(m, as.mass());
        CSC C = free_block(m, as.damping());
        for (int j = 0; j < m.nfree; ++j) {
            if (Mfree) Mfree[j] = 0.0;
            if (Cfree) Cfree[j] = 0.0;
            for (int k = M.colptr[j]; k < M.colptr[j + 1]; ++k)
                if (M.row[k] == j && Mfree) Mfree[j] = M.val[k];
            for (int k = C.colptr[j]; k < C.colptr[j + 1]; ++k)
                if (C.row[k] == j && Cfree) Cfree[j] = C.val[k];
        }
        return 0;
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return 3;
    }
}

}  // extern "C"
