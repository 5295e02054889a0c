// svlg_elem.hpp -- device routines of lin2DQuad4 / lin3DHexa8 with 2x2(x2) Gauss.
//
// Each routine follows the per-Gauss-point arithmetic of SVL's element classes:
//   Jacobian        lin2DQuad4.cpp:648-667, lin3DHexa8.cpp:756-785 (J_ij = dx_j/dxi_i)
//   B matrix        lin2DQuad4.cpp:688-710, lin3DHexa8.cpp:811-853 (Voigt xx,yy,(zz),xy,(yz,zx))
//   internal force  lin2DQuad4.cpp:383-412, lin3DHexa8.cpp:383-412 (sum w t |J| B^T sigma)
//   lumped mass     lin2DQuad4.cpp:243-285, lin3DHexa8.cpp:243-285 (row-sum of consistent)
//   elastic D       Elastic2DPlaneStrain.cpp:18-26, Elastic3DLinear.cpp:19-28
// with SVL's Gauss abscissa 0.577350269189626 (GaussQuadrature.cpp:312-315, 734-741).
// The inverse Jacobian is formed by cofactors (SVL uses an LU inverse; the two agree
// to rounding).
#pragma once

#include <hip/hip_runtime.h>

namespace svlg {

constexpr double kGP = 0.577350269189626;

template <int DIM>
struct Elem;

template <>
struct Elem<2> {
    static constexpr int nen = 4, nv = 3, ngp = 4, nd = 8;
};
template <>
struct Elem<3> {
    static constexpr int nen = 8, nv = 6, ngp = 8, nd = 24;
};

// Corner signs of the reference element in SVL node order.
__device__ __forceinline__ double sgx(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0; }
__device__ __forceinline__ double sgy(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0; }
__device__ __forceinline__ double sgz(int a) { return (a >= 4) ? 1.0 : -1.0; }

// Gauss point g: Quad (-,-),(+,-),(-,+),(+,+); Hexa x fastest then y then z.
__device__ __forceinline__ void gauss_point(int g, double &r, double &s, double &t) {
    r = (g & 1) ? kGP : -kGP;
    s = (g & 2) ? kGP : -kGP;
    t = (g & 4) ? kGP : -kGP;
}

// ---------------------------------------------------------------- Quad4 (2D)
// Shape-function derivatives in physical coordinates at (ri, si); returns |det J|.
__device__ __forceinline__ double quad_dn(const double (&x)[4][2], double ri, double si, double (&dn)[4][2]) {
    double J11 = -1.0 / 4.0 * (1.0 - si) * x[0][0] + 1.0 / 4.0 * (1.0 - si) * x[1][0] +
                 1.0 / 4.0 * (1.0 + si) * x[2][0] - 1.0 / 4.0 * (1.0 + si) * x[3][0];
    double J12 = -1.0 / 4.0 * (1.0 - si) * x[0][1] + 1.0 / 4.0 * (1.0 - si) * x[1][1] +
                 1.0 / 4.0 * (1.0 + si) * x[2][1] - 1.0 / 4.0 * (1.0 + si) * x[3][1];
    double J21 = -1.0 / 4.0 * (1.0 - ri) * x[0][0] - 1.0 / 4.0 * (1.0 + ri) * x[1][0] +
                 1.0 / 4.0 * (1.0 + ri) * x[2][0] + 1.0 / 4.0 * (1.0 - ri) * x[3][0];
    double J22 = -1.0 / 4.0 * (1.0 - ri) * x[0][1] - 1.0 / 4.0 * (1.0 + ri) * x[1][1] +
                 1.0 / 4.0 * (1.0 + ri) * x[2][1] + 1.0 / 4.0 * (1.0 - ri) * x[3][1];
    double det = J11 * J22 - J12 * J21;
    double inv = 1.0 / det;
    double I00 = J22 * inv, I01 = -J12 * inv, I10 = -J21 * inv, I11 = J11 * inv;
    const double dr[4] = {-(1.0 - si), (1.0 - si), (1.0 + si), -(1.0 + si)};
    const double ds[4] = {-(1.0 - ri), -(1.0 + ri), (1.0 + ri), (1.0 - ri)};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        dn[a][0] = 1.0 / 4.0 * I01 * ds[a] + 1.0 / 4.0 * I00 * dr[a];
        dn[a][1] = 1.0 / 4.0 * I11 * ds[a] + 1.0 / 4.0 * I10 * dr[a];
    }
    return fabs(det);
}

// f = sum_g w th |J| B^T D B u (general quadrilateral).
__device__ __forceinline__ void quad_fint(const double (&x)[4][2], const double (&u)[4][2], double lam, double mu,
                                          double th, double (&f)[4][2]) {
    const double c1 = lam + 2.0 * mu, c2 = lam, c3 = mu;
#pragma unroll
    for (int a = 0; a < 4; ++a) f[a][0] = f[a][1] = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        double r, s, t;
        gauss_point(g, r, s, t);
        double dn[4][2];
        double adet = quad_dn(x, r, s, dn);
        double exx = 0.0, eyy = 0.0, gxy = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            exx += dn[a][0] * u[a][0];
            eyy += dn[a][1] * u[a][1];
            gxy += dn[a][1] * u[a][0] + dn[a][0] * u[a][1];
        }
        double sxx = c1 * exx + c2 * eyy, syy = c2 * exx + c1 * eyy, sxy = c3 * gxy;
        double wj = 1.0 * th * adet;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            f[a][0] += wj * (dn[a][0] * sxx + dn[a][1] * sxy);
            f[a][1] += wj * (dn[a][1] * syy + dn[a][0] * sxy);
        }
    }
}

// Row-sum lumped mass per node (same for both components).
__device__ __forceinline__ void quad_lumped_mass(const double (&x)[4][2], double rho, double th, double (&ml)[4]) {
    double M[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) M[a][b] = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        double r, s, t;
        gauss_point(g, r, s, t);
        double dn[4][2];
        double adet = quad_dn(x, r, s, dn);
        double h[4] = {1.0 / 4.0 * (1.0 - r) * (1.0 - s), 1.0 / 4.0 * (1.0 + r) * (1.0 - s),
                       1.0 / 4.0 * (1.0 + r) * (1.0 + s), 1.0 / 4.0 * (1.0 - r) * (1.0 + s)};
        double sc = 1.0 * rho * th * adet;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) M[a][b] += sc * (h[a] * h[b]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double d = M[a][a];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (b != a) d += M[a][b];
        ml[a] = d;
    }
}

// ---------------------------------------------------------------- Hexa8 (3D)
__device__ __forceinline__ double hexa_dn(const double (&x)[8][3], double ri, double si, double ti, double (&dn)[8][3]) {
    double d[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        d[a][0] = 1.0 / 8.0 * sgx(a) * (1.0 + sgy(a) * si) * (1.0 + sgz(a) * ti);
        d[a][1] = 1.0 / 8.0 * sgy(a) * (1.0 + sgx(a) * ri) * (1.0 + sgz(a) * ti);
        d[a][2] = 1.0 / 8.0 * sgz(a) * (1.0 + sgx(a) * ri) * (1.0 + sgy(a) * si);
    }
    double J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) s += d[a][i] * x[a][j];
            J[i][j] = s;
        }
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double inv = 1.0 / det;
    double I[3][3];
    I[0][0] = c00 * inv;
    I[1][0] = c01 * inv;
    I[2][0] = c02 * inv;
    I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * inv;
    I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * inv;
    I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * inv;
    I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * inv;
    I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * inv;
    I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * inv;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) dn[a][i] = I[i][2] * d[a][2] + I[i][1] * d[a][1] + I[i][0] * d[a][0];
    return fabs(det);
}

__device__ __forceinline__ void hexa_fint(const double (&x)[8][3], const double (&u)[8][3], double lam, double mu,
                                          double (&f)[8][3]) {
    const double c1 = lam + 2.0 * mu, c2 = lam, c3 = mu;
#pragma unroll
    for (int a = 0; a < 8; ++a) f[a][0] = f[a][1] = f[a][2] = 0.0;
#pragma unroll 1
    for (int g = 0; g < 8; ++g) {
        double r, s, t;
        gauss_point(g, r, s, t);
        double dn[8][3];
        double adet = hexa_dn(x, r, s, t, dn);
        double e[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            e[0] += dn[a][0] * u[a][0];
            e[1] += dn[a][1] * u[a][1];
            e[2] += dn[a][2] * u[a][2];
            e[3] += dn[a][1] * u[a][0] + dn[a][0] * u[a][1];
            e[4] += dn[a][2] * u[a][1] + dn[a][1] * u[a][2];
            e[5] += dn[a][2] * u[a][0] + dn[a][0] * u[a][2];
        }
        double sg[6];
        sg[0] = c1 * e[0] + c2 * e[1] + c2 * e[2];
        sg[1] = c2 * e[0] + c1 * e[1] + c2 * e[2];
        sg[2] = c2 * e[0] + c2 * e[1] + c1 * e[2];
        sg[3] = c3 * e[3];
        sg[4] = c3 * e[4];
        sg[5] = c3 * e[5];
        double wj = 1.0 * adet;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            f[a][0] += wj * (dn[a][0] * sg[0] + dn[a][1] * sg[3] + dn[a][2] * sg[5]);
            f[a][1] += wj * (dn[a][1] * sg[1] + dn[a][0] * sg[3] + dn[a][2] * sg[4]);
            f[a][2] += wj * (dn[a][2] * sg[2] + dn[a][1] * sg[4] + dn[a][0] * sg[5]);
        }
    }
}

__device__ __forceinline__ void hexa_lumped_mass(const double (&x)[8][3], double rho, double (&ml)[8]) {
    double M[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) M[a][b] = 0.0;
#pragma unroll 1
    for (int g = 0; g < 8; ++g) {
        double r, s, t;
        gauss_point(g, r, s, t);
        double dn[8][3];
        double adet = hexa_dn(x, r, s, t, dn);
        double h[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) h[a] = 1.0 / 8.0 * (1.0 + sgx(a) * r) * (1.0 + sgy(a) * s) * (1.0 + sgz(a) * t);
        double sc = 1.0 * rho * adet;
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b) M[a][b] += sc * (h[a] * h[b]);
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double d = M[a][a];
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (b != a) d += M[a][b];
        ml[a] = d;
    }
}

}  // namespace svlg
