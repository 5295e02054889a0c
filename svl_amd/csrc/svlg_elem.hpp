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

// ------------------------------------------------- parallelepipeds: modal form
// For an affine element the Jacobian is constant, the displacement gradient is linear
// in the modal amplitudes of u, and the 2x2(x2) Gauss sum of B^T D B u separates by
// the orthogonality of the Gauss abscissae over the symmetric point set:
//   f_a = |J| sum_p s_p(a) F_p,  F_p from the 1-point stress of the linear modes and the
//   exact bilinear / trilinear (hourglass) mode stresses weighted by g^2 and g^4,
// g = 0.577350269189626. This equals SVL's Gauss-point loop to rounding for every
// parallelepiped (checked against the per-point routine and the oracle in the tests)
// at about half the arithmetic.

// Modal amplitudes (times 8) of 8 corner values in SVL corner order:
// mx = sum sx v, my, mz, mxy = sum sx sy v, myz, mzx, mxyz.
template <class T>
__device__ __forceinline__ void hexa_modes(const T (&v)[8], T &mx, T &my, T &mz, T &mxy, T &myz, T &mzx, T &mxyz) {
    T s00 = v[1] + v[0], d00 = v[1] - v[0];
    T s10 = v[2] + v[3], d10 = v[2] - v[3];
    T s01 = v[5] + v[4], d01 = v[5] - v[4];
    T s11 = v[6] + v[7], d11 = v[6] - v[7];
    T ss0 = s10 + s00, sd0 = s10 - s00, ds0 = d10 + d00, dd0 = d10 - d00;
    T ss1 = s11 + s01, sd1 = s11 - s01, ds1 = d11 + d01, dd1 = d11 - d01;
    mz = ss1 - ss0;
    my = sd1 + sd0;
    myz = sd1 - sd0;
    mx = ds1 + ds0;
    mzx = ds1 - ds0;
    mxy = dd1 + dd0;
    mxyz = dd1 - dd0;
}

// Corner values from modal amplitudes (transpose of hexa_modes, zero constant mode).
template <class T>
__device__ __forceinline__ void hexa_unmodes(T mx, T my, T mz, T mxy, T myz, T mzx, T mxyz, T (&v)[8]) {
    T ss0 = -mz, ss1 = mz;
    T sd0 = my - myz, sd1 = my + myz;
    T ds0 = mx - mzx, ds1 = mx + mzx;
    T dd0 = mxy - mxyz, dd1 = mxy + mxyz;
    T s00 = ss0 - sd0, s10 = ss0 + sd0, d00 = ds0 - dd0, d10 = ds0 + dd0;
    T s01 = ss1 - sd1, s11 = ss1 + sd1, d01 = ds1 - dd1, d11 = ds1 + dd1;
    v[0] = s00 - d00;
    v[1] = s00 + d00;
    v[3] = s10 - d10;
    v[2] = s10 + d10;
    v[4] = s01 - d01;
    v[5] = s01 + d01;
    v[7] = s11 - d11;
    v[6] = s11 + d11;
}

// P = L(G): H = Lam G, sigma = lam tr(H) I + mu (H + H^T), P = Lam^T sigma, restricted to
// the nonzero rows of G (GR) and the needed rows of P (PR).
template <int DIM, bool G0, bool G1, bool G2, bool P0, bool P1, bool P2, class T>
__device__ __forceinline__ void apply_L(const T (&G)[DIM][DIM], const T (&Lam)[DIM][DIM], T lam, T mu,
                                        T (&P)[DIM][DIM]) {
    constexpr bool gr[3] = {G0, G1, G2};
    constexpr bool pr[3] = {P0, P1, P2};
    T H[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            T s = T(0);
            bool first = true;
#pragma unroll
            for (int k = 0; k < DIM; ++k)
                if (gr[k]) {
                    s = first ? Lam[i][k] * G[k][c] : fma(Lam[i][k], G[k][c], s);
                    first = false;
                }
            H[i][c] = s;
        }
    T tr = H[0][0] + H[1][1];
    if constexpr (DIM == 3) tr += H[2][2];
    T S[DIM][DIM];
    const T lt = lam * tr, m2 = T(2) * mu;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        S[i][i] = fma(m2, H[i][i], lt);
#pragma unroll
        for (int j = i + 1; j < DIM; ++j) {
            T v = mu * (H[i][j] + H[j][i]);
            S[i][j] = v;
            S[j][i] = v;
        }
    }
#pragma unroll
    for (int k = 0; k < DIM; ++k)
        if (pr[k]) {
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                T s = Lam[0][k] * S[0][c];
#pragma unroll
                for (int i = 1; i < DIM; ++i) s = fma(Lam[i][k], S[i][c], s);
                P[k][c] = s;
            }
        }
}

// apply_L for a diagonal Lam (rectangular / brick elements aligned with the axes):
// H_ic = Lam_ii G_ic, P_kc = Lam_kk sigma_kc, with only the nonzero rows of G and the
// needed rows of P formed.
template <int DIM, bool G0, bool G1, bool G2, bool P0, bool P1, bool P2, class T>
__device__ __forceinline__ void apply_L_diag(const T (&G)[DIM][DIM], const T (&Lam)[DIM][DIM], T lam, T mu,
                                             T (&P)[DIM][DIM]) {
    constexpr bool gr[3] = {G0, G1, G2};
    constexpr bool pr[3] = {P0, P1, P2};
    T H[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
        if (gr[i]) {
#pragma unroll
            for (int c = 0; c < DIM; ++c) H[i][c] = Lam[i][i] * G[i][c];
        }
    T tr = T(0);
    bool first = true;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
        if (gr[i]) {
            tr = first ? H[i][i] : tr + H[i][i];
            first = false;
        }
    const T lt = lam * tr, m2 = T(2) * mu;
#pragma unroll
    for (int k = 0; k < DIM; ++k)
        if (pr[k]) {
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                T sv;
                if (k == c) {
                    sv = gr[k] ? fma(m2, H[k][k], lt) : lt;
                } else if (gr[k] && gr[c]) {
                    sv = mu * (H[k][c] + H[c][k]);
                } else if (gr[k]) {
                    sv = mu * H[k][c];
                } else if (gr[c]) {
                    sv = mu * H[c][k];
                } else {
                    sv = T(0);
                }
                P[k][c] = Lam[k][k] * sv;
            }
        }
}

// 3x3 inverse (cofactors) and |det|.
__device__ __forceinline__ double inv3(const double (&J)[3][3], double (&I)[3][3]) {
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double r = 1.0 / det;
    I[0][0] = c00 * r;
    I[1][0] = c01 * r;
    I[2][0] = c02 * r;
    I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
    I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
    I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
    I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    return fabs(det);
}

#define SVLG_APPLY_L(...) apply_L_sel<DIAG, __VA_ARGS__>

template <bool DIAG, int DIM, bool G0, bool G1, bool G2, bool P0, bool P1, bool P2, class T>
__device__ __forceinline__ void apply_L_sel(const T (&G)[DIM][DIM], const T (&Lam)[DIM][DIM], T lam, T mu,
                                            T (&P)[DIM][DIM]) {
    if constexpr (DIAG)
        apply_L_diag<DIM, G0, G1, G2, P0, P1, P2>(G, Lam, lam, mu, P);
    else
        apply_L<DIM, G0, G1, G2, P0, P1, P2>(G, Lam, lam, mu, P);
}

// Centre Jacobian of a Hexa8 (J[k][j] = d x_j / d xi_k = m_k(x_j) / 8), its inverse and |det J|.
__device__ __host__ inline double hexa_center_geometry(const double (&x)[8][3], double (&Lam)[3][3]);

// Modal internal force of an affine Hexa8: Lam = J^-1, l = lambda |J| / 8, m = mu |J| / 8.
// DIAG: Lam is diagonal (axis-aligned bricks).
template <bool DIAG = false, class T = double>
__device__ __forceinline__ void hexa_modal_force(const T (&Lam)[3][3], T l, T m, const T (&u)[8][3], T (&f)[8][3]) {
    T Mx[3], My[3], Mz[3], Mxy[3], Myz[3], Mzx[3], Mxyz[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] = u[a][c];
        hexa_modes(v, Mx[c], My[c], Mz[c], Mxy[c], Myz[c], Mzx[c], Mxyz[c]);
    }
    const T g2 = T(kGP * kGP), g4 = T(kGP * kGP * kGP * kGP);
    // The Gauss weights of the hourglass terms (g^2, g^4) scale the mode amplitudes
    // before the linear stress map.
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Mxy[c] *= g2;
        Myz[c] *= g2;
        Mzx[c] *= g2;
        Mxyz[c] *= g4;
    }
    T G[3][3], P[3][3];
    T Fx[3], Fy[3], Fz[3], Fxy[3], Fyz[3], Fzx[3], Fxyz[3];
    // Linear modes (1-point gradient): all rows.
#pragma unroll
    for (int c = 0; c < 3; ++c) { G[0][c] = Mx[c]; G[1][c] = My[c]; G[2][c] = Mz[c]; }
    SVLG_APPLY_L(3, true, true, true, true, true, true)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) { Fx[c] = P[0][c]; Fy[c] = P[1][c]; Fz[c] = P[2][c]; }
    // eta monomial: rows xi (mxy) and zeta (myz).
#pragma unroll
    for (int c = 0; c < 3; ++c) { G[0][c] = Mxy[c]; G[2][c] = Myz[c]; }
    SVLG_APPLY_L(3, true, false, true, true, false, true)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) { Fxy[c] = P[0][c]; Fyz[c] = P[2][c]; }
    // zeta monomial: rows xi (mzx) and eta (myz).
#pragma unroll
    for (int c = 0; c < 3; ++c) { G[0][c] = Mzx[c]; G[1][c] = Myz[c]; }
    SVLG_APPLY_L(3, true, true, false, true, true, false)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) { Fzx[c] = P[0][c]; Fyz[c] += P[1][c]; }
    // xi monomial: rows eta (mxy) and zeta (mzx).
#pragma unroll
    for (int c = 0; c < 3; ++c) { G[1][c] = Mxy[c]; G[2][c] = Mzx[c]; }
    SVLG_APPLY_L(3, false, true, true, false, true, true)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) { Fxy[c] += P[1][c]; Fzx[c] += P[2][c]; }
    // Trilinear mode: eta*zeta in row xi, xi*zeta in row eta, xi*eta in row zeta.
#pragma unroll
    for (int c = 0; c < 3; ++c) G[0][c] = Mxyz[c];
    SVLG_APPLY_L(3, true, false, false, true, false, false)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) Fxyz[c] = P[0][c];
#pragma unroll
    for (int c = 0; c < 3; ++c) G[1][c] = Mxyz[c];
    SVLG_APPLY_L(3, false, true, false, false, true, false)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) Fxyz[c] += P[1][c];
#pragma unroll
    for (int c = 0; c < 3; ++c) G[2][c] = Mxyz[c];
    SVLG_APPLY_L(3, false, false, true, false, false, true)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) Fxyz[c] += P[2][c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T v[8];
        hexa_unmodes(Fx[c], Fy[c], Fz[c], Fxy[c], Fyz[c], Fzx[c], Fxyz[c], v);
#pragma unroll
        for (int a = 0; a < 8; ++a) f[a][c] = v[a];
    }
}

// Internal force of an affine Hexa8 from corner coordinates and displacements.
__device__ __forceinline__ void hexa_affine_fint(const double (&x)[8][3], const double (&u)[8][3], double lam,
                                                 double mu, double (&f)[8][3]) {
    double Lam[3][3];
    const double s = 0.125 * hexa_center_geometry(x, Lam);
    hexa_modal_force(Lam, lam * s, mu * s, u, f);
}

// Quad4 corner order (-,-),(+,-),(+,+),(-,+): mx = sum sx v, my, mxy.
__device__ __forceinline__ void quad_modes(const double (&v)[4], double &mx, double &my, double &mxy) {
    double s0 = v[1] + v[0], d0 = v[1] - v[0];
    double s1 = v[2] + v[3], d1 = v[2] - v[3];
    my = s1 - s0;
    mx = d1 + d0;
    mxy = d1 - d0;
}

__device__ __forceinline__ void quad_unmodes(double mx, double my, double mxy, double (&v)[4]) {
    double s0 = -my, s1 = my;
    double d0 = mx - mxy, d1 = mx + mxy;
    v[0] = s0 - d0;
    v[1] = s0 + d0;
    v[3] = s1 - d1;
    v[2] = s1 + d1;
}

// Centre Jacobian of a Quad4 (J[k][j] = m_k(x_j) / 4), its inverse and |det J|.
__device__ __host__ inline double quad_center_geometry(const double (&x)[4][2], double (&Lam)[2][2]);

// Modal internal force of an affine Quad4: l = lambda |J| th / 4, m = mu |J| th / 4.
template <bool DIAG = false>
__device__ __forceinline__ void quad_modal_force(const double (&Lam)[2][2], double l, double m,
                                                 const double (&u)[4][2], double (&f)[4][2]) {
    double Mx[2], My[2], Mxy[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        double v[4] = {u[0][c], u[1][c], u[2][c], u[3][c]};
        quad_modes(v, Mx[c], My[c], Mxy[c]);
    }
    const double g2 = kGP * kGP;
#pragma unroll
    for (int c = 0; c < 2; ++c) Mxy[c] *= g2;
    double G[2][2], P[2][2], Fx[2], Fy[2], Fxy[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) { G[0][c] = Mx[c]; G[1][c] = My[c]; }
    SVLG_APPLY_L(2, true, true, false, true, true, false)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 2; ++c) { Fx[c] = P[0][c]; Fy[c] = P[1][c]; }
#pragma unroll
    for (int c = 0; c < 2; ++c) G[0][c] = Mxy[c];
    SVLG_APPLY_L(2, true, false, false, true, false, false)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 2; ++c) Fxy[c] = P[0][c];
#pragma unroll
    for (int c = 0; c < 2; ++c) G[1][c] = Mxy[c];
    SVLG_APPLY_L(2, false, true, false, false, true, false)(G, Lam, l, m, P);
#pragma unroll
    for (int c = 0; c < 2; ++c) Fxy[c] += P[1][c];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        double v[4];
        quad_unmodes(Fx[c], Fy[c], Fxy[c], v);
#pragma unroll
        for (int a = 0; a < 4; ++a) f[a][c] = v[a];
    }
}

__device__ __forceinline__ void quad_affine_fint(const double (&x)[4][2], const double (&u)[4][2], double lam,
                                                 double mu, double th, double (&f)[4][2]) {
    double Lam[2][2];
    const double s = 0.25 * quad_center_geometry(x, Lam) * th;
    quad_modal_force(Lam, lam * s, mu * s, u, f);
}

// Shared host/device geometry (also used on the host to build the geometry table).
__device__ __host__ inline double hexa_center_geometry(const double (&x)[8][3], double (&Lam)[3][3]) {
    double J[3][3];
    for (int j = 0; j < 3; ++j) {
        double v[8];
        for (int a = 0; a < 8; ++a) v[a] = x[a][j];
        double s00 = v[1] - v[0], s10 = v[2] - v[3], s01 = v[5] - v[4], s11 = v[6] - v[7];
        double mx = (s11 + s01) + (s10 + s00);
        double t00 = v[1] + v[0], t10 = v[2] + v[3], t01 = v[5] + v[4], t11 = v[6] + v[7];
        double my = (t11 - t01) + (t10 - t00);
        double mz = (t11 + t01) - (t10 + t00);
        J[0][j] = 0.125 * mx;
        J[1][j] = 0.125 * my;
        J[2][j] = 0.125 * mz;
    }
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double r = 1.0 / det;
    Lam[0][0] = c00 * r;
    Lam[1][0] = c01 * r;
    Lam[2][0] = c02 * r;
    Lam[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
    Lam[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
    Lam[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
    Lam[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    Lam[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    Lam[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    return fabs(det);
}

__device__ __host__ inline double quad_center_geometry(const double (&x)[4][2], double (&Lam)[2][2]) {
    double J[2][2];
    for (int j = 0; j < 2; ++j) {
        double s0 = x[1][j] + x[0][j], d0 = x[1][j] - x[0][j];
        double s1 = x[2][j] + x[3][j], d1 = x[2][j] - x[3][j];
        J[0][j] = 0.25 * (d1 + d0);
        J[1][j] = 0.25 * (s1 - s0);
    }
    double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double r = 1.0 / det;
    Lam[0][0] = J[1][1] * r;
    Lam[0][1] = -J[0][1] * r;
    Lam[1][0] = -J[1][0] * r;
    Lam[1][1] = J[0][0] * r;
    return fabs(det);
}

// ------------------------------------------------------------- consistent mass
// M_e w with SVL's consistent mass sum_g w rho (th) |J| H^T H (lin2DQuad4.cpp:243-285,
// lin3DHexa8.cpp:243-285, before the lumping branch), added to f.
// Affine elements: |J| is constant and the 2x2(x2) Gauss sum of N_a N_b factorises per
// direction into [[(1+g^2)/2, (1-g^2)/2], [(1-g^2)/2, (1+g^2)/2]], whose eigenvectors are
// the constant (eigenvalue 1) and the linear (eigenvalue g^2) corner patterns. The element
// matrix is therefore diagonal in the modal basis of the corner values: the constant mode
// keeps weight 1, each linear factor of a mode contributes g^2. rs = rho |J| / 8 (Hexa8) or
// rho th |J| / 4 (Quad4): the same scaling the stiffness table paths use for lambda, mu.
__device__ __forceinline__ void hexa_mass_modal(double rs, const double (&w)[8][3], double (&f)[8][3]) {
    const double g2 = kGP * kGP, g4 = g2 * g2, g6 = g4 * g2;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] = w[a][c];
        double mx, my, mz, mxy, myz, mzx, mxyz;
        hexa_modes(v, mx, my, mz, mxy, myz, mzx, mxyz);
        const double m0 = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
        double o[8];
        hexa_unmodes(g2 * mx, g2 * my, g2 * mz, g4 * mxy, g4 * myz, g4 * mzx, g6 * mxyz, o);
#pragma unroll
        for (int a = 0; a < 8; ++a) f[a][c] += rs * (o[a] + m0);
    }
}

__device__ __forceinline__ void quad_mass_modal(double rs, const double (&w)[4][2], double (&f)[4][2]) {
    const double g2 = kGP * kGP, g4 = g2 * g2;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        double v[4] = {w[0][c], w[1][c], w[2][c], w[3][c]};
        double mx, my, mxy;
        quad_modes(v, mx, my, mxy);
        const double m0 = (v[0] + v[1]) + (v[2] + v[3]);
        double o[4];
        quad_unmodes(g2 * mx, g2 * my, g4 * mxy, o);
#pragma unroll
        for (int a = 0; a < 4; ++a) f[a][c] += rs * (o[a] + m0);
    }
}

// General quadrilateral / hexahedron: the Gauss-point sum itself, f += sum_g rho (th) |J_g|
// N(g) (N(g) . w). rt = rho th (Quad4) or rho (Hexa8).
__device__ __forceinline__ void quad_mass_gauss(const double (&x)[4][2], double rt, const double (&w)[4][2],
                                                double (&f)[4][2]) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        double r, s, t;
        gauss_point(g, r, s, t);
        double dn[4][2];
        const double sc = rt * quad_dn(x, r, s, dn);
        const double h[4] = {1.0 / 4.0 * (1.0 - r) * (1.0 - s), 1.0 / 4.0 * (1.0 + r) * (1.0 - s),
                             1.0 / 4.0 * (1.0 + r) * (1.0 + s), 1.0 / 4.0 * (1.0 - r) * (1.0 + s)};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double hw = (h[0] * w[0][c] + h[1] * w[1][c]) + (h[2] * w[2][c] + h[3] * w[3][c]);
#pragma unroll
            for (int a = 0; a < 4; ++a) f[a][c] += sc * h[a] * hw;
        }
    }
}

__device__ __forceinline__ void hexa_mass_gauss(const double (&x)[8][3], double rt, const double (&w)[8][3],
                                                double (&f)[8][3]) {
#pragma unroll 1
    for (int g = 0; g < 8; ++g) {
        double r, s, t;
        gauss_point(g, r, s, t);
        double dn[8][3];
        const double sc = rt * hexa_dn(x, r, s, t, dn);
        double h[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) h[a] = 1.0 / 8.0 * (1.0 + sgx(a) * r) * (1.0 + sgy(a) * s) * (1.0 + sgz(a) * t);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double hw = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) hw += h[a] * w[a][c];
#pragma unroll
            for (int a = 0; a < 8; ++a) f[a][c] += sc * h[a] * hw;
        }
    }
}

// Consistent element mass of one tile element added to f: PATH 2, 3 take rs = rho |J| / 8 (Hexa8)
// or rho th |J| / 4 (Quad4) from the tile arrays; PATH 1 forms |J| from the staged coordinates;
// PATH 0 integrates at the Gauss points (rs = rho, times th in 2D).
template <int DIM, int PATH>
__device__ __forceinline__ void tile_elem_mass(const int (&lc)[Elem<DIM>::nen], const double *sX, double rs,
                                               const double (&w)[Elem<DIM>::nen][DIM],
                                               double (&f)[Elem<DIM>::nen][DIM]) {
    constexpr int NEN = Elem<DIM>::nen;
    if constexpr (PATH >= 2) {
        if constexpr (DIM == 3)
            hexa_mass_modal(rs, w, f);
        else
            quad_mass_modal(rs, w, f);
    } else {
        double x[NEN][DIM];
#pragma unroll
        for (int c = 0; c < NEN; ++c)
#pragma unroll
            for (int d = 0; d < DIM; ++d) x[c][d] = sX[lc[c] * DIM + d];
        if constexpr (PATH == 1) {
            double Lam[DIM][DIM];
            if constexpr (DIM == 3)
                hexa_mass_modal(rs * 0.125 * hexa_center_geometry(x, Lam), w, f);
            else
                quad_mass_modal(rs * 0.25 * quad_center_geometry(x, Lam), w, f);
        } else {
            if constexpr (DIM == 3)
                hexa_mass_gauss(x, rs, w, f);
            else
                quad_mass_gauss(x, rs, w, f);
        }
    }
}

// Element force f = K_e u for the operand u of one tile element; lam and mu carry any
// per-element factor (and |J|-scaling on the geometry-table paths). PATH: 0 Gauss points on
// staged coordinates sX, 1 modal form on staged coordinates, 2 modal form with the inverse
// Jacobian from the geometry table, 3 the same with a diagonal inverse Jacobian.
template <int DIM, int PATH>
__device__ __forceinline__ void tile_elem_force(const double *geo_table, double th, const int (&lc)[Elem<DIM>::nen],
                                                const double *sX,
                                           double lam, double mu, int gid, const double (&u)[Elem<DIM>::nen][DIM],
                                           double (&f)[Elem<DIM>::nen][DIM]) {
    constexpr int NEN = Elem<DIM>::nen;
    if constexpr (PATH >= 2) {
        const double *g = geo_table + size_t(gid) * (DIM * DIM);
        double Lam[DIM][DIM];
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int c = 0; c < DIM; ++c) Lam[r][c] = g[r * DIM + c];
        if constexpr (DIM == 3)
            hexa_modal_force<PATH == 3>(Lam, lam, mu, u, f);
        else
            quad_modal_force<PATH == 3>(Lam, lam, mu, u, f);
    } else {
        double x[NEN][DIM];
#pragma unroll
        for (int c = 0; c < NEN; ++c)
#pragma unroll
            for (int d = 0; d < DIM; ++d) x[c][d] = sX[lc[c] * DIM + d];
        if constexpr (DIM == 3) {
            if constexpr (PATH == 1)
                hexa_affine_fint(x, u, lam, mu, f);
            else
                hexa_fint(x, u, lam, mu, f);
        } else {
            if constexpr (PATH == 1)
                quad_affine_fint(x, u, lam, mu, th, f);
            else
                quad_fint(x, u, lam, mu, th, f);
        }
    }
}

}  // namespace svlg
