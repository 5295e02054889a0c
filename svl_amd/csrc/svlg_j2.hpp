// svlg_j2.hpp -- J2 plasticity at one Gauss point on the device: Plastic3DJ2 and
// PlasticPlaneStrainJ2 (RP/02-Materials/02-NonLinear/Plastic3DJ2.cpp:206-258,
// PlasticPlaneStrainJ2.cpp:227-279).
//
// State of a Gauss point, kJ2State fields (tensor components [11, 22, 33, 12, 23, 31]):
//   [0,6) plastic strain, [6,12) back stress, [12] alpha, [13,19) stress,
//   [19,25) unit normal n of the last plastic update, [25] a1, [26] a2,
// stored field-major: field f of Gauss point g of element e at ((f * ngp + g) * nelem + e), so
// the threads of neighbouring elements read neighbouring addresses (J2Ref below).
// where the consistent tangent of the last update is
//   C eps = K tr(eps) (1,1,1,0,0,0) + a1 Id eps - a2 (n . eps) n,
//   Id eps = (eps_i - tr/3 for i < 3, eps_i / 2 for shears), eps engineering,
// i.e. K D + (2G - c2) Id - (2G c1 - c2) n n^T with c1 = 1 / (1 + H / 3G),
// c2 = 4 G^2 dgamma / ||xi|| (SVL's TangentStiffness); a1 = 2G, a2 = 0 when elastic.
// As in SVL, the plastic strain, back stress and alpha advance from their current values on
// every call, so the Newton iterates of one step accumulate.
#pragma once

#include <hip/hip_runtime.h>

#include "svlg_elem.hpp"

namespace svlg {

constexpr int kJ2State = 27;

struct J2Par {
    double K, G, sy, H, beta;
};

// The kJ2State fields of one Gauss point: st[f] = p[f * stride].
template <typename T>
struct J2RefT {
    T *p;
    size_t stride;
    __device__ __host__ T &operator[](int f) const { return p[size_t(f) * stride]; }
};
using J2Ref = J2RefT<double>;
using J2CRef = J2RefT<const double>;

__device__ __host__ inline J2Ref j2_at(double *state, size_t nelem, int ngp, int e, int g) {
    return J2Ref{state + size_t(g) * nelem + e, size_t(ngp) * nelem};
}
__device__ __host__ inline J2CRef j2_at(const double *state, size_t nelem, int ngp, int e, int g) {
    return J2CRef{state + size_t(g) * nelem + e, size_t(ngp) * nelem};
}

__device__ __host__ inline void j2_init_state(const J2Par &p, J2Ref st) {
    for (int i = 0; i < kJ2State; ++i) st[i] = 0.0;
    st[25] = 2.0 * p.G;
}

// One UpdateState(strain, 1) with the tensor strain et.
__device__ inline void j2_update_gp(const J2Par &p, const double (&et)[6], J2Ref st) {
    const double tr = et[0] + et[1] + et[2];
    double ep[6], q[6], strial[6], xi[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        ep[i] = st[i];
        q[i] = st[6 + i];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double one = i < 3 ? 1.0 : 0.0;
        const double dev = et[i] - 1.0 / 3.0 * tr * one;
        strial[i] = 2.0 * p.G * (dev - ep[i]);
        xi[i] = strial[i] - q[i];
    }
    const double nx = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] +
                           2.0 * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]));
    const double F = nx - sqrt(2.0 / 3.0) * (p.sy + st[12] * p.beta * p.H);
    if (F <= 0) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double one = i < 3 ? 1.0 : 0.0;
            st[13 + i] = p.K * tr * one + strial[i];
            st[19 + i] = 0.0;
        }
        st[25] = 2.0 * p.G;
        st[26] = 0.0;
        return;
    }
    const double dg = F / (2.0 * p.G + 2.0 / 3.0 * p.H);
    st[12] += sqrt(2.0 / 3.0) * dg;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double one = i < 3 ? 1.0 : 0.0;
        const double n = xi[i] / nx;
        st[6 + i] = q[i] + 2.0 / 3.0 * (1.0 - p.beta) * p.H * dg * n;
        st[i] = ep[i] + dg * n;
        st[13 + i] = p.K * tr * one + strial[i] - 2.0 * p.G * dg * n;
        st[19 + i] = n;
    }
    const double c1 = 1.0 / (1.0 + p.H / p.G / 3.0);
    const double c2 = 4.0 * p.G * p.G * dg / nx;
    st[25] = 2.0 * p.G - c2;
    st[26] = 2.0 * p.G * c1 - c2;
}

// sigma' = C_t eps for an engineering strain eps (6 components) with the tangent of st.
__device__ inline void j2_tangent_apply(const J2Par &p, J2CRef st, const double (&e)[6], double (&s)[6]) {
    const double tr = e[0] + e[1] + e[2];
    double n[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) n[j] = st[19 + j];
    double ne = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) ne += n[j] * e[j];
    const double a1 = st[25], a2 = st[26];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double id = i < 3 ? e[i] - tr / 3.0 : 0.5 * e[i];
        s[i] = (i < 3 ? p.K * tr : 0.0) + a1 * id - a2 * ne * n[i];
    }
}

// Initial tangent K D + 2G (I - D/3) applied to an engineering strain (GetInitialTangentStiffness).
__device__ inline void j2_initial_apply(const J2Par &p, const double (&e)[6], double (&s)[6]) {
    const double tr = e[0] + e[1] + e[2];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double id = i < 3 ? e[i] - tr / 3.0 : 0.5 * e[i];
        s[i] = (i < 3 ? p.K * tr : 0.0) + 2.0 * p.G * id;
    }
}

// Engineering strain of the element's Voigt order -> tensor strain of the material and the
// 6-component engineering strain the tangent acts on (plane strain: slots 1, 2, 4).
template <int DIM>
__device__ inline void j2_embed(const double *ev, double (&et)[6], double (&e6)[6]) {
#pragma unroll
    for (int i = 0; i < 6; ++i) et[i] = e6[i] = 0.0;
    if constexpr (DIM == 3) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            e6[i] = ev[i];
            et[i] = i < 3 ? ev[i] : 0.5 * ev[i];
        }
    } else {
        e6[1] = ev[0];
        e6[2] = ev[1];
        e6[4] = ev[2];
        et[1] = ev[0];
        et[2] = ev[1];
        et[4] = 0.5 * ev[2];
    }
}

// Stress components of the element's Voigt order from a 6-vector.
template <int DIM>
__device__ inline void j2_extract(const double *s6, double *sv) {
    if constexpr (DIM == 3) {
#pragma unroll
        for (int i = 0; i < 6; ++i) sv[i] = s6[i];
    } else {
        sv[0] = s6[1];
        sv[1] = s6[2];
        sv[2] = s6[4];
    }
}


// Physical shape-function derivatives at Gauss point g of a quad / hexa with corner
// coordinates x; returns w |J| (times the thickness th for the plane-strain quad).
template <int DIM>
__device__ __forceinline__ double gp_dn(const double (&x)[Elem<DIM>::nen][DIM], int g, double th,
                                        double (&dn)[Elem<DIM>::nen][DIM]) {
    double r, s, t;
    gauss_point(g, r, s, t);
    if constexpr (DIM == 2)
        return th * quad_dn(x, r, s, dn);
    else
        return hexa_dn(x, r, s, t, dn);
}

// Engineering strain B u in the element's Voigt order ([xx,yy,xy] / [xx,yy,zz,xy,yz,zx]).
template <int DIM>
__device__ __forceinline__ void gp_strain(const double (&dn)[Elem<DIM>::nen][DIM],
                                          const double (&u)[Elem<DIM>::nen][DIM], double *ev) {
    constexpr int NEN = Elem<DIM>::nen;
    if constexpr (DIM == 2) {
        ev[0] = ev[1] = ev[2] = 0.0;
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            ev[0] += dn[a][0] * u[a][0];
            ev[1] += dn[a][1] * u[a][1];
            ev[2] += dn[a][1] * u[a][0] + dn[a][0] * u[a][1];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) ev[i] = 0.0;
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            ev[0] += dn[a][0] * u[a][0];
            ev[1] += dn[a][1] * u[a][1];
            ev[2] += dn[a][2] * u[a][2];
            ev[3] += dn[a][1] * u[a][0] + dn[a][0] * u[a][1];
            ev[4] += dn[a][2] * u[a][1] + dn[a][1] * u[a][2];
            ev[5] += dn[a][2] * u[a][0] + dn[a][0] * u[a][2];
        }
    }
}

// f += wj B^T sigma (sigma in the element's Voigt order).
template <int DIM>
__device__ __forceinline__ void gp_btsig(const double (&dn)[Elem<DIM>::nen][DIM], const double *sv, double wj,
                                         double (&f)[Elem<DIM>::nen][DIM]) {
    constexpr int NEN = Elem<DIM>::nen;
#pragma unroll
    for (int a = 0; a < NEN; ++a) {
        if constexpr (DIM == 2) {
            f[a][0] += wj * (dn[a][0] * sv[0] + dn[a][1] * sv[2]);
            f[a][1] += wj * (dn[a][1] * sv[1] + dn[a][0] * sv[2]);
        } else {
            f[a][0] += wj * (dn[a][0] * sv[0] + dn[a][1] * sv[3] + dn[a][2] * sv[5]);
            f[a][1] += wj * (dn[a][1] * sv[1] + dn[a][0] * sv[3] + dn[a][2] * sv[4]);
            f[a][2] += wj * (dn[a][2] * sv[2] + dn[a][1] * sv[4] + dn[a][0] * sv[5]);
        }
    }
}

// Element force of a J2 element from its Gauss-point states:
//  RHS (tangent = false): f = sum_g wj B^T (sigma_g + c D0 B w)          (c = -ak)
//  TAN (tangent = true):  f = sum_g wj B^T (C_t,g B u + c D0 B u)       (c = 2 ak / dt)
// D0 is the initial tangent; u is the staged operand (p for TAN, ignored for RHS), w the
// Rayleigh operand (V - 2/dt dU for RHS).
template <int DIM>
__device__ inline void j2_elem_force(const double (&x)[Elem<DIM>::nen][DIM], const double (&u)[Elem<DIM>::nen][DIM],
                                     const J2Par &p, const double *state, size_t nelem, int e, double c,
                                     bool tangent, double th, double (&f)[Elem<DIM>::nen][DIM]) {
    constexpr int NEN = Elem<DIM>::nen, NV = DIM == 2 ? 3 : 6, NGP = Elem<DIM>::ngp;
#pragma unroll
    for (int a = 0; a < NEN; ++a)
#pragma unroll
        for (int d = 0; d < DIM; ++d) f[a][d] = 0.0;
#pragma unroll 1
    for (int g = 0; g < NGP; ++g) {
        double dn[NEN][DIM];
        const double wj = gp_dn<DIM>(x, g, th, dn);
        double ev[NV], et[6], e6[6], s6[6], i6[6], sv[NV];
        gp_strain<DIM>(dn, u, ev);
        j2_embed<DIM>(ev, et, e6);
        j2_initial_apply(p, e6, i6);
        const J2CRef sg = j2_at(state, nelem, NGP, e, g);
        if (tangent) {
            j2_tangent_apply(p, sg, e6, s6);
#pragma unroll
            for (int i = 0; i < 6; ++i) s6[i] += c * i6[i];
        } else {
#pragma unroll
            for (int i = 0; i < 6; ++i) s6[i] = sg[13 + i] + c * i6[i];
        }
        j2_extract<DIM>(s6, sv);
        gp_btsig<DIM>(dn, sv, wj, f);
    }
}

// ---------------------------------------------------------------- axis-aligned Hexa8 bricks
// For a brick aligned with the axes (J diagonal, J^-1 = diag(lx, ly, lz), |J| = 1/(lx ly lz))
// the shape-function derivatives at a Gauss point follow from the corner sums
// S_m = sum_a m_a u_a over the sign patterns m = xi, eta, zeta, xi eta, eta zeta, zeta xi,
// xi eta zeta of SVL's corner order: du/dxi = (S_x + eta S_xy + zeta S_zx + eta zeta S_xyz) / 8
// and cyclically; the same patterns carry B^T sigma back to the corners. This is the Gauss
// loop of lin3DHexa8 (B from J^-1 at each point) evaluated without forming J or B.

// Corner sums S[m][c], m = x, y, z, xy, yz, zx, xyz (Walsh-Hadamard butterfly, 23 adds per c).
__device__ __forceinline__ void brick_sums(const double (&u)[8][3], double (&S)[7][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        // v[i][j][k]: corner with xi bit i, eta bit j, zeta bit k (SVL order 0..7).
        const double v000 = u[0][c], v100 = u[1][c], v110 = u[2][c], v010 = u[3][c];
        const double v001 = u[4][c], v101 = u[5][c], v111 = u[6][c], v011 = u[7][c];
        const double a00 = v100 + v000, b00 = v100 - v000, a10 = v110 + v010, b10 = v110 - v010;
        const double a01 = v101 + v001, b01 = v101 - v001, a11 = v111 + v011, b11 = v111 - v011;
        const double aa0 = a10 + a00, ab0 = a10 - a00, ba0 = b10 + b00, bb0 = b10 - b00;
        const double aa1 = a11 + a01, ab1 = a11 - a01, ba1 = b11 + b01, bb1 = b11 - b01;
        S[0][c] = ba1 + ba0;  // x
        S[1][c] = ab1 + ab0;  // y
        S[2][c] = aa1 - aa0;  // z
        S[3][c] = bb1 + bb0;  // xy
        S[4][c] = ab1 - ab0;  // yz
        S[5][c] = ba1 - ba0;  // zx
        S[6][c] = bb1 - bb0;  // xyz
    }
}

// Engineering strain [xx, yy, zz, xy, yz, zx] at the Gauss point (r, s, t) from the corner sums;
// l8 = (lx, ly, lz) / 8.
__device__ __forceinline__ void brick_strain(const double (&S)[7][3], const double (&l8)[3], double r, double s,
                                             double t, double (&ev)[6]) {
    double gx[3], gy[3], gz[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gx[c] = l8[0] * (((S[0][c] + s * S[3][c]) + t * S[5][c]) + (s * t) * S[6][c]);
        gy[c] = l8[1] * (((S[1][c] + r * S[3][c]) + t * S[4][c]) + (r * t) * S[6][c]);
        gz[c] = l8[2] * (((S[2][c] + s * S[4][c]) + r * S[5][c]) + (r * s) * S[6][c]);
    }
    ev[0] = gx[0];
    ev[1] = gy[1];
    ev[2] = gz[2];
    ev[3] = gy[0] + gx[1];
    ev[4] = gz[1] + gy[2];
    ev[5] = gz[0] + gx[2];
}

// f[a][c] = (|J| / 8) sum_m m_a F[m][c] (the transpose of brick_sums).
__device__ __forceinline__ void brick_corners(const double (&F)[7][3], double scale, double (&f)[8][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double aa1 = F[2][c], aa0 = -F[2][c];
        const double ab1 = F[1][c] + F[4][c], ab0 = F[1][c] - F[4][c];
        const double ba1 = F[0][c] + F[5][c], ba0 = F[0][c] - F[5][c];
        const double bb1 = F[3][c] + F[6][c], bb0 = F[3][c] - F[6][c];
        const double a00 = aa0 - ab0, a10 = aa0 + ab0, b00 = ba0 - bb0, b10 = ba0 + bb0;
        const double a01 = aa1 - ab1, a11 = aa1 + ab1, b01 = ba1 - bb1, b11 = ba1 + bb1;
        f[0][c] = scale * (a00 - b00);
        f[1][c] = scale * (a00 + b00);
        f[2][c] = scale * (a10 + b10);
        f[3][c] = scale * (a10 - b10);
        f[4][c] = scale * (a01 - b01);
        f[5][c] = scale * (a01 + b01);
        f[6][c] = scale * (a11 + b11);
        f[7][c] = scale * (a11 - b11);
    }
}

// j2_elem_force for an axis-aligned Hexa8 brick with J^-1 = diag(l): the same sum over the
// eight Gauss points (weights 1, |J| = 1 / (lx ly lz)).
__device__ inline void j2_brick_force(const double (&l)[3], const double (&u)[8][3], const J2Par &p,
                                      const double *state, size_t nelem, int e, double c, bool tangent,
                                      double (&f)[8][3]) {
    double S[7][3], F[7][3];
    brick_sums(u, S);
#pragma unroll
    for (int m = 0; m < 7; ++m)
#pragma unroll
        for (int k = 0; k < 3; ++k) F[m][k] = 0.0;
    const double l8[3] = {0.125 * l[0], 0.125 * l[1], 0.125 * l[2]};
#pragma unroll 1
    for (int g = 0; g < 8; ++g) {
        double r, s, t;
        gauss_point(g, r, s, t);
        double ev[6], et[6], e6[6], s6[6], i6[6];
        brick_strain(S, l8, r, s, t, ev);
        j2_embed<3>(ev, et, e6);
        j2_initial_apply(p, e6, i6);
        const J2CRef sg = j2_at(state, nelem, 8, e, g);
        if (tangent) {
            j2_tangent_apply(p, sg, e6, s6);
#pragma unroll
            for (int i = 0; i < 6; ++i) s6[i] += c * i6[i];
        } else {
#pragma unroll
            for (int i = 0; i < 6; ++i) s6[i] = sg[13 + i] + c * i6[i];
        }
        // Directional stresses of component k: (sigma_xk, sigma_yk, sigma_zk).
        const double tx[3] = {s6[0], s6[3], s6[5]};
        const double ty[3] = {s6[3], s6[1], s6[4]};
        const double tz[3] = {s6[5], s6[4], s6[2]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double X = l[0] * tx[k], Y = l[1] * ty[k], Z = l[2] * tz[k];
            F[0][k] += X;
            F[1][k] += Y;
            F[2][k] += Z;
            F[3][k] += s * X + r * Y;
            F[4][k] += t * Y + s * Z;
            F[5][k] += t * X + r * Z;
            F[6][k] += ((s * t) * X + (r * t) * Y) + (r * s) * Z;
        }
    }
    brick_corners(F, 0.125 / ((l[0] * l[1]) * l[2]), f);
}

}  // namespace svlg
