// svlg_pml.hip -- perfectly matched layers on the device: PML2DQuad4 / lin2DQuad4 (nodes of 5
// or 2 unknowns) or PML3DHexa8 / lin3DHexa8 (9 or 3), tied by EQUAL constraints, integrated
// with ExtendedNewmarkBeta + Linear, including the 3D history matrix G and the integrated
// displacement Ubar (the 2D element has none).
//
// Set-up: one thread per element-matrix entry computes M, C, K and G (node-major, nd x nd
// with nd = nodes x unknowns, stored per element at its own offset, the four values of an
// entry side by side); the free-space matrices T^T A T are summed per nonzero in
// element order from a host-built contribution list (pure index work), and Keff is scaled
// symmetrically by |diag|^-1/2. A step is one fused right-hand-side pass over the three
// matrices, a MINRES solve of the scaled system (three kernels per iteration, the scalar
// recurrence carried in device memory, convergence polled every few iterations) and the
// Newmark update. Keff is indefinite (the stress blocks are negative definite), so CG does
// not apply; SVL factorises it with Eigen (11-Solvers/EigenSolver.cpp).
#include "svlg_common.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace svlg {
namespace {

constexpr int VB = 256;             // block of the vector and row kernels
constexpr int MAXB = 1024;          // blocks of a reduction (partials per pass)
constexpr int POLL = 8;             // MINRES iterations between convergence checks
constexpr int LPR = 8;              // lanes per matrix row in the row kernels
constexpr double GA = 0.577350269189626;  // two-point Gauss abscissa (GaussQuadrature.cpp:303-316)

struct MrState {
    double beta, cp, sp, c, s, eta, beta1;
    int32_t done, iters, fail, pad;
};

__host__ __device__ inline int class_nun(int cls) { return cls == 0 ? 2 : cls == 1 ? 5 : cls == 2 ? 3 : 9; }
__host__ __device__ inline int class_nen(int cls) { return cls < 2 ? 4 : 8; }

struct EmatArgs {
    int nelem, lumped;
    const int32_t *cls, *conn;
    const int64_t *eoff;  // [nelem+1] first entry of each element (entries of nd x nd)
    const double *X, *E, *nu, *rho, *th, *pml;
    double *out;          // [entries][4]: M, C, K, G of each entry
};

struct Coef4 {
    double m, c, k, g;
};

// PML2DQuad4 / lin2DQuad4 entry (a p, b q) of M, C, K summed over the 2 x 2 Gauss points in
// SVL's order (r fastest). PML2DQuad4: PML2DQuad4.cpp:233-466 with the stretching factors of
// :655-693 (ax ay, ax by + ay bx, bx by on the mass-like blocks of M, C, K; ay / ax and by / bx
// on the displacement-stress coupling of C and K). lin2DQuad4: plane-strain stiffness and the
// consistent or row-sum mass (lin2DQuad4.cpp:243-318).
__device__ Coef4 entry2d(const EmatArgs &a, int e, bool pml, int na, int p, int nb, int q) {
    double x[4], y[4];
    for (int c = 0; c < 4; ++c) {
        const int n = a.conn[4 * e + c];
        x[c] = a.X[2 * n];
        y[c] = a.X[2 * n + 1];
    }
    const double E = a.E[e], nu = a.nu[e], rho = a.rho[e], th = a.th[e];
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    Coef4 r4{0.0, 0.0, 0.0, 0.0};
    for (int g = 0; g < 4; ++g) {
        const double r = (g & 1) ? GA : -GA, s = (g & 2) ? GA : -GA;
        const double N[4] = {0.25 * (1 - r) * (1 - s), 0.25 * (1 + r) * (1 - s), 0.25 * (1 + r) * (1 + s),
                             0.25 * (1 - r) * (1 + s)};
        const double Nr[4] = {-0.25 * (1 - s), 0.25 * (1 - s), 0.25 * (1 + s), -0.25 * (1 + s)};
        const double Ns[4] = {-0.25 * (1 - r), -0.25 * (1 + r), 0.25 * (1 + r), 0.25 * (1 - r)};
        double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
        for (int c = 0; c < 4; ++c) {
            J00 += Nr[c] * x[c]; J01 += Nr[c] * y[c];
            J10 += Ns[c] * x[c]; J11 += Ns[c] * y[c];
        }
        const double det = J00 * J11 - J01 * J10;
        const double w = th * fabs(det);
        const double da[2] = {(J11 * Nr[na] - J01 * Ns[na]) / det, (-J10 * Nr[na] + J00 * Ns[na]) / det};
        const double db[2] = {(J11 * Nr[nb] - J01 * Ns[nb]) / det, (-J10 * Nr[nb] + J00 * Ns[nb]) / det};
        if (!pml) {
            // K_ab,pq = lambda dNa_p dNb_q + mu dNa_q dNb_p + mu delta_pq grad Na . grad Nb
            const double gg = da[0] * db[0] + da[1] * db[1];
            r4.k += w * (lam * da[p] * db[q] + mu * da[q] * db[p] + (p == q ? mu * gg : 0.0));
            if (a.lumped) {
                if (na == nb && p == q) r4.m += rho * N[na] * w;
            } else if (p == q) {
                r4.m += rho * N[na] * N[nb] * w;
            }
            continue;
        }
        double xg = 0, yg = 0;
        for (int c = 0; c < 4; ++c) {
            xg += N[c] * x[c];
            yg += N[c] * y[c];
        }
        const double *pp = a.pml + 9 * (size_t)e;  // m, L, R, x0, y0, z0, nx, ny, nz
        const double m = pp[0], L = pp[1], R = pp[2];
        const double vp = sqrt((lam + 2.0 * mu) / rho);
        const double a0 = (m + 1.0) * (L / 10.0) / 2.0 / L * log(1.0 / R);
        const double b0 = (m + 1.0) * vp / 2.0 / L * log(1.0 / R);
        const double fx = pow((xg - pp[3]) * pp[6] / L, m), fy = pow((yg - pp[4]) * pp[7] / L, m);
        const double ax = 1.0 + a0 * fx, ay = 1.0 + a0 * fy, bx = b0 * fx, by = b0 * fy;
        if ((p < 2) == (q < 2)) {
            double bv = 0.0;
            if (p == q) bv = p < 2 ? rho : (p < 4 ? -(lam + 2.0 * mu) / (4.0 * mu * (lam + mu)) : -1.0 / mu);
            else if (p >= 2 && q >= 2 && p < 4 && q < 4) bv = lam / (4.0 * mu * (lam + mu));
            const double nn = bv * N[na] * N[nb] * w;
            r4.m += ax * ay * nn;
            r4.c += (ax * by + ay * bx) * nn;
            r4.k += bx * by * nn;
        } else {
            // u_x with xx (d/dx) and xy (d/dy), u_y with yy (d/dy) and xy (d/dx); the derivative
            // falls on the displacement node's shape function.
            const int pu = p < 2 ? p : q, ps = p < 2 ? q : p;
            const double *du = p < 2 ? da : db;
            const double ns = p < 2 ? N[nb] : N[na];
            if ((pu == 0 && ps == 2) || (pu == 1 && ps == 4)) {
                const double v = du[0] * ns * w;
                r4.c += ay * v;
                r4.k += by * v;
            } else if ((pu == 0 && ps == 4) || (pu == 1 && ps == 3)) {
                const double v = du[1] * ns * w;
                r4.c += ax * v;
                r4.k += bx * v;
            }
        }
    }
    return r4;
}

// PML3DHexa8 / lin3DHexa8 entry (a p, b q) of M, C, K, G over the 2 x 2 x 2 Gauss points.
// PML3DHexa8.cpp:214-715 with the stretching of :952-996: mass-like blocks (rho on the
// displacements, minus the compliance on xx, yy, zz and 1/mu on xy, yz, xz) times ax ay az (M),
// ax ay bz + ax by az + bx ay az (C), ax by bz + bx by az + bx ay bz (K), bx by bz (G); the
// displacement-stress coupling dN/dx_d N with, for d = x: ay az (C), ay bz + az by (K), by bz
// (G), and likewise for y and z. lin3DHexa8: isotropic stiffness and consistent / row-sum mass.
__device__ Coef4 entry3d(const EmatArgs &a, int e, bool pml, int na, int p, int nb, int q) {
    const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
              sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double X[8][3];
    for (int c = 0; c < 8; ++c) {
        const int n = a.conn[8 * e + c];
        for (int d = 0; d < 3; ++d) X[c][d] = a.X[3 * n + d];
    }
    const double E = a.E[e], nu = a.nu[e], rho = a.rho[e];
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    Coef4 r4{0.0, 0.0, 0.0, 0.0};
    for (int g = 0; g < 8; ++g) {
        const double r = (g & 1) ? GA : -GA, s = (g & 2) ? GA : -GA, t = (g & 4) ? GA : -GA;
        double N[8], D[3][8];
        for (int c = 0; c < 8; ++c) {
            const double fr = 1 + sx[c] * r, fs = 1 + sy[c] * s, ft = 1 + sz[c] * t;
            N[c] = 0.125 * fr * fs * ft;
            D[0][c] = 0.125 * sx[c] * fs * ft;
            D[1][c] = 0.125 * sy[c] * fr * ft;
            D[2][c] = 0.125 * sz[c] * fr * fs;
        }
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // J[k][j] = d x_j / d xi_k
        for (int c = 0; c < 8; ++c)
            for (int k = 0; k < 3; ++k)
                for (int j = 0; j < 3; ++j) J[k][j] += D[k][c] * X[c][j];
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                     c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        const double id = 1.0 / det;
        // inverse: Ji[j][k] = d xi_k / d x_j
        const double Ji[3][3] = {
            {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
            {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
            {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
        const double w = fabs(det);
        double da[3], db[3];
        for (int j = 0; j < 3; ++j) {
            da[j] = Ji[j][0] * D[0][na] + Ji[j][1] * D[1][na] + Ji[j][2] * D[2][na];
            db[j] = Ji[j][0] * D[0][nb] + Ji[j][1] * D[1][nb] + Ji[j][2] * D[2][nb];
        }
        if (!pml) {
            const double gg = da[0] * db[0] + da[1] * db[1] + da[2] * db[2];
            r4.k += w * (lam * da[p] * db[q] + mu * da[q] * db[p] + (p == q ? mu * gg : 0.0));
            if (a.lumped) {
                if (na == nb && p == q) r4.m += rho * N[na] * w;
            } else if (p == q) {
                r4.m += rho * N[na] * N[nb] * w;
            }
            continue;
        }
        double xg[3] = {0, 0, 0};
        for (int c = 0; c < 8; ++c)
            for (int d = 0; d < 3; ++d) xg[d] += N[c] * X[c][d];
        const double *pp = a.pml + 9 * (size_t)e;  // m, L, R, x0, y0, z0, nx, ny, nz
        const double m = pp[0], L = pp[1], R = pp[2];
        const double vp = sqrt((lam + 2.0 * mu) / rho);
        const double a0 = (m + 1.0) * (L / 10.0) / 2.0 / L * log(1.0 / R);
        const double b0 = (m + 1.0) * vp / 2.0 / L * log(1.0 / R);
        double sa[3], sb[3];
        for (int d = 0; d < 3; ++d) {
            const double f = pow((xg[d] - pp[3 + d]) * pp[6 + d] / L, m);
            sa[d] = 1.0 + a0 * f;
            sb[d] = b0 * f;
        }
        if ((p < 3) == (q < 3)) {
            double bv = 0.0;
            if (p == q) bv = p < 3 ? rho : (p < 6 ? -(lam + mu) / mu / (3.0 * lam + 2.0 * mu) : -1.0 / mu);
            else if (p >= 3 && q >= 3 && p < 6 && q < 6) bv = lam / 2.0 / mu / (3.0 * lam + 2.0 * mu);
            const double nn = bv * N[na] * N[nb] * w;
            r4.m += sa[0] * sa[1] * sa[2] * nn;
            r4.c += (sa[0] * sa[1] * sb[2] + sa[0] * sb[1] * sa[2] + sb[0] * sa[1] * sa[2]) * nn;
            r4.k += (sa[0] * sb[1] * sb[2] + sb[0] * sb[1] * sa[2] + sb[0] * sa[1] * sb[2]) * nn;
            r4.g += sb[0] * sb[1] * sb[2] * nn;
        } else {
            const int pu = p < 3 ? p : q, ps = p < 3 ? q : p;
            // derivative direction of the pair: x (0,3) (1,6) (2,8); y (0,6) (1,4) (2,7);
            // z (0,8) (1,7) (2,5)
            int d = -1;
            if ((pu == 0 && ps == 3) || (pu == 1 && ps == 6) || (pu == 2 && ps == 8)) d = 0;
            else if ((pu == 0 && ps == 6) || (pu == 1 && ps == 4) || (pu == 2 && ps == 7)) d = 1;
            else if ((pu == 0 && ps == 8) || (pu == 1 && ps == 7) || (pu == 2 && ps == 5)) d = 2;
            if (d >= 0) {
                const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;  // the two other directions
                const double v = (p < 3 ? da[d] : db[d]) * (p < 3 ? N[nb] : N[na]) * w;
                r4.c += sa[d1] * sa[d2] * v;
                r4.k += (sa[d1] * sb[d2] + sa[d2] * sb[d1]) * v;
                r4.g += sb[d1] * sb[d2] * v;
            }
        }
    }
    return r4;
}

__global__ void __launch_bounds__(VB) k_pml_emat(EmatArgs a, int64_t nent) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nent) return;
    int lo = 0, hi = a.nelem;  // the element whose entry range holds t
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (a.eoff[mid] <= t) lo = mid;
        else hi = mid;
    }
    const int e = lo, cls = a.cls[e], nun = class_nun(cls), nd = class_nen(cls) * nun;
    const int ij = (int)(t - a.eoff[e]), i = ij / nd, j = ij % nd;
    const bool pml = cls == SVLG_CLASS_PML2DQUAD4 || cls == SVLG_CLASS_PML3DHEXA8;
    const Coef4 v = cls < 2 ? entry2d(a, e, pml, i / nun, i % nun, j / nun, j % nun)
                            : entry3d(a, e, pml, i / nun, i % nun, j / nun, j % nun);
    double *o = a.out + 4 * t;
    o[0] = v.m;
    o[1] = v.c;
    o[2] = v.k;
    o[3] = v.g;
}

// One thread per free-space nonzero: its contributions (element-matrix offset, product of
// the two T entries) in element order, as Assembler sums its triplets.
__global__ void __launch_bounds__(VB) k_pml_assemble(int64_t nnz, const int64_t *cptr, const int64_t *cidx,
                                                     const double *cfac, const double *emat, double c0, double c1,
                                                     double c3, double *vM, double *vC, double *vK, double *vG,
                                                     double *vE) {
    for (int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (int64_t)gridDim.x * blockDim.x) {
        double m = 0.0, c = 0.0, k = 0.0, g = 0.0;
        for (int64_t t = cptr[z]; t < cptr[z + 1]; ++t) {
            const double *b = emat + 4 * cidx[t];
            const double f = cfac[t];
            m += f * b[0];
            c += f * b[1];
            k += f * b[2];
            g += f * b[3];
        }
        vM[z] = m;
        vC[z] = c;
        vK[z] = k;
        if (vG) vG[z] = g;
        vE[z] = k + c0 * m + c1 * c + c3 * g;
    }
}

__global__ void __launch_bounds__(VB) k_pml_scale(int n, const int32_t *rowptr, const int32_t *col, const double *vE,
                                                  double *S, int32_t *bad) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double d = 0.0;
        for (int z = rowptr[i]; z < rowptr[i + 1]; ++z)
            if (col[z] == i) d = vE[z];
        if (d == 0.0) {
            bad[0] = 1;
            S[i] = 1.0;
        } else {
            S[i] = 1.0 / sqrt(fabs(d));
        }
    }
}

__global__ void __launch_bounds__(VB) k_pml_scaled(int n, const int32_t *rowptr, const int32_t *col, const double *vE,
                                                   const double *S, double *vA) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int z = rowptr[i]; z < rowptr[i + 1]; ++z) vA[z] = S[i] * vE[z] * S[col[z]];
}

// Block sum of v into part[blockIdx.x] (deterministic tree).
__device__ inline void block_sum(double v, double *part) {
    __shared__ double sh[VB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = VB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Sum of the nb partials, the same value and order in every block.
__device__ inline double reduce_parts(const double *part, int nb) {
    __shared__ double sh[VB];
    __shared__ double res;
    double v = 0.0;
    for (int t = threadIdx.x; t < nb; t += VB) v += part[t];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = VB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) res = sh[0];
    __syncthreads();
    return res;
}

struct RhsArgs {
    int n, k, nt;
    const int32_t *rowptr, *col, *lptr, *lrow;
    const double *vM, *vC, *vK, *vG, *lfac, *series, *S, *U, *V, *A, *Ub;  // vG, Ub: NULL without G
    double c2, dt;
    double *r, *part;
};

// Scaled effective load S T^T (F_k - K U - G (Ubar + dt U + dt^2/6 V) + C V + M (4/dt V + A))
// (ExtendedNewmarkBeta.cpp:208-245) and the partial sums of its square.
__global__ void __launch_bounds__(VB) k_pml_rhs(RhsArgs a) {
    double acc = 0.0;
    const int lane = threadIdx.x % LPR;
    const int64_t nthr = (int64_t)a.n * LPR;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t - lane < nthr;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / LPR);
        double b = 0.0;
        for (int z = a.rowptr[i] + lane; z < a.rowptr[i + 1]; z += LPR) {
            const int c = a.col[z];
            b += -a.vK[z] * a.U[c] + a.vC[z] * a.V[c] + a.vM[z] * (a.c2 * a.V[c] + a.A[c]);
            if (a.vG) b -= a.vG[z] * (a.Ub[c] + a.dt * a.U[c] + a.dt * a.dt / 6.0 * a.V[c]);
        }
        for (int o = LPR / 2; o > 0; o >>= 1) b += __shfl_xor(b, o, LPR);
        if (lane == 0) {
            if (a.k < a.nt)
                for (int l = a.lptr[i]; l < a.lptr[i + 1]; ++l)
                    b += a.lfac[l] * a.series[(size_t)a.lrow[l] * a.nt + a.k];
            const double r = a.S[i] * b;
            a.r[i] = r;
            acc += r * r;
        }
    }
    block_sum(acc, a.part);
}

__global__ void __launch_bounds__(VB) k_mr_init(int n, int nb, const double *part, const double *r, double *v,
                                                double *vprev, double *w, double *wprev, double *x, double rtol,
                                                MrState *st) {
    const double beta1 = sqrt(reduce_parts(part, nb));
    const double inv = beta1 > 0.0 ? 1.0 / beta1 : 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[i] = r[i] * inv;
        vprev[i] = 0.0;
        w[i] = 0.0;
        wprev[i] = 0.0;
        x[i] = 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        MrState s{};
        s.beta = 0.0; s.cp = 1.0; s.sp = 0.0; s.c = 1.0; s.s = 0.0; s.eta = beta1; s.beta1 = beta1;
        s.done = beta1 == 0.0 ? 1 : 0;
        s.iters = 0;
        s.fail = 0;
        st[0] = s;
        st[1] = s;
    }
}

// MINRES (Paige-Saunders) iteration, part 1: Av = A v and the partials of v.Av; LPR lanes
// per row stride through its nonzeros and combine by shuffles.

__global__ void __launch_bounds__(VB) k_mr_spmv(int n, const int32_t *rowptr, const int32_t *col, const double *vA,
                                                const double *v, double *Av, double *part, const MrState *st) {
    if (st->done) return;
    double acc = 0.0;
    const int lane = threadIdx.x % LPR;
    const int64_t nthr = (int64_t)n * LPR;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t - lane < nthr;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / LPR);
        double s = 0.0;
        if (i < n)
            for (int z = rowptr[i] + lane; z < rowptr[i + 1]; z += LPR) s += vA[z] * v[col[z]];
        for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
        if (lane == 0 && i < n) {
            Av[i] = s;
            acc += v[i] * s;
        }
    }
    block_sum(acc, part);
}

// Part 2: the Lanczos vector Av - alpha v - beta vprev (in Av) and the partials of its square.
__global__ void __launch_bounds__(VB) k_mr_lanczos(int n, int nbs, const double *part1, const double *v,
                                                   const double *vprev, double *Av, double *part2, const MrState *st) {
    if (st->done) return;
    const double alpha = reduce_parts(part1, nbs), beta = st->beta;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double u = Av[i] - alpha * v[i] - beta * vprev[i];
        Av[i] = u;
        acc += u * u;
    }
    block_sum(acc, part2);
}

// Part 3: the Givens rotations, the new search direction (into wprev's buffer), the iterate,
// and the normalised Lanczos vector (into vprev's buffer); the host then swaps the pairs.
// Block 0 writes the scalar state of the next iteration into the other slot.
__global__ void __launch_bounds__(VB) k_mr_update(int n, int nbs, int nb, const double *part1, const double *part2,
                                                  const double *v, double *vprev, const double *Av, const double *w,
                                                  double *wprev, double *x, double rtol, const MrState *st,
                                                  MrState *nxt) {
    const MrState s = *st;
    if (s.done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = s;
        return;
    }
    const double alpha = reduce_parts(part1, nbs), bn = sqrt(reduce_parts(part2, nb));
    const double eps = s.sp * s.beta, dtmp = s.cp * s.beta;
    const double delta = s.c * dtmp + s.s * alpha, gt = -s.s * dtmp + s.c * alpha;
    const double gam = hypot(gt, bn);
    if (gam == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            MrState f = s;
            f.done = 1;
            f.fail = 1;
            *nxt = f;
        }
        return;
    }
    const double cn = gt / gam, sn = bn / gam, step = cn * s.eta, ig = 1.0 / gam;
    const double ib = bn > 0.0 ? 1.0 / bn : 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double wn = (v[i] - delta * w[i] - eps * wprev[i]) * ig;
        x[i] += step * wn;
        wprev[i] = wn;
        vprev[i] = Av[i] * ib;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        MrState t = s;
        t.beta = bn; t.cp = s.c; t.sp = s.s; t.c = cn; t.s = sn; t.eta = -sn * s.eta;
        t.iters = s.iters + 1;
        t.done = (fabs(t.eta) <= rtol * s.beta1 || bn == 0.0) ? 1 : 0;
        *nxt = t;
    }
}

// dU = S x, then ExtendedNewmarkBeta's update (ExtendedNewmarkBeta.cpp:123-154):
// Ubar += dt U + dt/3 dU + dt^2/6 V (old U, V; only with a history matrix), U += dU,
// A = 4/dt^2 dU - 4/dt V - A, V = 2/dt dU - V.
__global__ void __launch_bounds__(VB) k_pml_update(int n, const double *S, const double *x, double c0, double c1,
                                                   double c2, double dt, double *U, double *V, double *A,
                                                   double *Ub) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double du = S[i] * x[i], v = V[i];
        if (Ub) Ub[i] += dt * U[i] + dt / 3.0 * du + dt * dt / 6.0 * v;
        U[i] += du;
        A[i] = c0 * du - c2 * v - A[i];
        V[i] = c1 * du - v;
    }
}

// Total-DOF values (T u) of the recorded DOFs into row s of the histories.
__global__ void __launch_bounds__(VB) k_pml_record(int nrec, const int32_t *rec, const int32_t *dfree,
                                                   const double *dfac, const double *U, const double *V,
                                                   const double *A, double *hU, double *hV, double *hA) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int d = rec[r], f = dfree[d];
    const double g = f >= 0 ? dfac[d] : 0.0;
    hU[r] = f >= 0 ? g * U[f] : 0.0;
    hV[r] = f >= 0 ? g * V[f] : 0.0;
    hA[r] = f >= 0 ? g * A[f] : 0.0;
}

// Whether unknowns p, q of two nodes of an element of class cls can couple. PML elements couple
// a displacement only with itself (mass-like blocks) and with the stress unknowns of its
// derivative pairs; stresses couple among themselves only within the normal block or on the
// diagonal. Pairs outside this pattern are exact zeros in M, C, K and G and stay out of the
// sparse matrices (their contributions would add 0).
bool pair_can_couple(int cls, int p, int q) {
    if (cls == SVLG_CLASS_LIN2DQUAD4 || cls == SVLG_CLASS_LIN3DHEXA8) return true;
    const int nd = cls == SVLG_CLASS_PML2DQUAD4 ? 2 : 3;  // displacement unknowns
    const int nn = cls == SVLG_CLASS_PML2DQUAD4 ? 4 : 6;  // end of the normal-stress block
    const bool up = p < nd, uq = q < nd;
    if (up && uq) return p == q;
    if (!up && !uq) return p == q || (p < nn && q < nn);
    const int pu = up ? p : q, ps = up ? q : p;
    if (nd == 2) return (pu == 0 && (ps == 2 || ps == 4)) || (pu == 1 && (ps == 3 || ps == 4));
    return (pu == 0 && (ps == 3 || ps == 6 || ps == 8)) || (pu == 1 && (ps == 4 || ps == 6 || ps == 7)) ||
           (pu == 2 && (ps == 5 || ps == 7 || ps == 8));
}

int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(MAXB, (n + VB - 1) / VB)); }

}  // namespace
}  // namespace svlg

using namespace svlg;

struct svlg_pml {
    hipStream_t stream = nullptr;
    int nfree = 0, ntot = 0, nnodes = 0, nelem = 0, nb = 1, nbs = 1, nt = 0, maxit = 0;
    int64_t nnz = 0, ncontrib = 0;
    double dt = 0, rtol = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    bool has_g = false;  // a PML3DHexa8 element: history matrix G and Ubar
    DevBuf<int32_t> rowptr, col, lptr, lrow, dfree;
    DevBuf<double> dfac, lfac, series;
    DevBuf<double> vM, vC, vK, vG, vA, S, U, V, A, Ub, r, x, Av, vb0, vb1, wb0, wb1, part0, part1, part2;
    DevBuf<MrState> st;
    MrState *hst = nullptr;  // pinned copy of the polled state
    int last_iters = 0;              // MINRES iterations of the previous solve
    bool use_graph = true;           // replay the batches from a captured graph (else eager)
    hipGraphExec_t batch = nullptr;  // POLL iterations, captured once (the buffer pairs swap
                                     // every iteration, so an even batch ends where it began)
    ~svlg_pml() {
        if (batch) (void)hipGraphExecDestroy(batch);
        if (hst) (void)hipHostFree(hst);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static_assert(POLL % 2 == 0, "a batch must end with the buffer pairs where it began");

// POLL MINRES iterations starting from slot 0 and the buffer pairs (vb0, vb1), (wb0, wb1).
static void launch_batch(svlg_pml *p, hipStream_t s) {
    const int n = p->nfree, nb = p->nb, nbs = p->nbs;
    double *v = p->vb0.p, *vp = p->vb1.p, *w = p->wb0.p, *wp = p->wb1.p;
    for (int j = 0; j < POLL; ++j) {
        const MrState *cur = p->st.p + (j & 1);
        MrState *nxt = p->st.p + ((j + 1) & 1);
        hipLaunchKernelGGL(k_mr_spmv, dim3(nbs), dim3(VB), 0, s, n, p->rowptr.p, p->col.p, p->vA.p, v, p->Av.p,
                           p->part1.p, cur);
        hipLaunchKernelGGL(k_mr_lanczos, dim3(nb), dim3(VB), 0, s, n, nbs, p->part1.p, v, vp, p->Av.p, p->part2.p,
                           cur);
        hipLaunchKernelGGL(k_mr_update, dim3(nb), dim3(VB), 0, s, n, nbs, nb, p->part1.p, p->part2.p, v, vp, p->Av.p,
                           w, wp, p->x.p, p->rtol, cur, nxt);
        std::swap(v, vp);
        std::swap(w, wp);
    }
}

// Host-side checks of what launch_batch bakes into its launches (and a captured graph keeps):
// grids within [1, MAXB] blocks of VB threads, static LDS only (block_sum / reduce_parts: VB
// doubles plus one), and every pointer a live allocation of this handle of the sizes the
// kernels index (vectors nfree, partials MAXB, two scalar slots, the CSR arrays).
static int check_batch(const svlg_pml *p) {
    SVLG_CHECK(p->nb >= 1 && p->nb <= MAXB && p->nbs >= 1 && p->nbs <= MAXB, "MINRES grid out of range");
    SVLG_CHECK((int64_t)p->nbs * VB >= std::min<int64_t>((int64_t)p->nfree * LPR, (int64_t)MAXB * VB) &&
                   (int64_t)p->nb * VB >= std::min<int64_t>(p->nfree, (int64_t)MAXB * VB),
               "MINRES grid does not cover the rows");
    const DevBuf<double> *vec[] = {&p->vb0, &p->vb1, &p->wb0, &p->wb1, &p->x, &p->Av};
    for (const DevBuf<double> *b : vec) SVLG_CHECK(b->p && b->n == (size_t)p->nfree, "MINRES vector not allocated");
    SVLG_CHECK(p->part1.p && p->part1.n >= (size_t)std::max(p->nb, p->nbs) && p->part2.p &&
                   p->part2.n >= (size_t)std::max(p->nb, p->nbs),
               "MINRES partials not allocated");
    SVLG_CHECK(p->st.p && p->st.n == 2, "MINRES state slots not allocated");
    SVLG_CHECK(p->rowptr.p && p->rowptr.n == (size_t)p->nfree + 1 && p->col.p && p->col.n == (size_t)p->nnz &&
                   p->vA.p && p->vA.n == (size_t)p->nnz,
               "scaled matrix not allocated");
    return 0;
}

// Captures launch_batch into a graph; on any capture error the caller launches eagerly.
static void capture_batch(svlg_pml *p) {
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return;
    launch_batch(p, p->stream);
    const hipError_t ec = hipStreamEndCapture(p->stream, &g);
    if (ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return;
    }
    if (hipGraphInstantiate(&p->batch, g, nullptr, nullptr, 0) != hipSuccess) p->batch = nullptr;
    (void)hipGraphDestroy(g);
    (void)hipGetLastError();
}

extern "C" {

int svlg_pml_create(const svlg_mixed_desc *d, double dt, double rtol, int32_t maxit, svlg_pml **out) {
    SVLG_CHECK(d && out, "null argument");
    *out = nullptr;
    SVLG_CHECK(d->nnodes > 0 && d->nelem > 0 && d->coords && d->dof_ptr && d->dof_free && d->dof_factor &&
                   d->elem_class && d->conn && d->E && d->nu && d->rho && d->th,
               "incomplete mixed-model description");
    SVLG_CHECK(dt > 0.0 && rtol > 0.0 && maxit > 0, "dt, rtol and maxit must be positive");
    SVLG_CHECK(d->ndim == 2 || d->ndim == 3, "ndim must be 2 or 3");
    const int nn = d->nnodes, ne = d->nelem, dim = d->ndim;
    const int ntot = d->dof_ptr[nn];
    for (int n = 0; n < nn; ++n) {
        const int k = d->dof_ptr[n + 1] - d->dof_ptr[n];
        SVLG_CHECK(dim == 2 ? (k == 2 || k == 5) : (k == 3 || k == 9),
                   "every node carries 2 or 5 unknowns in 2D, 3 or 9 in 3D");
    }
    int nfree = 0;
    for (int t = 0; t < ntot; ++t) {
        SVLG_CHECK(d->dof_free[t] >= -1, "free DOF numbers are -1 or non-negative");
        nfree = std::max(nfree, d->dof_free[t] + 1);
    }
    SVLG_CHECK(nfree > 0, "the model has no free DOF");
    bool any_pml = false, has_g = false;
    const int nen = dim == 2 ? 4 : 8;
    std::vector<int64_t> eoff(ne + 1, 0);
    for (int e = 0; e < ne; ++e) {
        const int cls = d->elem_class[e];
        SVLG_CHECK(dim == 2 ? (cls == SVLG_CLASS_LIN2DQUAD4 || cls == SVLG_CLASS_PML2DQUAD4)
                            : (cls == SVLG_CLASS_LIN3DHEXA8 || cls == SVLG_CLASS_PML3DHEXA8),
                   "unknown element class for this dimension");
        const int nun = class_nun(cls), nd = nen * nun;
        const bool pml = cls == SVLG_CLASS_PML2DQUAD4 || cls == SVLG_CLASS_PML3DHEXA8;
        any_pml |= pml;
        has_g |= cls == SVLG_CLASS_PML3DHEXA8;
        eoff[e + 1] = eoff[e] + (int64_t)nd * nd;
        for (int c = 0; c < nen; ++c) {
            const int n = d->conn[nen * e + c];
            SVLG_CHECK(n >= 0 && n < nn, "element connectivity out of range");
            SVLG_CHECK(d->dof_ptr[n + 1] - d->dof_ptr[n] == nun,
                       "a node's unknown count does not match its element class");
        }
        SVLG_CHECK(d->E[e] > 0.0 && d->rho[e] > 0.0 && d->nu[e] > -1.0 && d->nu[e] < 0.5 &&
                       (dim == 3 || d->th[e] > 0.0),
                   "element material or thickness out of range");
    }
    if (any_pml) {
        SVLG_CHECK(d->pml != nullptr, "PML elements need their parameters");
        for (int e = 0; e < ne; ++e)
            if (d->elem_class[e] == SVLG_CLASS_PML2DQUAD4 || d->elem_class[e] == SVLG_CLASS_PML3DHEXA8)
                SVLG_CHECK(d->pml[9 * e + 1] > 0.0 && d->pml[9 * e + 2] > 0.0 && d->pml[9 * e + 2] <= 1.0,
                           "PML depth L must be positive and 0 < R <= 1");
    }
    SVLG_CHECK(d->nload == 0 || (d->load_dof && d->load_factor && d->load_series && d->load_nt > 0),
               "loads need DOFs, factors and series");
    for (int l = 0; l < d->nload; ++l) SVLG_CHECK(d->load_dof[l] >= 0 && d->load_dof[l] < ntot, "load DOF out of range");

    // Contributions of every element-matrix entry to the free-space matrices, in element order.
    struct Contrib { int64_t key, idx; double fac; };
    std::vector<Contrib> cl;
    cl.reserve((size_t)ne * 100);
    for (int e = 0; e < ne; ++e) {
        const int nun = class_nun(d->elem_class[e]), nd = nen * nun;
        int fr[72];
        double fa[72];
        for (int a = 0; a < nen; ++a)
            for (int p = 0; p < nun; ++p) {
                const int t = d->dof_ptr[d->conn[nen * e + a]] + p;
                fr[a * nun + p] = d->dof_free[t];
                fa[a * nun + p] = d->dof_factor[t];
            }
        const int cls = d->elem_class[e];
        for (int i = 0; i < nd; ++i) {
            if (fr[i] < 0) continue;
            for (int j = 0; j < nd; ++j) {
                if (fr[j] < 0 || !pair_can_couple(cls, i % nun, j % nun)) continue;
                cl.push_back({(int64_t)fr[i] * nfree + fr[j], eoff[e] + (int64_t)i * nd + j, fa[i] * fa[j]});
            }
        }
    }
    std::stable_sort(cl.begin(), cl.end(), [](const Contrib &a, const Contrib &b) { return a.key < b.key; });
    std::vector<int32_t> rowptr(nfree + 1, 0), col;
    std::vector<int64_t> cptr, cidx(cl.size());
    std::vector<double> cfac(cl.size());
    for (size_t t = 0; t < cl.size(); ++t) {
        if (t == 0 || cl[t].key != cl[t - 1].key) {
            cptr.push_back((int64_t)t);
            col.push_back((int32_t)(cl[t].key % nfree));
            rowptr[cl[t].key / nfree + 1] += 1;
        }
        cidx[t] = cl[t].idx;
        cfac[t] = cl[t].fac;
    }
    cptr.push_back((int64_t)cl.size());
    std::partial_sum(rowptr.begin(), rowptr.end(), rowptr.begin());
    SVLG_CHECK(col.size() < (size_t)INT32_MAX, "too many nonzeros");
    // Loads per free row: (series row, load factor x T entry), in load order.
    std::vector<std::pair<int, int>> lr;
    for (int l = 0; l < d->nload; ++l)
        if (d->dof_free[d->load_dof[l]] >= 0) lr.push_back({d->dof_free[d->load_dof[l]], l});
    std::stable_sort(lr.begin(), lr.end(), [](auto &a, auto &b) { return a.first < b.first; });
    std::vector<int32_t> lptr(nfree + 1, 0), lrow;
    std::vector<double> lfac;
    for (auto &[f, l] : lr) {
        lptr[f + 1] += 1;
        lrow.push_back(l);
        lfac.push_back(d->load_factor[l] * d->dof_factor[d->load_dof[l]]);
    }
    std::partial_sum(lptr.begin(), lptr.end(), lptr.begin());

    auto *p = new svlg_pml();
    auto fail = [&](int rc) { delete p; return rc; };
    p->nfree = nfree; p->ntot = ntot; p->nnodes = nn; p->nelem = ne;
    p->nnz = (int64_t)col.size(); p->ncontrib = (int64_t)cl.size();
    p->dt = dt; p->rtol = rtol; p->maxit = maxit;
    p->c0 = 4.0 / (dt * dt); p->c1 = 2.0 / dt; p->c2 = 4.0 / dt; p->c3 = dt / 3.0;
    p->has_g = has_g;
    p->nb = blocks_for(nfree);
    p->nbs = blocks_for((int64_t)nfree * LPR);
    p->nt = d->nload ? d->load_nt : 0;
    if (hipStreamCreate(&p->stream) != hipSuccess) { set_error("hipStreamCreate failed"); return fail(10); }
    if (hipHostMalloc((void **)&p->hst, sizeof(MrState)) != hipSuccess) { set_error("hipHostMalloc failed"); return fail(10); }
    hipStream_t s = p->stream;
    DevBuf<int32_t> cls, conn, bad;
    DevBuf<double> X, E, nu, rho, th, pml, emat, cf;
    DevBuf<int64_t> cp, ci, eo;
    std::vector<double> pml_h((size_t)9 * ne, 0.0);
    if (any_pml) std::copy(d->pml, d->pml + (size_t)9 * ne, pml_h.begin());
#define UP(buf, src, cnt)                                                              \
    if ((buf).upload((src), (cnt), s) != hipSuccess) {                                 \
        set_error("device allocation/upload failed for " #buf);                         \
        return fail(10);                                                               \
    }
    UP(cls, d->elem_class, ne);
    UP(conn, d->conn, (size_t)nen * ne);
    UP(X, d->coords, (size_t)dim * nn);
    UP(eo, eoff.data(), eoff.size());
    UP(E, d->E, ne);
    UP(nu, d->nu, ne);
    UP(rho, d->rho, ne);
    UP(th, d->th, ne);
    UP(pml, pml_h.data(), pml_h.size());
    UP(cp, cptr.data(), cptr.size());
    UP(ci, cidx.data(), cidx.size());
    UP(cf, cfac.data(), cfac.size());
    UP(p->rowptr, rowptr.data(), rowptr.size());
    UP(p->col, col.data(), col.size());
    UP(p->lptr, lptr.data(), lptr.size());
    UP(p->lrow, lrow.data(), lrow.size());
    UP(p->lfac, lfac.data(), lfac.size());
    UP(p->dfree, d->dof_free, ntot);
    UP(p->dfac, d->dof_factor, ntot);
    if (p->nt) UP(p->series, d->load_series, (size_t)d->nload * p->nt);
#undef UP
    int32_t zero = 0;
    DevBuf<double> *vecs[] = {&p->S, &p->U, &p->V, &p->A, &p->r, &p->x, &p->Av, &p->vb0, &p->vb1, &p->wb0, &p->wb1};
    for (DevBuf<double> *b : vecs)
        if (b->alloc(nfree) != hipSuccess) { set_error("device allocation failed"); return fail(10); }
    DevBuf<double> *mats[] = {&p->vM, &p->vC, &p->vK, &p->vA};
    for (DevBuf<double> *b : mats)
        if (b->alloc(p->nnz) != hipSuccess) { set_error("device allocation failed"); return fail(10); }
    if (has_g && (p->vG.alloc(p->nnz) != hipSuccess || p->Ub.alloc(nfree) != hipSuccess)) {
        set_error("device allocation failed");
        return fail(10);
    }
    const int64_t nent = eoff[ne];
    if (p->part0.alloc(MAXB) != hipSuccess || p->part1.alloc(MAXB) != hipSuccess ||
        p->part2.alloc(MAXB) != hipSuccess || p->st.alloc(2) != hipSuccess || emat.alloc((size_t)4 * nent) != hipSuccess ||
        bad.upload(&zero, 1, s) != hipSuccess) {
        set_error("device allocation failed");
        return fail(10);
    }
    DevBuf<double> vE;
    if (vE.alloc(p->nnz) != hipSuccess) { set_error("device allocation failed"); return fail(10); }
    EmatArgs ea{ne, d->lumped ? 1 : 0, cls.p, conn.p, eo.p, X.p, E.p, nu.p, rho.p, th.p, pml.p, emat.p};
    hipLaunchKernelGGL(k_pml_emat, dim3((unsigned)((nent + VB - 1) / VB)), dim3(VB), 0, s, ea, nent);
    hipLaunchKernelGGL(k_pml_assemble, dim3(blocks_for(p->nnz)), dim3(VB), 0, s, p->nnz, cp.p, ci.p, cf.p, emat.p,
                       p->c0, p->c1, p->c3, p->vM.p, p->vC.p, p->vK.p, has_g ? p->vG.p : nullptr, vE.p);
    hipLaunchKernelGGL(k_pml_scale, dim3(p->nb), dim3(VB), 0, s, nfree, p->rowptr.p, p->col.p, vE.p, p->S.p, bad.p);
    hipLaunchKernelGGL(k_pml_scaled, dim3(p->nb), dim3(VB), 0, s, nfree, p->rowptr.p, p->col.p, vE.p, p->S.p, p->vA.p);
    if (hipGetLastError() != hipSuccess) { set_error("PML set-up launch failed"); return fail(10); }
    int32_t hbad = 0;
    if (hipMemcpyAsync(&hbad, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_error("PML set-up failed");
        return fail(10);
    }
    if (hbad) { set_error("Keff has a zero diagonal entry"); return fail(11); }
    const size_t nbytes = sizeof(double) * nfree;
    if (hipMemsetAsync(p->U.p, 0, nbytes, s) != hipSuccess || hipMemsetAsync(p->V.p, 0, nbytes, s) != hipSuccess ||
        hipMemsetAsync(p->A.p, 0, nbytes, s) != hipSuccess ||
        (has_g && hipMemsetAsync(p->Ub.p, 0, nbytes, s) != hipSuccess) || hipStreamSynchronize(s) != hipSuccess) {
        set_error("PML state initialisation failed");
        return fail(10);
    }
    *out = p;
    return 0;
}

int svlg_pml_destroy(svlg_pml *p) {
    if (p) {
        (void)hipStreamSynchronize(p->stream);
        delete p;
    }
    return 0;
}

int svlg_pml_info(svlg_pml *p, int64_t *info4) {
    SVLG_CHECK(p && info4, "null argument");
    info4[0] = p->nfree;
    info4[1] = p->ntot;
    info4[2] = p->nnz;
    info4[3] = p->ncontrib;
    return 0;
}

int svlg_pml_steps(svlg_pml *p, int32_t k0, int32_t nsteps, int32_t nrec, const int32_t *rec_dofs, double *histU,
                   double *histV, double *histA, int32_t *iters) {
    SVLG_CHECK(p, "null handle");
    SVLG_CHECK(k0 >= 0 && nsteps >= 0 && nrec >= 0, "negative step or record count");
    SVLG_CHECK(nrec == 0 || rec_dofs, "null record list");
    for (int r = 0; r < nrec; ++r) SVLG_CHECK(rec_dofs[r] >= 0 && rec_dofs[r] < p->ntot, "recorded DOF out of range");
    hipStream_t s = p->stream;
    const int n = p->nfree, nb = p->nb;
    DevBuf<int32_t> rec;
    DevBuf<double> hist;
    const size_t hsz = (size_t)nsteps * nrec;
    if (nrec) {
        SVLG_HIP(rec.upload(rec_dofs, nrec, s));
        SVLG_HIP(hist.alloc(3 * hsz));
    }
    if (int rc = check_batch(p)) return rc;
    if (p->use_graph && !p->batch) capture_batch(p);
    RhsArgs ra{n, 0, p->nt, p->rowptr.p, p->col.p, p->lptr.p, p->lrow.p, p->vM.p, p->vC.p, p->vK.p,
               p->has_g ? p->vG.p : nullptr, p->lfac.p, p->series.p, p->S.p, p->U.p, p->V.p, p->A.p,
               p->has_g ? p->Ub.p : nullptr, p->c2, p->dt, p->r.p, p->part0.p};
    for (int st = 0; st < nsteps; ++st) {
        ra.k = k0 + st;
        hipLaunchKernelGGL(k_pml_rhs, dim3(p->nbs), dim3(VB), 0, s, ra);
        hipLaunchKernelGGL(k_mr_init, dim3(nb), dim3(VB), 0, s, n, p->nbs, p->part0.p, p->r.p, p->vb0.p, p->vb1.p,
                           p->wb0.p, p->wb1.p, p->x.p, p->rtol, p->st.p);
        // Batches of POLL iterations; once converged the remaining launches of a batch return
        // at once. A batch ends in slot 0 with the buffer pairs back in place.
        // The batches the previous solve needed, less one, go out before the first poll
        // (steps of a linear analysis converge in nearly the same number of iterations).
        for (int b = 1; b < p->last_iters / POLL; ++b) {
            if (p->batch) SVLG_HIP(hipGraphLaunch(p->batch, s));
            else launch_batch(p, s);
        }
        for (;;) {
            if (p->batch) SVLG_HIP(hipGraphLaunch(p->batch, s));
            else launch_batch(p, s);
            SVLG_HIP(hipGetLastError());
            SVLG_HIP(hipMemcpyAsync(p->hst, p->st.p, sizeof(MrState), hipMemcpyDeviceToHost, s));
            SVLG_HIP(hipStreamSynchronize(s));
            SVLG_CHECK(!p->hst->fail, "MINRES broke down (singular effective matrix)");
            if (p->hst->done) break;
            SVLG_CHECK(p->hst->iters < p->maxit, "MINRES did not converge within maxit iterations");
        }
        SVLG_CHECK(p->hst->iters <= p->maxit, "MINRES did not converge within maxit iterations");
        if (iters) iters[st] = p->hst->iters;
        p->last_iters = p->hst->iters;
        hipLaunchKernelGGL(k_pml_update, dim3(nb), dim3(VB), 0, s, n, p->S.p, p->x.p, p->c0, p->c1, p->c2, p->dt,
                           p->U.p, p->V.p, p->A.p, p->has_g ? p->Ub.p : nullptr);
        if (nrec)
            hipLaunchKernelGGL(k_pml_record, dim3((nrec + VB - 1) / VB), dim3(VB), 0, s, nrec, rec.p, p->dfree.p,
                               p->dfac.p, p->U.p, p->V.p, p->A.p, hist.p + (size_t)st * nrec,
                               hist.p + hsz + (size_t)st * nrec, hist.p + 2 * hsz + (size_t)st * nrec);
        SVLG_HIP(hipGetLastError());
    }
    if (nrec) {
        double *dst[3] = {histU, histV, histA};
        for (int c = 0; c < 3; ++c)
            if (dst[c]) SVLG_HIP(hipMemcpyAsync(dst[c], hist.p + c * hsz, sizeof(double) * hsz, hipMemcpyDeviceToHost, s));
    }
    SVLG_HIP(hipStreamSynchronize(s));
    return 0;
}

int svlg_pml_set_graph(svlg_pml *p, int32_t on) {
    SVLG_CHECK(p, "null handle");
    SVLG_HIP(hipStreamSynchronize(p->stream));
    if (!on && p->batch) {
        SVLG_HIP(hipGraphExecDestroy(p->batch));
        p->batch = nullptr;
    }
    p->use_graph = on != 0;
    return 0;
}

int svlg_pml_get_history(svlg_pml *p, double *Ubar) {
    SVLG_CHECK(p && Ubar, "null argument");
    const size_t nbytes = sizeof(double) * p->nfree;
    if (!p->has_g) {
        std::fill(Ubar, Ubar + p->nfree, 0.0);
        return 0;
    }
    SVLG_HIP(hipMemcpyAsync(Ubar, p->Ub.p, nbytes, hipMemcpyDeviceToHost, p->stream));
    SVLG_HIP(hipStreamSynchronize(p->stream));
    return 0;
}

int svlg_pml_get_state(svlg_pml *p, double *U, double *V, double *A) {
    SVLG_CHECK(p, "null handle");
    const size_t nbytes = sizeof(double) * p->nfree;
    if (U) SVLG_HIP(hipMemcpyAsync(U, p->U.p, nbytes, hipMemcpyDeviceToHost, p->stream));
    if (V) SVLG_HIP(hipMemcpyAsync(V, p->V.p, nbytes, hipMemcpyDeviceToHost, p->stream));
    if (A) SVLG_HIP(hipMemcpyAsync(A, p->A.p, nbytes, hipMemcpyDeviceToHost, p->stream));
    SVLG_HIP(hipStreamSynchronize(p->stream));
    return 0;
}

}  // extern "C"
