// svlg_nonlinear.hip -- NewmarkBeta with the Newton-Raphson algorithm on the device, for
// models with J2 (and elastic) elements.
//
// Replaces (RP = 02-Run_Process of the SVL source tree):
//   NewtonRaphson::ComputeNewIncrement        (RP/09-Algorithms/02-Newton/NewtonRaphson.cpp:24-75)
//   Algorithm::ComputeConvergence (tests 1-8) (RP/09-Algorithms/Algorithm.cpp:122-188)
//   Algorithm::UpdateStatesIncrements         (Algorithm.cpp:18-56) -> lin*::UpdateState ->
//     Plastic3DJ2 / PlasticPlaneStrainJ2::UpdateState (svlg_j2.hpp)
//   NewmarkBeta::ComputeEffectiveForce / ComputeEffectiveStiffness / ComputeNewStep
//     (RP/10-Integrators/03-Newmark/NewmarkBeta.cpp:115-144, 198-235)
// Per Newton iteration: the residual from the stored Gauss-point stresses (tile kernel), a
// conjugate-gradient solve of the tangent system K_t + 4/dt^2 M + 2/dt C applied matrix-free
// (Jacobi preconditioner of the initial effective stiffness, or the learned preconditioner of
// it in flexible CG), dU += du, the convergence test, then the Gauss-point update at U + dU. The mass is lumped; C = am M + ak K_initial + dashpots.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "svlg_common.hpp"
#include "svlg_elem.hpp"
#include "svlg_j2.hpp"
#include "svlg_solver.hpp"

namespace svlg {

int gather_free(Model &m, const double *dnode, double *dfree);
int record_positions(const Model &m, int nrec, const int32_t *rec_dofs, std::vector<int32_t> &pos);
int launch_nl_rhs(Model &m, int k, const double *un, const double *w, const double *dU, double *b);

static constexpr int kNlGrid = 256, kNlBlock = 512;

// Gauss-point states at rest: zero stresses and strains, elastic tangent.
template <int DIM>
__global__ void k_j2_init(int ne, const int32_t *__restrict__ mat, const double *__restrict__ j2p,
                          double *__restrict__ state, int32_t *__restrict__ plastic) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    plastic[e] = 0;
    if (mat[e] != SVLG_MAT_J2) return;
    const double *q = j2p + size_t(e) * 5;
    const J2Par p{q[0], q[1], q[2], q[3], q[4]};
    for (int g = 0; g < Elem<DIM>::ngp; ++g) j2_init_state(p, j2_at(state, size_t(ne), Elem<DIM>::ngp, e, g));
}

// lin*::UpdateState at U + dU for every J2 element: strain B (U + dU) at each Gauss point,
// then the material update. One thread per (element, Gauss point): a workgroup holds
// kJ2Block / ngp elements, thread t works on Gauss point t / epb of element t % epb, so the
// field-major states are read and written contiguously. plastic[e] = 1 when the update left
// some Gauss point of e with the elastoplastic tangent, 0 when every point holds the elastic
// one (the tangent operator then takes the element's elastic path). geo non-null (Hexa8 on
// the diagonal geometry table): the strain comes from the corner sums with J^-1 of the
// element's table entry egid[e] (brick_strain), without the coordinates.
constexpr int kJ2Block = 256;

template <int DIM>
__global__ __launch_bounds__(kJ2Block) void k_j2_update(int ne, const int32_t *__restrict__ conn,
                                                        const double *__restrict__ X, const double *__restrict__ U,
                                                        const double *__restrict__ dU,
                                                        const int32_t *__restrict__ mat,
                                                        const double *__restrict__ j2p, double th,
                                                        double *__restrict__ state, int32_t *__restrict__ plastic,
                                                        const double *__restrict__ geo,
                                                        const uint16_t *__restrict__ egid) {
    constexpr int NEN = Elem<DIM>::nen, NV = DIM == 2 ? 3 : 6, NGP = Elem<DIM>::ngp;
    constexpr int EPB = kJ2Block / NGP;
    __shared__ int flag[EPB];
    const int el = threadIdx.x % EPB, g = threadIdx.x / EPB;
    const int e = blockIdx.x * EPB + el;
    if (threadIdx.x < EPB) flag[threadIdx.x] = 0;
    __syncthreads();
    const bool j2 = e < ne && mat[e] == SVLG_MAT_J2;
    if (j2) {
        const double *q = j2p + size_t(e) * 5;
        const J2Par p{q[0], q[1], q[2], q[3], q[4]};
        double u[NEN][DIM];
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            const size_t nd = size_t(conn[size_t(e) * NEN + a]) * DIM;
#pragma unroll
            for (int d = 0; d < DIM; ++d) u[a][d] = U[nd + d] + dU[nd + d];
        }
        double ev[NV], et[6], e6[6];
        bool brick = false;
        if constexpr (DIM == 3) {
            if (geo) {
                brick = true;
                const double *gl = geo + size_t(egid[e]) * 9;
                const double l8[3] = {0.125 * gl[0], 0.125 * gl[4], 0.125 * gl[8]};
                double S[7][3], r, s, t;
                brick_sums(u, S);
                gauss_point(g, r, s, t);
                brick_strain(S, l8, r, s, t, ev);
            }
        }
        if (!brick) {
            double x[NEN][DIM];
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                const size_t nd = size_t(conn[size_t(e) * NEN + a]) * DIM;
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[a][d] = X[nd + d];
            }
            double dn[NEN][DIM];
            (void)gp_dn<DIM>(x, g, th, dn);
            gp_strain<DIM>(dn, u, ev);
        }
        j2_embed<DIM>(ev, et, e6);
        const J2Ref st = j2_at(state, size_t(ne), NGP, e, g);
        j2_update_gp(p, et, st);
        if (st[26] != 0.0) atomicOr(&flag[el], 1);
    }
    __syncthreads();
    if (g == 0 && j2) plastic[e] = flag[el];
}

// The six stress fields of every Gauss point of element e into out[ngp][6].
__global__ void k_gp_stress(const double *__restrict__ state, size_t nelem, int ngp, int e, double *__restrict__ out) {
    const int t = threadIdx.x;
    if (t >= ngp * 6) return;
    const int g = t / 6, i = t % 6;
    out[t] = j2_at(state, nelem, ngp, e, g)[13 + i];
}

// un = U + dU, w = V - 2/dt dU.
__global__ void k_nl_prep(int n, double dt, const double *__restrict__ U, const double *__restrict__ V,
                          const double *__restrict__ dU, double *__restrict__ un, double *__restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    un[i] = U[i] + dU[i];
    w[i] = V[i] - 2.0 / dt * dU[i];
}

// dUa += du; per-workgroup partials of du.du, dUa.dUa, sum (du b)^2, b.b.
__global__ __launch_bounds__(512) void k_nl_accum(int n, const double *__restrict__ du, double *__restrict__ dUa,
                                                  const double *__restrict__ b, double *__restrict__ part) {
    __shared__ double red[4 * 17];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double d = du[i];
        const double a = dUa[i] + d;
        dUa[i] = a;
        const double eb = d * b[i];
        v[0] += d * d;
        v[1] += a * a;
        v[2] += eb * eb;
        v[3] += b[i] * b[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) red[k * 17 + w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; ++k) {
            double r = 0.0;
            for (int i = 0; i < int(blockDim.x >> 6); ++i) r += red[k * 17 + i];
            part[4 * blockIdx.x + k] = r;
        }
    }
}

// NewmarkBeta::ComputeNewStep (NewmarkBeta.cpp:130-141): U += dU; A = 4/dt^2 dU - 4/dt V - A;
// V = 2/dt dU - V.
__global__ void k_nl_newmark(int n, double dt, const double *__restrict__ dU, const double *__restrict__ jac,
                             double *__restrict__ U, double *__restrict__ V, double *__restrict__ A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || jac[i] == 0.0) return;
    const double du = dU[i];
    U[i] += du;
    const double v = V[i];
    A[i] = 4.0 / dt / dt * du - 4.0 / dt * v - A[i];
    V[i] = 2.0 / dt * du - v;
}

__global__ void k_nl_record(int nrec, const int32_t *__restrict__ rec, const double *__restrict__ U,
                            double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec) out[r] = U[rec[r]];
}

// Algorithm::ComputeConvergence (Algorithm.cpp:122-188) from the norms of this iteration:
// s[0] = du.du, s[1] = dU.dU, s[2] = sum (du Feff)^2, s[3] = Feff.Feff.
static double convergence(int flag, const double (&s)[4], int k, double &NormFactor) {
    const double nF = std::sqrt(s[3]), nd = std::sqrt(s[0]), nD = std::sqrt(s[1]), nE = std::sqrt(s[2]);
    switch (flag) {
        case 1: return nF;
        case 2: return nd;
        case 3: return nE;
        case 4:
            if (k != 0) return nF / NormFactor;
            NormFactor = nF;
            return NormFactor;
        case 5:
            if (k != 0) return nd / NormFactor;
            NormFactor = nD;
            return nd;
        case 6:
            if (k != 0) return nE / NormFactor;
            NormFactor = nE;
            return NormFactor;
        case 7: return nd / nD;
        case 8: return -1.0;
        default: return 0.0;
    }
}

}  // namespace svlg

using namespace svlg;

extern "C" {

int svlg_nm_init(svlg_model *h, double dt, double rtol, int32_t maxit, int32_t precond);

int svlg_nl_init(svlg_model *h, double dt, double cg_rtol, int32_t cg_maxit, double cnvgtol, int32_t nmax,
                 int32_t cnvgtest) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(cnvgtol > 0.0 && nmax > 0, "cnvgtol and nmax must be positive");
    SVLG_CHECK(cnvgtest >= 1 && cnvgtest <= 8, "cnvgtest must be 1..8 (Algorithm::ComputeConvergence)");
    // NewtonRaphson subtracts K SupportMotion again in every iteration (NewtonRaphson.cpp:40-48).
    SVLG_CHECK(m.sm_ndofs == 0, "support motion with the Newton-Raphson algorithm is not on the device path");
    // Mass, damping, Jacobi preconditioner of the initial effective stiffness, work vectors.
    if (int rc = svlg_nm_init(h, dt, cg_rtol, cg_maxit, SVLG_PRECOND_JACOBI)) return rc;
    hipStream_t s = m.stream;
    const int n = m.ndofn, ngp = m.ndim == 3 ? 8 : 4;
    m.nl_ready = false;
    if (m.has_j2) {
        SVLG_HIP(m.gp_state.alloc(size_t(m.nelem) * ngp * kJ2State));
        SVLG_HIP(m.j2_plastic.alloc(m.nelem));
        SVLG_HIP(hipMemsetAsync(m.gp_state.p, 0, sizeof(double) * m.gp_state.n, s));
        const int nb = (m.nelem + 127) / 128;
        if (m.ndim == 3)
            hipLaunchKernelGGL(k_j2_init<3>, dim3(nb), dim3(128), 0, s, m.nelem, m.mat.p, m.j2p.p, m.gp_state.p,
                               m.j2_plastic.p);
        else
            hipLaunchKernelGGL(k_j2_init<2>, dim3(nb), dim3(128), 0, s, m.nelem, m.mat.p, m.j2p.p, m.gp_state.p,
                               m.j2_plastic.p);
    }
    for (DevBuf<double> *b : {&m.nl_dUa, &m.nl_un, &m.nl_w}) {
        SVLG_HIP(b->alloc(n));
        SVLG_HIP(hipMemsetAsync(b->p, 0, sizeof(double) * n, s));
    }
    SVLG_HIP(m.nl_part.alloc(4 * kNlGrid));
    SVLG_HIP(hipGetLastError());
    SVLG_HIP(hipStreamSynchronize(s));
    m.nl_tol = cnvgtol;
    m.nl_nmax = nmax;
    m.nl_test = cnvgtest;
    m.nl_normfactor = 0.0;
    m.nl_ready = true;
    return 0;
}

int svlg_nl_steps(svlg_model *h, int32_t k0, int32_t nsteps, int32_t nrec, const int32_t *rec_dofs, double *hist,
                  int32_t *newton_iters, double *newton_res, int32_t *cg_iters) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.nl_ready, "svlg_nl_init has not been called");
    SVLG_CHECK(nsteps >= 0 && nrec >= 0, "negative count");
    hipStream_t s = m.stream;
    const int n = m.ndofn;
    DevBuf<int32_t> rec;
    DevBuf<double> hbuf;
    if (nrec > 0 && nsteps > 0) {
        SVLG_CHECK(rec_dofs && hist, "recorder arrays missing");
        std::vector<int32_t> rn;
        if (int rc = record_positions(m, nrec, rec_dofs, rn)) return rc;
        SVLG_HIP(rec.upload(rn.data(), rn.size(), s));
        SVLG_HIP(hbuf.alloc(size_t(nrec) * nsteps));
    }
    const int epb = kJ2Block / (m.ndim == 3 ? 8 : 4);  // elements per k_j2_update workgroup
    const int nb = (n + 255) / 256, neb = (m.nelem + epb - 1) / epb;
    std::vector<double> hp(4 * kNlGrid);
    for (int st = 0; st < nsteps; ++st) {
        const int k = k0 + st;
        SVLG_HIP(hipMemsetAsync(m.nl_dUa.p, 0, sizeof(double) * n, s));
        int it = 0, cg_total = 0;
        double res = 0.0;
        do {
            hipLaunchKernelGGL(k_nl_prep, dim3(nb), dim3(256), 0, s, n, m.nm_dt, m.nU.p, m.nV.p, m.nl_dUa.p, m.nl_un.p,
                               m.nl_w.p);
            if (int rc = launch_nl_rhs(m, k, m.nl_un.p, m.nl_w.p, m.nl_dUa.p, m.cg_b.p)) return rc;
            SVLG_HIP(hipMemsetAsync(m.ndU.p, 0, sizeof(double) * n, s));
            int cit = 0;
            double rel = 0.0;
            if (int rc = cg_solve(m, m.cg_b.p, m.ndU.p, m.nm_rtol, m.nm_maxit, m.nm_precond, &cit, &rel, !m.cg_eager))
                return rc;
            cg_total += cit;
            hipLaunchKernelGGL(k_nl_accum, dim3(kNlGrid), dim3(kNlBlock), 0, s, n, m.ndU.p, m.nl_dUa.p, m.cg_b.p,
                               m.nl_part.p);
            SVLG_HIP(hipMemcpyAsync(hp.data(), m.nl_part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
            SVLG_HIP(hipStreamSynchronize(s));
            double sums[4] = {0.0, 0.0, 0.0, 0.0};
            for (int b = 0; b < kNlGrid; ++b)
                for (int c = 0; c < 4; ++c) sums[c] += hp[4 * b + c];
            res = convergence(m.nl_test, sums, it, m.nl_normfactor);
            if (newton_res) newton_res[size_t(st) * m.nl_nmax + it] = res;
            // UpdateStatesIncrements(U + dU).
            if (m.has_j2) {
                if (m.ndim == 3)
                    hipLaunchKernelGGL(k_j2_update<3>, dim3(neb), dim3(kJ2Block), 0, s, m.nelem, m.conn.p, m.X.p,
                                       m.nU.p, m.nl_dUa.p, m.mat.p, m.j2p.p, m.th, m.gp_state.p, m.j2_plastic.p,
                                       m.elem_path == 3 ? m.geo.p : nullptr, m.egid.p);
                else
                    hipLaunchKernelGGL(k_j2_update<2>, dim3(neb), dim3(kJ2Block), 0, s, m.nelem, m.conn.p, m.X.p,
                                       m.nU.p, m.nl_dUa.p, m.mat.p, m.j2p.p, m.th, m.gp_state.p, m.j2_plastic.p,
                                       nullptr, nullptr);
            }
            ++it;
        } while (res > m.nl_tol && it < m.nl_nmax);
        if (newton_iters) newton_iters[st] = it;
        if (cg_iters) cg_iters[st] = cg_total;
        hipLaunchKernelGGL(k_nl_newmark, dim3(nb), dim3(256), 0, s, n, m.nm_dt, m.nl_dUa.p, m.jac.p, m.nU.p, m.nV.p,
                           m.nA.p);
        if (nrec > 0)
            hipLaunchKernelGGL(k_nl_record, dim3((nrec + 63) / 64), dim3(64), 0, s, nrec, rec.p, m.nU.p,
                               hbuf.p + size_t(st) * nrec);
        m.nm_steps_done++;
    }
    SVLG_HIP(hipGetLastError());
    if (nrec > 0 && nsteps > 0)
        SVLG_HIP(hipMemcpyAsync(hist, hbuf.p, sizeof(double) * size_t(nrec) * nsteps, hipMemcpyDeviceToHost, s));
    SVLG_HIP(hipStreamSynchronize(s));
    return 0;
}

int svlg_gp_stress(svlg_model *h, int32_t elem, double *stress) {
    SVLG_CHECK(h && stress, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.nl_ready && m.has_j2, "no Gauss-point states (svlg_nl_init on a model with J2 elements)");
    SVLG_CHECK(elem >= 0 && elem < m.nelem, "element index out of range");
    const int ngp = m.ndim == 3 ? 8 : 4, nv = m.ndim == 3 ? 6 : 3;
    DevBuf<double> buf;
    SVLG_HIP(buf.alloc(size_t(ngp) * 6));
    hipLaunchKernelGGL(k_gp_stress, dim3(1), dim3(64), 0, m.stream, m.gp_state.p, size_t(m.nelem), ngp, elem, buf.p);
    std::vector<double> st(size_t(ngp) * 6);
    SVLG_HIP(hipMemcpyAsync(st.data(), buf.p, sizeof(double) * st.size(), hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    const int idx2[3] = {1, 2, 4};
    for (int g = 0; g < ngp; ++g)
        for (int i = 0; i < nv; ++i) stress[g * nv + i] = st[size_t(g) * 6 + (m.ndim == 3 ? i : idx2[i])];
    return 0;
}

}  // extern "C"
