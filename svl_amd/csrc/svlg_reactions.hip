// svlg_reactions.hip -- reactions at mesh nodes for the current state of the active
// integrator (no SVL entry point of its own: DynamicAnalysis::UpdateDomain computes them
// for every fixed node each step, DynamicAnalysis.cpp:130-150, through
// Integrator::ComputeReactionForce, CentralDifference.cpp:156-171):
//   R = sum over the node's elements of Fint_e + C_e V_e + M_e A_e  -  Fext(k)
// (Assembler::ComputeDynamicInternalForceVector, Assembler.cpp:273-287, 596-619;
// lin*::ComputeInternalDynamicForces, lin3DHexa8.cpp:416-437). With the lumped mass and
// C = am M + ak K + dashpots this is, per node,
//   R = sum_e [K_e (U_e + ak_e V_e)]_node + M_ii A_i + C_diag,ii V_i - Fext_i(k)
// (J2 elements: B^T sigma from the Gauss-point states + ak B^T D0 B V). With the consistent
// mass the inertia and the Rayleigh am term come from the element: [M_e (A_e + am_e V_e)]_node,
// and the node term keeps the dashpots alone. Only the requested nodes are computed; the
// restrained partner node of a dashpot carries -eta V of its node. With support motion the
// element forces use the displacement the strains hold (U_(k-1) at those DOFs: the strains lag
// one support increment, Algorithm.cpp:18-56 before CentralDifference.cpp:137).
#include <vector>

#include "svlg_common.hpp"
#include "svlg_drm.hpp"
#include "svlg_support.hpp"
#include "svlg_elem.hpp"
#include "svlg_j2.hpp"

namespace svlg {

struct ReacArgs {
    const int32_t *qnode;  // device node per request
    int nq;
    const int32_t *n2e_ptr, *n2e, *conn;
    const double *X, *lam, *mu, *ak;
    const double *U, *V, *A;           // Newmark state (mode 2)
    const double *Uk, *Uk1, *Uk2;      // CentralDifference ring U_k, U_(k-1), U_(k-2) (mode 1)
    const double *dtv;                 // per-DOF dt (null: dt)
    double dt;
    int mode, have_velocity;
    const double *mdiag, *cdiag;
    const int32_t *nl_ptr, *nl_idx;
    const double *load_dir, *load_factor, *load_series;
    int load_nt, k;
    double th;
    const int32_t *mat;
    const double *j2p, *state;  // null state: J2 elements as their initial elastic material
    int nelem;
    DrmRef drm;                 // domain reduction forces of every step (F null: none)
    const double *rho, *am;     // consistent mass: element density and Rayleigh am (null: lumped)
    SmRef sm;                   // support motion (slot null: none)
    const double *smlag;        // Newmark: U before the last support increment, per slot
    double *R;                  // [nq * DIM]
};

// SVL's V and A after the step: Newmark keeps them; CentralDifference forms them from the
// U ring (CentralDifference.cpp:141-148 with U = U_(k-1), Up = U_(k-2), dU = U_k - U_(k-1)).
__device__ __forceinline__ void state_va(const ReacArgs &a, size_t g, double &v, double &acc) {
    if (a.mode == 2) {
        v = a.V[g];
        acc = a.A[g];
        return;
    }
    if (!a.have_velocity) {
        v = acc = 0.0;
        return;
    }
    const double dt = a.dtv ? a.dtv[g] : a.dt;
    const double u = a.Uk1[g], up = a.Uk2[g], du = a.Uk[g] - a.Uk1[g];
    v = 1.0 / 2.0 / dt * ((u + du) - up);
    acc = 1.0 / dt / dt * ((du - u) + up);
}

template <int DIM>
__global__ void k_reactions(ReacArgs a) {
    constexpr int NEN = Elem<DIM>::nen;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nq) return;
    const int node = a.qnode[q];
    double r[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) r[d] = 0.0;
    const double *Ucur = a.mode == 2 ? a.U : a.Uk;
    for (int t = a.n2e_ptr[node]; t < a.n2e_ptr[node + 1]; ++t) {
        const int ea = a.n2e[t], e = ea / NEN, corner = ea % NEN;
        double x[NEN][DIM], u[NEN][DIM], v[NEN][DIM], f[NEN][DIM], w[NEN][DIM];
        const double ak = a.ak[e];
        const double am = a.rho ? a.am[e] : 0.0;
#pragma unroll
        for (int c = 0; c < NEN; ++c) {
            const size_t nd = size_t(a.conn[size_t(e) * NEN + c]) * DIM;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                double vv, aa;
                state_va(a, nd + d, vv, aa);
                x[c][d] = a.X[nd + d];
                v[c][d] = vv;
                double uc = Ucur[nd + d];
                if (a.sm.slot) {
                    const int sl = a.sm.slot[nd + d];
                    if (sl >= 0) uc = a.mode == 2 ? a.smlag[sl] : a.Uk1[nd + d];
                }
                u[c][d] = uc + ak * vv;
                w[c][d] = aa + am * vv;
            }
        }
        if (a.state && a.mat[e] == SVLG_MAT_J2) {
            const double *q5 = a.j2p + size_t(e) * 5;
            const J2Par jp{q5[0], q5[1], q5[2], q5[3], q5[4]};
            j2_elem_force<DIM>(x, v, jp, a.state, size_t(a.nelem), e, ak, false, a.th, f);
        } else if constexpr (DIM == 3) {
            hexa_fint(x, u, a.lam[e], a.mu[e], f);
        } else {
            quad_fint(x, u, a.lam[e], a.mu[e], a.th, f);
        }
        if (a.rho) {
            if constexpr (DIM == 3)
                hexa_mass_gauss(x, a.rho[e], w, f);
            else
                quad_mass_gauss(x, a.rho[e] * a.th, w, f);
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) r[d] += f[corner][d];
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const size_t g = size_t(node) * DIM + d;
        double vv, aa;
        state_va(a, g, vv, aa);
        double fe = 0.0;
        for (int q2 = a.nl_ptr[node]; q2 < a.nl_ptr[node + 1]; ++q2) {
            const int l = a.nl_idx[q2];
            const double amp = (a.k >= 0 && a.k < a.load_nt) ? a.load_series[size_t(l) * a.load_nt + a.k] : 0.0;
            fe += a.load_factor[l] * (amp * a.load_dir[size_t(l) * DIM + d]);
        }
        const int ds = drm_slot(a.drm, node);
        if (ds >= 0) fe += drm_force(a.drm, ds, a.k, DIM, d);
        a.R[size_t(q) * DIM + d] =
            a.rho ? (r[d] + a.cdiag[g] * vv) - fe : ((r[d] + a.cdiag[g] * vv) + a.mdiag[g] * aa) - fe;
    }
}

}  // namespace svlg

using namespace svlg;

extern "C" {

int svlg_reactions(svlg_model *h, int32_t n, const int32_t *nodes, int32_t k, double *R) {
    SVLG_CHECK(h && (n == 0 || (nodes && R)), "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.active == 1 || m.active == 2, "no integrator has been initialised");
    SVLG_CHECK(m.active != 1 || m.cd_ready, "svlg_cd_init has not been called");
    SVLG_CHECK(m.active != 2 || m.nm_ready, "svlg_nm_init has not been called");
    if (n == 0) return 0;
    std::vector<int32_t> dn(n);
    for (int i = 0; i < n; ++i) {
        SVLG_CHECK(nodes[i] >= 0 && nodes[i] < m.nnodes, "node index out of range");
        dn[i] = m.inv[nodes[i]];
    }
    hipStream_t s = m.stream;
    DevBuf<int32_t> qn;
    DevBuf<double> rb;
    SVLG_HIP(qn.upload(dn.data(), dn.size(), s));
    SVLG_HIP(rb.alloc(size_t(n) * m.ndim));
    ReacArgs a{};
    a.qnode = qn.p;
    a.nq = n;
    a.n2e_ptr = m.n2e_ptr.p;
    a.n2e = m.n2e.p;
    a.conn = m.conn.p;
    a.X = m.X.p;
    a.lam = m.lam.p;
    a.mu = m.mu.p;
    a.ak = m.ak.p;
    a.mode = m.active;
    if (m.active == 2) {
        a.U = m.nU.p;
        a.V = m.nV.p;
        a.A = m.nA.p;
    } else {
        a.Uk = m.U[m.ring].p;
        a.Uk1 = m.U[(m.ring + 2) % 3].p;
        a.Uk2 = m.U[(m.ring + 1) % 3].p;
        a.dtv = m.h_dt_dof.empty() ? nullptr : m.dt_dof.p;
        a.dt = m.dt;
        a.have_velocity = m.steps_done > 0;
    }
    a.mdiag = m.mdiag.p;
    a.cdiag = m.lumped ? m.cdiag.p : m.cdash.p;
    a.rho = m.lumped ? nullptr : m.rho.p;
    a.am = m.am.p;
    a.nl_ptr = m.nl_ptr.p;
    a.nl_idx = m.nl_idx.p;
    a.load_dir = m.load_dir.p;
    a.load_factor = m.load_factor.p;
    a.load_series = m.load_series.p;
    a.load_nt = m.load_nt;
    a.k = k;
    a.th = m.th;
    a.mat = m.mat.p;
    a.j2p = m.j2p.p;
    a.state = (m.active == 2 && m.has_j2 && m.nl_ready) ? m.gp_state.p : nullptr;
    a.nelem = m.nelem;
    a.drm = drm_ref(m);
    a.sm = sm_ref(m);
    a.smlag = m.sm_lag.p;
    a.R = rb.p;
    const int nb = (n + 63) / 64;
    if (m.ndim == 3)
        hipLaunchKernelGGL(k_reactions<3>, dim3(nb), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_reactions<2>, dim3(nb), dim3(64), 0, s, a);
    SVLG_HIP(hipGetLastError());
    SVLG_HIP(hipMemcpyAsync(R, rb.p, sizeof(double) * size_t(n) * m.ndim, hipMemcpyDeviceToHost, s));
    SVLG_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"
