// svlg_implicit.hip -- ExtendedNewmarkBeta (gamma = 1/2, beta = 1/4) with the Linear
// algorithm and a matrix-free, Jacobi-preconditioned conjugate-gradient solve of the
// effective-stiffness system, all on the device in FP64.
//
// Replaces (RP = 02-Run_Process of the SVL source tree):
//   ExtendedNewmarkBeta::ComputeEffectiveForce      (RP/10-Integrators/03-Newmark/ExtendedNewmarkBeta.cpp:208-229)
//   ExtendedNewmarkBeta::ComputeEffectiveStiffness  (:233-245)
//   ExtendedNewmarkBeta::ComputeNewStep             (:123-154)
//   Linear::ComputeNewIncrement                     (RP/09-Algorithms/01-Linear/Linear.cpp:22-56)
//   LinearSystem::SolveSystem / GetSolution         (RP/11-Solvers/LinearSystem.hpp:44-66): SVL
//     factors Keff (EigenSolver LDL^T, MUMPS) or calls PETSc KSP; here CG to a relative
//     residual tolerance ||b - Keff x|| <= rtol ||b||.
// With Rayleigh damping C_e = am M_e + ak K_e and dashpots, and G = 0:
//   Keff = sum_e (1 + 2 ak_e/dt) K_e + (4/dt^2) M + (2/dt) (sum_e am_e M_e + C_dash)
//   Feff = Fext - sum_e K_e (U_e - ak_e V_e) + C_diag V + M (4/dt V + A)
// The lumped mass enters through per-DOF diagonals; the consistent mass (SVL's default
// massform) element by element, beside the stiffness, in the same tile pass.
// The element operator reuses the node tiles of the explicit corner-slot kernel: a
// workgroup stages the operand of its owned + halo nodes in LDS, recomputes every element
// touching its nodes and sums the corner slots of each owned node in a fixed order, so a
// solve is bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "svlg_common.hpp"
#include "svlg_drm.hpp"
#include "svlg_support.hpp"
#include "svlg_elem.hpp"
#include "svlg_j2.hpp"
#include "svlg_solver.hpp"

namespace svlg {

int gather_free(Model &m, const double *dnode, double *dfree);
int record_positions(const Model &m, int nrec, const int32_t *rec_dofs, std::vector<int32_t> &pos);
int scatter_free(Model &m, const double *dfree, double *dnode);

enum { OP_CG = 0, OP_RHS = 1 };

struct OpArgs {
    const int32_t *own_ptr, *halo_ptr, *halo_node, *elem_ptr;
    const uint16_t *lconn;
    const double *elam, *emu;
    const uint16_t *egeo;
    const double *geo, *X;
    const double *coef;   // OP_CG: 1 + 2 ak/dt per tile element; OP_RHS: -ak per tile element
    int ntiles, max_stage, max_own;
    double th;
    const double *jac;    // 0 marks a restrained DOF
    // OP_CG
    const double *z, *p_old;
    double *p_new;
    const double *dgn;
    double *y;
    double *scal;
    CgParts cg;
    int reduce;           // OP_CG: 1 = iteration `iter` of a solve (beta, stopping test, p.q partials)
    int iter;
    // OP_RHS
    const double *U, *V, *A, *mdiag, *cdiag;
    double dt;
    double *b;
    const int32_t *nl_ptr, *nl_idx;
    const double *load_dir, *load_factor, *load_series;
    int load_nt, k;
    // Models with J2 elements (NL kernels): element -> material, parameters, Gauss-point
    // states; for the Newton residual the staged operands are U + dU and V - 2/dt dU, and the
    // mass term uses the step's V and dU.
    const int32_t *teid;   // tile element -> element tag index
    const int32_t *mat;    // [nelem]
    const double *j2p;     // [nelem * 5]
    const double *state;   // [nelem * ngp * kJ2State]
    const int32_t *plastic;  // [nelem] elastoplastic tangent somewhere in the element
    const double *Vm, *dUn;
    const int *jdev;       // non-null: the iteration is *jdev + iter (graph replays)
    int nelem;             // elements (the Gauss-point state stride)
    DrmRef drm;            // OP_RHS: domain reduction forces of every step (F null: none)
    // Consistent mass (kernels with CM): per tile element the mass scale (OP_CG: times
    // 4/dt^2 + 2 am/dt; OP_RHS: alone) and, for OP_RHS, am. cdiag then holds the dashpots alone
    // and dgn (2/dt) times them.
    const double *cmass, *cam;
};

// One node tile of the effective-stiffness operator (one thread per tile element in the
// element phase, one thread per owned node in the sum phase).
//  OP_CG : stage p = z + beta p_old (owned values are written to p_new), y = Keff p on the
//          owned DOFs (0 on restrained DOFs), per-tile partial of p . y.
//  OP_RHS: stage U and V, y = sum_e K_e (U_e - ak_e V_e), then on the free DOFs
//          b = (((Fext - y) + C_diag V) + M (4/dt V + A)).
//  NL: elements of J2 material use their Gauss-point states instead: the internal force from
//      the stored stresses (OP_RHS) or the consistent tangent (OP_CG), each plus the Rayleigh
//      term of the initial stiffness (coordinates are staged on every path).
// MOC > 0: every tile owns at most MOC nodes; the slot planes then have the compile-time
// stride MOC + 1 with column MOC as the trash column (branch-free corner writes, immediate
// plane offsets).
// CM: consistent element mass (MassFormulation CONSISTENT). Every element adds its mass
// matrix times the operand: OP_CG (4/dt^2 + 2 am/dt) M_e p; OP_RHS stages the mass operand
// w = 4/dt V + A (Newton: 4/dt V + A - 4/dt^2 dU) and subtracts M_e (w_e + am V_e) from its
// slot, so that the node sum gives Feff = Fext - sum_e [K_e (U - ak V) - M_e (w + am V)]_e
// + C_dash V.
template <int DIM, int PATH, int MODE, bool NL, int MOC = 0, bool CM = false>
__global__ __launch_bounds__(512) void k_nm_op(OpArgs a) {
    constexpr int NEN = Elem<DIM>::nen;
    // J2 Gauss-point loops need the coordinates, except on axis-aligned Hexa8 bricks (the
    // diagonal geometry table), where they run on the corner sums (j2_brick_force).
    constexpr bool BRICK = PATH == 3 && DIM == 3;
    constexpr bool STAGE_X = PATH < 2 || (NL && !BRICK);
    constexpr bool RHS = MODE == OP_RHS;
    constexpr int NPRE = RHS ? (CM ? 3 : 2) : 1;  // staged state arrays before the coordinates
    extern __shared__ double smem[];
    __shared__ double red[3 * 17];
    double beta = 0.0;
    if constexpr (!RHS) {
        if (a.reduce) {
            // The DONE flag and *jdev are loaded together.
            const double flag = a.scal[SC_DONE];
            const int j = iter_of(a.jdev, a.iter);
            if (flag != 0.0) return;
            // Stopping test and beta of iteration j from the partials of the previous kernels;
            // scal holds r.z and p.q of iteration j-1 (left by the vector kernel). The flexible
            // solve's test on r_j was applied by the first layer of its preconditioner
            // (cg_stop_test, the same sums), which set DONE if it held: only beta remains.
            double v[3] = {0.0, 0.0, 0.0};
            bool done = false;
            if (!a.cg.flexible) {
                cg_vec(a.cg, j, red, v);
                done = v[1] <= a.scal[SC_TOL2] * v[2] || v[2] == 0.0;
            }
            if (j > 0) {
                if (a.cg.flexible) {
                    double zz[2];
                    cg_pre(a.cg, j, red, zz);
                    const double pq = a.scal[SC_PQ];
                    beta = pq != 0.0 ? -zz[1] / pq : 0.0;
                } else {
                    const double rz0 = a.scal[SC_RHO];
                    beta = rz0 != 0.0 ? v[0] / rz0 : 0.0;
                }
            }
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                if (!a.cg.flexible) {
                    a.scal[SC_RR] = v[1];
                    a.scal[SC_BB] = v[2];
                }
                a.scal[SC_ITERS] = double(j);
                a.scal[SC_BETA] = beta;
                if (done) a.scal[SC_DONE] = 1.0;
            }
            if (done) return;
        }
    }
    const int nt = a.ntiles;
    // XCD-aware mapping: consecutive workgroups go to different XCDs, each XCD walks a
    // contiguous range of tiles.
    const int bI = blockIdx.x, xcd = bI & 7, j = bI >> 3;
    const int q = nt >> 3, r = nt & 7;
    const int tile = xcd * q + min(xcd, r) + j;
    const int o0 = a.own_ptr[tile], nown = a.own_ptr[tile + 1] - o0;
    const int h0 = a.halo_ptr[tile], nh = a.halo_ptr[tile + 1] - h0;
    const int e0 = a.elem_ptr[tile], ne = a.elem_ptr[tile + 1] - e0;
    const int MS = a.max_stage, MO = MOC > 0 ? MOC + 1 : a.max_own;
    double *sU = smem;
    double *sV = smem + size_t(MS) * DIM;
    double *sW = smem + size_t(2 * MS) * DIM;  // CM && RHS: mass operand
    double *sX = smem + size_t(NPRE * MS) * DIM;
    double *sIn = smem + size_t((NPRE + (STAGE_X ? 1 : 0)) * MS) * DIM;
    const int tid = threadIdx.x, B = blockDim.x;
    const size_t g0 = size_t(o0) * DIM;
    const bool has_elem = tid < ne;
    int lc[NEN];
    double lam = 0.0, mu = 0.0, c2 = 0.0;
    int gid = 0;
    if (has_elem) {
        const uint16_t *l = a.lconn + size_t(e0 + tid) * NEN;
#pragma unroll
        for (int c = 0; c < NEN; ++c) lc[c] = l[c];
        lam = a.elam[e0 + tid];
        mu = a.emu[e0 + tid];
        if constexpr (RHS) {
            c2 = a.coef[e0 + tid];
        } else {
            const double cf = a.coef[e0 + tid];
            lam *= cf;
            mu *= cf;
        }
        if constexpr (PATH >= 2) gid = a.egeo[e0 + tid];
    } else {
#pragma unroll
        for (int c = 0; c < NEN; ++c) lc[c] = 0;
    }
    for (int i = tid; i < nown * DIM; i += B) {
        const size_t g = g0 + i;
        if constexpr (RHS) {
            sU[i] = a.U[g];
            sV[i] = a.V[g];
            if constexpr (CM)
                sW[i] = a.Vm ? 4.0 / a.dt * a.Vm[g] + a.A[g] - 4.0 / a.dt / a.dt * a.dUn[g]
                             : 4.0 / a.dt * a.V[g] + a.A[g];
        } else {
            const double v = a.z[g] + beta * a.p_old[g];
            sU[i] = v;
            a.p_new[g] = v;
        }
        if constexpr (STAGE_X) sX[i] = a.X[g];
    }
    for (int sl = tid; sl < nh; sl += B) {
        const size_t n = size_t(a.halo_node[h0 + sl]) * DIM;
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            const int i = (nown + sl) * DIM + c;
            if constexpr (RHS) {
                sU[i] = a.U[n + c];
                sV[i] = a.V[n + c];
                if constexpr (CM)
                    sW[i] = a.Vm ? 4.0 / a.dt * a.Vm[n + c] + a.A[n + c] - 4.0 / a.dt / a.dt * a.dUn[n + c]
                                 : 4.0 / a.dt * a.V[n + c] + a.A[n + c];
            } else {
                sU[i] = a.z[n + c] + beta * a.p_old[n + c];
            }
            if constexpr (STAGE_X) sX[i] = a.X[n + c];
        }
    }
    for (int i = tid; i < NEN * DIM * MO; i += B) sIn[i] = 0.0;
    __syncthreads();
    if (has_elem) {
        double u[NEN][DIM], f[NEN][DIM];
        bool j2 = false;
        if constexpr (NL) {
            // Tangent operator: a J2 element whose points all hold the elastic tangent is the
            // elastic element of its initial constants (lam, mu), taken on the fast path below.
            const int tag = a.teid[e0 + tid];
            if (a.mat[tag] == SVLG_MAT_J2 && (RHS || a.plastic[tag])) {
                j2 = true;
                const double *q5 = a.j2p + size_t(tag) * 5;
                const J2Par jp{q5[0], q5[1], q5[2], q5[3], q5[4]};
#pragma unroll
                for (int c = 0; c < NEN; ++c)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) u[c][d] = RHS ? sV[lc[c] * DIM + d] : sU[lc[c] * DIM + d];
                const double cR = RHS ? c2 : a.coef[e0 + tid] - 1.0;
                if constexpr (BRICK) {
                    const double *gl = a.geo + size_t(gid) * 9;
                    const double l[3] = {gl[0], gl[4], gl[8]};
                    j2_brick_force(l, u, jp, a.state, size_t(a.nelem), tag, cR, !RHS, f);
                } else {
                    double x[NEN][DIM];
#pragma unroll
                    for (int c = 0; c < NEN; ++c)
#pragma unroll
                        for (int d = 0; d < DIM; ++d) x[c][d] = sX[lc[c] * DIM + d];
                    j2_elem_force<DIM>(x, u, jp, a.state, size_t(a.nelem), tag, cR, !RHS, a.th, f);
                }
            }
        }
        if (!j2) {
#pragma unroll
            for (int c = 0; c < NEN; ++c)
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    u[c][d] = sU[lc[c] * DIM + d];
                    if constexpr (RHS) u[c][d] += c2 * sV[lc[c] * DIM + d];
                }
            tile_elem_force<DIM, PATH>(a.geo, a.th, lc, sX, lam, mu, gid, u, f);
        }
        if constexpr (CM) {
            double w[NEN][DIM];
            double rs = a.cmass[e0 + tid];
            if constexpr (RHS) {
                const double am = a.cam[e0 + tid];
                rs = -rs;
#pragma unroll
                for (int c = 0; c < NEN; ++c)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) w[c][d] = sW[lc[c] * DIM + d] + am * sV[lc[c] * DIM + d];
            } else {
#pragma unroll
                for (int c = 0; c < NEN; ++c)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) w[c][d] = sU[lc[c] * DIM + d];
            }
            tile_elem_mass<DIM, PATH>(lc, sX, rs, w, f);
        }
        if constexpr (MOC > 0) {
#pragma unroll
            for (int c = 0; c < NEN; ++c) {
                const int col = lc[c] < nown ? lc[c] : MOC;
#pragma unroll
                for (int d = 0; d < DIM; ++d) sIn[(c * DIM + d) * MO + col] = f[c][d];
            }
        } else {
#pragma unroll
            for (int c = 0; c < NEN; ++c)
                if (lc[c] < nown) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) sIn[(c * DIM + d) * MO + lc[c]] = f[c][d];
                }
        }
    }
    __syncthreads();
    double acc = 0.0;
    if (tid < nown) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            double fi = 0.0;
#pragma unroll
            for (int c = 0; c < NEN; ++c) fi += sIn[(c * DIM + d) * MO + tid];
            const size_t g = g0 + size_t(tid) * DIM + d;
            const bool freedof = a.jac[g] != 0.0;
            if constexpr (RHS) {
                double fe = 0.0;
                const int node = o0 + tid;
                for (int q2 = a.nl_ptr[node]; q2 < a.nl_ptr[node + 1]; ++q2) {
                    const int l = a.nl_idx[q2];
                    const double amp =
                        (a.k >= 0 && a.k < a.load_nt) ? a.load_series[size_t(l) * a.load_nt + a.k] : 0.0;
                    fe += a.load_factor[l] * (amp * a.load_dir[size_t(l) * DIM + d]);
                }
                const int ds = drm_slot(a.drm, node);
                if (ds >= 0) fe += drm_force(a.drm, ds, a.k, DIM, d);
                const double v = sV[tid * DIM + d];
                // Newton residual (Vm set): the staged V is V - 2/dt dU, the mass term uses
                // the step's V and dU.
                const double mt = a.Vm ? 4.0 / a.dt * a.Vm[g] + a.A[g] - 4.0 / a.dt / a.dt * a.dUn[g]
                                       : 4.0 / a.dt * v + a.A[g];
                if constexpr (CM)
                    a.b[g] = freedof ? (fe - fi) + a.cdiag[g] * v : 0.0;
                else
                    a.b[g] = freedof ? ((fe - fi) + a.cdiag[g] * v) + a.mdiag[g] * mt : 0.0;
            } else {
                const double x = sU[tid * DIM + d];
                const double yv = freedof ? fi + a.dgn[g] * x : 0.0;
                a.y[g] = yv;
                acc += x * yv;
            }
        }
    }
    if constexpr (!RHS) {
        if (!a.reduce) return;
        const double sp = block_sum(acc, red);
        if (tid == 0) a.cg.pq[size_t(iter_of(a.jdev, a.iter) & 1) * nt + tile] = sp;
    }
}

__global__ void k_solver_reset(double *scal, double tol2, double maxit) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        for (int i = 0; i < SC_COUNT; ++i) scal[i] = 0.0;
        scal[SC_TOL2] = tol2;
        scal[SC_MAXIT] = maxit;
    }
}

// r_0 = b - q and, for the Jacobi preconditioner, z_0 = jac r_0; partials (r.z, r.r, b.b)
// into vp[0] and b.b into ip.
__global__ __launch_bounds__(512) void k_cg_init(int n, CgParts c, const double *__restrict__ b,
                                                 const double *__restrict__ q, const double *__restrict__ jac,
                                                 double *__restrict__ r, double *__restrict__ z, GcnInput g) {
    __shared__ double red[17];
    double rz = 0.0, rr = 0.0, bb = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double ri = jac[i] != 0.0 ? b[i] - q[i] : 0.0;
        r[i] = ri;
        if (!c.flexible) {
            const double zi = jac[i] * ri;
            z[i] = zi;
            rz += ri * zi;
        } else {
            gcn_input(g, n, i, ri);
        }
        rr += ri * ri;
        bb += b[i] * b[i];
    }
    rz = block_sum(rz, red);
    rr = block_sum(rr, red);
    bb = block_sum(bb, red);
    if (threadIdx.x == 0) {
        c.vp[3 * blockIdx.x + 0] = rz;
        c.vp[3 * blockIdx.x + 1] = rr;
        c.vp[3 * blockIdx.x + 2] = bb;
        c.ip[blockIdx.x] = bb;
    }
}

// Iteration j: alpha_j = r_j.z_j / p_j.q_j; x += alpha p; r -= alpha q; for Jacobi also
// z = jac r. Partials (r.z (Jacobi), r.r, b.b) of iteration j+1 into vp[(j+1)&1]; the first
// workgroup leaves r_j.z_j and p_j.q_j in scal for the next operator kernel.
__global__ __launch_bounds__(512) void k_cg_update(int n, int j0, CgParts c, double *scal, double *__restrict__ x,
                                                   const double *__restrict__ p, const double *__restrict__ q,
                                                   double *__restrict__ r, double *__restrict__ z,
                                                   const double *__restrict__ jac, const int *jdev, GcnInput g) {
    __shared__ double red[3 * 17];
    const double flag = scal[SC_DONE];
    const int j = iter_of(jdev, j0);
    if (flag != 0.0) return;
    const double pq = cg_pq(c, j, red);
    double rzj;
    if (c.flexible) {
        double zz[2];
        cg_pre(c, j, red, zz);
        rzj = zz[0];
    } else {
        double v[3];
        cg_vec(c, j, red, v);
        rzj = v[0];
    }
    const double alpha = pq != 0.0 ? rzj / pq : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[SC_RHO] = rzj;
        scal[SC_PQ] = pq;
        scal[SC_ALPHA] = alpha;
    }
    double rz = 0.0, rr = 0.0;
    // Grid-stride loop, four strides per pass with every load issued first (the per-thread
    // summation order is that of the plain grid-stride loop).
    constexpr int NB = 4;
    const int S = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x + threadIdx.x; base < n; base += NB * S) {
        double xv[NB], pv[NB], qv[NB], rv[NB], jv[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int i = base + k * S;
            const bool in = i < n;
            xv[k] = in ? x[i] : 0.0;
            pv[k] = in ? p[i] : 0.0;
            qv[k] = in ? q[i] : 0.0;
            rv[k] = in ? r[i] : 0.0;
            jv[k] = (in && !c.flexible) ? jac[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int i = base + k * S;
            if (i >= n) break;
            x[i] = xv[k] + alpha * pv[k];
            const double ri = rv[k] - alpha * qv[k];
            r[i] = ri;
            if (!c.flexible) {
                const double zi = jv[k] * ri;
                z[i] = zi;
                rz += ri * zi;
            } else {
                gcn_input(g, n, i, ri);
            }
            rr += ri * ri;
        }
    }
    rz = block_sum(rz, red);
    rr = block_sum(rr, red);
    if (threadIdx.x == 0) {
        double *v = c.vp + size_t((j + 1) & 1) * 3 * c.nvb;
        v[3 * blockIdx.x + 0] = rz;
        v[3 * blockIdx.x + 1] = rr;
        v[3 * blockIdx.x + 2] = c.ip[blockIdx.x];
    }
}

// After `niter` iterations without convergence: the stopping test on r_niter.
__global__ __launch_bounds__(256) void k_cg_final(int niter, CgParts c, double *scal) {
    __shared__ double red[3 * 17];
    if (scal[SC_DONE] != 0.0) return;
    double v[3];
    cg_vec(c, niter, red, v);
    if (threadIdx.x == 0) {
        scal[SC_RR] = v[1];
        scal[SC_BB] = v[2];
        scal[SC_ITERS] = double(niter);
        scal[SC_DONE] = (v[1] <= scal[SC_TOL2] * v[2] || v[2] == 0.0) ? 1.0 : 0.0;
    }
}

// A zero right-hand side has the zero solution whatever the initial guess.
__global__ void k_cg_zero_rhs(int n, const double *__restrict__ scal, double *__restrict__ x) {
    if (scal[SC_BB] != 0.0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = 0.0;
}

// U += dU; A = 4/dt^2 dU - 4/dt V - A; V = 2/dt dU - V   (ExtendedNewmarkBeta.cpp:141-150).
__global__ void k_nm_update(int n, double dt, const double *__restrict__ dU, const double *__restrict__ jac,
                            double *__restrict__ U, double *__restrict__ V, double *__restrict__ A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || jac[i] == 0.0) return;
    const double du = dU[i];
    U[i] += du;
    const double v = V[i];
    A[i] = 4.0 / dt / dt * du - 4.0 / dt * v - A[i];
    V[i] = 2.0 / dt * du - v;
}

// b -= y (the support-motion term of the effective force; y is 0 on the restrained DOFs).
__global__ void k_sub(int n, const double *__restrict__ y, double *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] -= y[i];
}

// Diagonal of sum_e (1 + 2 ak_e/dt) K_e by unit probes through the Gauss-point routine
// (valid for every element shape), one colour class per launch so no two threads of a
// launch touch the same node. rho non-null (consistent mass): plus (4/dt^2 + 2 am_e/dt) times
// the diagonal of the consistent element mass.
template <int DIM>
__global__ void k_kdiag_colored(int e0, int e1, const int32_t *__restrict__ conn_c, const int32_t *__restrict__ eid_c,
                                const double *__restrict__ X, const double *__restrict__ lam,
                                const double *__restrict__ mu, const double *__restrict__ ak, double inv_dt, double th,
                                const double *__restrict__ rho, const double *__restrict__ am,
                                double *__restrict__ kd) {
    const int e = e0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    constexpr int nen = Elem<DIM>::nen;
    const int tag = eid_c[e];
    int nodes[nen];
    double x[nen][DIM];
#pragma unroll
    for (int a = 0; a < nen; ++a) {
        nodes[a] = conn_c[size_t(e) * nen + a];
#pragma unroll
        for (int d = 0; d < DIM; ++d) x[a][d] = X[size_t(nodes[a]) * DIM + d];
    }
    const double cf = 1.0 + 2.0 * ak[tag] * inv_dt;
#pragma unroll 1
    for (int a = 0; a < nen; ++a)
#pragma unroll 1
        for (int d = 0; d < DIM; ++d) {
            double u[nen][DIM], f[nen][DIM];
#pragma unroll
            for (int b = 0; b < nen; ++b)
#pragma unroll
                for (int c = 0; c < DIM; ++c) u[b][c] = (b == a && c == d) ? 1.0 : 0.0;
            if constexpr (DIM == 2)
                quad_fint(x, u, lam[tag], mu[tag], th, f);
            else
                hexa_fint(x, u, lam[tag], mu[tag], f);
            double v = cf * f[a][d];
            if (rho) {
#pragma unroll
                for (int b = 0; b < nen; ++b)
#pragma unroll
                    for (int c = 0; c < DIM; ++c) f[b][c] = 0.0;
                if constexpr (DIM == 2)
                    quad_mass_gauss(x, rho[tag] * th, u, f);
                else
                    hexa_mass_gauss(x, rho[tag], u, f);
                v += (4.0 * inv_dt * inv_dt + 2.0 * am[tag] * inv_dt) * f[a][d];
            }
            kd[size_t(nodes[a]) * DIM + d] += v;
        }
}

__global__ void k_jacobi(int n, const int32_t *__restrict__ fdof, const double *__restrict__ kd,
                         const double *__restrict__ dgn, double *__restrict__ jac, int *__restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fdof[i] < 0) {
        jac[i] = 0.0;
        return;
    }
    const double d = kd[i] + dgn[i];
    if (!(d > 0.0)) atomicOr(bad, 1);
    jac[i] = d > 0.0 ? 1.0 / d : 0.0;
}

// M null (consistent mass, applied by the elements): the dashpot term alone.
__global__ void k_dgn(int n, double dt, const double *__restrict__ M, const double *__restrict__ C,
                      const int32_t *__restrict__ fdof, double *__restrict__ dgn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dgn[i] = fdof[i] < 0 ? 0.0 : (M ? 4.0 / dt / dt * M[i] + 2.0 / dt * C[i] : 2.0 / dt * C[i]);
}

__global__ void k_record_nm(int nrec, const int32_t *__restrict__ rec, const double *__restrict__ U,
                            double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec) out[r] = U[rec[r]];
}

// ------------------------------------------------------------------ host side
static constexpr int kVecGrid = 256, kVecBlock = 512;

static OpArgs op_args(const Model &m) {
    const TileSet &t = m.tiles;
    OpArgs a;
    std::memset(&a, 0, sizeof(a));
    a.own_ptr = t.own_ptr.p;
    a.halo_ptr = t.halo_ptr.p;
    a.halo_node = t.halo_node.p;
    a.elem_ptr = t.elem_ptr.p;
    a.lconn = t.elem_lconn.p;
    a.elam = t.elem_lam.p;
    a.emu = t.elem_mu.p;
    a.egeo = t.elem_geo.p;
    a.geo = m.geo.p;
    a.X = m.X.p;
    a.ntiles = t.ntiles;
    a.max_stage = t.max_stage;
    a.max_own = t.max_own;
    a.th = m.th;
    a.jac = m.jac.p;
    a.dgn = m.dgn.p;
    a.scal = m.scal.p;
    a.teid = m.tile_eid.p;
    a.mat = m.mat.p;
    a.j2p = m.j2p.p;
    a.state = m.gp_state.p;
    a.plastic = m.j2_plastic.p;
    a.nelem = m.nelem;
    return a;
}

// Compile-time slot stride of the CG operator on axis-aligned Hexa8 tiles of up to 168 or 336
// owned nodes (0: the stride read from the tile set).
static int op_moc(const Model &m, int mode) {
    (void)mode;  // both the CG operator and the right-hand side / residual
    if (m.ndim != 3 || m.elem_path != 3 || !m.lumped) return 0;
    const int mo = m.tiles.max_own;
    return mo <= 168 ? 168 : mo <= 336 ? 336 : 0;
}

static size_t op_lds(const Model &m, int mode) {
    const bool brick = m.elem_path == 3 && m.ndim == 3;  // J2 points on the corner sums: no coordinates
    const int staged = (mode == OP_RHS ? (m.lumped ? 2 : 3) : 1) + (m.elem_path < 2 || (m.has_j2 && !brick) ? 1 : 0);
    const int moc = op_moc(m, mode);
    const size_t stride = moc ? moc + 1 : m.tiles.max_own;
    return (size_t(staged) * m.tiles.max_stage * m.ndim + size_t(m.nen) * m.ndim * stride) * sizeof(double);
}

template <int DIM, int MODE, bool NL, bool CM>
static const void *op_fn_p(int path) {
    switch (path) {
        case 3: return reinterpret_cast<const void *>(&k_nm_op<DIM, 3, MODE, NL, 0, CM>);
        case 2: return reinterpret_cast<const void *>(&k_nm_op<DIM, 2, MODE, NL, 0, CM>);
        case 1: return reinterpret_cast<const void *>(&k_nm_op<DIM, 1, MODE, NL, 0, CM>);
        default: return reinterpret_cast<const void *>(&k_nm_op<DIM, 0, MODE, NL, 0, CM>);
    }
}

template <int DIM, int MODE>
static const void *op_fn_t(int path, bool nl, bool cm) {
    if (nl) return cm ? op_fn_p<DIM, MODE, true, true>(path) : op_fn_p<DIM, MODE, true, false>(path);
    return cm ? op_fn_p<DIM, MODE, false, true>(path) : op_fn_p<DIM, MODE, false, false>(path);
}

static const void *op_fn_sel(const Model &m, int mode, bool nl) {
    switch (op_moc(m, mode) * (mode == OP_CG ? 1 : -1)) {
        case 168:
            return nl ? reinterpret_cast<const void *>(&k_nm_op<3, 3, OP_CG, true, 168>)
                      : reinterpret_cast<const void *>(&k_nm_op<3, 3, OP_CG, false, 168>);
        case 336:
            return nl ? reinterpret_cast<const void *>(&k_nm_op<3, 3, OP_CG, true, 336>)
                      : reinterpret_cast<const void *>(&k_nm_op<3, 3, OP_CG, false, 336>);
        case -168:
            return nl ? reinterpret_cast<const void *>(&k_nm_op<3, 3, OP_RHS, true, 168>)
                      : reinterpret_cast<const void *>(&k_nm_op<3, 3, OP_RHS, false, 168>);
        case -336:
            return nl ? reinterpret_cast<const void *>(&k_nm_op<3, 3, OP_RHS, true, 336>)
                      : reinterpret_cast<const void *>(&k_nm_op<3, 3, OP_RHS, false, 336>);
        default: break;
    }
    const bool cm = !m.lumped;
    if (m.ndim == 3)
        return mode == OP_CG ? op_fn_t<3, OP_CG>(m.elem_path, nl, cm) : op_fn_t<3, OP_RHS>(m.elem_path, nl, cm);
    return mode == OP_CG ? op_fn_t<2, OP_CG>(m.elem_path, nl, cm) : op_fn_t<2, OP_RHS>(m.elem_path, nl, cm);
}

static const void *op_fn(const Model &m, int mode) {
    // J2 elements from their Gauss-point states, unless the initial (elastic) operator is
    // asked for (training pairs of the learned preconditioner) or no states exist yet (before
    // svlg_nl_init every point is at rest, where the tangent is the initial one).
    const bool nl = m.has_j2 && !m.nm_initial && m.gp_state.p != nullptr && m.j2_plastic.p != nullptr;
    return op_fn_sel(m, mode, nl);
}

static int launch_op(Model &m, int mode, const OpArgs &a0) {
    OpArgs a = a0;
    if (!m.lumped) {
        // Consistent mass: the element mass coefficients of this mode; the node terms keep the
        // dashpots alone.
        a.cmass = mode == OP_CG ? m.cmass_keff.p : m.cmass_rhs.p;
        a.cam = m.cmass_am.p;
        if (mode == OP_RHS) a.cdiag = m.cdash.p;
    }
    void *args[] = {&a};
    SVLG_HIP(hipLaunchKernel(op_fn(m, mode), dim3(m.tiles.ntiles), dim3(m.tiles.block), args, op_lds(m, mode),
                             m.stream));
    return 0;
}

// y = Keff x (node-DOF layout) through the CG operator with beta = 0.
int keff_apply_node(Model &m, const double *x, double *y) {
    hipLaunchKernelGGL(k_solver_reset, dim3(1), dim3(64), 0, m.stream, m.scal.p, 0.0, 0.0);
    OpArgs a = op_args(m);
    a.coef = m.coef_keff.p;
    a.z = x;
    a.p_old = x;
    a.p_new = m.cg_p[1].p;
    a.y = y;
    a.reduce = 0;
    return launch_op(m, OP_CG, a);
}

// b -= Keff a_k over the free DOFs (ExtendedNewmarkBeta::ComputeSupportMotionVector, :191-204:
// Feff -= T^T (K SupportMotion) with K the total effective stiffness).
int sm_keff_subtract(Model &m, int k, double *b) {
    if (m.sm_ndofs == 0) return 0;
    if (int rc = sm_vector(m, k, m.sm_vec.p)) return rc;
    if (int rc = keff_apply_node(m, m.sm_vec.p, m.sm_y.p)) return rc;
    hipLaunchKernelGGL(k_sub, dim3((m.ndofn + 255) / 256), dim3(256), 0, m.stream, m.ndofn, m.sm_y.p, b);
    SVLG_HIP(hipGetLastError());
    return 0;
}

// Partial-sum buffers of a solve inside Model::part.
static CgParts cg_parts(Model &m, int flexible) {
    CgParts c;
    c.ntiles = m.tiles.ntiles;
    c.nvb = kVecGrid;
    c.flexible = flexible;
    c.pq = m.part.p;
    c.vp = c.pq + 2 * size_t(c.ntiles);
    c.tp = c.vp + 6 * size_t(c.nvb);
    c.ip = c.tp + 6 * size_t(c.ntiles);
    return c;
}

static size_t cg_parts_size(const Model &m) { return 8 * size_t(m.tiles.ntiles) + 7 * size_t(kVecGrid); }

// Solves Keff x = b (node-DOF layout; x holds the initial guess on entry) by CG with the
// Jacobi preconditioner (precond = SVLG_PRECOND_JACOBI) or by flexible CG with the learned
// preconditioner (SVLG_PRECOND_LEARNED; beta = -z.q / p.q, Notay's FCG(1)). Per iteration:
// one operator launch, one vector launch (plus the preconditioner's); the host polls the
// done flag every 16 iterations.
__global__ void k_iter_set(int *jdev, int j) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *jdev = j;
}

__global__ void k_iter_advance(int *jdev, int by) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *jdev += by;
}

// Launches of CG iteration j (jdev null: j is the iteration; else *jdev + j is, with j the
// parity-fixed position inside a captured pair). p ping-pongs: even iterations stage
// p = z + beta cg_p[1] into cg_p[0], odd ones the reverse.
static int cg_iteration(Model &m, const CgParts &c, OpArgs &a, double *x, int j, const int *jdev) {
    const int cur = (j & 1) ? 0 : 1;
    a.iter = j;
    a.jdev = jdev;
    a.p_old = m.cg_p[cur].p;
    a.p_new = m.cg_p[1 - cur].p;
    if (int rc = launch_op(m, OP_CG, a)) return rc;
    hipLaunchKernelGGL(k_cg_update, dim3(kVecGrid), dim3(kVecBlock), 0, m.stream, m.ndofn, j, c, m.scal.p, x,
                       m.cg_p[1 - cur].p, m.cg_q.p, m.cg_r.p, m.cg_z.p, m.jac.p, jdev,
                       c.flexible ? gcn_input_args(m) : GcnInput{});
    if (c.flexible)
        if (int rc = gcn_precond(m, c, j + 1, m.cg_r.p, m.cg_q.p, m.cg_z.p, jdev)) return rc;
    return 0;
}

// The executable graph of `count` iterations *jdev .. *jdev + count - 1 (starting at an even
// iteration) followed by *jdev += count, captured on first use for (x, flexible, operator
// kernel, count).
static hipGraphExec_t cg_run_graph(Model &m, const CgParts &c, const OpArgs &a0, double *x, int count) {
    const void *op = op_fn(m, OP_CG);
    for (auto &g : m.cg_graphs)
        if (g.x == x && g.flexible == c.flexible && g.op == op && g.count == count) return g.exec;
    if (!m.jdev.p && m.jdev.alloc(1) != hipSuccess) return nullptr;
    if (hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    OpArgs a = a0;
    int rc = 0;
    for (int j = 0; j < count && rc == 0; ++j) rc = cg_iteration(m, c, a, x, j, m.jdev.p);
    hipLaunchKernelGGL(k_iter_advance, dim3(1), dim3(64), 0, m.stream, m.jdev.p, count);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(m.stream, &g);
    if (ec != hipSuccess || rc != 0 || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return nullptr;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    m.cg_graphs.push_back({exec, x, c.flexible, op, count});
    return exec;
}

int cg_solve(Model &m, const double *b, double *x, double rtol, int maxit, int precond, int *iters, double *relres,
             bool graph) {
    hipStream_t s = m.stream;
    const int n = m.ndofn;
    const int flexible = precond == SVLG_PRECOND_LEARNED ? 1 : 0;
    const CgParts c = cg_parts(m, flexible);
    hipLaunchKernelGGL(k_solver_reset, dim3(1), dim3(64), 0, s, m.scal.p, rtol * rtol, double(maxit));
    OpArgs a = op_args(m);
    a.coef = m.coef_keff.p;
    a.cg = c;
    a.y = m.cg_q.p;
    // q = Keff x0 (x0 staged as z with beta = 0; cg_p[1] receives a copy of x0).
    a.z = x;
    a.p_old = x;
    a.p_new = m.cg_p[1].p;
    a.reduce = 0;
    if (int rc = launch_op(m, OP_CG, a)) return rc;
    hipLaunchKernelGGL(k_cg_init, dim3(kVecGrid), dim3(kVecBlock), 0, s, n, c, b, m.cg_q.p, m.jac.p, m.cg_r.p,
                       m.cg_z.p, flexible ? gcn_input_args(m) : GcnInput{});
    if (flexible)
        if (int rc = gcn_precond(m, c, 0, m.cg_r.p, m.cg_q.p, m.cg_z.p)) return rc;
    a.z = m.cg_z.p;
    a.reduce = 1;
    hipGraphExec_t pair = graph && maxit >= 2 ? cg_run_graph(m, c, a, x, 2) : nullptr;
    hipGraphExec_t run8 = pair && maxit >= 8 ? cg_run_graph(m, c, a, x, 8) : nullptr;
    hipGraphExec_t one = pair ? cg_run_graph(m, c, a, x, 1) : nullptr;
    if (pair) hipLaunchKernelGGL(k_iter_set, dim3(1), dim3(64), 0, s, m.jdev.p, 0);
    // The polls copy the scalars into pinned host memory (a direct copy, no staging).
    if (!m.scal_host && hipHostMalloc((void **)&m.scal_host, sizeof(double) * SC_COUNT) != hipSuccess) {
        m.scal_host = nullptr;
        set_error("hipHostMalloc failed");
        return 10;
    }
    double *hs = m.scal_host;
    int launched = 0;  // iterations 0 .. launched-1 are queued
    bool done = false;
    // The solve that converges at iteration j sets DONE in the operator kernel of iteration
    // j (Jacobi) or in the first preconditioner layer of iteration j - 1 (learned: the
    // converged residual is not preconditioned), i.e. once `det` iterations are queued. With
    // the previous solve's count as the prediction (successive time steps need nearly the same
    // number) the flag is first read back then, and after every launch from there on; every
    // 16 iterations while there is no prediction. Iterations queued past convergence exit at
    // once on the device.
    const int expect = m.cg_last_iters[flexible];
    const int det = expect > 0 ? expect + 1 - flexible : 0;
    while (launched < maxit && !done) {
        const int before = launched;
        // A pair that starts one short of `det` would queue one iteration past the one
        // expected to see DONE: that iteration goes alone.
        const bool tail = det > 0 && launched == det - 1;
        if (pair && launched + 2 <= maxit && !tail) {
            // The graphs start at even iterations: *jdev tracks `launched`. Runs of eight while
            // they stay within the prediction (every graph launch costs a few microseconds of
            // dispatch), then the rest of the prediction as one graph (captured on first use
            // for its count; after an odd count further iterations are launched eagerly), then
            // pairs.
            const int lim = det > 0 ? std::min(maxit, det) : maxit;
            const int rem = lim - launched;
            hipGraphExec_t rest = det > 0 && rem > 2 && rem < 8 ? cg_run_graph(m, c, a, x, rem) : nullptr;
            if (rest) {
                SVLG_HIP(hipGraphLaunch(rest, s));
                launched += rem;
                if (rem & 1) pair = nullptr;
            } else if (run8 && launched + 8 <= lim) {
                SVLG_HIP(hipGraphLaunch(run8, s));
                launched += 8;
            } else {
                SVLG_HIP(hipGraphLaunch(pair, s));
                launched += 2;
            }
        } else if (pair && tail && one) {
            // The predicted last iteration alone; the count is odd after it, so any further
            // iteration is launched eagerly.
            SVLG_HIP(hipGraphLaunch(one, s));
            launched += 1;
            pair = nullptr;
        } else {
            if (pair) pair = nullptr;  // odd remainder: eager from here on
            if (int rc = cg_iteration(m, c, a, x, launched, nullptr)) return rc;
            launched += 1;
        }
        if ((det > 0 && launched >= det) || (before >> 4) != (launched >> 4) || launched >= maxit) {
            SVLG_HIP(hipMemcpyAsync(hs, m.scal.p, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, s));
            SVLG_HIP(hipStreamSynchronize(s));
            done = hs[SC_DONE] != 0.0;
        }
    }
    if (!done) {
        // The stopping test on r_maxit, then the scalars.
        hipLaunchKernelGGL(k_cg_final, dim3(1), dim3(256), 0, s, launched, c, m.scal.p);
        SVLG_HIP(hipGetLastError());
        SVLG_HIP(hipMemcpyAsync(hs, m.scal.p, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, s));
        SVLG_HIP(hipStreamSynchronize(s));
    }
    // The poll that saw DONE holds the final scalars (written with the flag); the work
    // queued after this call needs no further synchronisation.
    if (hs[SC_BB] == 0.0) hipLaunchKernelGGL(k_cg_zero_rhs, dim3(kVecGrid), dim3(kVecBlock), 0, s, n, m.scal.p, x);
    SVLG_HIP(hipGetLastError());
    if (iters) *iters = int(hs[SC_ITERS]);
    m.cg_last_iters[flexible] = int(hs[SC_ITERS]);
    const double rel = hs[SC_BB] > 0.0 ? std::sqrt(hs[SC_RR] / hs[SC_BB]) : 0.0;
    if (relres) *relres = rel;
    if (hs[SC_DONE] == 0.0 || !(rel <= rtol)) {
        set_error("conjugate gradient did not reach the tolerance within the iteration limit");
        return 12;
    }
    return 0;
}

// NewmarkBeta residual of a Newton iterate (NewmarkBeta.cpp:212-228): b = Fext(k) - Fint(U + dU) + M (4/dt V + A - 4/dt^2 dU) + C (V - 2/dt dU), with
// un = U + dU and w = V - 2/dt dU already formed.
int launch_nl_rhs(Model &m, int k, const double *un, const double *w, const double *dU, double *b) {
    OpArgs a = op_args(m);
    a.coef = m.coef_rhs.p;
    a.U = un;
    a.V = w;
    a.Vm = m.nV.p;
    a.A = m.nA.p;
    a.dUn = dU;
    a.mdiag = m.mdiag.p;
    a.cdiag = m.cdiag.p;
    a.dt = m.nm_dt;
    a.b = b;
    a.nl_ptr = m.nl_ptr.p;
    a.nl_idx = m.nl_idx.p;
    a.load_dir = m.load_dir.p;
    a.load_factor = m.load_factor.p;
    a.load_series = m.load_series.p;
    a.load_nt = m.load_nt;
    a.k = k;
    a.drm = drm_ref(m);
    return launch_op(m, OP_RHS, a);
}

}  // namespace svlg

using namespace svlg;

extern "C" {

int svlg_nm_init(svlg_model *h, double dt, double rtol, int32_t maxit, int32_t precond) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(dt > 0.0 && rtol > 0.0 && maxit > 0, "dt, rtol and maxit must be positive");
    SVLG_CHECK(m.lumped || m.ndrm == 0,
               "domain reduction loads on the device need MassFormulation = LUMPED (their element forces are "
               "evaluated with a diagonal mass)");
    SVLG_CHECK(m.variant == 3, "the device Newmark path needs corner-slot node tiles (a corner-unique mesh)");
    SVLG_CHECK(precond == SVLG_PRECOND_JACOBI,
               "svlg_nm_init takes SVLG_PRECOND_JACOBI; switch with svlg_nm_set_precond after svlg_gnn_init");
    for (int mode : {OP_CG, OP_RHS})
        SVLG_CHECK(op_lds(m, mode) <= 160 * 1024, "Newmark operator tile needs more than 160 KB of LDS; use smaller tiles");
    hipStream_t s = m.stream;
    const int n = m.ndofn;
    m.nm_ready = false;
    m.gcn_ready = false;
    m.drop_cg_graphs();
    m.nm_dt = dt;
    m.nm_rtol = rtol;
    m.nm_maxit = maxit;
    m.nm_precond = precond;
    for (DevBuf<double> *b : {&m.nU, &m.nV, &m.nA, &m.ndU, &m.dgn, &m.jac, &m.cg_r, &m.cg_z, &m.cg_q, &m.cg_b,
                              &m.cg_p[0], &m.cg_p[1]}) {
        SVLG_HIP(b->alloc(n));
        SVLG_HIP(hipMemsetAsync(b->p, 0, sizeof(double) * n, s));
    }
    if (m.sm_ndofs > 0) {
        for (DevBuf<double> *b : {&m.sm_vec, &m.sm_y}) {
            SVLG_HIP(b->alloc(n));
            SVLG_HIP(hipMemsetAsync(b->p, 0, sizeof(double) * n, s));
        }
        SVLG_HIP(hipMemsetAsync(m.sm_lag.p, 0, sizeof(double) * m.sm_ndofs, s));
    }
    SVLG_HIP(m.part.alloc(cg_parts_size(m)));
    SVLG_HIP(m.scal.alloc(SC_COUNT));
    SVLG_HIP(hipMemsetAsync(m.scal.p, 0, sizeof(double) * SC_COUNT, s));
    // Per tile-element coefficients: 1 + 2 ak/dt (operator) and -ak (right-hand side).
    std::vector<double> ak_tag(m.nelem);
    SVLG_HIP(hipMemcpyAsync(ak_tag.data(), m.ak.p, sizeof(double) * m.nelem, hipMemcpyDeviceToHost, s));
    SVLG_HIP(hipStreamSynchronize(s));
    std::vector<double> ck(m.tile_elem_id.size()), cr(m.tile_elem_id.size());
    for (size_t i = 0; i < ck.size(); ++i) {
        const double ak = ak_tag[m.tile_elem_id[i]];
        ck[i] = 1.0 + 2.0 * ak / dt;
        cr[i] = -ak;
    }
    SVLG_HIP(m.coef_keff.upload(ck.data(), ck.size(), s));
    SVLG_HIP(m.coef_rhs.upload(cr.data(), cr.size(), s));
    if (!m.lumped) {
        // Consistent mass: rs = rho |J|/8 (rho th |J|/4) on the table paths, rho (rho th) where the
        // element routine forms |J| itself.
        std::vector<double> mk(ck.size()), mr(ck.size()), ma(ck.size());
        for (size_t i = 0; i < ck.size(); ++i) {
            const int e = m.tile_elem_id[i];
            const double rs = m.elem_path >= 2 ? m.h_rho[e] * m.h_escale[e] : m.h_rho[e] * (m.ndim == 2 ? m.th : 1.0);
            mk[i] = rs * (4.0 / dt / dt + 2.0 * m.h_am[e] / dt);
            mr[i] = rs;
            ma[i] = m.h_am[e];
        }
        SVLG_HIP(m.cmass_keff.upload(mk.data(), mk.size(), s));
        SVLG_HIP(m.cmass_rhs.upload(mr.data(), mr.size(), s));
        SVLG_HIP(m.cmass_am.upload(ma.data(), ma.size(), s));
    }
    // dgn = (4/dt^2) M + (2/dt) C_diag (lumped), (2/dt) C_dash (consistent); Jacobi = 1 / diag(Keff).
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_dgn, dim3(nb), dim3(256), 0, s, n, dt, m.lumped ? m.mdiag.p : nullptr,
                       m.lumped ? m.cdiag.p : m.cdash.p, m.fdof.p, m.dgn.p);
    DevBuf<double> kd;
    SVLG_HIP(kd.alloc(n));
    SVLG_HIP(hipMemsetAsync(kd.p, 0, sizeof(double) * n, s));
    for (int c = 0; c < m.ncolors; ++c) {
        const int e0 = m.color_off[c], e1 = m.color_off[c + 1];
        if (e1 <= e0) continue;
        if (m.ndim == 3)
            hipLaunchKernelGGL(k_kdiag_colored<3>, dim3((e1 - e0 + 127) / 128), dim3(128), 0, s, e0, e1, m.conn_c.p,
                               m.eid_c.p, m.X.p, m.lam.p, m.mu.p, m.ak.p, 1.0 / dt, m.th,
                               m.lumped ? nullptr : m.rho.p, m.am.p, kd.p);
        else
            hipLaunchKernelGGL(k_kdiag_colored<2>, dim3((e1 - e0 + 127) / 128), dim3(128), 0, s, e0, e1, m.conn_c.p,
                               m.eid_c.p, m.X.p, m.lam.p, m.mu.p, m.ak.p, 1.0 / dt, m.th,
                               m.lumped ? nullptr : m.rho.p, m.am.p, kd.p);
    }
    DevBuf<int> bad;
    SVLG_HIP(bad.alloc(1));
    SVLG_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_jacobi, dim3(nb), dim3(256), 0, s, n, m.fdof.p, kd.p, m.dgn.p, m.jac.p, bad.p);
    SVLG_HIP(hipGetLastError());
    int hb = 0;
    SVLG_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SVLG_HIP(hipStreamSynchronize(s));
    SVLG_CHECK(hb == 0, "effective stiffness has a non-positive diagonal entry");
    for (int mode : {OP_CG, OP_RHS}) {
        const size_t lds = op_lds(m, mode);
        if (lds > 64 * 1024)
            for (bool nl : {false, true}) {
                if (nl && !m.has_j2) continue;
                SVLG_HIP(hipFuncSetAttribute(op_fn_sel(m, mode, nl), hipFuncAttributeMaxDynamicSharedMemorySize,
                                             int(lds)));
            }
    }
    m.nm_steps_done = 0;
    m.nm_ready = true;
    m.active = 2;
    return 0;
}

int svlg_nm_steps(svlg_model *h, int32_t k0, int32_t nsteps, int32_t nrec, const int32_t *rec_dofs, double *hist,
                  int32_t *iters, double *relres) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    SVLG_CHECK(!m.has_j2 || m.nm_initial, "models with J2 materials step with svlg_nl_steps (Newton-Raphson)");
    SVLG_CHECK(nsteps >= 0 && nrec >= 0, "negative count");
    hipStream_t s = m.stream;
    DevBuf<int32_t> rec;
    DevBuf<double> hbuf;
    if (nrec > 0 && nsteps > 0) {
        SVLG_CHECK(rec_dofs && hist, "recorder arrays missing");
        std::vector<int32_t> rn;
        if (int rc = record_positions(m, nrec, rec_dofs, rn)) return rc;
        SVLG_HIP(rec.upload(rn.data(), rn.size(), s));
        SVLG_HIP(hbuf.alloc(size_t(nrec) * nsteps));
    }
    OpArgs a = op_args(m);
    a.coef = m.coef_rhs.p;
    a.U = m.nU.p;
    a.V = m.nV.p;
    a.A = m.nA.p;
    a.mdiag = m.mdiag.p;
    a.cdiag = m.cdiag.p;
    a.dt = m.nm_dt;
    a.b = m.cg_b.p;
    a.nl_ptr = m.nl_ptr.p;
    a.nl_idx = m.nl_idx.p;
    a.load_dir = m.load_dir.p;
    a.load_factor = m.load_factor.p;
    a.load_series = m.load_series.p;
    a.load_nt = m.load_nt;
    const int nb = (m.ndofn + 255) / 256;
    for (int st = 0; st < nsteps; ++st) {
        // Feff (ExtendedNewmarkBeta::ComputeEffectiveForce with the Linear algorithm's dU = 0).
        a.k = k0 + st;
        a.drm = drm_ref(m);
        // Support motion: the internal force sees U before the last support increment (the
        // strains lag one increment), and Feff -= Keff a_k.
        if (int rc = sm_lag_in(m, m.nU.p, nullptr, m.sm_lag.p)) return rc;
        if (int rc = launch_op(m, OP_RHS, a)) return rc;
        if (int rc = sm_keff_subtract(m, a.k, m.cg_b.p)) return rc;
        // Keff dU = Feff; the previous increment is the initial guess.
        int it = 0;
        double rel = 0.0;
        const int rc = cg_solve(m, m.cg_b.p, m.ndU.p, m.nm_rtol, m.nm_maxit, m.nm_precond, &it, &rel, !m.cg_eager);
        if (iters) iters[st] = it;
        if (relres) relres[st] = rel;
        if (rc) {
            (void)sm_finish(m, 0, a.k, m.nU.p, nullptr);
            return rc;
        }
        hipLaunchKernelGGL(k_nm_update, dim3(nb), dim3(256), 0, s, m.ndofn, m.nm_dt, m.ndU.p, m.jac.p, m.nU.p, m.nV.p,
                           m.nA.p);
        // U_k = U_(k-1) + a_k at the support-motion DOFs; U_(k-1) becomes the next lag.
        if (int rc = sm_finish(m, 2, a.k, m.nU.p, nullptr)) return rc;
        if (nrec > 0)
            hipLaunchKernelGGL(k_record_nm, dim3((nrec + 63) / 64), dim3(64), 0, s, nrec, rec.p, m.nU.p,
                               hbuf.p + size_t(st) * nrec);
        m.nm_steps_done++;
    }
    SVLG_HIP(hipGetLastError());
    if (nrec > 0 && nsteps > 0)
        SVLG_HIP(hipMemcpyAsync(hist, hbuf.p, sizeof(double) * size_t(nrec) * nsteps, hipMemcpyDeviceToHost, s));
    SVLG_HIP(hipStreamSynchronize(s));
    return 0;
}

int svlg_nm_effective_force(svlg_model *h, int32_t k, double factor, double *Feff) {
    (void)factor;  // unused by ExtendedNewmarkBeta::ComputeEffectiveForce as well
    SVLG_CHECK(h && Feff, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    SVLG_CHECK(!m.has_j2 || m.nm_initial, "models with J2 materials form their residual in svlg_nl_steps");
    OpArgs a = op_args(m);
    a.coef = m.coef_rhs.p;
    a.U = m.nU.p;
    a.V = m.nV.p;
    a.A = m.nA.p;
    a.mdiag = m.mdiag.p;
    a.cdiag = m.cdiag.p;
    a.dt = m.nm_dt;
    a.b = m.cg_b.p;
    a.nl_ptr = m.nl_ptr.p;
    a.nl_idx = m.nl_idx.p;
    a.load_dir = m.load_dir.p;
    a.load_factor = m.load_factor.p;
    a.load_series = m.load_series.p;
    a.load_nt = m.load_nt;
    a.k = k;
    a.drm = drm_ref(m);
    if (int rc = sm_lag_in(m, m.nU.p, nullptr, m.sm_lag.p)) return rc;
    if (int rc = launch_op(m, OP_RHS, a)) return rc;
    if (int rc = sm_finish(m, 0, k, m.nU.p, nullptr)) return rc;
    DevBuf<double> ff;
    SVLG_HIP(ff.alloc(std::max(m.nfree, 1)));
    if (int rc = gather_free(m, m.cg_b.p, ff.p)) return rc;
    SVLG_HIP(hipMemcpyAsync(Feff, ff.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_nm_apply_increment(svlg_model *h, int32_t k, const double *dU) {
    SVLG_CHECK(h && dU, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    DevBuf<double> ff;
    SVLG_HIP(ff.upload(dU, std::max(m.nfree, 1), m.stream));
    if (int rc = scatter_free(m, ff.p, m.ndU.p)) return rc;
    const int nb = (m.ndofn + 255) / 256;
    hipLaunchKernelGGL(k_nm_update, dim3(nb), dim3(256), 0, m.stream, m.ndofn, m.nm_dt, m.ndU.p, m.jac.p, m.nU.p,
                       m.nV.p, m.nA.p);
    SVLG_HIP(hipGetLastError());
    // DU = dU + SupportMotion (:134-135): the lag takes U_(k-1), U gains a_k.
    if (int rc = sm_lag_in(m, m.nU.p, m.nU.p, nullptr)) return rc;
    if (int rc = sm_finish(m, 2, k, m.nU.p, nullptr)) return rc;
    m.nm_steps_done++;
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_support_motion_vector(svlg_model *h, int32_t k, double factor, double *Feff) {
    (void)factor;  // unused by the integrators as well
    SVLG_CHECK(h && Feff, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.active == 1 || m.active == 2, "no integrator has been initialised");
    // CentralDifference: K = M/dt^2 + C/2dt is diagonal on the device (lumped M, Rayleigh ak = 0,
    // dashpots), so T^T (K a_k) has no free entries and Feff is unchanged.
    if (m.sm_ndofs == 0 || m.active == 1) return 0;
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    DevBuf<double> bn, ff;
    SVLG_HIP(bn.alloc(m.ndofn));
    SVLG_HIP(ff.alloc(std::max(m.nfree, 1)));
    SVLG_HIP(hipMemsetAsync(bn.p, 0, sizeof(double) * m.ndofn, m.stream));
    if (int rc = sm_keff_subtract(m, k, bn.p)) return rc;
    if (int rc = gather_free(m, bn.p, ff.p)) return rc;
    std::vector<double> hb(m.nfree);
    SVLG_HIP(hipMemcpyAsync(hb.data(), ff.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    for (int i = 0; i < m.nfree; ++i) Feff[i] += hb[i];  // hb = -(Keff a_k) at the free DOFs
    return 0;
}

int svlg_nm_set_precond(svlg_model *h, int32_t precond) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    SVLG_CHECK(precond == SVLG_PRECOND_JACOBI || precond == SVLG_PRECOND_LEARNED, "unknown preconditioner");
    SVLG_CHECK(precond != SVLG_PRECOND_LEARNED || m.gcn_ready, "svlg_gnn_init has not been called");
    m.nm_precond = precond;
    return 0;
}

int svlg_nm_get_state(svlg_model *h, double *U, double *V, double *A) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    DevBuf<double> tmp;
    SVLG_HIP(tmp.alloc(std::max(m.nfree, 1)));
    double *outs[3] = {U, V, A};
    const double *src[3] = {m.nU.p, m.nV.p, m.nA.p};
    for (int i = 0; i < 3; ++i) {
        if (!outs[i]) continue;
        if (int rc = gather_free(m, src[i], tmp.p)) return rc;
        SVLG_HIP(hipMemcpyAsync(outs[i], tmp.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
        SVLG_HIP(hipStreamSynchronize(m.stream));
    }
    return 0;
}

int svlg_keff_apply(svlg_model *h, const double *xfree, double *yfree) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    SVLG_CHECK(xfree && yfree, "null vector");
    DevBuf<double> xf, xn, yn;
    SVLG_HIP(xf.upload(xfree, m.nfree, m.stream));
    SVLG_HIP(xn.alloc(m.ndofn));
    SVLG_HIP(yn.alloc(m.ndofn));
    SVLG_HIP(hipMemsetAsync(xn.p, 0, sizeof(double) * m.ndofn, m.stream));
    if (int rc = scatter_free(m, xf.p, xn.p)) return rc;
    if (int rc = keff_apply_node(m, xn.p, yn.p)) return rc;
    if (int rc = gather_free(m, yn.p, xf.p)) return rc;
    SVLG_HIP(hipMemcpyAsync(yfree, xf.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_solve(svlg_model *h, const double *bfree, double *xfree, double rtol, int32_t maxit, int32_t *iters,
               double *relres) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    SVLG_CHECK(bfree && xfree, "null vector");
    SVLG_CHECK(rtol > 0.0 && maxit > 0, "rtol and maxit must be positive");
    DevBuf<double> f, bn, xn;
    SVLG_HIP(f.upload(bfree, m.nfree, m.stream));
    SVLG_HIP(bn.alloc(m.ndofn));
    SVLG_HIP(xn.alloc(m.ndofn));
    SVLG_HIP(hipMemsetAsync(bn.p, 0, sizeof(double) * m.ndofn, m.stream));
    SVLG_HIP(hipMemsetAsync(xn.p, 0, sizeof(double) * m.ndofn, m.stream));
    if (int rc = scatter_free(m, f.p, bn.p)) return rc;
    int it = 0;
    double rel = 0.0;
    const int rc = cg_solve(m, bn.p, xn.p, rtol, maxit, m.nm_precond, &it, &rel);
    if (iters) *iters = it;
    if (relres) *relres = rel;
    if (rc) return rc;
    if (int rc2 = gather_free(m, xn.p, f.p)) return rc2;
    SVLG_HIP(hipMemcpyAsync(xfree, f.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

}  // extern "C"
