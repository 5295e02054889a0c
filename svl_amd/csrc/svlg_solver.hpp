// svlg_solver.hpp -- device helpers shared by the conjugate-gradient solver and the
// learned preconditioner: the solver's device scalars, deterministic workgroup sums and the
// partial-sum buffers through which the kernels of one iteration hand over their dot products.
#pragma once

#include "svlg_common.hpp"

namespace svlg {

// Device scalars of one solve (Model::scal); written by the first workgroup of the kernels
// for the host (progress, stopping) and read by later kernels (TOL2, DONE).
enum {
    SC_RHO = 0,  // r . z of the current residual
    SC_PQ,       // p . Keff p of the current direction
    SC_RR,       // r . r
    SC_BB,       // b . b
    SC_ALPHA,
    SC_BETA,
    SC_DONE,     // 1 once converged
    SC_ITERS,
    SC_TOL2,     // rtol^2
    SC_MAXIT,
    SC_COUNT = 16
};

// Deterministic sum over a workgroup (blockDim a multiple of 64, at most 1024): a fixed
// shuffle tree inside each wave, then the wave sums in wave order. buf holds 17 doubles.
__device__ __forceinline__ double block_sum(double v, double *buf) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) buf[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int i = 0; i < int(blockDim.x >> 6); ++i) r += buf[i];
        buf[16] = r;
    }
    __syncthreads();
    const double r = buf[16];
    __syncthreads();
    return r;
}

// Fixed-order sum of count partials (stride apart) by one workgroup.
__device__ __forceinline__ double sum_partials(const double *part, int count, int stride, double *buf) {
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) s += part[size_t(i) * stride];
    return block_sum(s, buf);
}

// The workgroup part of sum_records: per-thread sums v (fixed order) to the totals, by a
// fixed shuffle tree inside each wave and the wave sums in wave order. buf holds 17 * K doubles.
template <int K>
__device__ __forceinline__ void reduce_records(double (&v)[K], double *buf, double (&out)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) buf[k * 17 + w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) {
            double r = 0.0;
            for (int i = 0; i < int(blockDim.x >> 6); ++i) r += buf[k * 17 + i];
            buf[k * 17 + 16] = r;
        }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = buf[k * 17 + 16];
    __syncthreads();
}

// Fixed-order sums of the first K components of count records (stride doubles apart), one
// pass and one workgroup round. buf holds 17 * K doubles.
template <int K>
__device__ __forceinline__ void sum_records(const double *part, int count, int stride, double *buf, double (&out)[K]) {
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += part[size_t(i) * stride + k];
    reduce_records<K>(v, buf, out);
}

// Iteration index of a kernel launch: the launch's own index, plus the device-resident base
// *jdev when the launch belongs to a replayed graph of iterations (jdev non-null).
__device__ __forceinline__ int iter_of(const int *jdev, int j) { return jdev ? *jdev + j : j; }

// Partial sums of one solve. Iteration j works on r_j, z_j, p_j, q_j = Keff p_j. Every kernel
// writes its per-workgroup partials into the half of a ping-pong pair picked by the parity of
// the iteration it produces, and the kernels that need a total re-sum the partials in
// workgroup order at their start: the kernel boundary makes them visible, and every workgroup
// derives bitwise the same scalars, with no fence, atomic or reduction launch.
//   vp[j&1]  [nvb][3]       (r_j . z_j (Jacobi) or 0, r_j . r_j, b . b)  (vector kernel of j-1 / init)
//   tp[j&1]  [ntiles][3]    (z_j . r_j, z_j . q_(j-1), 0) per tile        (preconditioner after j-1)
//   pq[j&1]  [ntiles]       p_j . q_j per tile                            (operator kernel of j)
//   ip       [nvb]          b . b per vector workgroup                    (init)
// The totals a kernel needs from the previous iteration (r_(j-1).z_(j-1), p_(j-1).q_(j-1)) are
// left in Model::scal by the first workgroup of the vector kernel that used them.
struct CgParts {
    double *pq, *vp, *tp, *ip;
    int ntiles, nvb, flexible;
};

// Input layer of the learned preconditioner, formed by the CG vector kernels of the flexible
// solve from the residual they produce: x0 = s r and H^1 = x0 c_1^T (C channels, FP32), with s
// = D^-1/2, or D^-1 for the one-channel kernel's scaled states.
struct GcnInput {
    const double *s;
    const float *c1;
    float *x0, *h1;
    int C;
};

__device__ __forceinline__ void gcn_input(const GcnInput &g, int n, int i, double r) {
    const float xv = float(g.s[i] * r);
    g.x0[i] = xv;
    for (int k = 0; k < g.C; ++k) g.h1[size_t(k) * n + i] = xv * g.c1[k];
}

// (r.z, r.r, b.b) of iteration j from the vector partials.
__device__ __forceinline__ void cg_vec(const CgParts &c, int j, double *buf, double (&o)[3]) {
    sum_records<3>(c.vp + size_t(j & 1) * 3 * c.nvb, c.nvb, 3, buf, o);
}
// CG's stopping test on r_j, j = iter_of(jdev, j0) (r.r <= rtol^2 b.b, or b = 0), from the
// vector partials as the Jacobi solve's operator kernel applies it: the r.r and b.b sums are
// bitwise cg_vec's components 1 and 2 (per-thread order and reduction tree of sum_records),
// so every workgroup reaches the same answer. In the flexible solve this is the only test
// (the operator kernel relies on it). Returns true (and the first workgroup records the scalars and DONE) if the
// solve has converged, or if DONE was already set.
__device__ __forceinline__ bool cg_stop_test(const CgParts &c, const int *jdev, int j0, double *scal, double *buf) {
    const double flag = scal[SC_DONE];
    const int j = iter_of(jdev, j0);
    if (flag != 0.0) return true;
    double v[2];
    sum_records<2>(c.vp + size_t(j & 1) * 3 * c.nvb + 1, c.nvb, 3, buf, v);
    const bool done = v[0] <= scal[SC_TOL2] * v[1] || v[1] == 0.0;
    if (done && blockIdx.x == 0 && threadIdx.x == 0) {
        scal[SC_RR] = v[0];
        scal[SC_BB] = v[1];
        scal[SC_ITERS] = double(j);
        scal[SC_DONE] = 1.0;
    }
    return done;
}
__device__ __forceinline__ double cg_pq(const CgParts &c, int j, double *buf) {
    return sum_partials(c.pq + size_t(j & 1) * c.ntiles, c.ntiles, 1, buf);
}
// (z.r, z.q) of iteration j from the preconditioner partials.
__device__ __forceinline__ void cg_pre(const CgParts &c, int j, double *buf, double (&o)[2]) {
    sum_records<2>(c.tp + size_t(j & 1) * 3 * c.ntiles, c.ntiles, 3, buf, o);
}

// Preconditioner hook of the flexible CG loop (svlg_gcn.hip): z = P(r) on the node-DOF
// layout for iteration j (z = z_j; r = r_j; q = q_(j-1)) from the input layer the vector
// kernel already formed (GcnInput); writes the tp[j&1] partials (z.r, z.q per tile). With
// jdev the iteration is *jdev + j (graph replays). Its first layer applies the stopping test
// on r_j (cg_stop_test), so that a converged residual is not preconditioned: DONE is then set
// one launch after the vector kernel of iteration j - 1.
int gcn_precond(Model &m, const CgParts &c, int j, const double *r, const double *q, double *z,
                const int *jdev = nullptr);
// The input layer of the solver's preconditioner (x0, H^1 buffers and c_1).
GcnInput gcn_input_args(Model &m);

// Preconditioned CG (Jacobi) or flexible CG (learned) on Keff x = b from the given x, to
// ||b - Keff x|| <= rtol ||b|| (svlg_implicit.hip). graph: replay captured pairs of
// iterations (for solution buffers reused across calls, such as the integrators' increment).
int cg_solve(Model &m, const double *b, double *x, double rtol, int maxit, int precond, int *iters, double *relres,
             bool graph = false);

}  // namespace svlg
