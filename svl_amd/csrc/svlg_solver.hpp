// svlg_solver.hpp -- device helpers shared by the conjugate-gradient solver and the
// learned preconditioner: the solver's device scalars, deterministic workgroup sums and
// the last-workgroup reduction.
#pragma once

#include "svlg_common.hpp"

namespace svlg {

// Device scalars of one solve (Model::scal).
enum {
    SC_RHO = 0,  // r . z of the current residual
    SC_PQ,       // p . Keff p of the current direction
    SC_RR,       // r . r
    SC_BB,       // b . b
    SC_ALPHA,
    SC_BETA,
    SC_DONE,     // 1 once converged or out of iterations
    SC_ITERS,
    SC_TOL2,     // rtol^2
    SC_MAXIT,
    SC_COUNT = 16
};

// Counters of the last-workgroup reductions (Model::counters), one per kernel kind.
enum { CNT_OP = 0, CNT_VEC, CNT_PRE, CNT_COUNT = 8 };

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

// Called by every thread of every workgroup of a grid after thread 0 has stored the
// workgroup's partial sums: returns true in exactly one workgroup, the last to arrive,
// after which the partials of all workgroups are visible to it. The winner resets the
// counter (release_counter) once it is done with the partials.
__device__ __forceinline__ bool arrive_last(unsigned *counter, unsigned nblocks, int *flag) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned t = atomicAdd(counter, 1u);
        *flag = (t == nblocks - 1) ? 1 : 0;
    }
    __syncthreads();
    const bool last = *flag != 0;
    if (last) __threadfence();
    return last;
}

__device__ __forceinline__ void release_counter(unsigned *counter) {
    if (threadIdx.x == 0) atomicExch(counter, 0u);
}

// Coherent read of a partial written by another workgroup.
__device__ __forceinline__ double load_partial(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Fixed-order sum of count partials (stride apart) by one workgroup.
__device__ __forceinline__ double sum_partials(const double *part, int count, int stride, double *buf) {
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) s += load_partial(part + size_t(i) * stride);
    return block_sum(s, buf);
}

// Preconditioner hook of the flexible CG loop (svlg_gcn.hip): z = P(r) on the node-DOF
// layout; the last kernel it launches also forms z.r and z.q and sets the CG scalars
// RHO and BETA (BETA = 0 when first).
int gcn_precond(Model &m, const double *r, const double *q, double *z, bool first);

}  // namespace svlg
