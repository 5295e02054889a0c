// svlg_support.hpp -- support motion (prescribed displacements of restrained DOFs) shared by
// the step kernels.
//
// SVL: Assembler::ComputeSupportMotionIncrement (Assembler.cpp:493-532) forms, for every
// POINTLOAD_SUPPORT_MOTION load of the combination and every DOF of its nodes,
// factor (g(j) - g(j-1)) (g(0) alone at j = 0) with g(j) = Node::GetSupportMotion(j)
// (Node.cpp:228-249: the series value at j, or its first value past the end). Step k uses
// j = k (CentralDifference.cpp:195, ExtendedNewmarkBeta.cpp:197) or j = k - 1 (NewmarkBeta.cpp:
// 187): the step's applied increment a_k. The integrators add a_k to the increment of the
// restrained DOFs (CentralDifference.cpp:137, ExtendedNewmarkBeta.cpp:134-135, NewmarkBeta.cpp:
// 132) and subtract T^T K a_k from the effective force (:190-203 / :191-204 / :181-195).
//
// The element strains lag one support increment: Algorithm::UpdateStatesIncrements
// (Algorithm.cpp:18-56) updates the strains from the node displacements plus the free
// increment before ComputeNewStep adds a_k, and DynamicAnalysis::UpdateDomain only commits
// them (DynamicAnalysis.cpp:69-128; Elastic3DLinear.cpp:133-158). The internal force of step
// k + 1 and the reactions after step k therefore see U_k - a_k at those DOFs. The device keeps
// the true U_k and subtracts a_k where an internal force is formed.
#pragma once

#include "svlg_common.hpp"

namespace svlg {

// Device view: DOFs with support motion in ascending device node-DOF order (slot s) with
// dof[s] their node-DOF, the entries of slot s in description order [dptr[s], dptr[s+1]),
// each with its factor and values [vptr[e], vptr[e+1]); slot[g] maps a node-DOF to its slot
// (-1: none); offset: step k applies the increment of sample k - offset. slot null: none.
struct SmRef {
    const int32_t *slot, *dof, *dptr, *vptr;
    const double *fac, *vals;
    int ndofs, offset;
};

inline SmRef sm_ref(const Model &m) {
    if (m.sm_ndofs == 0) return SmRef{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0};
    return SmRef{m.sm_slot.p, m.sm_dof.p, m.sm_dptr.p, m.sm_vptr.p, m.sm_fac.p, m.sm_vals.p, m.sm_ndofs,
                 m.sm_offset};
}

__device__ __forceinline__ double sm_value(const SmRef &r, int e, int j) {
    const int v0 = r.vptr[e], n = r.vptr[e + 1] - v0;
    return r.vals[v0 + (j < n ? j : 0)];
}

// a_k of slot s: the increment step k applies (0 before the first step).
__device__ __forceinline__ double sm_applied(const SmRef &r, int s, int k) {
    const int j = k - r.offset;
    if (k < 1 || j < 0) return 0.0;
    double d = 0.0;
    for (int e = r.dptr[s]; e < r.dptr[s + 1]; ++e) {
        const double g = sm_value(r, e, j);
        d += r.fac[e] * (j > 0 ? g - sm_value(r, e, j - 1) : g);
    }
    return d;
}

// a_k of node-DOF g (0 without support motion there).
__device__ __forceinline__ double sm_applied_dof(const SmRef &r, size_t g, int k) {
    if (!r.slot) return 0.0;
    const int s = r.slot[g];
    return s >= 0 ? sm_applied(r, s, k) : 0.0;
}

// Uploads the support motion of the description (after the device node numbering exists).
int sm_setup(Model &m, const svlg_mesh_desc *d);

// Lags U at the support-motion DOFs for an internal force: their values go to the save
// buffer and U takes Usrc there (CentralDifference: U_(k-2) from the ring) or, with Usrc
// null, the per-slot lag buffer (Newmark: U before the last support increment). Plain copies,
// so undoing it restores the same bits.
int sm_lag_in(Model &m, double *U, const double *Usrc, const double *lag);

// After the lagged kernels of step k (kdev non-null: step *kdev + k, as in graph replays):
// mode 0 restores U; mode 1 (CentralDifference) restores U and sets Unext = U + a_k; mode 2
// (Newmark) keeps the restored value as the next lag and sets U += a_k.
int sm_finish(Model &m, int mode, int k, double *U, double *Unext, const int *kdev = nullptr);

// vec = a_k at the support-motion DOFs (other entries untouched).
int sm_vector(Model &m, int k, double *vec);

// Unext = U + a_k at the support-motion DOFs (an integrator's increment of step k).
int sm_add_increment(Model &m, int k, const double *U, double *Unext);

}  // namespace svlg
