// svlg_support.hip -- support motion on the device: set-up and the small per-step kernels
// around the step kernels (one thread per support-motion DOF; the step kernels themselves are
// unchanged and leave the restrained DOFs at 0).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "svlg_support.hpp"

namespace svlg {

// (Usrc may be U itself: no restrict on the pair.)
__global__ void k_sm_lag(SmRef r, double *U, const double *Usrc, const double *__restrict__ lag,
                         double *__restrict__ save) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= r.ndofs) return;
    const int g = r.dof[s];
    save[s] = U[g];
    U[g] = Usrc ? Usrc[g] : lag[s];
}

// mode 0: U = save (undo the lag); 1 (CentralDifference): U = save, Unext = save + a_k;
// 2 (Newmark): lag = save, U = save + a_k.
__global__ void k_sm_finish(SmRef r, int mode, int k, const int *__restrict__ kdev, double *__restrict__ U,
                            double *__restrict__ Unext, const double *__restrict__ save, double *__restrict__ lag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= r.ndofs) return;
    const int g = r.dof[s];
    const double u = save[s];
    const int kk = kdev ? *kdev + k : k;
    if (mode == 0) {
        U[g] = u;
    } else if (mode == 1) {
        U[g] = u;
        Unext[g] = u + sm_applied(r, s, kk);
    } else {
        lag[s] = u;
        U[g] = u + sm_applied(r, s, kk);
    }
}

__global__ void k_sm_vector(SmRef r, int k, double *__restrict__ vec) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < r.ndofs) vec[r.dof[s]] = sm_applied(r, s, k);
}

__global__ void k_sm_add_increment(SmRef r, int k, const double *__restrict__ U, double *__restrict__ Unext) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= r.ndofs) return;
    const int g = r.dof[s];
    Unext[g] = U[g] + sm_applied(r, s, k);
}

static unsigned sm_blocks(const Model &m) { return unsigned((m.sm_ndofs + 63) / 64); }

int sm_setup(Model &m, const svlg_mesh_desc *d) {
    m.sm_ndofs = 0;
    if (d->nsm <= 0) return 0;
    const int ne = d->nsm, ndim = m.ndim;
    SVLG_CHECK(d->sm_node && d->sm_dir && d->sm_factor && d->sm_ptr && d->sm_values, "support motion arrays missing");
    SVLG_CHECK(d->sm_offset == 0 || d->sm_offset == 1, "sm_offset must be 0 or 1");
    SVLG_CHECK(d->sm_ptr[0] == 0, "sm_ptr[0] must be 0");
    std::vector<int64_t> dofs(ne);
    for (int e = 0; e < ne; ++e) {
        const int n = d->sm_node[e], c = d->sm_dir[e];
        SVLG_CHECK(n >= 0 && n < m.nnodes, "support motion node out of range");
        SVLG_CHECK(c >= 0 && c < ndim, "support motion direction out of range");
        SVLG_CHECK(d->free_dof[size_t(n) * ndim + c] < 0, "support motion needs a restrained DOF");
        SVLG_CHECK(d->sm_ptr[e + 1] > d->sm_ptr[e], "a support motion needs at least one value");
        SVLG_CHECK(std::isfinite(d->sm_factor[e]), "support motion factor is not finite");
        dofs[e] = int64_t(m.inv[n]) * ndim + c;
    }
    // Slots in ascending device DOF order; entries of a slot in description order.
    std::vector<int32_t> order(ne);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return dofs[a] < dofs[b]; });
    std::vector<int32_t> slot(m.ndofn, -1), sdof, dptr{0}, vptr{0};
    std::vector<double> fac, vals;
    for (int i = 0; i < ne; ++i) {
        const int e = order[i];
        if (sdof.empty() || sdof.back() != dofs[e]) {
            if (!sdof.empty()) dptr.push_back(i);
            sdof.push_back(int32_t(dofs[e]));
            slot[dofs[e]] = int32_t(sdof.size() - 1);
        }
        fac.push_back(d->sm_factor[e]);
        for (int v = d->sm_ptr[e]; v < d->sm_ptr[e + 1]; ++v) {
            SVLG_CHECK(std::isfinite(d->sm_values[v]), "support motion value is not finite");
            vals.push_back(d->sm_values[v]);
        }
        vptr.push_back(int32_t(vals.size()));
    }
    dptr.push_back(ne);
    hipStream_t st = m.stream;
    SVLG_HIP(m.sm_slot.upload(slot.data(), slot.size(), st));
    SVLG_HIP(m.sm_dof.upload(sdof.data(), sdof.size(), st));
    SVLG_HIP(m.sm_dptr.upload(dptr.data(), dptr.size(), st));
    SVLG_HIP(m.sm_vptr.upload(vptr.data(), vptr.size(), st));
    SVLG_HIP(m.sm_fac.upload(fac.data(), fac.size(), st));
    SVLG_HIP(m.sm_vals.upload(vals.data(), vals.size(), st));
    SVLG_HIP(m.sm_save.alloc(sdof.size()));
    SVLG_HIP(m.sm_lag.alloc(sdof.size()));
    SVLG_HIP(hipMemsetAsync(m.sm_lag.p, 0, sizeof(double) * sdof.size(), st));
    SVLG_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    m.sm_ndofs = int(sdof.size());
    m.sm_offset = d->sm_offset;
    return 0;
}

int sm_lag_in(Model &m, double *U, const double *Usrc, const double *lag) {
    if (m.sm_ndofs == 0) return 0;
    hipLaunchKernelGGL(k_sm_lag, dim3(sm_blocks(m)), dim3(64), 0, m.stream, sm_ref(m), U, Usrc, lag, m.sm_save.p);
    SVLG_HIP(hipGetLastError());
    return 0;
}

int sm_finish(Model &m, int mode, int k, double *U, double *Unext, const int *kdev) {
    if (m.sm_ndofs == 0) return 0;
    hipLaunchKernelGGL(k_sm_finish, dim3(sm_blocks(m)), dim3(64), 0, m.stream, sm_ref(m), mode, k, kdev, U, Unext,
                       m.sm_save.p, m.sm_lag.p);
    SVLG_HIP(hipGetLastError());
    return 0;
}

int sm_vector(Model &m, int k, double *vec) {
    if (m.sm_ndofs == 0) return 0;
    hipLaunchKernelGGL(k_sm_vector, dim3(sm_blocks(m)), dim3(64), 0, m.stream, sm_ref(m), k, vec);
    SVLG_HIP(hipGetLastError());
    return 0;
}

int sm_add_increment(Model &m, int k, const double *U, double *Unext) {
    if (m.sm_ndofs == 0) return 0;
    hipLaunchKernelGGL(k_sm_add_increment, dim3(sm_blocks(m)), dim3(64), 0, m.stream, sm_ref(m), k, U, Unext);
    SVLG_HIP(hipGetLastError());
    return 0;
}

}  // namespace svlg

using namespace svlg;

extern "C" {

int svlg_get_displacements(svlg_model *h, double *U) {
    SVLG_CHECK(h && U, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.active == 1 || m.active == 2, "no integrator has been initialised");
    SVLG_CHECK(m.active != 1 || m.cd_ready, "svlg_cd_init has not been called");
    SVLG_CHECK(m.active != 2 || m.nm_ready, "svlg_nm_init has not been called");
    const double *src = m.active == 1 ? m.U[m.ring].p : m.nU.p;
    std::vector<double> hu(m.ndofn);
    SVLG_HIP(hipMemcpyAsync(hu.data(), src, sizeof(double) * m.ndofn, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    for (int n = 0; n < m.nnodes; ++n)
        for (int c = 0; c < m.ndim; ++c) U[size_t(n) * m.ndim + c] = hu[size_t(m.inv[n]) * m.ndim + c];
    return 0;
}

}  // extern "C"
