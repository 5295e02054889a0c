// svlg_drm.hip -- domain reduction load (SVL ElementLoad GENERALWAVE) on the device.
//
// Per listed element and step k, SVL computes (lin2DQuad4.cpp:563-614, lin3DHexa8.cpp:
// 660-718) F_e = M' A_0 + C' V_0 + K' U_0, where M', C', K' are the element matrices with
// every entry (i, j) whose two DOFs share the exterior flag zeroed and U_0, V_0, A_0 are
// row k of the node signals (exterior ones negated when read, Driver.hpp:1669-1720); the
// assembler adds F_e to the load's vector in element-list order, and the vector enters
// Fext times the load's factor (Assembler.cpp:460-478, 485).
//
// With the lumped mass every mass entry couples a DOF with itself, so M' = 0, and with
// Rayleigh damping C = am M + ak K, C' = ak K'. Then F_e = K' (U_0 + ak V_0), and with
// P_b, P_x the selections of interior and exterior DOFs, K' = P_x K P_b + P_b K P_x:
// exterior rows of K (P_b w), interior rows of K (P_x w). One thread per DRM element
// evaluates the two internal forces with the Gauss-point element kernel and keeps the
// matching rows; the step kernels add the corners of a node in element-list order.
#include <algorithm>
#include <vector>

#include "svlg_drm.hpp"
#include "svlg_elem.hpp"

namespace svlg {

struct DrmArgs {
    int n, nt, nnode, k;
    const int *kdev;
    const int32_t *elem, *cslot, *ext, *conn;
    const double *X, *lam, *mu, *ak, *sig;
    double th;
    double *fe;
};

template <int DIM>
__global__ __launch_bounds__(64) void k_drm_forces(DrmArgs a) {
    constexpr int NEN = Elem<DIM>::nen, NF = 3 * DIM;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.n) return;
    const int ks = a.kdev ? *a.kdev + a.k : a.k;
    const bool have = ks >= 0 && ks < a.nt;
    const int e = a.elem[j];
    const double ak = a.ak[e];
    double x[NEN][DIM], wb[NEN][DIM], wx[NEN][DIM], f1[NEN][DIM], f2[NEN][DIM];
    bool ext[NEN];
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
        const size_t nd = size_t(a.conn[size_t(e) * NEN + c]) * DIM;
        const int slot = a.cslot[size_t(j) * NEN + c];
        ext[c] = a.ext[slot] != 0;
        const double *row = a.sig + (size_t(have ? ks : 0) * a.nnode + slot) * NF;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            x[c][d] = a.X[nd + d];
            const double w = have ? row[d] + ak * row[DIM + d] : 0.0;
            wb[c][d] = ext[c] ? 0.0 : w;
            wx[c][d] = ext[c] ? w : 0.0;
        }
    }
    if constexpr (DIM == 3) {
        hexa_fint(x, wb, a.lam[e], a.mu[e], f1);
        hexa_fint(x, wx, a.lam[e], a.mu[e], f2);
    } else {
        quad_fint(x, wb, a.lam[e], a.mu[e], a.th, f1);
        quad_fint(x, wx, a.lam[e], a.mu[e], a.th, f2);
    }
#pragma unroll
    for (int c = 0; c < NEN; ++c)
#pragma unroll
        for (int d = 0; d < DIM; ++d) a.fe[(size_t(j) * NEN + c) * DIM + d] = ext[c] ? f1[c][d] : f2[c][d];
}

template <int DIM>
__global__ void k_drm_add(int nnode, const int32_t *__restrict__ dnode, const int32_t *__restrict__ nl_ptr,
                          DrmRef r, double *__restrict__ Fext) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnode) return;
    const int node = dnode[s];
    const bool loaded = nl_ptr[node + 1] > nl_ptr[node];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const size_t g = size_t(node) * DIM + d;
        const double base = loaded ? Fext[g] : 0.0;
        Fext[g] = base + drm_node_force<DIM>(r, node, d);
    }
}

int drm_setup(Model &m, const svlg_mesh_desc *d) {
    m.ndrm = 0;
    if (d->ndrm_elem <= 0) return 0;
    const int nen = m.nen, ndim = m.ndim, nf = 3 * ndim;
    const int ne = d->ndrm_elem, nnode = d->ndrm_node, nt = d->drm_nt;
    SVLG_CHECK(d->drm_elem && d->drm_node && d->drm_cond && d->drm_signal, "DRM arrays missing");
    SVLG_CHECK(nnode > 0 && nt > 0, "a DRM load needs nodes and at least one signal row");
    std::vector<int32_t> slot(m.nnodes, -1);
    for (int i = 0; i < nnode; ++i) {
        const int n = d->drm_node[i];
        SVLG_CHECK(n >= 0 && n < m.nnodes, "DRM node out of range");
        SVLG_CHECK(slot[n] < 0, "DRM node listed twice");
        slot[n] = i;
    }
    std::vector<int32_t> cslot(size_t(ne) * nen), ext(nnode), dnode(nnode);
    std::vector<int32_t> cnt(m.nnodes + 1, 0);
    for (int j = 0; j < ne; ++j) {
        const int e = d->drm_elem[j];
        SVLG_CHECK(e >= 0 && e < m.nelem, "DRM element out of range");
        SVLG_CHECK(!d->mat_type || d->mat_type[e] == SVLG_MAT_ELASTIC,
                   "DRM elements must be of an elastic material");
        for (int a = 0; a < nen; ++a) {
            const int n = d->conn[size_t(e) * nen + a];
            SVLG_CHECK(slot[n] >= 0, "a node of a DRM element has no signal");
            cslot[size_t(j) * nen + a] = slot[n];
            cnt[m.inv[n] + 1]++;
        }
    }
    for (int i = 0; i < nnode; ++i) {
        ext[i] = d->drm_cond[i] != 0;
        dnode[i] = m.inv[d->drm_node[i]];
    }
    for (int n = 0; n < m.nnodes; ++n) cnt[n + 1] += cnt[n];
    std::vector<int32_t> idx(cnt[m.nnodes]), fill(cnt.begin(), cnt.end() - 1);
    for (int j = 0; j < ne; ++j)
        for (int a = 0; a < nen; ++a) idx[fill[m.inv[d->conn[size_t(d->drm_elem[j]) * nen + a]]]++] = j * nen + a;
    // Time-major signals, exterior ones negated (SVL negates them when it reads the files).
    std::vector<double> sig(size_t(nt) * nnode * nf);
    for (int i = 0; i < nnode; ++i) {
        const double sg = ext[i] ? -1.0 : 1.0;
        for (int k = 0; k < nt; ++k)
            for (int f = 0; f < nf; ++f)
                sig[(size_t(k) * nnode + i) * nf + f] = sg * d->drm_signal[(size_t(i) * nt + k) * nf + f];
    }
    hipStream_t s = m.stream;
    SVLG_HIP(m.drm_elem.upload(d->drm_elem, ne, s));
    SVLG_HIP(m.drm_cslot.upload(cslot.data(), cslot.size(), s));
    SVLG_HIP(m.drm_ext.upload(ext.data(), ext.size(), s));
    SVLG_HIP(m.drm_dnode.upload(dnode.data(), dnode.size(), s));
    SVLG_HIP(m.drm_ptr.upload(cnt.data(), cnt.size(), s));
    SVLG_HIP(m.drm_idx.upload(idx.data(), idx.size(), s));
    SVLG_HIP(m.drm_sig.upload(sig.data(), sig.size(), s));
    SVLG_HIP(m.drm_fe.alloc(size_t(ne) * nen * ndim));
    SVLG_HIP(hipMemsetAsync(m.drm_fe.p, 0, sizeof(double) * size_t(ne) * nen * ndim, s));
    SVLG_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    m.ndrm = ne;
    m.ndrm_node = nnode;
    m.drm_nt = nt;
    m.drm_factor = d->drm_factor;
    return 0;
}

int drm_forces(Model &m, int k, const int *kdev) {
    if (m.ndrm == 0) return 0;
    DrmArgs a{};
    a.n = m.ndrm;
    a.nt = m.drm_nt;
    a.nnode = m.ndrm_node;
    a.k = k;
    a.kdev = kdev;
    a.elem = m.drm_elem.p;
    a.cslot = m.drm_cslot.p;
    a.ext = m.drm_ext.p;
    a.conn = m.conn.p;
    a.X = m.X.p;
    a.lam = m.lam.p;
    a.mu = m.mu.p;
    a.ak = m.ak.p;
    a.sig = m.drm_sig.p;
    a.th = m.th;
    a.fe = m.drm_fe.p;
    const int nb = (m.ndrm + 63) / 64;
    if (m.ndim == 3)
        hipLaunchKernelGGL(k_drm_forces<3>, dim3(nb), dim3(64), 0, m.stream, a);
    else
        hipLaunchKernelGGL(k_drm_forces<2>, dim3(nb), dim3(64), 0, m.stream, a);
    SVLG_HIP(hipGetLastError());
    return 0;
}

int drm_add_fext(Model &m, double *Fext) {
    if (m.ndrm == 0) return 0;
    const int nb = (m.ndrm_node + 63) / 64;
    if (m.ndim == 3)
        hipLaunchKernelGGL(k_drm_add<3>, dim3(nb), dim3(64), 0, m.stream, m.ndrm_node, m.drm_dnode.p, m.nl_ptr.p,
                           drm_ref(m), Fext);
    else
        hipLaunchKernelGGL(k_drm_add<2>, dim3(nb), dim3(64), 0, m.stream, m.ndrm_node, m.drm_dnode.p, m.nl_ptr.p,
                           drm_ref(m), Fext);
    SVLG_HIP(hipGetLastError());
    return 0;
}

}  // namespace svlg
