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
// exterior rows of K (P_b w), interior rows of K (P_x w).
//
// The signals are fixed inputs and the DRM elements are elastic, so the nodal forces of
// every step depend on nothing the time loop computes. The set-up evaluates all of them
// once, a chunk of signal rows per pass: the rows of the chunk are copied from the caller's
// array (no host-side copy), then two threads per (row, DRM element) form the two masked
// element forces and one thread per (row, DRM node) the sum of its corners in element-list
// order times the load factor. The result stays as F[k][slot][dim] (slots in device node
// order; a third of the signals' size); a step kernel adds F[k][slot] at a DRM node where
// the point loads enter Fext: no launch in the time loop.
#include <algorithm>
#include <numeric>
#include <vector>

#include "svlg_drm.hpp"
#include "svlg_elem.hpp"

namespace svlg {

struct DrmArgs {
    int64_t nthreads;
    int n, kc;
    const int32_t *elem, *cslot, *ext, *uidx, *conn;
    const double *X, *lam, *mu, *ak, *sig;  // sig: rows k0 .. k0+kc-1 of every node, [node][kc][3 dim]
    double th;
    double *fe;
};

// Thread t: step k0 + t / (2 n), DRM element j, side. Side 0 evaluates K (P_b w) and keeps
// the exterior rows, side 1 evaluates K (P_x w) and keeps the interior rows (disjoint rows
// of fe). AFFINE: every element is a parallelepiped (parallelogram), so the one-point modal
// form of the element force is exact; otherwise the Gauss-point loop.
template <int DIM, bool AFFINE>
__global__ __launch_bounds__(64) void k_drm_forces(DrmArgs a) {
    constexpr int NEN = Elem<DIM>::nen, NF = 3 * DIM;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= a.nthreads) return;
    const int64_t kk = t / (2 * a.n);
    const int r = int(t - kk * 2 * a.n), j = r >> 1, side = r & 1;
    const int e = a.elem[j];
    const double ak = a.ak[e];
    double x[NEN][DIM], w[NEN][DIM], f[NEN][DIM];
    bool keep[NEN];
#pragma unroll
    for (int c = 0; c < NEN; ++c) {
        const size_t nd = size_t(a.conn[size_t(e) * NEN + c]) * DIM;
        const int slot = a.cslot[size_t(j) * NEN + c];
        const bool ext = a.ext[slot] != 0;
        const bool staged = side == 0 ? !ext : ext;
        keep[c] = !staged;
        // Exterior signals enter negated, as SVL negates them when it reads the node files.
        const double sg = ext ? -1.0 : 1.0;
        const double *row = a.sig + (size_t(a.uidx[slot]) * a.kc + size_t(kk)) * NF;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            x[c][d] = a.X[nd + d];
            w[c][d] = staged ? sg * row[d] + ak * (sg * row[DIM + d]) : 0.0;
        }
    }
    if constexpr (DIM == 3) {
        if constexpr (AFFINE) hexa_affine_fint(x, w, a.lam[e], a.mu[e], f);
        else hexa_fint(x, w, a.lam[e], a.mu[e], f);
    } else {
        if constexpr (AFFINE) quad_affine_fint(x, w, a.lam[e], a.mu[e], a.th, f);
        else quad_fint(x, w, a.lam[e], a.mu[e], a.th, f);
    }
    double *o = a.fe + (size_t(kk) * a.n + j) * NEN * DIM;
#pragma unroll
    for (int c = 0; c < NEN; ++c)
        if (keep[c])
#pragma unroll
            for (int d = 0; d < DIM; ++d) o[c * DIM + d] = f[c][d];
}

// Thread t: step k0 + t / ns, slot s = t % ns: factor * (its corners in element-list order).
template <int DIM>
__global__ __launch_bounds__(64) void k_drm_gather(int64_t nthreads, int ns, int nen, int n, int k0,
                                                   const int32_t *__restrict__ ptr, const int32_t *__restrict__ idx,
                                                   const double *__restrict__ fe, double factor,
                                                   double *__restrict__ F) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nthreads) return;
    const int64_t kk = t / ns;
    const int s = int(t - kk * ns);
    const double *base = fe + size_t(kk) * n * nen * DIM;
    double acc[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) acc[d] = 0.0;
    for (int q = ptr[s]; q < ptr[s + 1]; ++q) {
        const size_t o = size_t(idx[q]) * DIM;
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc[d] += base[o + d];
    }
    double *out = F + ((size_t(k0) + size_t(kk)) * ns + s) * DIM;
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = factor * acc[d];
}

template <int DIM>
__global__ void k_drm_add(int ns, int k, const int32_t *__restrict__ dnode, const int32_t *__restrict__ nl_ptr,
                          DrmRef r, double *__restrict__ Fext) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int node = dnode[s];
    const bool loaded = nl_ptr[node + 1] > nl_ptr[node];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const size_t g = size_t(node) * DIM + d;
        const double base = loaded ? Fext[g] : 0.0;
        Fext[g] = base + drm_force(r, s, k, DIM, d);
    }
}

int drm_setup(Model &m, const svlg_mesh_desc *d) {
    m.ndrm = 0;
    if (d->ndrm_elem <= 0) return 0;
    const int nen = m.nen, ndim = m.ndim, nf = 3 * ndim;
    const int ne = d->ndrm_elem, ns = d->ndrm_node, nt = d->drm_nt;
    SVLG_CHECK(d->drm_elem && d->drm_node && d->drm_cond && d->drm_signal, "DRM arrays missing");
    SVLG_CHECK(ns > 0 && nt > 0, "a DRM load needs nodes and at least one signal row");
    // Internal slots in ascending device node order (coalesced reads in the step kernels).
    std::vector<int32_t> order(ns);
    std::iota(order.begin(), order.end(), 0);
    for (int i = 0; i < ns; ++i) {
        const int n = d->drm_node[i];
        SVLG_CHECK(n >= 0 && n < m.nnodes, "DRM node out of range");
    }
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return m.inv[d->drm_node[a]] < m.inv[d->drm_node[b]]; });
    std::vector<int32_t> slot(m.nnodes, -1), dnode(ns), ext(ns);
    for (int s = 0; s < ns; ++s) {
        const int n = d->drm_node[order[s]];
        SVLG_CHECK(slot[n] < 0, "DRM node listed twice");
        slot[n] = s;
        dnode[s] = m.inv[n];
        ext[s] = d->drm_cond[order[s]] != 0;
    }
    std::vector<int32_t> cslot(size_t(ne) * nen), cnt(ns + 1, 0);
    for (int j = 0; j < ne; ++j) {
        const int e = d->drm_elem[j];
        SVLG_CHECK(e >= 0 && e < m.nelem, "DRM element out of range");
        SVLG_CHECK(!d->mat_type || d->mat_type[e] == SVLG_MAT_ELASTIC,
                   "DRM elements must be of an elastic material");
        for (int a = 0; a < nen; ++a) {
            const int n = d->conn[size_t(e) * nen + a];
            SVLG_CHECK(slot[n] >= 0, "a node of a DRM element has no signal");
            cslot[size_t(j) * nen + a] = slot[n];
            cnt[slot[n] + 1]++;
        }
    }
    for (int s = 0; s < ns; ++s) cnt[s + 1] += cnt[s];
    std::vector<int32_t> idx(cnt[ns]), fill(cnt.begin(), cnt.end() - 1);
    for (int j = 0; j < ne; ++j)
        for (int a = 0; a < nen; ++a) idx[fill[cslot[size_t(j) * nen + a]]++] = j * nen + a;
    std::vector<int32_t> nslot(m.nnodes, -1);
    for (int s = 0; s < ns; ++s) nslot[dnode[s]] = s;
    hipStream_t st = m.stream;
    DevBuf<int32_t> delem, dcslot, dext, duidx, dptr, didx;
    DevBuf<double> dsig, fe;
    SVLG_HIP(delem.upload(d->drm_elem, ne, st));
    SVLG_HIP(dcslot.upload(cslot.data(), cslot.size(), st));
    SVLG_HIP(dext.upload(ext.data(), ext.size(), st));
    SVLG_HIP(duidx.upload(order.data(), order.size(), st));
    SVLG_HIP(dptr.upload(cnt.data(), cnt.size(), st));
    SVLG_HIP(didx.upload(idx.data(), idx.size(), st));
    SVLG_HIP(m.drm_F.alloc(size_t(nt) * ns * ndim));
    // Signal rows per pass: element forces and signal rows of at most 2^25 doubles (256 MiB)
    // each at a time. The caller's signals ([node][nt][3 dim]) are copied a block of rows per
    // pass, with no host-side copy.
    const size_t per_step = std::max(size_t(ne) * nen * ndim, size_t(ns) * nf);
    const int chunk = int(std::max<size_t>(1, std::min<size_t>(size_t(nt), (size_t(1) << 25) / per_step)));
    SVLG_HIP(fe.alloc(size_t(chunk) * ne * nen * ndim));
    SVLG_HIP(hipMemsetAsync(fe.p, 0, sizeof(double) * size_t(chunk) * ne * nen * ndim, st));
    SVLG_HIP(dsig.alloc(size_t(chunk) * ns * nf));
    for (int k0 = 0; k0 < nt; k0 += chunk) {
        const int kc = std::min(chunk, nt - k0);
        SVLG_HIP(hipMemcpy2DAsync(dsig.p, sizeof(double) * size_t(kc) * nf, d->drm_signal + size_t(k0) * nf,
                                  sizeof(double) * size_t(nt) * nf, sizeof(double) * size_t(kc) * nf, size_t(ns),
                                  hipMemcpyHostToDevice, st));
        DrmArgs a{};
        a.nthreads = int64_t(kc) * 2 * ne;
        a.n = ne;
        a.kc = kc;
        a.elem = delem.p;
        a.cslot = dcslot.p;
        a.ext = dext.p;
        a.uidx = duidx.p;
        a.conn = m.conn.p;
        a.X = m.X.p;
        a.lam = m.lam.p;
        a.mu = m.mu.p;
        a.ak = m.ak.p;
        a.sig = dsig.p;
        a.th = m.th;
        a.fe = fe.p;
        const unsigned nb = unsigned((a.nthreads + 63) / 64);
        const int64_t ng64 = int64_t(kc) * ns;
        const unsigned ng = unsigned((ng64 + 63) / 64);
        if (ndim == 3) {
            if (m.affine) hipLaunchKernelGGL((k_drm_forces<3, true>), dim3(nb), dim3(64), 0, st, a);
            else hipLaunchKernelGGL((k_drm_forces<3, false>), dim3(nb), dim3(64), 0, st, a);
            hipLaunchKernelGGL(k_drm_gather<3>, dim3(ng), dim3(64), 0, st, ng64, ns, nen, ne, k0, dptr.p, didx.p,
                               fe.p, d->drm_factor, m.drm_F.p);
        } else {
            if (m.affine) hipLaunchKernelGGL((k_drm_forces<2, true>), dim3(nb), dim3(64), 0, st, a);
            else hipLaunchKernelGGL((k_drm_forces<2, false>), dim3(nb), dim3(64), 0, st, a);
            hipLaunchKernelGGL(k_drm_gather<2>, dim3(ng), dim3(64), 0, st, ng64, ns, nen, ne, k0, dptr.p, didx.p,
                               fe.p, d->drm_factor, m.drm_F.p);
        }
        SVLG_HIP(hipGetLastError());
    }
    SVLG_HIP(m.drm_slot.upload(nslot.data(), nslot.size(), st));
    SVLG_HIP(m.drm_dnode.upload(dnode.data(), dnode.size(), st));
    SVLG_HIP(hipStreamSynchronize(st));  // the host vectors and the temporaries go out of scope
    m.ndrm = ne;
    m.ndrm_node = ns;
    m.drm_nt = nt;
    return 0;
}

int drm_add_fext(Model &m, int k, double *Fext) {
    if (m.ndrm == 0) return 0;
    const int nb = (m.ndrm_node + 63) / 64;
    if (m.ndim == 3)
        hipLaunchKernelGGL(k_drm_add<3>, dim3(nb), dim3(64), 0, m.stream, m.ndrm_node, k, m.drm_dnode.p, m.nl_ptr.p,
                           drm_ref(m), Fext);
    else
        hipLaunchKernelGGL(k_drm_add<2>, dim3(nb), dim3(64), 0, m.stream, m.ndrm_node, k, m.drm_dnode.p, m.nl_ptr.p,
                           drm_ref(m), Fext);
    SVLG_HIP(hipGetLastError());
    return 0;
}

}  // namespace svlg
