// svlg_model.hip -- model upload, element coloring, mass/damping assembly, Fint(U).
//
// Replaces, behind the C ABI in include/svlg.h:
//   Mesh::Initialize / Element::SetDomain           (RP/06-Mesh/Mesh.cpp:18-49)
//   Assembler::ComputeMassMatrix (lumped diagonal)   (RP/07-Assembler/Assembler.cpp:48-67, 661-697)
//   Assembler::ComputeDampingMatrix (dashpots, am M) (Assembler.cpp:117-158)
//   Assembler::ComputeInternalForceVector            (Assembler.cpp:240-269)
// RP = 02-Run_Process of the SVL source tree.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "svlg_common.hpp"
#include "svlg_elem.hpp"

namespace svlg {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

int build_tiles(Model &m);  // svlg_explicit.hip

// ---------------------------------------------------------------- kernels
// Per-element row-sum lumped mass, scattered later in ascending element order.
template <int DIM>
__global__ void k_elem_lumped_mass(int nelem, const int32_t *__restrict__ conn, const double *__restrict__ X,
                                   const double *__restrict__ rho, double th, double *__restrict__ out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    constexpr int nen = Elem<DIM>::nen;
    double x[nen][DIM];
#pragma unroll
    for (int a = 0; a < nen; ++a) {
        int n = conn[size_t(e) * nen + a];
#pragma unroll
        for (int d = 0; d < DIM; ++d) x[a][d] = X[size_t(n) * DIM + d];
    }
    double ml[nen];
    if constexpr (DIM == 2)
        quad_lumped_mass(x, rho[e], th, ml);
    else
        hexa_lumped_mass(x, rho[e], ml);
#pragma unroll
    for (int a = 0; a < nen; ++a) out[size_t(e) * nen + a] = ml[a];
}

// M_ii = sum over adjacent elements in ascending tag of the lumped entries whose
// magnitude exceeds mtol (triplet filter of Assembler.cpp:687 and setFromTriplets'
// duplicate summation in triplet order). Rayleigh am*M_e adds to C the same way.
__global__ void k_node_gather_mass(int nnodes, int ndim, const int32_t *__restrict__ ptr,
                                   const int32_t *__restrict__ idx, const double *__restrict__ me,
                                   const double *__restrict__ am, int nen, double mtol,
                                   double *__restrict__ mdiag, double *__restrict__ cdiag) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnodes) return;
    double msum = 0.0, csum = 0.0;
    for (int k = ptr[n]; k < ptr[n + 1]; ++k) {
        int ea = idx[k];
        double v = me[ea];
        if (fabs(v) > mtol) msum += v;
        double c = am[ea / nen] * v;
        if (fabs(c) > mtol) csum += c;
    }
    for (int d = 0; d < ndim; ++d) {
        mdiag[size_t(n) * ndim + d] = msum;
        cdiag[size_t(n) * ndim + d] = csum;
    }
}

// Dashpots in element-tag order; one thread keeps the summation order of SVL.
__global__ void k_dashpots(int ndash, int ndim, const int32_t *__restrict__ node, const int32_t *__restrict__ dir,
                           const double *__restrict__ eta, double ktol, double *__restrict__ cdiag) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = 0; i < ndash; ++i) {
        double v = eta[i];
        if (fabs(v) > ktol) cdiag[size_t(node[i]) * ndim + dir[i]] += v;
    }
}

// Colored Fint: every element of one color adds into F with no conflicts.
template <int DIM>
__global__ void k_fint_colored(int e0, int e1, const int32_t *__restrict__ conn_c, const int32_t *__restrict__ eid_c,
                               const double *__restrict__ X, const double *__restrict__ U,
                               const double *__restrict__ lam, const double *__restrict__ mu, double th,
                               double ftol, double *__restrict__ F) {
    int e = e0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    constexpr int nen = Elem<DIM>::nen;
    int tag = eid_c[e];
    int nodes[nen];
    double x[nen][DIM], u[nen][DIM], f[nen][DIM];
#pragma unroll
    for (int a = 0; a < nen; ++a) {
        nodes[a] = conn_c[size_t(e) * nen + a];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            x[a][d] = X[size_t(nodes[a]) * DIM + d];
            u[a][d] = U[size_t(nodes[a]) * DIM + d];
        }
    }
    if constexpr (DIM == 2)
        quad_fint(x, u, lam[tag], mu[tag], th, f);
    else
        hexa_fint(x, u, lam[tag], mu[tag], f);
#pragma unroll
    for (int a = 0; a < nen; ++a)
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            double v = f[a][d];
            if (fabs(v) > ftol) F[size_t(nodes[a]) * DIM + d] += v;
        }
}

__global__ void k_scatter_free(int ndofn, const int32_t *__restrict__ fdof, const double *__restrict__ vfree,
                               double *__restrict__ vnode) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndofn) return;
    int f = fdof[i];
    vnode[i] = f >= 0 ? vfree[f] : 0.0;
}

__global__ void k_gather_free(int ndofn, const int32_t *__restrict__ fdof, const double *__restrict__ vnode,
                              double *__restrict__ vfree) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndofn) return;
    int f = fdof[i];
    if (f >= 0) vfree[f] = vnode[i];
}

// Launch helpers shared with the integrators.
int launch_fint_colored(Model &m, const double *U, double ftol, double *F) {
    for (int c = 0; c < m.ncolors; ++c) {
        int e0 = m.color_off[c], e1 = m.color_off[c + 1];
        if (e1 <= e0) continue;
        int nb = (e1 - e0 + 127) / 128;
        if (m.ndim == 2)
            hipLaunchKernelGGL(k_fint_colored<2>, dim3(nb), dim3(128), 0, m.stream, e0, e1, m.conn_c.p, m.eid_c.p,
                               m.X.p, U, m.lam.p, m.mu.p, m.th, ftol, F);
        else
            hipLaunchKernelGGL(k_fint_colored<3>, dim3(nb), dim3(128), 0, m.stream, e0, e1, m.conn_c.p, m.eid_c.p,
                               m.X.p, U, m.lam.p, m.mu.p, m.th, ftol, F);
    }
    SVLG_HIP(hipGetLastError());
    return 0;
}

int scatter_free(Model &m, const double *dfree, double *dnode) {
    int nb = (m.ndofn + 255) / 256;
    hipLaunchKernelGGL(k_scatter_free, dim3(nb), dim3(256), 0, m.stream, m.ndofn, m.fdof.p, dfree, dnode);
    SVLG_HIP(hipGetLastError());
    return 0;
}
int gather_free(Model &m, const double *dnode, double *dfree) {
    int nb = (m.ndofn + 255) / 256;
    hipLaunchKernelGGL(k_gather_free, dim3(nb), dim3(256), 0, m.stream, m.ndofn, m.fdof.p, dnode, dfree);
    SVLG_HIP(hipGetLastError());
    return 0;
}

// Greedy coloring in ascending element tag: each element takes the smallest color not
// used by an earlier element sharing one of its nodes.
static int color_elements(int nelem, int nen, int nnodes, const int32_t *conn, std::vector<int32_t> &color,
                          int &ncolors) {
    std::vector<uint64_t> used(nnodes, 0);
    color.assign(nelem, 0);
    ncolors = 0;
    for (int e = 0; e < nelem; ++e) {
        uint64_t u = 0;
        for (int a = 0; a < nen; ++a) u |= used[conn[size_t(e) * nen + a]];
        if (u == ~uint64_t(0)) return 1;
        int c = __builtin_ctzll(~u);
        color[e] = c;
        ncolors = std::max(ncolors, c + 1);
        for (int a = 0; a < nen; ++a) used[conn[size_t(e) * nen + a]] |= (uint64_t(1) << c);
    }
    return 0;
}

// A parallelepiped has X(xi) affine: x_a = x_0 + sum_i sign_i(a) * g_i for every corner.
static bool all_affine(int dim, int nelem, int nen, const int32_t *conn, const double *X) {
    for (int e = 0; e < nelem; ++e) {
        const int32_t *c = conn + size_t(e) * nen;
        double scale = 0.0;
        for (int a = 0; a < nen; ++a)
            for (int d = 0; d < dim; ++d)
                scale = std::max(scale, std::fabs(X[size_t(c[a]) * dim + d] - X[size_t(c[0]) * dim + d]));
        double tol = 1e-12 * (scale > 0 ? scale : 1.0);
        for (int d = 0; d < dim; ++d) {
            auto x = [&](int a) { return X[size_t(c[a]) * dim + d]; };
            if (dim == 2) {
                if (std::fabs((x(2) - x(3)) - (x(1) - x(0))) > tol) return false;
            } else {
                double gx = x(1) - x(0), gy = x(3) - x(0), gz = x(4) - x(0);
                if (std::fabs(x(2) - x(0) - gx - gy) > tol) return false;
                if (std::fabs(x(5) - x(0) - gx - gz) > tol) return false;
                if (std::fabs(x(7) - x(0) - gy - gz) > tol) return false;
                if (std::fabs(x(6) - x(0) - gx - gy - gz) > tol) return false;
            }
        }
    }
    return true;
}

}  // namespace svlg

using namespace svlg;

extern "C" {

const char *svlg_last_error(void) { return svlg::g_error.c_str(); }

const char *svlg_version(void) { return "svlg 0.1 (gfx950, FP64 explicit + implicit)"; }

int svlg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int svlg_model_create(const svlg_mesh_desc *d, svlg_model **out) {
    SVLG_CHECK(d && out, "svlg_model_create: null argument");
    *out = nullptr;
    SVLG_CHECK(d->ndim == 2 || d->ndim == 3, "ndim must be 2 or 3");
    SVLG_CHECK((d->ndim == 2 && d->elem_type == SVLG_ELEM_QUAD4) || (d->ndim == 3 && d->elem_type == SVLG_ELEM_HEXA8),
               "element type must be lin2DQuad4 in 2D or lin3DHexa8 in 3D");
    SVLG_CHECK(d->nnodes > 0 && d->nelem > 0, "empty mesh");
    SVLG_CHECK(d->coords && d->conn && d->free_dof && d->E && d->nu && d->rho, "missing mesh arrays");
    int ndev = 0;
    SVLG_CHECK(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, "no HIP device available");
    auto *h = new svlg_model();
    Model &m = h->m;
    m.ndim = d->ndim;
    m.nen = d->ndim == 2 ? 4 : 8;
    m.nnodes = d->nnodes;
    m.nelem = d->nelem;
    m.nfree = d->nfree;
    m.ndofn = d->nnodes * d->ndim;
    m.th = d->thickness;
    m.lumped = d->lumped != 0;
    auto fail = [&](int code) {
        delete h;
        return code;
    };
    for (size_t i = 0; i < size_t(d->nelem) * m.nen; ++i)
        if (d->conn[i] < 0 || d->conn[i] >= d->nnodes) {
            set_error("connectivity references a node out of range");
            return fail(11);
        }
    if (d->mat_type)
        for (int e = 0; e < d->nelem; ++e)
            if (d->mat_type[e] != SVLG_MAT_ELASTIC) {
                set_error("only Elastic2DPlaneStrain / Elastic3DLinear materials are supported by this build");
                return fail(11);
            }
    m.h_free.assign(d->free_dof, d->free_dof + m.ndofn);
    {
        int cnt = 0;
        for (int i = 0; i < m.ndofn; ++i) {
            if (m.h_free[i] >= d->nfree) { set_error("free_dof index >= nfree"); return fail(11); }
            if (m.h_free[i] >= 0) cnt++;
        }
        if (cnt != d->nfree) { set_error("nfree does not match free_dof"); return fail(11); }
    }
    if (hipStreamCreateWithFlags(&m.own_stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(10);
    }
    m.stream = m.own_stream;
    hipStream_t s = m.stream;

    if (color_elements(d->nelem, m.nen, d->nnodes, d->conn, m.h_color, m.ncolors)) {
        set_error("element coloring needs more than 64 colors");
        return fail(11);
    }
    m.affine = all_affine(m.ndim, d->nelem, m.nen, d->conn, d->coords);
    // Color-sorted order (stable: ascending tag inside each color).
    m.color_off.assign(m.ncolors + 1, 0);
    for (int e = 0; e < d->nelem; ++e) m.color_off[m.h_color[e] + 1]++;
    for (int c = 0; c < m.ncolors; ++c) m.color_off[c + 1] += m.color_off[c];
    std::vector<int32_t> pos(m.color_off.begin(), m.color_off.end() - 1);
    std::vector<int32_t> conn_c(size_t(d->nelem) * m.nen), eid_c(d->nelem);
    for (int e = 0; e < d->nelem; ++e) {
        int p = pos[m.h_color[e]]++;
        eid_c[p] = e;
        std::memcpy(&conn_c[size_t(p) * m.nen], &d->conn[size_t(e) * m.nen], sizeof(int32_t) * m.nen);
    }
    // Lame constants of the isotropic D (c2 = lambda, c3 = mu; c1 = c2 + 2 c3).
    std::vector<double> lam(d->nelem), mu(d->nelem), am(d->nelem, 0.0), ak(d->nelem, 0.0);
    for (int e = 0; e < d->nelem; ++e) {
        double E = d->E[e], nu = d->nu[e];
        lam[e] = E * nu / (1.0 - 2.0 * nu) / (1.0 + nu);
        mu[e] = E / (2.0 * (1.0 + nu));
        if (d->ray_am) am[e] = d->ray_am[e];
        if (d->ray_ak) ak[e] = d->ray_ak[e];
        if (ak[e] != 0.0) m.has_ak = true;
    }
    // Node -> (element, local corner) in ascending element order.
    std::vector<int32_t> n2e_ptr(d->nnodes + 1, 0), n2e(size_t(d->nelem) * m.nen);
    for (size_t i = 0; i < size_t(d->nelem) * m.nen; ++i) n2e_ptr[d->conn[i] + 1]++;
    for (int n = 0; n < d->nnodes; ++n) n2e_ptr[n + 1] += n2e_ptr[n];
    {
        std::vector<int32_t> p(n2e_ptr.begin(), n2e_ptr.end() - 1);
        for (int e = 0; e < d->nelem; ++e)
            for (int a = 0; a < m.nen; ++a) n2e[p[d->conn[size_t(e) * m.nen + a]]++] = e * m.nen + a;
    }
#define UP(buf, src, cnt)                                                   \
    if (buf.upload(src, cnt, s) != hipSuccess) {                            \
        set_error("device allocation/upload failed for " #buf);             \
        return fail(10);                                                    \
    }
    UP(m.X, d->coords, size_t(d->nnodes) * m.ndim);
    UP(m.conn, d->conn, size_t(d->nelem) * m.nen);
    UP(m.conn_c, conn_c.data(), conn_c.size());
    UP(m.eid_c, eid_c.data(), eid_c.size());
    UP(m.lam, lam.data(), lam.size());
    UP(m.mu, mu.data(), mu.size());
    UP(m.rho, d->rho, size_t(d->nelem));
    UP(m.am, am.data(), am.size());
    UP(m.ak, ak.data(), ak.size());
    UP(m.fdof, m.h_free.data(), m.h_free.size());
    UP(m.n2e_ptr, n2e_ptr.data(), n2e_ptr.size());
    UP(m.n2e, n2e.data(), n2e.size());
    // Loads.
    m.nload = d->nload;
    m.load_nt = d->load_nt;
    if (d->nload > 0) {
        SVLG_CHECK(d->load_node && d->load_dir && d->load_factor && d->load_series && d->load_nt > 0,
                   "incomplete load arrays");
        UP(m.load_node, d->load_node, size_t(d->nload));
        UP(m.load_dir, d->load_dir, size_t(d->nload) * m.ndim);
        UP(m.load_factor, d->load_factor, size_t(d->nload));
        UP(m.load_series, d->load_series, size_t(d->nload) * d->load_nt);
        std::vector<int32_t> dofs;
        for (int l = 0; l < d->nload; ++l) {
            SVLG_CHECK(d->load_node[l] >= 0 && d->load_node[l] < d->nnodes, "load node out of range");
            for (int c = 0; c < m.ndim; ++c) dofs.push_back(d->load_node[l] * m.ndim + c);
        }
        std::sort(dofs.begin(), dofs.end());
        dofs.erase(std::unique(dofs.begin(), dofs.end()), dofs.end());
        m.h_load_dofs = dofs;
        UP(m.load_dofs, dofs.data(), dofs.size());
    }
    // Mass and damping diagonals.
    if (m.mdiag.alloc(m.ndofn) != hipSuccess || m.cdiag.alloc(m.ndofn) != hipSuccess ||
        m.scratch.alloc(size_t(d->nelem) * m.nen) != hipSuccess) {
        set_error("device allocation failed (diagonals)");
        return fail(10);
    }
    {
        int nb = (d->nelem + 127) / 128;
        if (m.ndim == 2)
            hipLaunchKernelGGL(k_elem_lumped_mass<2>, dim3(nb), dim3(128), 0, s, d->nelem, m.conn.p, m.X.p, m.rho.p,
                               m.th, m.scratch.p);
        else
            hipLaunchKernelGGL(k_elem_lumped_mass<3>, dim3(nb), dim3(128), 0, s, d->nelem, m.conn.p, m.X.p, m.rho.p,
                               m.th, m.scratch.p);
        int nbn = (d->nnodes + 127) / 128;
        hipLaunchKernelGGL(k_node_gather_mass, dim3(nbn), dim3(128), 0, s, d->nnodes, m.ndim, m.n2e_ptr.p, m.n2e.p,
                           m.scratch.p, m.am.p, m.nen, 1e-12, m.mdiag.p, m.cdiag.p);
    }
    if (d->ndash > 0) {
        SVLG_CHECK(d->dash_node && d->dash_dir && d->dash_eta, "incomplete dashpot arrays");
        for (int i = 0; i < d->ndash; ++i) {
            if (d->dash_node[i] < 0 || d->dash_node[i] >= d->nnodes || d->dash_dir[i] < 0 || d->dash_dir[i] >= m.ndim) {
                set_error("dashpot node or direction out of range");
                return fail(11);
            }
        }
        DevBuf<int32_t> dn, dd;
        DevBuf<double> de;
        UP(dn, d->dash_node, size_t(d->ndash));
        UP(dd, d->dash_dir, size_t(d->ndash));
        UP(de, d->dash_eta, size_t(d->ndash));
        hipLaunchKernelGGL(k_dashpots, dim3(1), dim3(64), 0, s, d->ndash, m.ndim, dn.p, dd.p, de.p, 1e-12, m.cdiag.p);
        if (hipStreamSynchronize(s) != hipSuccess) { set_error("dashpot assembly failed"); return fail(10); }
    }
#undef UP
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        set_error("model set-up kernels failed");
        return fail(10);
    }
    int rc = build_tiles(m);
    if (rc) return fail(rc);
    *out = h;
    return 0;
}

int svlg_model_destroy(svlg_model *h) {
    if (!h) return 0;
    if (h->m.stream) (void)hipStreamSynchronize(h->m.stream);
    hipStream_t own = h->m.own_stream;
    delete h;
    if (own) (void)hipStreamDestroy(own);
    return 0;
}

int svlg_model_set_stream(svlg_model *h, void *stream) {
    SVLG_CHECK(h, "null model");
    SVLG_HIP(hipStreamSynchronize(h->m.stream));
    h->m.stream = stream ? (hipStream_t)stream : h->m.own_stream;
    return 0;
}

void *svlg_model_stream(svlg_model *h) { return h ? (void *)h->m.stream : nullptr; }

int svlg_model_info(svlg_model *h, int64_t *info) {
    SVLG_CHECK(h && info, "null argument");
    Model &m = h->m;
    info[0] = m.ncolors;
    info[1] = m.nelem;
    info[2] = m.nnodes;
    info[3] = m.nfree;
    info[4] = m.variant;
    info[5] = m.tiles.ntiles;
    info[6] = m.affine ? 1 : 0;
    info[7] = 0;
    return 0;
}

int svlg_element_colors(svlg_model *h, int32_t *colors) {
    SVLG_CHECK(h && colors, "null argument");
    std::memcpy(colors, h->m.h_color.data(), sizeof(int32_t) * h->m.nelem);
    return 0;
}

int svlg_mass_diagonal(svlg_model *h, double *mfree) {
    SVLG_CHECK(h && mfree, "null argument");
    Model &m = h->m;
    DevBuf<double> tmp;
    SVLG_HIP(tmp.alloc(m.nfree));
    if (int rc = gather_free(m, m.mdiag.p, tmp.p)) return rc;
    SVLG_HIP(hipMemcpyAsync(mfree, tmp.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_damping_diagonal(svlg_model *h, double *cfree) {
    SVLG_CHECK(h && cfree, "null argument");
    Model &m = h->m;
    DevBuf<double> tmp;
    SVLG_HIP(tmp.alloc(m.nfree));
    if (int rc = gather_free(m, m.cdiag.p, tmp.p)) return rc;
    SVLG_HIP(hipMemcpyAsync(cfree, tmp.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_internal_force(svlg_model *h, const double *Ufree, double ftol, double *Ffree) {
    SVLG_CHECK(h && Ufree && Ffree, "null argument");
    Model &m = h->m;
    DevBuf<double> uf, un, fn;
    SVLG_HIP(uf.upload(Ufree, m.nfree, m.stream));
    SVLG_HIP(un.alloc(m.ndofn));
    SVLG_HIP(fn.alloc(m.ndofn));
    SVLG_HIP(hipMemsetAsync(fn.p, 0, sizeof(double) * m.ndofn, m.stream));
    if (int rc = scatter_free(m, uf.p, un.p)) return rc;
    if (int rc = launch_fint_colored(m, un.p, ftol, fn.p)) return rc;
    if (int rc = gather_free(m, fn.p, uf.p)) return rc;
    SVLG_HIP(hipMemcpyAsync(Ffree, uf.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_sync(svlg_model *h) {
    SVLG_CHECK(h, "null model");
    SVLG_HIP(hipStreamSynchronize(h->m.stream));
    return 0;
}

}  // extern "C"
