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
#include <map>
#include <mutex>
#include <numeric>
#include <stdexcept>

#include "svlg_common.hpp"
#include "svlg_drm.hpp"
#include "svlg_support.hpp"
#include "svlg_elem.hpp"
#include "svlg_tiles.hpp"

namespace svlg {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

int prepare_fused(Model &m);                           // svlg_explicit.hip
int build_hexa_kmat(Model &m, DevBuf<float> &kmat);    // svlg_gcn.hip

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

// Dashpots: one thread per damped DOF adds its dashpots in element-tag order (the
// triplet order of Assembler.cpp:148-154), after the Rayleigh terms of the continuum
// elements (whose tags come first).
__global__ void k_dashpots(int ndofs, const int32_t *__restrict__ dof, const int32_t *__restrict__ ptr,
                           const double *__restrict__ eta, double ktol, double *__restrict__ cdiag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    double c = cdiag[dof[i]];
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
        double v = eta[k];
        if (fabs(v) > ktol) c += v;
    }
    cdiag[dof[i]] = c;
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
// Recorder positions: rec_dofs[r] >= 0 is a free DOF index, -1 - t the total DOF t = node *
// ndim + c of a mesh node (a restrained DOF, e.g. one with support motion); pos[r] is its
// device node-DOF position.
int record_positions(const Model &m, int nrec, const int32_t *rec_dofs, std::vector<int32_t> &pos) {
    std::vector<int32_t> f2n(m.nfree, -1);
    for (int i = 0; i < m.ndofn; ++i)
        if (m.h_free[i] >= 0) f2n[m.h_free[i]] = i;
    pos.assign(nrec, 0);
    for (int r = 0; r < nrec; ++r) {
        const int32_t v = rec_dofs[r];
        if (v >= 0) {
            SVLG_CHECK(v < m.nfree, "recorded DOF is not a free DOF");
            pos[r] = f2n[v];
        } else {
            const int64_t t = -1 - int64_t(v);
            SVLG_CHECK(t < int64_t(m.nnodes) * m.ndim, "recorded total DOF out of range");
            pos[r] = m.inv[t / m.ndim] * m.ndim + int32_t(t % m.ndim);
        }
    }
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

static int create_impl(const svlg_mesh_desc *d, const svlg_exec_opts *o, svlg_model **out) {
    SVLG_CHECK(d && out, "svlg_model_create: null argument");
    *out = nullptr;
    SVLG_CHECK(d->ndim == 2 || d->ndim == 3, "ndim must be 2 or 3");
    SVLG_CHECK((d->ndim == 2 && d->elem_type == SVLG_ELEM_QUAD4) || (d->ndim == 3 && d->elem_type == SVLG_ELEM_HEXA8),
               "element type must be lin2DQuad4 in 2D or lin3DHexa8 in 3D");
    SVLG_CHECK(d->nnodes > 0 && d->nelem > 0, "empty mesh");
    SVLG_CHECK(d->coords && d->conn && d->free_dof && d->E && d->nu && d->rho, "missing mesh arrays");
    const int ndim = d->ndim, nen = ndim == 2 ? 4 : 8, nn = d->nnodes, ne = d->nelem;
    for (size_t i = 0; i < size_t(ne) * nen; ++i)
        SVLG_CHECK(d->conn[i] >= 0 && d->conn[i] < nn, "connectivity references a node out of range");
    bool any_j2 = false;
    if (d->mat_type)
        for (int e = 0; e < ne; ++e) {
            SVLG_CHECK(d->mat_type[e] == SVLG_MAT_ELASTIC || d->mat_type[e] == SVLG_MAT_J2, "unknown material type");
            if (d->mat_type[e] == SVLG_MAT_J2) {
                SVLG_CHECK(d->j2 != nullptr, "J2 elements need the j2 parameter array");
                const double *q = d->j2 + size_t(e) * 5;
                SVLG_CHECK(q[0] > 0.0 && q[1] > 0.0 && q[2] > 0.0 && q[3] >= 0.0 && q[4] >= 0.0 && q[4] <= 1.0,
                           "J2 parameters must satisfy K > 0, G > 0, sigma_y > 0, H >= 0, 0 <= beta <= 1");
                any_j2 = true;
            }
        }
    {
        int cnt = 0;
        for (int i = 0; i < nn * ndim; ++i) {
            SVLG_CHECK(d->free_dof[i] >= -1 && d->free_dof[i] < d->nfree, "free_dof out of range");
            if (d->free_dof[i] >= 0) cnt++;
        }
        SVLG_CHECK(cnt == d->nfree, "nfree does not match free_dof");
    }
    if (d->ndash > 0) {
        SVLG_CHECK(d->dash_node && d->dash_dir && d->dash_eta, "incomplete dashpot arrays");
        for (int i = 0; i < d->ndash; ++i)
            SVLG_CHECK(d->dash_node[i] >= 0 && d->dash_node[i] < nn && d->dash_dir[i] >= 0 && d->dash_dir[i] < ndim,
                       "dashpot node or direction out of range");
    }
    if (d->nload > 0) {
        SVLG_CHECK(d->load_node && d->load_dir && d->load_factor && d->load_series && d->load_nt > 0,
                   "incomplete load arrays");
        for (int l = 0; l < d->nload; ++l)
            SVLG_CHECK(d->load_node[l] >= 0 && d->load_node[l] < nn, "load node out of range");
    }
    if (d->ndrm_elem > 0) {
        SVLG_CHECK(d->drm_elem && d->drm_node && d->drm_cond && d->drm_signal, "incomplete DRM arrays");
        for (int i = 0; i < d->ndrm_elem; ++i)
            SVLG_CHECK(d->drm_elem[i] >= 0 && d->drm_elem[i] < ne, "DRM element out of range");
    }
    int ndev = 0;
    SVLG_CHECK(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, "no HIP device available");

    auto *h = new svlg_model();
    Model &m = h->m;
    auto fail = [&](int code) {
        hipStream_t own = m.own_stream;
        delete h;
        if (own) (void)hipStreamDestroy(own);
        return code;
    };
    if (hipGetDevice(&m.device) != hipSuccess) {
        set_error("hipGetDevice failed");
        return fail(10);
    }
    m.ndim = ndim;
    m.nen = nen;
    m.nnodes = nn;
    m.nelem = ne;
    m.nfree = d->nfree;
    m.ndofn = nn * ndim;
    m.th = d->thickness;
    m.lumped = d->lumped != 0;
    int tile_nodes = (o && o->tile_nodes > 0) ? o->tile_nodes : (ndim == 3 ? 160 : 448);
    int block = (o && o->block > 0) ? o->block : 512;
    const int req_variant = (o && o->variant > 0) ? o->variant : 0;
    if (req_variant < 0 || req_variant > 3) { set_error("variant must be 0 (auto), 1, 2 or 3"); return fail(11); }
    if (block % 64 != 0 || block > 512) { set_error("block must be a multiple of 64, at most 512"); return fail(11); }
    if (hipStreamCreateWithFlags(&m.own_stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(10);
    }
    m.stream = m.own_stream;
    hipStream_t s = m.stream;

    // Coloring and affinity in SVL's tag order.
    if (color_elements(ne, nen, nn, d->conn, m.h_color, m.ncolors)) {
        set_error("element coloring needs more than 64 colors");
        return fail(11);
    }
    m.affine = all_affine(ndim, ne, nen, d->conn, d->coords);
    std::vector<double> am(ne, 0.0), ak(ne, 0.0);
    for (int e = 0; e < ne; ++e) {
        if (d->ray_am) am[e] = d->ray_am[e];
        if (d->ray_ak) ak[e] = d->ray_ak[e];
        if (ak[e] != 0.0) m.has_ak = true;
    }
    // Nodes that need the general update.
    std::vector<uint8_t> special(nn, 0);
    for (int n = 0; n < nn; ++n)
        for (int c = 0; c < ndim; ++c)
            if (d->free_dof[size_t(n) * ndim + c] < 0) special[n] = 1;
    for (int i = 0; i < d->ndash; ++i) special[d->dash_node[i]] = 1;
    for (int l = 0; l < d->nload; ++l) special[d->load_node[l]] = 1;
    for (int i = 0; i < std::max(d->ndrm_elem, 0); ++i)
        for (int a = 0; a < nen; ++a) special[d->conn[size_t(d->drm_elem[i]) * nen + a]] = 1;
    for (int e = 0; e < ne; ++e)
        if (am[e] != 0.0)
            for (int a = 0; a < nen; ++a) special[d->conn[size_t(e) * nen + a]] = 1;

    // Corner-slot scatter needs every (node, corner position) pair to occur at most once.
    bool corner_unique = true;
    {
        std::vector<uint8_t> used(nn, 0);
        for (int e = 0; e < ne && corner_unique; ++e)
            for (int a = 0; a < nen; ++a) {
                uint8_t &u = used[d->conn[size_t(e) * nen + a]];
                if (u & (1u << a)) { corner_unique = false; break; }
                u |= uint8_t(1u << a);
            }
    }
    HostTiles ht;
    try {
        make_tiles(ndim, nn, ne, nen, d->coords, d->conn, m.h_color.data(), m.ncolors, special.data(), tile_nodes, ht);
    } catch (const std::exception &ex) {
        set_error(std::string("tiling failed: ") + ex.what());
        return fail(11);
    }
    m.perm = ht.perm;
    m.inv = ht.inv;
    // Device (tile-ordered) copies of every node-indexed array.
    std::vector<double> Xp(size_t(nn) * ndim);
    m.h_free.resize(size_t(nn) * ndim);
    for (int p = 0; p < nn; ++p)
        for (int c = 0; c < ndim; ++c) {
            Xp[size_t(p) * ndim + c] = d->coords[size_t(m.perm[p]) * ndim + c];
            m.h_free[size_t(p) * ndim + c] = d->free_dof[size_t(m.perm[p]) * ndim + c];
        }
    std::vector<int32_t> connp(size_t(ne) * nen);
    for (size_t i = 0; i < connp.size(); ++i) connp[i] = m.inv[d->conn[i]];
    // Color-sorted order (stable: ascending tag inside each color).
    m.color_off.assign(m.ncolors + 1, 0);
    for (int e = 0; e < ne; ++e) m.color_off[m.h_color[e] + 1]++;
    for (int c = 0; c < m.ncolors; ++c) m.color_off[c + 1] += m.color_off[c];
    std::vector<int32_t> pos(m.color_off.begin(), m.color_off.end() - 1);
    std::vector<int32_t> conn_c(size_t(ne) * nen), eid_c(ne);
    for (int e = 0; e < ne; ++e) {
        int p = pos[m.h_color[e]]++;
        eid_c[p] = e;
        std::memcpy(&conn_c[size_t(p) * nen], &connp[size_t(e) * nen], sizeof(int32_t) * nen);
    }
    // Lame constants of the isotropic D (c2 = lambda, c3 = mu; c1 = c2 + 2 c3).
    // J2 elements: the initial tangent K D + 2G (I - D/3) has lambda = K - 2G/3, mu = G.
    std::vector<double> lam(ne), mu(ne);
    std::vector<int32_t> mat(ne, SVLG_MAT_ELASTIC);
    std::vector<double> j2p(size_t(ne) * 5, 0.0);
    for (int e = 0; e < ne; ++e) {
        if (any_j2 && d->mat_type[e] == SVLG_MAT_J2) {
            mat[e] = SVLG_MAT_J2;
            const double *q = d->j2 + size_t(e) * 5;
            for (int i = 0; i < 5; ++i) j2p[size_t(e) * 5 + i] = q[i];
            lam[e] = q[0] - 2.0 / 3.0 * q[1];
            mu[e] = q[1];
            continue;
        }
        double E = d->E[e], nu = d->nu[e];
        lam[e] = E * nu / (1.0 - 2.0 * nu) / (1.0 + nu);
        mu[e] = E / (2.0 * (1.0 + nu));
    }
    m.has_j2 = any_j2;
    // Node -> (element, local corner) in ascending element order.
    std::vector<int32_t> n2e_ptr(nn + 1, 0), n2e(size_t(ne) * nen);
    for (size_t i = 0; i < connp.size(); ++i) n2e_ptr[connp[i] + 1]++;
    for (int n = 0; n < nn; ++n) n2e_ptr[n + 1] += n2e_ptr[n];
    {
        std::vector<int32_t> p(n2e_ptr.begin(), n2e_ptr.end() - 1);
        for (int e = 0; e < ne; ++e)
            for (int a = 0; a < nen; ++a) n2e[p[connp[size_t(e) * nen + a]]++] = e * nen + a;
    }
#define UP(buf, src, cnt)                                                   \
    if (buf.upload(src, cnt, s) != hipSuccess) {                            \
        set_error("device allocation/upload failed for " #buf);             \
        return fail(10);                                                    \
    }
    UP(m.X, Xp.data(), Xp.size());
    UP(m.conn, connp.data(), connp.size());
    UP(m.conn_c, conn_c.data(), conn_c.size());
    UP(m.eid_c, eid_c.data(), eid_c.size());
    UP(m.lam, lam.data(), lam.size());
    UP(m.mu, mu.data(), mu.size());
    UP(m.rho, d->rho, size_t(ne));
    m.h_rho.assign(d->rho, d->rho + ne);
    m.h_am = am;
    UP(m.am, am.data(), am.size());
    UP(m.ak, ak.data(), ak.size());
    UP(m.fdof, m.h_free.data(), m.h_free.size());
    UP(m.n2e_ptr, n2e_ptr.data(), n2e_ptr.size());
    UP(m.n2e, n2e.data(), n2e.size());
    UP(m.mat, mat.data(), mat.size());
    if (any_j2) UP(m.j2p, j2p.data(), j2p.size());
    // Element path of the fused kernel and, for parallelepipeds, the table of distinct
    // (J^-1, |J|) geometries (one entry for a uniform structured grid).
    int geo_opt = o ? o->geo_table : 0;
    m.elem_path = m.affine ? 1 : 0;
    std::vector<double> escale(ne, 1.0), gtab;
    std::vector<uint16_t> gid(ne, 0);
    if (m.affine && geo_opt != -1) {
        std::map<std::vector<uint64_t>, int> classes;
        bool ok = true;
        for (int e = 0; e < ne && ok; ++e) {
            std::vector<double> lam_e(ndim * ndim);
            double adet;
            if (ndim == 3) {
                double x[8][3], L[3][3];
                for (int a = 0; a < 8; ++a)
                    for (int c = 0; c < 3; ++c) x[a][c] = d->coords[size_t(d->conn[size_t(e) * 8 + a]) * 3 + c];
                adet = hexa_center_geometry(x, L);
                for (int i = 0; i < 9; ++i) lam_e[i] = L[i / 3][i % 3];
                escale[e] = 0.125 * adet;
            } else {
                double x[4][2], L[2][2];
                for (int a = 0; a < 4; ++a)
                    for (int c = 0; c < 2; ++c) x[a][c] = d->coords[size_t(d->conn[size_t(e) * 4 + a]) * 2 + c];
                adet = quad_center_geometry(x, L);
                for (int i = 0; i < 4; ++i) lam_e[i] = L[i / 2][i % 2];
                escale[e] = 0.25 * adet * m.th;
            }
            std::vector<uint64_t> key(lam_e.size() + 1);
            std::memcpy(key.data(), lam_e.data(), sizeof(double) * lam_e.size());
            std::memcpy(&key[lam_e.size()], &adet, sizeof(double));
            auto it = classes.find(key);
            if (it == classes.end()) {
                if (classes.size() >= 65535) { ok = false; break; }
                int id = int(classes.size());
                classes.emplace(key, id);
                gtab.insert(gtab.end(), lam_e.begin(), lam_e.end());
                gid[e] = uint16_t(id);
            } else {
                gid[e] = uint16_t(it->second);
            }
        }
        if (ok) {
            m.elem_path = 2;
            m.ngeo = int(classes.size());
            // Axis-aligned bricks / rectangles: every J^-1 in the table is diagonal.
            bool diag = true;
            for (int g = 0; g < m.ngeo && diag; ++g)
                for (int i = 0; i < ndim; ++i)
                    for (int k = 0; k < ndim; ++k)
                        if (i != k && gtab[size_t(g) * ndim * ndim + i * ndim + k] != 0.0) diag = false;
            if (diag) m.elem_path = 3;
        }
    }
    if (m.elem_path >= 2) {
        UP(m.geo, gtab.data(), gtab.size());
        UP(m.egid, gid.data(), gid.size());
    }
    m.h_escale = escale;
    // Tiles.
    {
        TileSet &t = m.tiles;
        t.ntiles = ht.ntiles;
        t.ncolors = ht.ncolors;
        t.max_stage = ht.max_stage;
        t.max_own = ht.max_own;
        t.max_elem = ht.max_elem;
        t.block = block;
        const int staged_arrays = m.elem_path >= 2 ? 1 : 2;  // U (and X unless the table is used)
        // Fused-kernel flavour: corner slots when the mesh allows it and one round of
        // <= 512 threads covers every tile element and owned node; colour phases otherwise.
        const int need = std::max(ht.max_elem, ht.max_own);
        const int cblock = std::max(64, (need + 63) / 64 * 64);
        const size_t corner_lds =  // + one trash cell for the branch-free corner scatter
            (size_t(staged_arrays * ht.max_stage) * ndim + size_t(nen) * ndim * ht.max_own + 1) * sizeof(double);
        const bool corner_ok = corner_unique && cblock <= 512 && corner_lds <= 160 * 1024;
        if (req_variant == 0) m.variant = corner_ok ? 3 : 1;
        else m.variant = req_variant;
        if (m.variant == 3 && !corner_ok) {
            set_error("corner-slot kernel needs a corner-unique mesh and tiles of at most 512 elements");
            return fail(11);
        }
        if (m.variant == 3) {
            t.block = cblock;
            t.lds_bytes = corner_lds;
            // Corner slots make the sums independent of the element order, so the tile's
            // elements go in the order of their first corner's staged slot instead of by
            // colour: neighbouring lanes then touch neighbouring LDS slots (fewer bank
            // conflicts in the gathers and the corner-slot writes). The colour offsets are
            // not used by the corner-slot kernels.
            std::vector<int32_t> ord, id2;
            std::vector<uint16_t> lc2;
            for (int ti = 0; ti < ht.ntiles; ++ti) {
                const int e0 = ht.elem_ptr[ti], e1 = ht.elem_ptr[ti + 1];
                ord.resize(e1 - e0);
                std::iota(ord.begin(), ord.end(), e0);
                std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
                    return ht.elem_lconn[size_t(a) * nen] < ht.elem_lconn[size_t(b) * nen];
                });
                id2.resize(e1 - e0);
                lc2.resize(size_t(e1 - e0) * nen);
                for (int i = 0; i < e1 - e0; ++i) {
                    id2[i] = ht.elem_id[ord[i]];
                    for (int a = 0; a < nen; ++a) lc2[size_t(i) * nen + a] = ht.elem_lconn[size_t(ord[i]) * nen + a];
                }
                std::copy(id2.begin(), id2.end(), ht.elem_id.begin() + e0);
                std::copy(lc2.begin(), lc2.end(), ht.elem_lconn.begin() + size_t(e0) * nen);
            }
        } else {
            t.lds_bytes = size_t(staged_arrays * ht.max_stage + ht.max_own) * ndim * sizeof(double);
        }
        std::vector<double> el(ht.elem_id.size()), em(ht.elem_id.size());
        std::vector<uint16_t> eg(ht.elem_id.size(), 0);
        for (size_t i = 0; i < ht.elem_id.size(); ++i) {
            int e = ht.elem_id[i];
            if (m.elem_path >= 2) {
                el[i] = lam[e] * escale[e];
                em[i] = mu[e] * escale[e];
                eg[i] = gid[e];
            } else {
                el[i] = lam[e];
                em[i] = mu[e];
            }
        }
        UP(t.elem_geo, eg.data(), eg.size());
        m.tile_elem_id = ht.elem_id;
        UP(m.tile_eid, ht.elem_id.data(), ht.elem_id.size());
        UP(t.own_ptr, ht.own_ptr.data(), ht.own_ptr.size());
        UP(t.nplain, ht.nplain.data(), ht.nplain.size());
        UP(t.halo_ptr, ht.halo_ptr.data(), ht.halo_ptr.size());
        UP(t.halo_node, ht.halo_node.data(), ht.halo_node.size());
        UP(t.elem_ptr, ht.elem_ptr.data(), ht.elem_ptr.size());
        UP(t.elem_lconn, ht.elem_lconn.data(), ht.elem_lconn.size());
        UP(t.elem_lam, el.data(), el.size());
        UP(t.elem_mu, em.data(), em.size());
        UP(t.color_ptr, ht.color_ptr.data(), ht.color_ptr.size());
        if (m.variant == 3 && m.elem_path >= 2) {
            int hmax = 0;
            for (int ti = 0; ti < ht.ntiles; ++ti) hmax = std::max(hmax, ht.halo_ptr[ti + 1] - ht.halo_ptr[ti]);
            const int emax = ht.max_elem;
            // One staging pass of two block strides must cover the owned DOFs and the halo.
            if (std::max(ht.max_own * ndim, hmax) <= 2 * t.block && emax <= t.block && hmax > 0) {
                const size_t nt = size_t(ht.ntiles);
                std::vector<int32_t> hp(nt * hmax, 0);
                std::vector<uint16_t> lp(nt * emax * nen, 0), gp(nt * emax, 0);
                std::vector<double> lap(nt * emax, 0.0), mup(nt * emax, 0.0);
                for (int ti = 0; ti < ht.ntiles; ++ti) {
                    for (int i = ht.halo_ptr[ti]; i < ht.halo_ptr[ti + 1]; ++i)
                        hp[size_t(ti) * hmax + (i - ht.halo_ptr[ti])] = ht.halo_node[i];
                    for (int i = ht.elem_ptr[ti]; i < ht.elem_ptr[ti + 1]; ++i) {
                        const size_t o = size_t(ti) * emax + (i - ht.elem_ptr[ti]);
                        for (int a = 0; a < nen; ++a) lp[o * nen + a] = ht.elem_lconn[size_t(i) * nen + a];
                        lap[o] = el[i];
                        mup[o] = em[i];
                        gp[o] = eg[i];
                    }
                }
                UP(t.hpad, hp.data(), hp.size());
                UP(t.lcpad, lp.data(), lp.size());
                UP(t.geopad, gp.data(), gp.size());
                UP(t.lampad, lap.data(), lap.size());
                UP(t.mupad, mup.data(), mup.size());
                t.hstride = hmax;
                t.estride = emax;
            }
        }
    }
    // Loads (device node numbering).
    m.nload = d->nload;
    m.load_nt = d->load_nt;
    if (d->nload > 0) {
        std::vector<int32_t> lnode(d->nload);
        for (int l = 0; l < d->nload; ++l) lnode[l] = m.inv[d->load_node[l]];
        UP(m.load_node, lnode.data(), lnode.size());
        UP(m.load_dir, d->load_dir, size_t(d->nload) * ndim);
        UP(m.load_factor, d->load_factor, size_t(d->nload));
        UP(m.load_series, d->load_series, size_t(d->nload) * d->load_nt);
        std::vector<int32_t> dofs;
        for (int l = 0; l < d->nload; ++l)
            for (int c = 0; c < ndim; ++c) dofs.push_back(lnode[l] * ndim + c);
        std::sort(dofs.begin(), dofs.end());
        dofs.erase(std::unique(dofs.begin(), dofs.end()), dofs.end());
        m.h_load_dofs = dofs;
        UP(m.load_dofs, dofs.data(), dofs.size());
        std::vector<int32_t> nlp(nn + 1, 0), nli(d->nload);
        for (int l = 0; l < d->nload; ++l) nlp[lnode[l] + 1]++;
        for (int n = 0; n < nn; ++n) nlp[n + 1] += nlp[n];
        std::vector<int32_t> q(nlp.begin(), nlp.end() - 1);
        for (int l = 0; l < d->nload; ++l) nli[q[lnode[l]]++] = l;
        UP(m.nl_ptr, nlp.data(), nlp.size());
        UP(m.nl_idx, nli.data(), nli.size());
    } else {
        std::vector<int32_t> nlp(nn + 1, 0);
        UP(m.nl_ptr, nlp.data(), nlp.size());
    }
    // Mass and damping diagonals.
    if (m.mdiag.alloc(m.ndofn) != hipSuccess || m.cdiag.alloc(m.ndofn) != hipSuccess ||
        m.cdash.alloc(m.ndofn) != hipSuccess || m.scratch.alloc(size_t(ne) * nen) != hipSuccess ||
        hipMemsetAsync(m.cdash.p, 0, sizeof(double) * m.ndofn, s) != hipSuccess) {
        set_error("device allocation failed (diagonals)");
        return fail(10);
    }
    {
        int nb = (ne + 127) / 128;
        if (ndim == 2)
            hipLaunchKernelGGL(k_elem_lumped_mass<2>, dim3(nb), dim3(128), 0, s, ne, m.conn.p, m.X.p, m.rho.p, m.th,
                               m.scratch.p);
        else
            hipLaunchKernelGGL(k_elem_lumped_mass<3>, dim3(nb), dim3(128), 0, s, ne, m.conn.p, m.X.p, m.rho.p,
                               m.th, m.scratch.p);
        int nbn = (nn + 127) / 128;
        hipLaunchKernelGGL(k_node_gather_mass, dim3(nbn), dim3(128), 0, s, nn, ndim, m.n2e_ptr.p, m.n2e.p,
                           m.scratch.p, m.am.p, nen, 1e-12, m.mdiag.p, m.cdiag.p);
    }
    if (d->ndash > 0) {
        // Group the dashpots by device DOF, keeping their element-tag order inside a DOF.
        std::vector<std::pair<int32_t, int32_t>> key(d->ndash);
        for (int i = 0; i < d->ndash; ++i) key[i] = {m.inv[d->dash_node[i]] * ndim + d->dash_dir[i], i};
        std::stable_sort(key.begin(), key.end(),
                         [](const auto &x, const auto &y) { return x.first < y.first; });
        std::vector<int32_t> dofs, ptr{0};
        std::vector<double> eta(d->ndash);
        for (int i = 0; i < d->ndash; ++i) {
            if (dofs.empty() || dofs.back() != key[i].first) {
                if (!dofs.empty()) ptr.push_back(i);
                dofs.push_back(key[i].first);
            }
            eta[i] = d->dash_eta[key[i].second];
        }
        ptr.push_back(d->ndash);
        DevBuf<int32_t> dd, dp;
        DevBuf<double> de;
        UP(dd, dofs.data(), dofs.size());
        UP(dp, ptr.data(), ptr.size());
        UP(de, eta.data(), eta.size());
        const int nd = int(dofs.size());
        hipLaunchKernelGGL(k_dashpots, dim3((nd + 255) / 256), dim3(256), 0, s, nd, dd.p, dp.p, de.p, 1e-12,
                           m.cdiag.p);
        hipLaunchKernelGGL(k_dashpots, dim3((nd + 255) / 256), dim3(256), 0, s, nd, dd.p, dp.p, de.p, 1e-12,
                           m.cdash.p);
        if (hipStreamSynchronize(s) != hipSuccess) { set_error("dashpot assembly failed"); return fail(10); }
    }
#undef UP
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        set_error("model set-up kernels failed");
        return fail(10);
    }
    {
        const int prec = o ? o->elem_precision : 0;
        if (prec == 1) {
            if (!(ndim == 3 && m.variant == 3 && m.elem_path >= 2 && m.ngeo == 1)) {
                set_error("elem_precision 1 (BF16 MFMA) needs a Hexa8 mesh with one element geometry and "
                          "corner-slot tiles");
                return fail(11);
            }
            if (int rc = build_hexa_kmat(m, m.cd_kmat)) return fail(rc);
            m.elem_bf16 = 1;
        } else if (prec == 2) {
            if (!(ndim == 3 && m.variant == 3 && m.elem_path == 3)) {
                set_error("elem_precision 2 (FP32 element step) needs axis-aligned Hexa8 bricks (the diagonal "
                          "geometry table) and corner-slot tiles");
                return fail(11);
            }
            m.elem_fp32 = 1;
        } else if (prec != 0) {
            set_error("elem_precision must be 0 (FP64), 1 (BF16 MFMA) or 2 (FP32)");
            return fail(11);
        }
    }
    if (o && (o->cg_launch < 0 || o->cg_launch > 1)) {
        set_error("cg_launch must be 0 (graph replays) or 1 (eager launches)");
        return fail(11);
    }
    m.cg_eager = o ? o->cg_launch : 0;
    if (int rc = drm_setup(m, d)) return fail(rc);
    if (int rc = sm_setup(m, d)) return fail(rc);
    if (int rc = prepare_fused(m)) return fail(rc);
    *out = h;
    return 0;
}

int svlg_model_create(const svlg_mesh_desc *d, svlg_model **out) { return create_impl(d, nullptr, out); }

int svlg_plan_tiles(const svlg_mesh_desc *d, int32_t tile_nodes, int64_t *stats) {
    SVLG_CHECK(d && stats, "null argument");
    SVLG_CHECK(d->ndim == 2 || d->ndim == 3, "ndim must be 2 or 3");
    const int ndim = d->ndim, nen = ndim == 2 ? 4 : 8, nn = d->nnodes, ne = d->nelem;
    for (size_t i = 0; i < size_t(ne) * nen; ++i)
        SVLG_CHECK(d->conn[i] >= 0 && d->conn[i] < nn, "connectivity references a node out of range");
    std::vector<int32_t> color;
    int ncolors = 0;
    SVLG_CHECK(color_elements(ne, nen, nn, d->conn, color, ncolors) == 0, "coloring failed");
    std::vector<uint8_t> special(nn, 0);
    for (int n = 0; n < nn; ++n)
        for (int c = 0; c < ndim; ++c)
            if (d->free_dof[size_t(n) * ndim + c] < 0) special[n] = 1;
    for (int i = 0; i < d->ndash; ++i) special[d->dash_node[i]] = 1;
    for (int l = 0; l < d->nload; ++l) special[d->load_node[l]] = 1;
    for (int i = 0; i < std::max(d->ndrm_elem, 0); ++i) {
        SVLG_CHECK(d->drm_elem && d->drm_elem[i] >= 0 && d->drm_elem[i] < ne, "DRM element out of range");
        for (int a = 0; a < nen; ++a) special[d->conn[size_t(d->drm_elem[i]) * nen + a]] = 1;
    }
    HostTiles t;
    try {
        make_tiles(ndim, nn, ne, nen, d->coords, d->conn, color.data(), ncolors, special.data(),
                   tile_nodes > 0 ? tile_nodes : (ndim == 3 ? 160 : 448), t);
    } catch (const std::exception &ex) {
        set_error(std::string("tiling failed: ") + ex.what());
        return 11;
    }
    // Invariants: every node owned once, owned ranges contiguous, plain nodes first;
    // each tile lists exactly the elements touching its nodes, colour-sorted; every
    // corner maps to a staged slot holding that corner's node.
    std::vector<int> owner(nn, -1);
    std::vector<int32_t> seen(ne, -1);
    int64_t tot_elem = 0, tot_halo = 0;
    for (int ti = 0; ti < t.ntiles; ++ti) {
        int o0 = t.own_ptr[ti], o1 = t.own_ptr[ti + 1];
        SVLG_CHECK(o0 <= o1 && o1 - o0 <= t.max_own, "bad owned range");
        for (int p = o0; p < o1; ++p) {
            int n = t.perm[p];
            SVLG_CHECK(owner[n] < 0 && t.inv[n] == p, "node owned twice or bad permutation");
            owner[n] = ti;
            SVLG_CHECK((p - o0 < t.nplain[ti]) == (special[n] == 0), "plain/special order broken");
        }
        int e0 = t.elem_ptr[ti], e1 = t.elem_ptr[ti + 1];
        int nh = t.halo_ptr[ti + 1] - t.halo_ptr[ti];
        int nstage = (o1 - o0) + nh;
        tot_elem += e1 - e0;
        tot_halo += nh;
        const int32_t *cp = &t.color_ptr[size_t(ti) * (ncolors + 1)];
        SVLG_CHECK(cp[0] == 0 && cp[ncolors] == e1 - e0, "bad colour offsets");
        for (int k = e0; k < e1; ++k) {
            int e = t.elem_id[k];
            int c = 0;
            while (cp[c + 1] <= k - e0) ++c;
            SVLG_CHECK(color[e] == c, "element outside its colour segment");
            seen[e] = ti;
            bool touches = false;
            for (int a = 0; a < nen; ++a) {
                int sl = t.elem_lconn[size_t(k) * nen + a];
                SVLG_CHECK(sl < nstage, "corner slot beyond the staged nodes");
                int dev = sl < o1 - o0 ? o0 + sl : t.halo_node[t.halo_ptr[ti] + sl - (o1 - o0)];
                SVLG_CHECK(t.perm[dev] == d->conn[size_t(e) * nen + a], "corner slot holds the wrong node");
                if (sl < o1 - o0) touches = true;
            }
            SVLG_CHECK(touches, "tile element touches no owned node");
        }
        for (int p = o0; p < o1; ++p) {
            (void)p;
        }
    }
    for (int n = 0; n < nn; ++n) SVLG_CHECK(owner[n] >= 0, "node not owned by any tile");
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < nen; ++a) {
            int ti = owner[d->conn[size_t(e) * nen + a]];
            bool found = false;
            for (int k = t.elem_ptr[ti]; k < t.elem_ptr[ti + 1] && !found; ++k) found = t.elem_id[k] == e;
            SVLG_CHECK(found, "element missing from the tile owning one of its nodes");
        }
    stats[0] = t.ntiles;
    stats[1] = t.max_own;
    stats[2] = t.max_stage;
    stats[3] = t.max_elem;
    stats[4] = tot_elem;
    stats[5] = tot_halo;
    stats[6] = int64_t(size_t(2 * t.max_stage + t.max_own) * ndim * sizeof(double));
    stats[7] = ncolors;
    return 0;
}

int svlg_model_create_ex(const svlg_mesh_desc *d, const svlg_exec_opts *o, svlg_model **out) {
    return create_impl(d, o, out);
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
    info[6] = m.elem_path;
    info[7] = int64_t(m.tiles.lds_bytes);
    info[8] = m.elem_bf16 ? 1 : m.elem_fp32 ? 2 : 0;
    info[9] = m.ngeo;
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
    SVLG_CHECK(!m.has_j2, "svlg_internal_force restates elastic elements; J2 forces come from their states");
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

__global__ void k_stream_copy(double *__restrict__ dst, const double *__restrict__ src, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int svlg_stream_copy(double *dst, const double *src, int64_t n, void *stream) {
    SVLG_CHECK(dst && src && n >= 0, "bad copy arguments");
    if (n == 0) return 0;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_stream_copy, dim3(unsigned(blocks)), dim3(256), 0, (hipStream_t)stream, dst, src, n);
    SVLG_HIP(hipGetLastError());
    return 0;
}

int svlg_sync(svlg_model *h) {
    SVLG_CHECK(h, "null model");
    SVLG_HIP(hipStreamSynchronize(h->m.stream));
    return 0;
}

}  // extern "C"
