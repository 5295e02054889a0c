// svlg_common.hpp -- shared host/device declarations of libsvlg (MI355X, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "svlg.h"

namespace svlg {

void set_error(const std::string &msg);

#define SVLG_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::svlg::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
            return 10;                                                              \
        }                                                                           \
    } while (0)

#define SVLG_CHECK(cond, msg)                                                       \
    do {                                                                            \
        if (!(cond)) {                                                              \
            ::svlg::set_error(msg);                                                 \
            return 11;                                                              \
        }                                                                           \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&p, count * sizeof(T));
    }
    hipError_t upload(const T *h, size_t count, hipStream_t s) {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0) return e;
        return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

// Node tiles of the fused explicit kernel (device copies of HostTiles): one workgroup
// owns the contiguous device nodes [own_ptr[t], own_ptr[t+1]) and recomputes every
// element touching them from an LDS-staged copy of U and X of owned + halo nodes.
struct TileSet {
    int ntiles = 0, ncolors = 0;
    int max_stage = 0, max_elem = 0, max_own = 0;
    int block = 256;
    size_t lds_bytes = 0;
    DevBuf<int32_t> own_ptr, nplain;         // [ntiles+1], [ntiles]
    DevBuf<int32_t> halo_ptr, halo_node;     // [ntiles+1], device node of each halo slot
    DevBuf<int32_t> elem_ptr;                // [ntiles+1]
    DevBuf<uint16_t> elem_lconn;             // [sum elems * nen] staged slot of each corner
    DevBuf<double> elem_lam, elem_mu;        // [sum elems] Lame constants per tile element
                                             // (times |J|/8, or |J| th/4 in 2D, on the table path)
    DevBuf<uint16_t> elem_geo;               // [sum elems] geometry-table entry (table path)
    DevBuf<int32_t> color_ptr;               // [ntiles * (ncolors+1)]
    // Corner-slot step: fixed-stride copies of the halo list and of the tile-element data
    // (tile t at t * hstride / t * estride; padding: node 0, zero elements), so that a
    // workgroup can load them with its tile index alone, before its tile bounds arrive.
    // Built when the element path needs no staged coordinates and one staging pass covers
    // every tile; hstride = 0 otherwise.
    int hstride = 0, estride = 0;
    DevBuf<int32_t> hpad;
    DevBuf<uint16_t> lcpad, geopad;
    DevBuf<double> lampad, mupad;
};

struct Model {
    int ndim = 0, nen = 0, nnodes = 0, nelem = 0, nfree = 0, ndofn = 0;  // ndofn = nnodes*ndim
    int device = 0;                   // HIP device current at creation (the model's memory lives there)
    double th = 1.0;
    bool lumped = true;
    bool affine = false;
    bool has_ak = false;
    hipStream_t own_stream = nullptr, stream = nullptr;

    // Host copies needed by later set-up calls.
    std::vector<int32_t> h_free;      // [ndofn]
    std::vector<int32_t> h_color;     // [nelem]
    std::vector<int32_t> color_off;   // [ncolors+1] offsets into color-sorted order
    int ncolors = 0;

    // Geometry and materials (device).
    DevBuf<double> X;                 // [nnodes*ndim]
    DevBuf<int32_t> conn;             // [nelem*nen] tag order
    DevBuf<int32_t> conn_c;           // [nelem*nen] color-sorted order
    DevBuf<int32_t> eid_c;            // [nelem] tag index of color-sorted slot
    DevBuf<double> lam, mu, rho;      // [nelem] tag order
    DevBuf<double> am, ak;            // [nelem] Rayleigh
    DevBuf<int32_t> fdof;             // [ndofn] free index or -1
    DevBuf<int32_t> n2e_ptr, n2e;     // node -> (element*nen + local) in ascending element

    // Diagonals (node-DOF layout).
    DevBuf<double> mdiag, cdiag;      // lumped mass, damping
    DevBuf<double> cdash;             // damping of the dashpots alone (consistent-mass operators)
    std::vector<double> h_rho, h_am;  // [nelem] density, Rayleigh am (host)
    std::vector<double> h_escale;     // [nelem] |J|/8 (Hexa8) or |J| th/4 (Quad4) on the table paths
    DevBuf<double> scratch;           // [nelem*nen] per-element lumped masses

    // Loads.
    int nload = 0, load_nt = 0;
    DevBuf<int32_t> load_node;
    DevBuf<double> load_dir, load_factor, load_series;
    std::vector<int32_t> h_load_dofs; // distinct loaded node-DOFs
    DevBuf<int32_t> load_dofs;
    DevBuf<int32_t> nl_ptr, nl_idx;   // node -> loads acting on it (ascending load index)

    // Domain reduction load (ELEMENTLOAD_DOMAIN_REDUCTION): the factored nodal forces of every
    // signal row, drm_F [nt][slot][ndim] (slots in ascending device node order), the slot of
    // every device node (-1 outside the layer) and the device node of every slot.
    int ndrm = 0, ndrm_node = 0, drm_nt = 0;
    DevBuf<double> drm_F;
    DevBuf<int32_t> drm_slot, drm_dnode;

    // Support motion (svlg_support.hpp): the restrained node-DOFs with a prescribed motion.
    int sm_ndofs = 0, sm_offset = 0;
    DevBuf<int32_t> sm_slot, sm_dof, sm_dptr, sm_vptr;
    DevBuf<double> sm_fac, sm_vals;
    DevBuf<double> sm_vec, sm_y;      // Newmark: a_k as a node-DOF vector, Keff times it
    DevBuf<double> sm_save, sm_lag;   // [sm_ndofs] U while lagged; Newmark: U_(k-1) there

    // Device node numbering (tile order): perm[device] = original, inv[original] = device.
    std::vector<int32_t> perm, inv;

    // Central difference state.
    bool cd_ready = false;
    int active = 0;                   // integrator of the last init: 1 CentralDifference, 2 Newmark
    double dt = 0.0, ftol = 1e-12;
    int ring = 0;                     // slot of U_k (current)
    long steps_done = 0;
    DevBuf<double> U[3];              // ring: U_k, U_{k-1}, U_{k-2}
    DevBuf<double> F;                 // internal-force accumulator
    DevBuf<double> Fext;              // dense external force (loaded DOFs only written)
    DevBuf<double> keff, kforce;      // diag(M/dt^2 + C/2dt), diag(M/dt^2 - C/2dt)
    DevBuf<double> kinv;              // 1 / keff (0 where keff = 0: the restrained DOFs)
    DevBuf<uint16_t> egid;            // [nelem] geometry-table entry per element (table paths)
    DevBuf<double> hist;              // recorder buffer
    DevBuf<int32_t> rec;              // recorded node-DOFs
    // Independent meshes advanced together (one model whose parts share no node): each part
    // keeps its own time step.
    std::vector<double> h_dt_dof;     // [ndofn] time step of every DOF (empty: m.dt everywhere)
    DevBuf<double> dt_dof;
    // Graph-captured stepping: cd_graph_steps launches (a multiple of 3, so the U ring comes
    // back to the captured slot) plus the advance of the device step index kdev.
    hipGraphExec_t cd_graph = nullptr;
    int cd_graph_steps = 0, cd_graph_ring = -1;
    DevBuf<int> kdev;
    // Captured runs of CG iterations (2 or 8, starting at an even iteration) for the solution
    // buffers the integrators reuse, keyed by (x, flexible, operator kernel, count); the
    // iteration base lives in jdev.
    struct CgGraph {
        hipGraphExec_t exec;
        const double *x;
        int flexible;
        const void *op;
        int count;
    };
    std::vector<CgGraph> cg_graphs;
    DevBuf<int> jdev;
    void drop_cg_graphs() {
        for (auto &g : cg_graphs) (void)hipGraphExecDestroy(g.exec);
        cg_graphs.clear();
    }

    // Implicit (Extended)NewmarkBeta state and Jacobi / learned PCG work vectors.
    bool nm_ready = false;
    bool nm_initial = false;          // operators use the initial stiffness of J2 elements
    double nm_dt = 0.0, nm_rtol = 1e-8;
    int nm_maxit = 1000, nm_precond = 0;
    long nm_steps_done = 0;
    DevBuf<double> nU, nV, nA, ndU;   // state and last increment (node-DOF layout)
    DevBuf<double> dgn;               // (4/dt^2) M + (2/dt) C_diag   (lumped, per DOF)
    DevBuf<double> jac;               // 1 / diag(Keff), 0 on restrained DOFs
    DevBuf<double> coef_keff, coef_rhs;  // per tile element: 1 + 2 ak/dt, -ak
    // Consistent mass (MassFormulation CONSISTENT), per tile element: the element-mass scale
    // rs (rho |J|/8, rho th |J|/4 on the table paths, else rho (th)) times 4/dt^2 + 2 am/dt for
    // the operator, rs alone and am for the right-hand side.
    DevBuf<double> cmass_keff, cmass_rhs, cmass_am;
    DevBuf<double> cg_r, cg_z, cg_q, cg_b, cg_p[2];
    DevBuf<double> part;              // partial sums of a solve (CgParts)
    DevBuf<double> scal;              // device scalars of the solver
    int cg_last_iters[2] = {0, 0};    // iterations of the last solve (Jacobi, flexible)
    double *scal_host = nullptr;      // pinned host copy of the solver scalars (convergence polls)
    std::vector<int32_t> tile_elem_id;   // tag index of every tile element (host)
    DevBuf<int32_t> tile_eid;         // the same on the device

    // J2 materials (Plastic3DJ2 / PlasticPlaneStrainJ2) and the Newton-Raphson path.
    bool has_j2 = false;
    DevBuf<int32_t> mat;              // [nelem] material type
    DevBuf<double> j2p;               // [nelem * 5] K, G, sigma_y, H, beta
    DevBuf<double> gp_state;          // [nelem * ngp * kJ2State] Gauss-point states
    DevBuf<int32_t> j2_plastic;       // [nelem] 1: some Gauss point holds the elastoplastic tangent
    bool nl_ready = false;
    double nl_tol = 1e-6;
    int nl_nmax = 20, nl_test = 5;
    double nl_normfactor = 0.0;       // Algorithm::NormFactor, kept across steps
    DevBuf<double> nl_dUa, nl_un, nl_w;  // Newton increment, U + dU, V - 2/dt dU
    DevBuf<double> nl_part;           // reduction partials

    // Learned graph-convolution preconditioner (svlg_gcn.hip). Hidden states are FP32
    // planes [channel][ndofn]; parameters theta live on the host (FP64) and the device (FP32).
    bool gcn_ready = false;
    int gcn_C = 1, gcn_L = 1, gcn_mfma = 0;
    double gcn_lmin = 0.0, gcn_lmax = 0.0;
    int gcn_val_iters[2] = {0, 0};    // validation iterations before / after the last training
    std::vector<double> gcn_theta;
    DevBuf<float> gcn_w;              // theta in FP32
    DevBuf<double> gcn_s;             // D^-1/2 (0 on restrained DOFs)
    DevBuf<float> gcn_sf;             // the same in FP32
    DevBuf<float> gcn_jacf;           // D^-1 in FP32 (one-channel kernel: states scaled by D^-1/2)
    DevBuf<double> gcn_elam, gcn_emu; // per tile element (lambda, mu) * (1 + 2 ak/dt)
    DevBuf<float> gcn_elamf, gcn_emuf, gcn_dgnf;  // FP32 copies read by the one-channel layer kernel
    DevBuf<float> gcn_elamf_pad, gcn_emuf_pad;    // the same at the tiles' fixed element stride
    DevBuf<float> gcn_kmat;           // [2][24][24] K_lambda, K_mu of the single geometry class (MFMA path)
    DevBuf<uint4> gcn_kimg;           // [3][64] per-lane BF16 A operands of the MFMA element step
    DevBuf<float> gcn_x0;             // s * r
    DevBuf<float> gcn_h[2];           // inference ping-pong hidden states
    DevBuf<float> gcn_H, gcn_Y;       // training: H^l and Ehat H^(l-1) for every layer
    DevBuf<float> gcn_G[2];           // training: backward ping-pong
    DevBuf<double> gcn_part;          // training: per-tile gradient partials
    DevBuf<double> gcn_tmp[3];        // training: FP64 work vectors (node-DOF layout)

    TileSet tiles;
    int variant = 1;                  // 1 fused node tiles, 2 element-colored
    int elem_path = 0;                // 0 Gauss points, 1 modal + staged X, 2 modal + geometry table
    int ngeo = 0;
    DevBuf<double> geo;               // [ngeo * ndim * ndim] inverse Jacobians of distinct geometries
    int elem_bf16 = 0;                // explicit element step on BF16 MFMA (single geometry, Hexa8)
    int cg_eager = 0;                 // integrator CG solves launched eagerly instead of from graphs
    int elem_fp32 = 0;                // explicit element step in FP32 (axis-aligned Hexa8 bricks)
    DevBuf<float> cd_kmat;            // its K_lambda, K_mu

    Model() = default;
    Model(const Model &) = delete;
    Model &operator=(const Model &) = delete;
    ~Model() {
        if (cd_graph) (void)hipGraphExecDestroy(cd_graph);
        drop_cg_graphs();
        if (scal_host) (void)hipHostFree(scal_host);
    }
};

}  // namespace svlg

struct svlg_model {
    svlg::Model m;
};
