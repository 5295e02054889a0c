// svlg_abi_check.cpp -- a C++ caller of libsvlg built with g++ against include/svlg.h alone
// (test infrastructure: it proves the header matches the library for a compiled caller, not
// only for the Python ctypes table).
//
// It builds a small lin2DQuad4 layer (Lysmer dashpots on the sides and base, a vertical
// Ricker point load at the surface), runs it through
//   A. the oracle (oracle/libsvl_oracle.so, the FP64 restatement of SVL's algorithm);
//   B. svlg_cd_steps / svlg_nm_steps (the fused device steps);
//   C. the Integrator methods one by one, in the order SVL's Linear algorithm calls them
//      (Linear::ComputeNewIncrement, Linear.cpp:22-56): ComputeEffectiveForce,
//      ComputeEffectiveStiffness, the linear solve (CentralDifference: the diagonal division
//      of the LDL^T solve; ExtendedNewmarkBeta: svlg_solve), then ComputeNewStep's update;
// and requires B and C to match A to 1e-10 relative L2 (CG at 1e-13 on the implicit path).
// Each model runs twice: as described, and with a vertical support motion of the base nodes
// (ComputeSupportMotionVector between the stiffness and the solve; the recorded DOFs then
// include base nodes, read through svlg_get_displacements).
// Exit status 0 = every check passed.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "svlg.h"

extern "C" {
// The oracle's C API (oracle/svl_oracle.cpp).
typedef struct oracle_opts {
    int32_t integrator, nt;
    double dt, mtol, ktol, ftol;
    int32_t algorithm, nmax, cnvgtest, pad;
    double cnvgtol;
} oracle_opts;
int oracle_run(const svlg_mesh_desc *d, const oracle_opts *o, int32_t nrec, const int32_t *rec_dofs, double *hist_U,
               double *hist_V, double *hist_A, double *seconds, double *final_U, int32_t *newton_iters,
               double *newton_res, int32_t nreac, const int32_t *reac_tot, double *hist_R);
const char *oracle_last_error(void);
}

namespace {

struct Layer {
    int nx, ny, nn, ne, nfree = 0;
    std::vector<double> coords, E, nu, rho, am, ak, dash_eta, load_dir, load_factor, load_series;
    std::vector<int32_t> free_dof, conn, dash_node, dash_dir, load_node;
    std::vector<int32_t> sm_node, sm_dir, sm_ptr;
    std::vector<double> sm_factor, sm_values;
    svlg_mesh_desc d;
};

// nx x ny Quad4 of 1 m, rho 2000, Vs 200, nu 0.25; base nodes restrained vertically; Lysmer
// dashpots (eta = rho V h) on the sides (both directions) and the base (horizontal).
// support: every base node moves vertically by 2 mm (1 - cos) over the first 0.1 s, the
// left half with factor 1, the right half with factor 0.5.
void make_layer(Layer &L, int nx, int ny, int nt, double dt, bool rayleigh, bool support) {
    L.nx = nx;
    L.ny = ny;
    L.nn = (nx + 1) * (ny + 1);
    L.ne = nx * ny;
    const double rho = 2000.0, vs = 200.0, nu = 0.25, G = rho * vs * vs, E = 2.0 * G * (1.0 + nu);
    const double vp = vs * std::sqrt(2.0 * (1.0 - nu) / (1.0 - 2.0 * nu));
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            L.coords.push_back(i - nx / 2.0);
            L.coords.push_back(j - ny);
        }
    for (int n = 0; n < L.nn; ++n) {
        const bool base = n / (nx + 1) == 0;
        L.free_dof.push_back(L.nfree++);
        L.free_dof.push_back(base ? -1 : L.nfree++);
    }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int a = j * (nx + 1) + i;
            L.conn.insert(L.conn.end(), {a, a + 1, a + nx + 2, a + nx + 1});
        }
    L.E.assign(L.ne, E);
    L.nu.assign(L.ne, nu);
    L.rho.assign(L.ne, rho);
    L.am.assign(L.ne, rayleigh ? 0.5712 : 0.0);
    L.ak.assign(L.ne, rayleigh ? 1.447e-3 : 0.0);
    for (int n = 0; n < L.nn; ++n) {
        const int i = n % (nx + 1), j = n / (nx + 1);
        const bool side = i == 0 || i == nx;
        if (j == 0) {
            L.dash_node.push_back(n);
            L.dash_dir.push_back(0);
            L.dash_eta.push_back(rho * vs);
        } else if (side) {
            L.dash_node.insert(L.dash_node.end(), {n, n});
            L.dash_dir.insert(L.dash_dir.end(), {0, 1});
            L.dash_eta.insert(L.dash_eta.end(), {rho * vs, rho * vp});
        }
    }
    const int top = ny * (nx + 1) + nx / 2;
    L.load_node.push_back(top);
    L.load_dir = {0.0, -1.0};
    L.load_factor = {1.0};
    const double f0 = 12.0, t0 = 0.1;
    for (int k = 0; k < nt; ++k) {
        const double b = std::pow(M_PI * f0 * (k * dt - t0), 2.0);
        L.load_series.push_back(1e5 * (1.0 - 2.0 * b) * std::exp(-b));
    }
    std::memset(&L.d, 0, sizeof(L.d));
    L.d.ndim = 2;
    L.d.elem_type = SVLG_ELEM_QUAD4;
    L.d.nnodes = L.nn;
    L.d.nelem = L.ne;
    L.d.coords = L.coords.data();
    L.d.free_dof = L.free_dof.data();
    L.d.nfree = L.nfree;
    L.d.conn = L.conn.data();
    L.d.E = L.E.data();
    L.d.nu = L.nu.data();
    L.d.rho = L.rho.data();
    L.d.thickness = 1.0;
    L.d.lumped = 1;
    L.d.ray_am = L.am.data();
    L.d.ray_ak = L.ak.data();
    L.d.ndash = (int32_t)L.dash_node.size();
    L.d.dash_node = L.dash_node.data();
    L.d.dash_dir = L.dash_dir.data();
    L.d.dash_eta = L.dash_eta.data();
    L.d.nload = 1;
    L.d.load_nt = nt;
    L.d.load_node = L.load_node.data();
    L.d.load_dir = L.load_dir.data();
    L.d.load_factor = L.load_factor.data();
    L.d.load_series = L.load_series.data();
    if (support) {
        std::vector<double> g;
        for (int k = 0; k < nt; ++k) {
            const double t = k * dt;
            g.push_back(t < 0.1 ? 1e-3 * (1.0 - std::cos(2.0 * M_PI * t / 0.1)) : 0.0);
        }
        L.sm_ptr.push_back(0);
        for (int i = 0; i <= nx; ++i) {
            L.sm_node.push_back(i);
            L.sm_dir.push_back(1);
            L.sm_factor.push_back(i <= nx / 2 ? 1.0 : 0.5);
            L.sm_values.insert(L.sm_values.end(), g.begin(), g.end());
            L.sm_ptr.push_back((int32_t)L.sm_values.size());
        }
        L.d.nsm = (int32_t)L.sm_node.size();
        L.d.sm_node = L.sm_node.data();
        L.d.sm_dir = L.sm_dir.data();
        L.d.sm_factor = L.sm_factor.data();
        L.d.sm_ptr = L.sm_ptr.data();
        L.d.sm_values = L.sm_values.data();
        L.d.sm_offset = 0;
    }
}

// Recorded DOFs: every 7th free DOF, and with support motion the vertical DOF of base nodes
// 0, 3, 6, ... (oracle: -1 - total DOF).
void recorded(const Layer &L, int stride, std::vector<int32_t> &rec_free, std::vector<int32_t> &rec_base,
              std::vector<int32_t> &rec_oracle) {
    for (int f = stride / 2; f < L.nfree; f += stride) rec_free.push_back(f);
    if (L.d.nsm > 0)
        for (int i = 0; i <= L.nx; i += 3) rec_base.push_back(i);
    rec_oracle = rec_free;
    for (int n : rec_base) rec_oracle.push_back(-1 - (n * 2 + 1));
}

// The recorded values of the current state: free DOFs from get_state, base DOFs from the
// total displacement vector.
int gather(svlg_model *m, const Layer &L, const std::vector<int32_t> &rec_free, const std::vector<int32_t> &rec_base,
           bool nm, double *row) {
    std::vector<double> U(L.nfree), Ut((size_t)L.nn * 2);
    if ((nm ? svlg_nm_get_state(m, U.data(), nullptr, nullptr) : svlg_cd_get_state(m, U.data(), nullptr, nullptr)) != 0)
        return 2;
    if (svlg_get_displacements(m, Ut.data()) != 0) return 2;
    size_t r = 0;
    for (int f : rec_free) row[r++] = U[f];
    for (int n : rec_base) row[r++] = Ut[(size_t)n * 2 + 1];
    return 0;
}

double rel_l2(const std::vector<double> &a, const std::vector<double> &b) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / (den > 0.0 ? den : 1e-300));
}

int failures = 0;

void expect(bool ok, const char *what, double err) {
    std::printf("%-58s rel L2 %.2e  %s\n", what, err, ok ? "ok" : "FAILED");
    if (!ok) ++failures;
}

#define CHECK_RC(call)                                                                   \
    do {                                                                                 \
        if ((call) != 0) {                                                               \
            std::printf("%s failed: %s\n", #call, svlg_last_error());                    \
            return 2;                                                                    \
        }                                                                                \
    } while (0)

int run_central_difference(bool support) {
    const int nt = 301;
    Layer L;
    const double vp = 200.0 * std::sqrt(3.0);
    const double dt = 0.5 / vp;
    make_layer(L, 24, 12, nt, dt, false, support);
    std::vector<int32_t> rec, base, rec_or;
    recorded(L, 7, rec, base, rec_or);
    const int nrec = (int)rec.size(), nall = (int)rec_or.size(), ns = nt - 1;
    std::vector<double> ref((size_t)ns * nall), ref_free((size_t)ns * nrec);
    oracle_opts o{0, nt, dt, 1e-12, 1e-12, 1e-12, 0, 20, 5, 0, 1e-6};
    if (oracle_run(&L.d, &o, nall, rec_or.data(), ref.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                   nullptr, nullptr) != 0) {
        std::printf("oracle_run failed: %s\n", oracle_last_error());
        return 2;
    }
    for (int s = 0; s < ns; ++s)
        for (int r = 0; r < nrec; ++r) ref_free[(size_t)s * nrec + r] = ref[(size_t)s * nall + r];
    svlg_model *m = nullptr;
    CHECK_RC(svlg_model_create(&L.d, &m));
    // B: fused steps.
    std::vector<double> hb((size_t)ns * nrec);
    CHECK_RC(svlg_cd_init(m, dt, 1e-12));
    CHECK_RC(svlg_cd_steps(m, 1, ns, nrec, rec.data(), hb.data()));
    expect(rel_l2(hb, ref_free) < 1e-10,
           support ? "CentralDifference svlg_cd_steps vs oracle (support motion)"
                   : "CentralDifference svlg_cd_steps vs oracle",
           rel_l2(hb, ref_free));
    // C: the Integrator methods one by one.
    std::vector<double> hc((size_t)ns * nall), feff(L.nfree), kd(L.nfree), du(L.nfree);
    CHECK_RC(svlg_cd_init(m, dt, 1e-12));
    CHECK_RC(svlg_cd_effective_stiffness(m, kd.data()));
    for (int k = 1; k < nt; ++k) {
        CHECK_RC(svlg_cd_effective_force(m, k, 1.0, feff.data()));
        CHECK_RC(svlg_support_motion_vector(m, k, 1.0, feff.data()));
        for (int f = 0; f < L.nfree; ++f) du[f] = feff[f] / kd[f];
        CHECK_RC(svlg_cd_apply_increment(m, k, du.data()));
        CHECK_RC(gather(m, L, rec, base, false, &hc[(size_t)(k - 1) * nall]));
    }
    expect(rel_l2(hc, ref) < 1e-10,
           support ? "CentralDifference Integrator methods vs oracle (support)"
                   : "CentralDifference Integrator methods vs oracle",
           rel_l2(hc, ref));
    if (!support) {
        // Progressive force: the point load of step k at the loaded node.
        std::vector<double> F((size_t)L.nn * 2);
        CHECK_RC(svlg_external_force(m, 20, F.data()));
        const double want = -L.load_series[20];
        const double got = F[(size_t)L.load_node[0] * 2 + 1];
        expect(got == want, "svlg_external_force at the loaded node (exact)", std::fabs(got - want));
    }
    CHECK_RC(svlg_model_destroy(m));
    return 0;
}

int run_newmark(bool support) {
    const int nt = 121;
    const double dt = 2e-3;
    Layer L;
    make_layer(L, 20, 10, nt, dt, true, support);
    std::vector<int32_t> rec, base, rec_or;
    recorded(L, 5, rec, base, rec_or);
    const int nrec = (int)rec.size(), nall = (int)rec_or.size(), ns = nt - 1;
    std::vector<double> ref((size_t)ns * nall), ref_free((size_t)ns * nrec);
    oracle_opts o{1, nt, dt, 1e-12, 1e-12, 1e-12, 0, 20, 5, 0, 1e-6};
    for (int lumped : {1, 0}) {
        L.d.lumped = lumped;
        if (oracle_run(&L.d, &o, nall, rec_or.data(), ref.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       0, nullptr, nullptr) != 0) {
            std::printf("oracle_run failed: %s\n", oracle_last_error());
            return 2;
        }
        for (int s = 0; s < ns; ++s)
            for (int r = 0; r < nrec; ++r) ref_free[(size_t)s * nrec + r] = ref[(size_t)s * nall + r];
        char what[96];
        svlg_model *m = nullptr;
        CHECK_RC(svlg_model_create(&L.d, &m));
        std::vector<double> hb((size_t)ns * nrec);
        CHECK_RC(svlg_nm_init(m, dt, 1e-13, 5000, SVLG_PRECOND_JACOBI));
        CHECK_RC(svlg_nm_steps(m, 1, ns, nrec, rec.data(), hb.data(), nullptr, nullptr));
        std::snprintf(what, sizeof(what), "ExtendedNewmarkBeta svlg_nm_steps vs oracle (%s%s)",
                      lumped ? "lumped" : "consistent", support ? ", support" : "");
        expect(rel_l2(hb, ref_free) < 1e-10, what, rel_l2(hb, ref_free));
        std::vector<double> hc((size_t)ns * nall), feff(L.nfree), du(L.nfree);
        CHECK_RC(svlg_nm_init(m, dt, 1e-13, 5000, SVLG_PRECOND_JACOBI));
        for (int k = 1; k < nt; ++k) {
            int32_t it = 0;
            double rr = 0.0;
            CHECK_RC(svlg_nm_effective_force(m, k, 1.0, feff.data()));
            CHECK_RC(svlg_support_motion_vector(m, k, 1.0, feff.data()));
            CHECK_RC(svlg_solve(m, feff.data(), du.data(), 1e-13, 5000, &it, &rr));
            CHECK_RC(svlg_nm_apply_increment(m, k, du.data()));
            CHECK_RC(gather(m, L, rec, base, true, &hc[(size_t)(k - 1) * nall]));
        }
        std::snprintf(what, sizeof(what), "ExtendedNewmarkBeta Integrator methods vs oracle (%s%s)",
                      lumped ? "lumped" : "consistent", support ? ", support" : "");
        expect(rel_l2(hc, ref) < 1e-10, what, rel_l2(hc, ref));
        CHECK_RC(svlg_model_destroy(m));
    }
    return 0;
}

}  // namespace

int main() {
    std::printf("libsvlg %s, %d HIP device(s)\n", svlg_version(), svlg_device_count());
    if (svlg_device_count() < 1) {
        std::printf("no HIP device\n");
        return 3;
    }
    for (bool support : {false, true}) {
        if (int rc = run_central_difference(support)) return rc;
        if (int rc = run_newmark(support)) return rc;
    }
    std::printf("%s\n", failures ? "FAILED" : "all checks passed");
    return failures ? 1 : 0;
}
