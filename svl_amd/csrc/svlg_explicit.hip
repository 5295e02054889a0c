// svlg_explicit.hip -- CentralDifference with lumped mass, matrix-free, on the device.
//
// Replaces CentralDifference::Initialize / ComputeEffectiveForce / ComputeEffectiveStiffness
// / ComputeNewStep (RP/10-Integrators/02-CentralDifference/CentralDifference.cpp:36-71,
// 124-152, 206-231) driven by Linear::ComputeNewIncrement (RP/09-Algorithms/01-Linear/
// Linear.cpp:22-56) inside DynamicAnalysis::Analyze (RP/08-Analysis/02-Dynamic/
// DynamicAnalysis.cpp:25-64). With lumped mass and dashpot-only damping Keff is
// diagonal, so EigenSolver's LDL^T solve is the division dU_i = Feff_i / Keff_ii.
//
// Per step k (load sample k, displacements U_{k-1}, U_{k-2}):
//   Feff = (Fext(k) - Fint(U_{k-1})) + (M/dt^2 - C/2dt)(U_{k-1} - U_{k-2})
//   U_k  = U_{k-1} + Feff / (M/dt^2 + C/2dt)
// The three displacement vectors rotate through a ring so V and A (which need
// U_{k-2}) can be formed on demand with SVL's expressions.
#include <algorithm>
#include <cstring>

#include "svlg_common.hpp"
#include "svlg_drm.hpp"
#include "svlg_support.hpp"
#include "svlg_elem.hpp"
#include "svlg_mfma.hpp"

namespace svlg {

int launch_fint_colored(Model &m, const double *U, double ftol, double *F);
int gather_free(Model &m, const double *dnode, double *dfree);
int record_positions(const Model &m, int nrec, const int32_t *rec_dofs, std::vector<int32_t> &pos);
int scatter_free(Model &m, const double *dfree, double *dnode);

// keff = M/dt^2 + C/2dt, kforce = M/dt^2 - C/2dt per DOF (CentralDifference.cpp:64-70); dtv
// (per DOF, may be null) gives every DOF its own time step (independent batched meshes).
__global__ void k_cd_coeffs(int n, double dt0, const double *__restrict__ dtv, const double *__restrict__ M,
                            const double *__restrict__ C, const int32_t *__restrict__ fdof,
                            double *__restrict__ keff, double *__restrict__ kforce, double *__restrict__ kinv,
                            int *__restrict__ singular) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fdof[i] < 0) {
        keff[i] = 0.0;
        kforce[i] = 0.0;
        kinv[i] = 0.0;
        return;
    }
    const double dt = dtv ? dtv[i] : dt0;
    double k = 1.0 / dt / dt * M[i] + 1.0 / 2.0 / dt * C[i];
    keff[i] = k;
    kforce[i] = 1.0 / dt / dt * M[i] - 1.0 / 2.0 / dt * C[i];
    kinv[i] = k != 0.0 ? 1.0 / k : 0.0;
    if (!(k != 0.0)) atomicOr(singular, 1);
}

// Fext(k) = sum_l factor_l * (amp_l[k] * dir_l) by assignment per load, then summed in
// load order (Assembler.cpp:334-350, 485). One thread keeps that order.
__global__ void k_loads(int k, int nload, int nt, int ndim, int nldofs, const int32_t *__restrict__ ldofs,
                        const int32_t *__restrict__ node, const double *__restrict__ dir,
                        const double *__restrict__ factor, const double *__restrict__ series,
                        double *__restrict__ Fext) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = 0; i < nldofs; ++i) Fext[ldofs[i]] = 0.0;
    for (int l = 0; l < nload; ++l) {
        double amp = (k >= 0 && k < nt) ? series[size_t(l) * nt + k] : 0.0;
        for (int c = 0; c < ndim; ++c) Fext[size_t(node[l]) * ndim + c] += factor[l] * (amp * dir[size_t(l) * ndim + c]);
    }
}

// U_next = U + (((Fext - F) + kforce (U - Up)) / keff); resets F for the next step.
__global__ void k_cd_update(int n, const double *__restrict__ U, const double *__restrict__ Up,
                            double *__restrict__ Unext, double *__restrict__ F, const double *__restrict__ Fext,
                            const double *__restrict__ keff, const double *__restrict__ kforce) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double k = keff[i];
    double fi = F[i];
    F[i] = 0.0;
    if (k == 0.0) {
        Unext[i] = 0.0;
        return;
    }
    double u = U[i];
    double feff = (Fext[i] - fi) + kforce[i] * (u - Up[i]);
    Unext[i] = u + feff / k;
}

// Feff = (Fext - F) + kforce (U - Up) on every node DOF (0 where keff = 0, the restrained
// DOFs); resets F for the next step (CentralDifference::ComputeEffectiveForce, :206-221).
__global__ void k_cd_feff(int n, const double *__restrict__ U, const double *__restrict__ Up, double *__restrict__ F,
                          const double *__restrict__ Fext, const double *__restrict__ keff,
                          const double *__restrict__ kforce, double *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double fi = F[i];
    F[i] = 0.0;
    out[i] = keff[i] == 0.0 ? 0.0 : (Fext[i] - fi) + kforce[i] * (U[i] - Up[i]);
}

// U_next = U + dU on the free DOFs, 0 on the restrained ones (CentralDifference::ComputeNewStep,
// :137-148, with the algorithm's increment; support motion is added by its own kernel).
__global__ void k_cd_increment(int n, const double *__restrict__ U, const double *__restrict__ dU,
                               const double *__restrict__ keff, double *__restrict__ Unext) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Unext[i] = keff[i] == 0.0 ? 0.0 : U[i] + dU[i];
}

__global__ void k_record(int nrec, const int32_t *__restrict__ rec, const double *__restrict__ U,
                         double *__restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec) out[r] = U[rec[r]];
}

// Arguments of the fused step kernel.
struct FusedArgs {
    const int32_t *own_ptr, *nplain, *halo_ptr, *halo_node, *elem_ptr, *color_ptr;
    const uint16_t *lconn;
    const double *elam, *emu;
    const uint16_t *egeo;
    const double *geo;
    const double *X, *Ucur, *Uprev;
    double *Unext;
    const double *keff, *kforce, *kinv;
    const int32_t *nl_ptr, *nl_idx;
    const double *load_dir, *load_factor, *load_series;
    int load_nt, k;
    double th, ftol;
    int ntiles, ncolors, max_stage, max_own;
    const float *kmat;  // PATH 4: K_lambda, K_mu of the single element geometry
    const int *kdev;    // non-null: load sample = *kdev + k (graph replays advance *kdev)
    DrmRef drm;         // domain reduction forces of every step (F null: none)
    // Corner-slot step: fixed-stride tile data (TileSet::hpad ...; hstride 0: not built).
    const int32_t *hpad;
    const uint16_t *lcpad, *geopad;
    const double *lampad, *mupad;
    int hstride, estride;
};

__device__ __forceinline__ int color_of(const int32_t *__restrict__ cp, int i) {
    int c = 0;
    while (cp[c + 1] <= i) ++c;
    return c;
}

// One CentralDifference step for one node tile (one workgroup per tile):
//  1. stage U_{k-1} and X of owned + halo nodes in LDS;
//  2. every element touching an owned node computes its internal force (affine modal
//     form or the Gauss-point loop) and adds the owned corners into LDS, one element
//     color at a time (no two elements of a color share a node: race-free and the
//     summation order is fixed);
//  3. owned DOFs: U_k = U_{k-1} + ((Fext - Fint) + kforce (U_{k-1} - U_{k-2})) / keff,
//     with kforce == keff and Fext == 0 for the plain nodes at the head of the range.
// Workgroups b and b+8 run on the same XCD; tiles are handed out so that each XCD
// works on a contiguous range of neighbouring tiles (shared halo lines in its L2).
template <int DIM, int PATH>
__global__ __launch_bounds__(512) void k_cd_fused(FusedArgs a) {
    constexpr int NEN = Elem<DIM>::nen;
    constexpr int NB = 4;  // loads kept in flight per thread in the staging and update loops
    extern __shared__ double smem[];
    __shared__ int32_t cps[65];
    const int nt = a.ntiles;
    const int kstep = a.kdev ? *a.kdev + a.k : a.k;
    const int b = blockIdx.x, xcd = b & 7, j = b >> 3;
    const int q = nt >> 3, r = nt & 7;
    const int tile = xcd * q + min(xcd, r) + j;
    const int o0 = a.own_ptr[tile], nown = a.own_ptr[tile + 1] - o0;
    const int h0 = a.halo_ptr[tile], nh = a.halo_ptr[tile + 1] - h0;
    const int e0 = a.elem_ptr[tile], ne = a.elem_ptr[tile + 1] - e0;
    constexpr bool STAGE_X = PATH < 2;
    double *sU = smem;
    double *sX = smem + size_t(a.max_stage) * DIM;
    double *sF = smem + size_t((STAGE_X ? 2 : 1) * a.max_stage) * DIM;
    const int tid = threadIdx.x, B = blockDim.x;
    const int nU = nown * DIM, nH = nh * DIM;
    for (int c = tid; c <= a.ncolors; c += B) cps[c] = a.color_ptr[size_t(tile) * (a.ncolors + 1) + c];
    // Stage owned (contiguous) and halo (gathered) U, and X when needed, NB loads in flight.
    {
        const double *Ug = a.Ucur + size_t(o0) * DIM;
        const double *Xg = a.X + size_t(o0) * DIM;
        for (int base = 0; base < nU; base += NB * B) {
            double v[NB], w[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int i = base + k * B + tid;
                v[k] = i < nU ? Ug[i] : 0.0;
                if constexpr (STAGE_X) w[k] = i < nU ? Xg[i] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int i = base + k * B + tid;
                if (i < nU) {
                    sU[i] = v[k];
                    if constexpr (STAGE_X) sX[i] = w[k];
                    sF[i] = 0.0;
                }
            }
        }
        // Halo: one index load per node, then its DIM components.
        for (int base = 0; base < nh; base += NB * B) {
            size_t n[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int sl = base + k * B + tid;
                n[k] = sl < nh ? size_t(a.halo_node[h0 + sl]) * DIM : 0;
            }
            double v[NB][DIM], w[NB][DIM];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int sl = base + k * B + tid;
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    v[k][c] = sl < nh ? a.Ucur[n[k] + c] : 0.0;
                    if constexpr (STAGE_X) w[k][c] = sl < nh ? a.X[n[k] + c] : 0.0;
                }
            }
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int sl = base + k * B + tid;
                if (sl < nh) {
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        sU[nU + sl * DIM + c] = v[k][c];
                        if constexpr (STAGE_X) sX[nU + sl * DIM + c] = w[k][c];
                    }
                }
            }
        }
    }
    __syncthreads();
    const int32_t *cp = cps;
    for (int base = 0; base < ne; base += B) {
        const int i = base + tid;
        const bool valid = i < ne;
        double f[NEN][DIM];
        int lc[NEN];
        int mycol = -1;
        if (valid) {
            // Corner slots in one 8- or 16-byte load.
            if constexpr (NEN == 8) {
                const uint4 pk = *reinterpret_cast<const uint4 *>(a.lconn + size_t(e0 + i) * 8);
                const uint32_t w4[4] = {pk.x, pk.y, pk.z, pk.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    lc[2 * c] = int(w4[c] & 0xffffu);
                    lc[2 * c + 1] = int(w4[c] >> 16);
                }
            } else {
                const uint2 pk = *reinterpret_cast<const uint2 *>(a.lconn + size_t(e0 + i) * 4);
                lc[0] = int(pk.x & 0xffffu);
                lc[1] = int(pk.x >> 16);
                lc[2] = int(pk.y & 0xffffu);
                lc[3] = int(pk.y >> 16);
            }
            const double lam = a.elam[e0 + i], mu = a.emu[e0 + i];
            double u[NEN][DIM];
#pragma unroll
            for (int c = 0; c < NEN; ++c)
#pragma unroll
                for (int d = 0; d < DIM; ++d) u[c][d] = sU[lc[c] * DIM + d];
            if constexpr (PATH >= 2) {
                const double *g = a.geo + size_t(a.egeo[e0 + i]) * (DIM * DIM);
                double Lam[DIM][DIM];
#pragma unroll
                for (int r2 = 0; r2 < DIM; ++r2)
#pragma unroll
                    for (int c2 = 0; c2 < DIM; ++c2) Lam[r2][c2] = g[r2 * DIM + c2];
                if constexpr (DIM == 3)
                    hexa_modal_force<PATH == 3>(Lam, lam, mu, u, f);
                else
                    quad_modal_force<PATH == 3>(Lam, lam, mu, u, f);
            } else {
                double x[NEN][DIM];
#pragma unroll
                for (int c = 0; c < NEN; ++c)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) x[c][d] = sX[lc[c] * DIM + d];
                if constexpr (DIM == 3) {
                    if constexpr (PATH == 1)
                        hexa_affine_fint(x, u, lam, mu, f);
                    else
                        hexa_fint(x, u, lam, mu, f);
                } else {
                    if constexpr (PATH == 1)
                        quad_affine_fint(x, u, lam, mu, a.th, f);
                    else
                        quad_fint(x, u, lam, mu, a.th, f);
                }
            }
            mycol = color_of(cp, i);
        }
        const int clo = color_of(cp, base);
        const int chi = color_of(cp, min(base + B, ne) - 1);
        for (int c = clo; c <= chi; ++c) {
            __syncthreads();
            if (mycol == c) {
#pragma unroll
                for (int q2 = 0; q2 < NEN; ++q2)
                    if (lc[q2] < nown) {
#pragma unroll
                        for (int d = 0; d < DIM; ++d) {
                            double v = f[q2][d];
                            if (fabs(v) > a.ftol) sF[lc[q2] * DIM + d] += v;
                        }
                    }
            }
        }
    }
    // Owned DOFs: load U_{k-2} and keff with NB loads in flight, then update.
    const int npl = a.nplain[tile];
    const size_t g0 = size_t(o0) * DIM;
    for (int base = 0; base < nU; base += NB * B) {
        double up[NB], kd[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int i = base + k * B + tid;
            up[k] = i < nU ? a.Uprev[g0 + i] : 0.0;
            kd[k] = i < nU ? a.keff[g0 + i] : 0.0;
        }
        if (base == 0) __syncthreads();
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int i = base + k * B + tid;
            if (i >= nU) continue;
            const int sl = i / DIM;
            const double u = sU[i], fi = sF[i];
            if (sl < npl) {
                a.Unext[g0 + i] = u + ((0.0 - fi) + kd[k] * (u - up[k])) / kd[k];
            } else if (kd[k] == 0.0) {
                a.Unext[g0 + i] = 0.0;
            } else {
                const int c = i - sl * DIM, node = o0 + sl;
                double fe = 0.0;
                for (int q2 = a.nl_ptr[node]; q2 < a.nl_ptr[node + 1]; ++q2) {
                    const int l = a.nl_idx[q2];
                    const double amp =
                        (kstep >= 0 && kstep < a.load_nt) ? a.load_series[size_t(l) * a.load_nt + kstep] : 0.0;
                    fe += a.load_factor[l] * (amp * a.load_dir[size_t(l) * DIM + c]);
                }
                const int ds = drm_slot(a.drm, node);
                if (ds >= 0) fe += drm_force(a.drm, ds, kstep, DIM, c);
                a.Unext[g0 + i] = u + ((fe - fi) + a.kforce[g0 + i] * (u - up[k])) / kd[k];
            }
        }
    }
}

// Corner-slot variant for meshes in which every (node, corner position) pair belongs
// to at most one element (conforming structured meshes): one thread per tile element
// and one thread per owned node, all tile elements in a single round. Each element
// writes its corner forces to the unique LDS slot (corner a, component c, node); each
// owned node then adds its NEN slots in corner order. No color phases; the corners go
// through the slot planes in NP passes of NEN / NP corners (same summation order as one
// pass, 1/NP of the slot LDS, one barrier per pass).
// MOC > 0: every tile owns at most MOC nodes; the slot planes then have the compile-time
// stride MOC + 1 with column MOC as the trash column, so that every slot address is one
// per-corner base plus an immediate offset.
// LATE: the update operands of a node (U_{k-2}, 1/keff) are loaded after the element
// phase instead of before the first barrier: 14 fewer registers live through the element
// step, more resident waves (on the Hexa8 table path, 86 -> 72 VGPRs).
// The update multiplies by the stored 1/keff (kinv) instead of dividing by keff: the
// diagonal solve as Eigen's LDL^T applies it (the inverse of D times the vector).
template <int DIM, int PATH, int MOC, bool LATE, int NP, bool PAD>
__device__ __forceinline__ void cd_corner_body(const FusedArgs &a) {
    constexpr int NEN = Elem<DIM>::nen;
    // Corner-slot planes per pass: the MFMA element step writes all corners at once.
    constexpr int NPASS = PATH == 4 ? 1 : NP, NH = NEN / NPASS;
    extern __shared__ double smem[];
    const int nt = a.ntiles;
    const int kstep = a.kdev ? *a.kdev + a.k : a.k;
    const int b = blockIdx.x, xcd = b & 7, j = b >> 3;
    const int q = nt >> 3, r = nt & 7;
    const int tile = xcd * q + min(xcd, r) + j;
    constexpr bool STAGE_X = PATH < 2;
    const int tid = threadIdx.x, B = blockDim.x;
    // Fixed-stride tile data: the halo node indices and the element data are loaded with
    // the tile index alone, in the same round trip as the tile bounds below.
    constexpr int NS = 2;  // strides of the one-pass staging
    // PAD: the fixed-stride copies exist (the host selects this instantiation only then).
    constexpr bool pad = PAD && !STAGE_X && PATH != 4;
    int hn[NS];  // halo node indices (raw: the address is formed where it is used)
    int lc[NEN];
    double lam = 0.0, mu = 0.0;
    int gid = 0;
    // Unconditional, in-bounds loads (indices clamped to the last entry; what lies past the
    // tile's counts is not used), unpacked only after the staging loads are issued.
    uint32_t pkw[NEN / 2];
    if constexpr (pad) {
        const int32_t *hp = a.hpad + size_t(tile) * a.hstride;
#pragma unroll
        for (int k = 0; k < NS; ++k) hn[k] = hp[min(k * B + tid, a.hstride - 1)];
        const size_t ei = size_t(tile) * a.estride + min(tid, a.estride - 1);
        if constexpr (NEN == 8) {
            const uint4 pk = *reinterpret_cast<const uint4 *>(a.lcpad + ei * 8);
            pkw[0] = pk.x;
            pkw[1] = pk.y;
            pkw[2] = pk.z;
            pkw[3] = pk.w;
        } else {
            const uint2 pk = *reinterpret_cast<const uint2 *>(a.lcpad + ei * 4);
            pkw[0] = pk.x;
            pkw[1] = pk.y;
        }
        lam = a.lampad[ei];
        mu = a.mupad[ei];
        if constexpr (PATH >= 2) gid = a.geopad[ei];
    }
    const int o0 = a.own_ptr[tile], nown = a.own_ptr[tile + 1] - o0;
    const int h0 = a.halo_ptr[tile], nh = a.halo_ptr[tile + 1] - h0;
    const int e0 = a.elem_ptr[tile], ne = a.elem_ptr[tile + 1] - e0;
    const int npl = a.nplain[tile];  // read with the tile bounds (one scalar round trip)
    const int MO = MOC > 0 ? MOC + 1 : a.max_own;
    double *sU = smem;
    double *sX = smem + size_t(a.max_stage) * DIM;
    double *sIn = smem + size_t((STAGE_X ? 2 : 1) * a.max_stage) * DIM;  // [NH*DIM][MO]
    const int nU = nown * DIM, nH = nh * DIM;
    const size_t g0 = size_t(o0) * DIM;
    // Prefetch: update operands of this thread's node, element data of this thread's element.
    const bool has_node = tid < nown;
    double up[DIM], kd[DIM];
    if constexpr (!LATE)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            up[c] = has_node ? a.Uprev[g0 + size_t(tid) * DIM + c] : 0.0;
            kd[c] = has_node ? a.kinv[g0 + size_t(tid) * DIM + c] : 0.0;
        }
    const bool has_elem = PATH != 4 && tid < ne;
    if constexpr (!pad) {
        // (with PAD the element data are in flight and unpacked after the staging loads)
        if (has_elem) {
            if constexpr (NEN == 8) {
                const uint4 pk = *reinterpret_cast<const uint4 *>(a.lconn + size_t(e0 + tid) * 8);
                const uint32_t w4[4] = {pk.x, pk.y, pk.z, pk.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    lc[2 * c] = int(w4[c] & 0xffffu);
                    lc[2 * c + 1] = int(w4[c] >> 16);
                }
            } else {
                const uint2 pk = *reinterpret_cast<const uint2 *>(a.lconn + size_t(e0 + tid) * 4);
                lc[0] = int(pk.x & 0xffffu);
                lc[1] = int(pk.x >> 16);
                lc[2] = int(pk.y & 0xffffu);
                lc[3] = int(pk.y >> 16);
            }
            lam = a.elam[e0 + tid];
            mu = a.emu[e0 + tid];
            if constexpr (PATH >= 2) gid = a.egeo[e0 + tid];
        } else {
#pragma unroll
            for (int c = 0; c < NEN; ++c) lc[c] = 0;
        }
    }
    // Stage U (and X), zero the corner slots.
    {
        const double *Ug = a.Ucur + g0;
        if constexpr (pad) {
            // One pass (the host checked that two strides cover the owned DOFs and the
            // halo): owned loads with the halo loads through the prefetched indices.
            double v[NS], hv[NS][DIM];
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const int i = k * B + tid;
                v[k] = i < nU ? Ug[i] : 0.0;
#pragma unroll
                for (int c = 0; c < DIM; ++c) hv[k][c] = i < nh ? a.Ucur[size_t(hn[k]) * DIM + c] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const int i = k * B + tid;
                if (i < nU) sU[i] = v[k];
                if (i < nh)
#pragma unroll
                    for (int c = 0; c < DIM; ++c) sU[nU + i * DIM + c] = hv[k][c];
            }
#pragma unroll
            for (int c = 0; c < NEN / 2; ++c) {
                lc[2 * c] = int(pkw[c] & 0xffffu);
                lc[2 * c + 1] = int(pkw[c] >> 16);
            }
        } else if constexpr (!STAGE_X) {
            // One loop covers the owned DOFs and the halo nodes: every pass issues the halo
            // index loads and the owned loads together, then the halo loads, then the stores —
            // two dependent global round trips per pass instead of three.
            const int nrange = nU > nh ? nU : nh;
            for (int base = 0; base < nrange; base += NS * B) {
                size_t n[NS];
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const int sl = base + k * B + tid;
                    n[k] = sl < nh ? size_t(a.halo_node[h0 + sl]) * DIM : 0;
                }
                double v[NS];
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const int i = base + k * B + tid;
                    v[k] = i < nU ? Ug[i] : 0.0;
                }
                double hv[NS][DIM];
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const int sl = base + k * B + tid;
#pragma unroll
                    for (int c = 0; c < DIM; ++c) hv[k][c] = sl < nh ? a.Ucur[n[k] + c] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const int i = base + k * B + tid;
                    if (i < nU) sU[i] = v[k];
                    if (i < nh)
#pragma unroll
                        for (int c = 0; c < DIM; ++c) sU[nU + i * DIM + c] = hv[k][c];
                }
            }
        } else {
            // Coordinates staged too (Gauss-point and affine paths): owned values, then the
            // halo (one index load per node, then its DIM components).
            constexpr int NB = 4;  // loads kept in flight per thread
            const double *Xg = a.X + g0;
            for (int base = 0; base < nU; base += NB * B) {
                double v[NB], w[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int i = base + k * B + tid;
                    v[k] = i < nU ? Ug[i] : 0.0;
                    w[k] = i < nU ? Xg[i] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int i = base + k * B + tid;
                    if (i < nU) {
                        sU[i] = v[k];
                        sX[i] = w[k];
                    }
                }
            }
            for (int base = 0; base < nh; base += NB * B) {
                size_t n[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int sl = base + k * B + tid;
                    n[k] = sl < nh ? size_t(a.halo_node[h0 + sl]) * DIM : 0;
                }
                double v[NB][DIM], w[NB][DIM];
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int sl = base + k * B + tid;
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        v[k][c] = sl < nh ? a.Ucur[n[k] + c] : 0.0;
                        w[k][c] = sl < nh ? a.X[n[k] + c] : 0.0;
                    }
                }
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int sl = base + k * B + tid;
                    if (sl < nh) {
#pragma unroll
                        for (int c = 0; c < DIM; ++c) {
                            sU[nU + sl * DIM + c] = v[k][c];
                            sX[nU + sl * DIM + c] = w[k][c];
                        }
                    }
                }
            }
        }
        for (int i = tid; i < NH * DIM * MO; i += B) sIn[i] = 0.0;
    }
    __syncthreads();
    double fi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) fi[d] = 0.0;
    if constexpr (PATH == 4) {
        // BF16 MFMA element step (single element geometry), FP64 slots and update.
        bf16x8 Ak[3];
        mfma_load_kblocks(a.kmat, Ak);
        mfma_hexa_elements<double>(a.lconn, a.elam, a.emu, e0, ne, nown, sU, sIn, MO, Ak, a.ftol);
        __syncthreads();
        if (has_node)
#pragma unroll
            for (int d = 0; d < DIM; ++d)
#pragma unroll
                for (int c = 0; c < NEN; ++c) fi[d] += sIn[(c * DIM + d) * MO + tid];
    } else {
        double f[NEN][DIM];
        if (has_elem) {
            double u[NEN][DIM];
#pragma unroll
            for (int c = 0; c < NEN; ++c)
#pragma unroll
                for (int d = 0; d < DIM; ++d) u[c][d] = sU[lc[c] * DIM + d];
            if constexpr (PATH == 5) {
                // FP32 element step (axis-aligned bricks): the modal force in single precision
                // from the FP64 staged displacements; slots, sums and the update stay FP64.
                static_assert(DIM == 3, "the FP32 element step is a Hexa8 path");
                const double *g = a.geo + size_t(gid) * 9;
                float Lf[3][3], uf[8][3], ff[8][3];
#pragma unroll
                for (int r2 = 0; r2 < 3; ++r2)
#pragma unroll
                    for (int c2 = 0; c2 < 3; ++c2) Lf[r2][c2] = float(g[r2 * 3 + c2]);
#pragma unroll
                for (int c = 0; c < 8; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) uf[c][d] = float(u[c][d]);
                hexa_modal_force<true, float>(Lf, float(lam), float(mu), uf, ff);
#pragma unroll
                for (int c = 0; c < 8; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) f[c][d] = double(ff[c][d]);
            } else if constexpr (PATH >= 2) {
                const double *g = a.geo + size_t(gid) * (DIM * DIM);
                double Lam[DIM][DIM];
#pragma unroll
                for (int r2 = 0; r2 < DIM; ++r2)
#pragma unroll
                    for (int c2 = 0; c2 < DIM; ++c2) Lam[r2][c2] = g[r2 * DIM + c2];
                if constexpr (DIM == 3)
                    hexa_modal_force<PATH == 3>(Lam, lam, mu, u, f);
                else
                    quad_modal_force<PATH == 3>(Lam, lam, mu, u, f);
            } else {
                double x[NEN][DIM];
#pragma unroll
                for (int c = 0; c < NEN; ++c)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) x[c][d] = sX[lc[c] * DIM + d];
                if constexpr (DIM == 3) {
                    if constexpr (PATH == 1)
                        hexa_affine_fint(x, u, lam, mu, f);
                    else
                        hexa_fint(x, u, lam, mu, f);
                } else {
                    if constexpr (PATH == 1)
                        quad_affine_fint(x, u, lam, mu, a.th, f);
                    else
                        quad_fint(x, u, lam, mu, a.th, f);
                }
            }
        }
        // NPASS passes of NH corners through NH slot planes: elements write corners
        // pass*NH .. pass*NH+NH-1 (branch-free: corners of nodes this tile does not own go to
        // the trash cell after the slots), then every owned node adds its planes in corner
        // order and clears them for the next pass. Same summation order as one pass over
        // all corners, 1/NPASS of the LDS.
        const int trash = NH * DIM * MO;
        // LATE: the owned node's update operands are loaded at the start of the middle pass,
        // once half of the element forces are written out (their registers are free), so
        // that they arrive during the remaining passes.
#pragma unroll
        for (int pass = 0; pass < NPASS; ++pass) {
            if (pass > 0) __syncthreads();
            if constexpr (LATE)
                if (pass == NPASS / 2 && has_node)
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        up[c] = a.Uprev[g0 + size_t(tid) * DIM + c];
                        kd[c] = a.kinv[g0 + size_t(tid) * DIM + c];
                    }
            if (has_elem)
#pragma unroll
                for (int c = 0; c < NH; ++c) {
                    const int cc = pass * NH + c;
                    const bool own = lc[cc] < nown;
                    const int col = own ? lc[cc] : MOC;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        const double v = f[cc][d];
                        const double fv = fabs(v) > a.ftol ? v : 0.0;
                        if constexpr (MOC > 0)
                            sIn[(c * DIM + d) * MO + col] = fv;
                        else
                            sIn[own ? (c * DIM + d) * MO + lc[cc] : trash] = fv;
                    }
                }
            __syncthreads();
            if (has_node)
#pragma unroll
                for (int d = 0; d < DIM; ++d)
#pragma unroll
                    for (int c = 0; c < NH; ++c) {
                        double &slot = sIn[(c * DIM + d) * MO + tid];
                        fi[d] += slot;
                        if (pass + 1 < NPASS) slot = 0.0;
                    }
        }
    }
    if (!has_node) return;
    if constexpr (LATE && PATH == 4)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            up[c] = a.Uprev[g0 + size_t(tid) * DIM + c];
            kd[c] = a.kinv[g0 + size_t(tid) * DIM + c];
        }
    const int ds = tid >= npl ? drm_slot(a.drm, o0 + tid) : -1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double u = sU[tid * DIM + d];
        const size_t g = g0 + size_t(tid) * DIM + d;
        if (tid < npl) {
            // Plain nodes: kforce = keff (no damping), so Feff / keff = (u - up) - fi / keff.
            a.Unext[g] = u + ((u - up[d]) - fi[d] * kd[d]);
        } else if (kd[d] == 0.0) {
            a.Unext[g] = 0.0;
        } else {
            const int node = o0 + tid;
            double fe = 0.0;
            for (int q2 = a.nl_ptr[node]; q2 < a.nl_ptr[node + 1]; ++q2) {
                const int l = a.nl_idx[q2];
                const double amp =
                    (kstep >= 0 && kstep < a.load_nt) ? a.load_series[size_t(l) * a.load_nt + kstep] : 0.0;
                fe += a.load_factor[l] * (amp * a.load_dir[size_t(l) * DIM + d]);
            }
            if (ds >= 0) fe += drm_force(a.drm, ds, kstep, DIM, d);
            a.Unext[g] = u + ((fe - fi[d]) + a.kforce[g] * (u - up[d])) * kd[d];
        }
    }
}

template <int DIM, int PATH, int MOC = 0, bool LATE = false, int NP = 2, bool PAD = false>
__global__ __launch_bounds__(512) void k_cd_corner(FusedArgs a) {
    cd_corner_body<DIM, PATH, MOC, LATE, NP, PAD>(a);
}

template <int DIM, int PATH>
static const void *fused_fn(bool corner) {
    return corner ? reinterpret_cast<const void *>(&k_cd_corner<DIM, PATH>)
                  : reinterpret_cast<const void *>(&k_cd_fused<DIM, PATH>);
}

// Corner passes of the corner-slot step (the slot planes hold NEN / passes corners): four on
// the compile-time stride Hexa8 paths, two elsewhere (one for the MFMA element step).
static int use_moc(const Model &m);
static int cd_passes(const Model &m) {
    if (m.elem_bf16) return 1;
    return m.ndim == 3 && use_moc(m) ? 4 : 2;
}

// Owned-node bound of the compile-time slot stride of the corner-slot step on axis-aligned
// meshes (0: the stride read from the tile set): 160 (the default 3D tile size) for Hexa8,
// 192 and 448 (the batched and the default 2D tile sizes) for Quad4.
static int use_moc(const Model &m) {
    if (m.variant != 3 || m.elem_bf16 || m.elem_path != 3) return 0;
    if (m.elem_fp32) return m.tiles.max_own <= 160 ? 160 : 0;
    const int mo = m.tiles.max_own;
    if (m.ndim == 3) return mo <= 160 ? 160 : 0;
    return mo <= 192 ? 192 : mo <= 448 ? 448 : 0;
}

static const void *moc_kernel(const Model &m) {
    const bool pad = m.tiles.hstride > 0;  // fixed-stride tile data built
    if (m.elem_fp32)
        return pad ? reinterpret_cast<const void *>(&k_cd_corner<3, 5, 160, true, 4, true>)
                   : reinterpret_cast<const void *>(&k_cd_corner<3, 5, 160, true, 4>);
    switch (m.ndim * 1000 + use_moc(m)) {
        case 3160:
            return pad ? reinterpret_cast<const void *>(&k_cd_corner<3, 3, 160, true, 4, true>)
                       : reinterpret_cast<const void *>(&k_cd_corner<3, 3, 160, true, 4>);
        case 2192:
            return pad ? reinterpret_cast<const void *>(&k_cd_corner<2, 3, 192, false, 2, true>)
                       : reinterpret_cast<const void *>(&k_cd_corner<2, 3, 192>);
        case 2448:
            return pad ? reinterpret_cast<const void *>(&k_cd_corner<2, 3, 448, false, 2, true>)
                       : reinterpret_cast<const void *>(&k_cd_corner<2, 3, 448>);
        default: return nullptr;
    }
}

static const void *fused_kernel(const Model &m) {
    const bool cr = m.variant == 3;
    if (m.elem_bf16) return reinterpret_cast<const void *>(&k_cd_corner<3, 4>);
    if (use_moc(m)) return moc_kernel(m);
    if (m.elem_fp32) return reinterpret_cast<const void *>(&k_cd_corner<3, 5>);
    if (m.ndim == 3)
        switch (m.elem_path) {
            case 3: return fused_fn<3, 3>(cr);
            case 2: return fused_fn<3, 2>(cr);
            case 1: return fused_fn<3, 1>(cr);
            default: return fused_fn<3, 0>(cr);
        }
    switch (m.elem_path) {
        case 3: return fused_fn<2, 3>(cr);
        case 2: return fused_fn<2, 2>(cr);
        case 1: return fused_fn<2, 1>(cr);
        default: return fused_fn<2, 0>(cr);
    }
}

int prepare_fused(Model &m) {
    if (m.variant != 1 && m.variant != 3) return 0;
    if (m.variant == 3) {
        // Staged U (and X unless the geometry table is used), the corner-slot planes of one
        // pass (half the corners; all of them for the MFMA element step) and the trash cell.
        const int staged = m.elem_path >= 2 ? 1 : 2;
        const int planes = m.nen / cd_passes(m);
        m.tiles.lds_bytes =
            use_moc(m) ? (size_t(staged) * m.tiles.max_stage * m.ndim + size_t(planes) * m.ndim * (use_moc(m) + 1)) *
                             sizeof(double)
                       : (size_t(staged) * m.tiles.max_stage * m.ndim + size_t(planes) * m.ndim * m.tiles.max_own + 1) *
                             sizeof(double);
    }
    SVLG_CHECK(m.tiles.block <= 512, "fused kernel block must be at most 512 threads");
    SVLG_CHECK(m.tiles.lds_bytes <= 160 * 1024, "node tile needs more than 160 KB of LDS; use smaller tiles");
    if (m.tiles.lds_bytes > 64 * 1024)
        SVLG_HIP(hipFuncSetAttribute(fused_kernel(m), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(m.tiles.lds_bytes)));
    return 0;
}

static int launch_fused(Model &m, int k, const double *Ucur, const double *Uprev, double *Unext,
                        const int *kdev = nullptr) {
    const TileSet &t = m.tiles;
    FusedArgs a;
    a.kdev = kdev;
    a.own_ptr = t.own_ptr.p;
    a.nplain = t.nplain.p;
    a.halo_ptr = t.halo_ptr.p;
    a.halo_node = t.halo_node.p;
    a.elem_ptr = t.elem_ptr.p;
    a.color_ptr = t.color_ptr.p;
    a.lconn = t.elem_lconn.p;
    a.elam = t.elem_lam.p;
    a.emu = t.elem_mu.p;
    a.X = m.X.p;
    a.Ucur = Ucur;
    a.Uprev = Uprev;
    a.Unext = Unext;
    a.keff = m.keff.p;
    a.kforce = m.kforce.p;
    a.kinv = m.kinv.p;
    a.nl_ptr = m.nl_ptr.p;
    a.nl_idx = m.nl_idx.p;
    a.load_dir = m.load_dir.p;
    a.load_factor = m.load_factor.p;
    a.load_series = m.load_series.p;
    a.load_nt = m.load_nt;
    a.k = k;
    a.th = m.th;
    a.ftol = m.ftol;
    a.ntiles = t.ntiles;
    a.ncolors = t.ncolors;
    a.max_stage = t.max_stage;
    a.max_own = t.max_own;
    a.egeo = t.elem_geo.p;
    a.geo = m.geo.p;
    a.drm = drm_ref(m);
    a.hpad = t.hpad.p;
    a.lcpad = t.lcpad.p;
    a.geopad = t.geopad.p;
    a.lampad = t.lampad.p;
    a.mupad = t.mupad.p;
    a.hstride = t.hstride;
    a.estride = t.estride;
    dim3 grid(t.ntiles), block(t.block);
#define SVLG_LAUNCH_FUSED(D, P)                                                                        \
    do {                                                                                              \
        if (m.variant == 3)                                                                           \
            hipLaunchKernelGGL((k_cd_corner<D, P>), grid, block, t.lds_bytes, m.stream, a);          \
        else                                                                                          \
            hipLaunchKernelGGL((k_cd_fused<D, P>), grid, block, t.lds_bytes, m.stream, a);           \
    } while (0)
    a.kmat = m.cd_kmat.p;
    if (m.elem_bf16) {
        hipLaunchKernelGGL((k_cd_corner<3, 4>), grid, block, t.lds_bytes, m.stream, a);
    } else if (use_moc(m)) {
        void *args[] = {&a};
        SVLG_HIP(hipLaunchKernel(moc_kernel(m), grid, block, args, t.lds_bytes, m.stream));
    } else if (m.elem_fp32) {
        hipLaunchKernelGGL((k_cd_corner<3, 5>), grid, block, t.lds_bytes, m.stream, a);
    } else if (m.ndim == 3) {
        if (m.elem_path == 3) SVLG_LAUNCH_FUSED(3, 3);
        else if (m.elem_path == 2) SVLG_LAUNCH_FUSED(3, 2);
        else if (m.elem_path == 1) SVLG_LAUNCH_FUSED(3, 1);
        else SVLG_LAUNCH_FUSED(3, 0);
    } else {
        if (m.elem_path == 3) SVLG_LAUNCH_FUSED(2, 3);
        else if (m.elem_path == 2) SVLG_LAUNCH_FUSED(2, 2);
        else if (m.elem_path == 1) SVLG_LAUNCH_FUSED(2, 1);
        else SVLG_LAUNCH_FUSED(2, 0);
    }
#undef SVLG_LAUNCH_FUSED
    return 0;
}

static int cd_one_step(Model &m, int k) {
    double *Ucur = m.U[m.ring].p;
    double *Uprev = m.U[(m.ring + 2) % 3].p;
    double *Unext = m.U[(m.ring + 1) % 3].p;
    // Support motion: the internal force sees U_(k-2) at its DOFs (the strains lag one
    // increment); afterwards U_(k-1) returns and U_k = U_(k-1) + a_k there.
    if (int rc = sm_lag_in(m, Ucur, Uprev, nullptr)) return rc;
    if (m.variant == 1 || m.variant == 3) {
        if (int rc = launch_fused(m, k, Ucur, Uprev, Unext)) return rc;
        if (int rc = sm_finish(m, 1, k, Ucur, Unext)) return rc;
        m.ring = (m.ring + 1) % 3;
        m.steps_done++;
        return 0;
    }
    if (m.nload > 0)
        hipLaunchKernelGGL(k_loads, dim3(1), dim3(64), 0, m.stream, k, m.nload, m.load_nt, m.ndim,
                           (int)m.h_load_dofs.size(), m.load_dofs.p, m.load_node.p, m.load_dir.p, m.load_factor.p,
                           m.load_series.p, m.Fext.p);
    if (int rc = drm_add_fext(m, k, m.Fext.p)) return rc;
    if (int rc = launch_fint_colored(m, Ucur, m.ftol, m.F.p)) return rc;
    int nb = (m.ndofn + 255) / 256;
    hipLaunchKernelGGL(k_cd_update, dim3(nb), dim3(256), 0, m.stream, m.ndofn, Ucur, Uprev, Unext, m.F.p, m.Fext.p,
                       m.keff.p, m.kforce.p);
    if (int rc = sm_finish(m, 1, k, Ucur, Unext)) return rc;
    m.ring = (m.ring + 1) % 3;
    m.steps_done++;
    return 0;
}

// CentralDifference::Initialize (CentralDifference.cpp:36-71) with the per-DOF time steps
// dtd (device DOF order) or one dt.
static int cd_init_impl(Model &m, double dt, const std::vector<double> *dtd, double ftol) {
    SVLG_CHECK(dt > 0.0, "dt must be positive");
    SVLG_CHECK(m.lumped, "CentralDifference on the device needs MassFormulation = LUMPED");
    SVLG_CHECK(!m.has_ak, "CentralDifference on the device needs a diagonal C (Rayleigh ak must be 0)");
    SVLG_CHECK(!m.has_j2, "the explicit kernels restate elastic materials only; J2 models use svlg_nl_steps");
    m.dt = dt;
    m.ftol = ftol;
    m.cd_ready = false;
    if (m.cd_graph) {
        (void)hipGraphExecDestroy(m.cd_graph);
        m.cd_graph = nullptr;
        m.cd_graph_steps = 0;
        m.cd_graph_ring = -1;
    }
    for (auto &b : m.U) {
        SVLG_HIP(b.alloc(m.ndofn));
        SVLG_HIP(hipMemsetAsync(b.p, 0, sizeof(double) * m.ndofn, m.stream));
    }
    SVLG_HIP(m.F.alloc(m.ndofn));
    SVLG_HIP(hipMemsetAsync(m.F.p, 0, sizeof(double) * m.ndofn, m.stream));
    SVLG_HIP(m.Fext.alloc(m.ndofn));
    SVLG_HIP(hipMemsetAsync(m.Fext.p, 0, sizeof(double) * m.ndofn, m.stream));
    SVLG_HIP(m.keff.alloc(m.ndofn));
    SVLG_HIP(m.kforce.alloc(m.ndofn));
    SVLG_HIP(m.kinv.alloc(m.ndofn));
    if (dtd) {
        m.h_dt_dof = *dtd;
        SVLG_HIP(m.dt_dof.upload(dtd->data(), dtd->size(), m.stream));
    } else {
        m.h_dt_dof.clear();
    }
    DevBuf<int> sing;
    SVLG_HIP(sing.alloc(1));
    SVLG_HIP(hipMemsetAsync(sing.p, 0, sizeof(int), m.stream));
    int nb = (m.ndofn + 255) / 256;
    hipLaunchKernelGGL(k_cd_coeffs, dim3(nb), dim3(256), 0, m.stream, m.ndofn, dt, dtd ? m.dt_dof.p : nullptr,
                       m.mdiag.p, m.cdiag.p, m.fdof.p, m.keff.p, m.kforce.p, m.kinv.p, sing.p);
    SVLG_HIP(hipGetLastError());
    int hs = 0;
    SVLG_HIP(hipMemcpyAsync(&hs, sing.p, sizeof(int), hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    // EigenSolver::SolveSystem reports failure when the factorization breaks down
    // (EigenSolver.cpp:52-55): a free DOF with no mass and no damping.
    SVLG_CHECK(hs == 0, "EigenSolver::SolveSystem() : Solution for the linear system failed (zero Keff diagonal)");
    m.ring = 0;
    m.steps_done = 0;
    m.cd_ready = true;
    m.active = 1;
    return 0;
}

__global__ void k_step_advance(int *kdev, int by) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *kdev += by;
}

__global__ void k_step_set(int *kdev, int k) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *kdev = k;
}

// Captures cd_graph_steps fused launches (load sample *kdev + s for launch s) and the advance
// of *kdev into one executable graph, starting from the current U ring slot.
static int cd_capture(Model &m, int steps) {
    if (m.cd_graph) {
        (void)hipGraphExecDestroy(m.cd_graph);
        m.cd_graph = nullptr;
    }
    if (!m.kdev.p) SVLG_HIP(m.kdev.alloc(1));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    SVLG_HIP(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal));
    int ring = m.ring, rc = 0;
    for (int s = 0; s < steps && rc == 0; ++s) {
        double *Ucur = m.U[ring].p, *Uprev = m.U[(ring + 2) % 3].p, *Unext = m.U[(ring + 1) % 3].p;
        rc = sm_lag_in(m, Ucur, Uprev, nullptr);
        if (rc == 0) rc = launch_fused(m, s, Ucur, Uprev, Unext, m.kdev.p);
        if (rc == 0) rc = sm_finish(m, 1, s, Ucur, Unext, m.kdev.p);
        ring = (ring + 1) % 3;
    }
    if (rc == 0) hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, m.stream, m.kdev.p, steps);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(m.stream, &g);
    if (rc != 0) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    SVLG_HIP(ec);
    const hipError_t ei = hipGraphInstantiate(&m.cd_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SVLG_HIP(ei);
    m.cd_graph_steps = steps;
    m.cd_graph_ring = m.ring;
    return 0;
}

}  // namespace svlg

using namespace svlg;

extern "C" {

int svlg_cd_init(svlg_model *h, double dt, double ftol) {
    SVLG_CHECK(h, "null model");
    return cd_init_impl(h->m, dt, nullptr, ftol);
}

int svlg_cd_init_batch(svlg_model *h, int32_t nparts, const int32_t *part_node_ptr, const double *dts, double ftol) {
    SVLG_CHECK(h && part_node_ptr && dts, "null argument");
    Model &m = h->m;
    SVLG_CHECK(nparts >= 1, "nparts must be positive");
    SVLG_CHECK(part_node_ptr[0] == 0 && part_node_ptr[nparts] == m.nnodes, "part ranges must cover the nodes");
    for (int p = 0; p < nparts; ++p) {
        SVLG_CHECK(part_node_ptr[p + 1] >= part_node_ptr[p], "part ranges must be ascending");
        SVLG_CHECK(dts[p] > 0.0, "dt must be positive");
    }
    // Per device DOF: the time step of the part that holds the node (description order).
    std::vector<double> dtd(m.ndofn);
    int part = 0;
    std::vector<double> dt_node(m.nnodes);
    for (int n = 0; n < m.nnodes; ++n) {
        while (n >= part_node_ptr[part + 1]) ++part;
        dt_node[n] = dts[part];
    }
    for (int p = 0; p < m.nnodes; ++p)
        for (int c = 0; c < m.ndim; ++c) dtd[size_t(p) * m.ndim + c] = dt_node[m.perm[p]];
    return cd_init_impl(m, dts[0], &dtd, ftol);
}

int svlg_cd_steps_async(svlg_model *h, int32_t k0, int32_t nsteps) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.cd_ready, "svlg_cd_init has not been called");
    for (int s = 0; s < nsteps; ++s)
        if (int rc = cd_one_step(m, k0 + s)) return rc;
    SVLG_HIP(hipGetLastError());
    return 0;
}

int svlg_cd_graph_steps(svlg_model *h, int32_t k0, int32_t nsteps, int32_t graph_steps) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.cd_ready, "svlg_cd_init has not been called");
    SVLG_CHECK(nsteps >= 0, "negative count");
    SVLG_CHECK(m.variant == 1 || m.variant == 3, "graph stepping needs the fused kernel (variant 1 or 3)");
    SVLG_CHECK(graph_steps >= 3 && graph_steps % 3 == 0 && graph_steps <= 3000,
               "graph_steps must be a multiple of 3 in [3, 3000]");
    int done = 0;
    if (nsteps >= graph_steps) {
        if (!m.cd_graph || m.cd_graph_steps != graph_steps || m.cd_graph_ring != m.ring)
            if (int rc = cd_capture(m, graph_steps)) return rc;
        hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, m.stream, m.kdev.p, int(k0));
        for (; done + graph_steps <= nsteps; done += graph_steps) SVLG_HIP(hipGraphLaunch(m.cd_graph, m.stream));
        m.steps_done += done;  // the ring returns to its slot after every replay
    }
    for (; done < nsteps; ++done)
        if (int rc = cd_one_step(m, k0 + done)) return rc;
    SVLG_HIP(hipGetLastError());
    return 0;
}

int svlg_autotune_tiles(const svlg_mesh_desc *desc, const svlg_exec_opts *base, const int32_t *cand, int32_t ncand,
                        int32_t nparts, const int32_t *part_node_ptr, const double *dts, int32_t steps,
                        int32_t *best_tile_nodes, double *us_per_step) {
    SVLG_CHECK(desc && cand && dts && best_tile_nodes, "null argument");
    SVLG_CHECK(ncand >= 1 && steps >= 1 && nparts >= 1, "empty search");
    SVLG_CHECK(nparts == 1 || part_node_ptr, "part ranges missing");
    double best = -1.0;
    *best_tile_nodes = 0;
    for (int c = 0; c < ncand; ++c) {
        if (us_per_step) us_per_step[c] = -1.0;
        svlg_exec_opts o{};
        if (base) o = *base;
        o.tile_nodes = cand[c];
        svlg_model *h = nullptr;
        if (svlg_model_create_ex(desc, &o, &h) != 0) continue;  // e.g. the tile does not fit the LDS
        Model &m = h->m;
        int rc = nparts > 1 ? svlg_cd_init_batch(h, nparts, part_node_ptr, dts, 1e-12) : svlg_cd_init(h, dts[0], 1e-12);
        if (rc == 0 && (m.variant == 1 || m.variant == 3)) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
            for (int s = 0; ok && s < 3; ++s) ok = cd_one_step(m, 1 + s) == 0;
            ok = ok && hipEventRecord(e0, m.stream) == hipSuccess;
            for (int s = 0; ok && s < steps; ++s) ok = cd_one_step(m, 4 + s) == 0;
            ok = ok && hipEventRecord(e1, m.stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float ms = 0.0f;
            if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
                const double us = 1e3 * double(ms) / steps;
                if (us_per_step) us_per_step[c] = us;
                if (best < 0.0 || us < best) {
                    best = us;
                    *best_tile_nodes = cand[c];
                }
            }
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
        svlg_model_destroy(h);
    }
    SVLG_CHECK(best > 0.0, "no candidate tile size could be built and timed");
    return 0;
}

int svlg_cd_steps(svlg_model *h, int32_t k0, int32_t nsteps, int32_t nrec, const int32_t *rec_dofs, double *hist) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.cd_ready, "svlg_cd_init has not been called");
    SVLG_CHECK(nsteps >= 0 && nrec >= 0, "negative count");
    if (nrec > 0) {
        SVLG_CHECK(rec_dofs && hist, "recorder arrays missing");
        std::vector<int32_t> rn;
        if (int rc = record_positions(m, nrec, rec_dofs, rn)) return rc;
        SVLG_HIP(m.rec.upload(rn.data(), rn.size(), m.stream));
        SVLG_HIP(m.hist.alloc(size_t(nrec) * std::max(nsteps, 1)));
    }
    for (int s = 0; s < nsteps; ++s) {
        if (int rc = cd_one_step(m, k0 + s)) return rc;
        if (nrec > 0)
            hipLaunchKernelGGL(k_record, dim3((nrec + 63) / 64), dim3(64), 0, m.stream, nrec, m.rec.p, m.U[m.ring].p,
                               m.hist.p + size_t(s) * nrec);
    }
    SVLG_HIP(hipGetLastError());
    if (nrec > 0 && nsteps > 0)
        SVLG_HIP(hipMemcpyAsync(hist, m.hist.p, sizeof(double) * size_t(nrec) * nsteps, hipMemcpyDeviceToHost,
                                m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_cd_effective_force(svlg_model *h, int32_t k, double factor, double *Feff) {
    (void)factor;  // unused by CentralDifference::ComputeEffectiveForce as well
    SVLG_CHECK(h && Feff, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.cd_ready, "svlg_cd_init has not been called");
    double *Ucur = m.U[m.ring].p;
    const double *Up = m.U[(m.ring + 2) % 3].p;
    if (int rc = sm_lag_in(m, Ucur, Up, nullptr)) return rc;
    if (m.nload > 0)
        hipLaunchKernelGGL(k_loads, dim3(1), dim3(64), 0, m.stream, k, m.nload, m.load_nt, m.ndim,
                           (int)m.h_load_dofs.size(), m.load_dofs.p, m.load_node.p, m.load_dir.p, m.load_factor.p,
                           m.load_series.p, m.Fext.p);
    if (int rc = drm_add_fext(m, k, m.Fext.p)) return rc;
    if (int rc = launch_fint_colored(m, Ucur, m.ftol, m.F.p)) return rc;
    DevBuf<double> fn, ff;
    SVLG_HIP(fn.alloc(m.ndofn));
    SVLG_HIP(ff.alloc(std::max(m.nfree, 1)));
    const int nb = (m.ndofn + 255) / 256;
    hipLaunchKernelGGL(k_cd_feff, dim3(nb), dim3(256), 0, m.stream, m.ndofn, Ucur, Up, m.F.p, m.Fext.p, m.keff.p,
                       m.kforce.p, fn.p);
    if (int rc = sm_finish(m, 0, k, Ucur, nullptr)) return rc;
    if (int rc = gather_free(m, fn.p, ff.p)) return rc;
    SVLG_HIP(hipMemcpyAsync(Feff, ff.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_external_force(svlg_model *h, int32_t k, double *F) {
    SVLG_CHECK(h && F, "null argument");
    Model &m = h->m;
    DevBuf<double> fe;
    SVLG_HIP(fe.alloc(m.ndofn));
    SVLG_HIP(hipMemsetAsync(fe.p, 0, sizeof(double) * m.ndofn, m.stream));
    if (m.nload > 0)
        hipLaunchKernelGGL(k_loads, dim3(1), dim3(64), 0, m.stream, k, m.nload, m.load_nt, m.ndim,
                           (int)m.h_load_dofs.size(), m.load_dofs.p, m.load_node.p, m.load_dir.p, m.load_factor.p,
                           m.load_series.p, fe.p);
    if (int rc = drm_add_fext(m, k, fe.p)) return rc;
    SVLG_HIP(hipGetLastError());
    std::vector<double> h_fe(m.ndofn);
    SVLG_HIP(hipMemcpyAsync(h_fe.data(), fe.p, sizeof(double) * m.ndofn, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    // Device node order -> description order.
    for (int n = 0; n < m.nnodes; ++n)
        for (int c = 0; c < m.ndim; ++c) F[size_t(n) * m.ndim + c] = h_fe[size_t(m.inv[n]) * m.ndim + c];
    return 0;
}

int svlg_cd_effective_stiffness(svlg_model *h, double *Kdiag) {
    SVLG_CHECK(h && Kdiag, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.cd_ready, "svlg_cd_init has not been called");
    DevBuf<double> ff;
    SVLG_HIP(ff.alloc(std::max(m.nfree, 1)));
    if (int rc = gather_free(m, m.keff.p, ff.p)) return rc;
    SVLG_HIP(hipMemcpyAsync(Kdiag, ff.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_cd_apply_increment(svlg_model *h, int32_t k, const double *dU) {
    SVLG_CHECK(h && dU, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.cd_ready, "svlg_cd_init has not been called");
    DevBuf<double> ff, dn;
    SVLG_HIP(ff.upload(dU, std::max(m.nfree, 1), m.stream));
    SVLG_HIP(dn.alloc(m.ndofn));
    if (int rc = scatter_free(m, ff.p, dn.p)) return rc;
    const int nb = (m.ndofn + 255) / 256;
    hipLaunchKernelGGL(k_cd_increment, dim3(nb), dim3(256), 0, m.stream, m.ndofn, m.U[m.ring].p, dn.p, m.keff.p,
                       m.U[(m.ring + 1) % 3].p);
    SVLG_HIP(hipGetLastError());
    if (int rc = sm_add_increment(m, k, m.U[m.ring].p, m.U[(m.ring + 1) % 3].p)) return rc;
    m.ring = (m.ring + 1) % 3;
    m.steps_done++;
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int svlg_cd_get_state(svlg_model *h, double *U, double *V, double *A) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.cd_ready, "svlg_cd_init has not been called");
    std::vector<double> u0(m.ndofn), u1(m.ndofn), u2(m.ndofn);
    SVLG_HIP(hipMemcpyAsync(u0.data(), m.U[m.ring].p, sizeof(double) * m.ndofn, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipMemcpyAsync(u1.data(), m.U[(m.ring + 2) % 3].p, sizeof(double) * m.ndofn, hipMemcpyDeviceToHost,
                            m.stream));
    SVLG_HIP(hipMemcpyAsync(u2.data(), m.U[(m.ring + 1) % 3].p, sizeof(double) * m.ndofn, hipMemcpyDeviceToHost,
                            m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    for (int i = 0; i < m.ndofn; ++i) {
        int f = m.h_free[i];
        if (f < 0) continue;
        const double dt = m.h_dt_dof.empty() ? m.dt : m.h_dt_dof[i];
        if (U) U[f] = u0[i];
        // CentralDifference::ComputeNewStep (:141-145) with U = U_{k-1}, Up = U_{k-2},
        // dU = U_k - U_{k-1}.
        double dU = u0[i] - u1[i];
        if (V) V[f] = m.steps_done > 0 ? 1.0 / 2.0 / dt * (u0[i] - u2[i]) : 0.0;
        if (A) A[f] = m.steps_done > 0 ? 1.0 / dt / dt * ((dU - u1[i]) + u2[i]) : 0.0;
    }
    return 0;
}

}  // extern "C"
