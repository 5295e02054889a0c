// svlg_gcn.hip -- learned preconditioner of the effective-stiffness solve: a linear graph
// convolution network on the mesh's DOF graph whose propagation operator is the
// Jacobi-normalised effective stiffness Ehat = D^-1/2 Keff D^-1/2 (D = diag Keff):
//
//   x0  = s * r / sigma                    s = D^-1/2 (0 on restrained DOFs), sigma = rms(r)
//   H^1 = x0 c_1^T
//   H^l = H^(l-1) A_l + (Ehat H^(l-1)) B_l + x0 c_l^T          l = 2..L
//   z   = sigma s * (H^L w)
//
// H^l holds C channels per DOF. With C = 1, A_l = 1, B_l = -omega_l, c_l = omega_l, w = 1 the
// network is L steps of Richardson iteration on Ehat y = x0 with step sizes omega_l; with the
// omega_l the reciprocal Chebyshev nodes of [lmin, lmax] it is the Chebyshev polynomial
// preconditioner, which is the initial state. Training (the forward and backward kernels
// below) minimises the energy-norm error ||z - e||^2_Keff / ||e||^2_Keff over
// residual/solution pairs (r, e), Keff e = r, that the model's own Newmark solves produce.
// z is linear in r, but the map is not symmetric for C > 1, so the solver runs flexible CG.
//
// Ehat is applied on the node tiles of the Newmark operator: one workgroup stages s * H of
// its owned + halo nodes, recomputes every element touching its nodes and sums corner slots
// in a fixed order. The element step is either the FP64 element routine (any mesh) or, for
// Hexa8 meshes with a single element geometry, a BF16 MFMA product with the element's
// K_lambda and K_mu (FP32 accumulation): f_e = lambda_e K_lambda u_e + mu_e K_mu u_e.
// Hidden states are FP32; the diagonal part, z and every CG scalar stay FP64.
//
// Parameter vector theta (FP64 on the host, FP32 on the device): for l = 1..L the block
// [A_l (C x C, row = input channel) | B_l (C x C) | c_l (C)], then w (C). A_1 and B_1 are
// unused (H^0 = 0).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "svlg_common.hpp"
#include "svlg_elem.hpp"
#include "svlg_mfma.hpp"
#include "svlg_solver.hpp"

namespace svlg {

int gather_free(Model &m, const double *dnode, double *dfree);
int scatter_free(Model &m, const double *dfree, double *dnode);
int keff_apply_node(Model &m, const double *x, double *y);

static constexpr int kMaxC = 4;
static constexpr int kMaxL = 16;

static inline int theta_layer(int C) { return 2 * C * C + C; }
static inline int theta_size(int C, int L) { return L * theta_layer(C) + C; }

enum { G_FWD = 0, G_BWD = 1, G_OUT = 2 };

struct GcnArgs {
    const int32_t *own_ptr, *halo_ptr, *halo_node, *elem_ptr;
    const uint16_t *lconn, *egeo;
    const double *elam, *emu;  // (lambda, mu) * (1 + 2 ak/dt) per tile element
    const double *geo, *X;
    double th;
    int ntiles, max_stage, max_own;
    const float *kmat;         // [2][24][24] (MFMA path); the fast kernel takes the A-operand image
    const double *dgn, *s;
    const float *sf;
    int C, ndofn;
    const float *in;           // [C][ndofn]
    float *out;                // [C][ndofn]
    float *ystore;             // G_FWD: Ehat in, if not null
    const float *x0;
    const float *M1, *M2, *v;  // out = in M1 + (Ehat in) M2 + x0 v^T (v may be null)
    int trans;                 // use M1^T and M2^T
    // G_BWD: weight-gradient partials per tile [2 C^2 + C]
    const float *hprev, *yl;
    double *gpart;
    // G_OUT
    const float *w;
    const double *r, *q;
    double *z, *scal;
    CgParts cg;
    int iter, normalize;
    const int *jdev;           // non-null: the iteration is *jdev + iter (graph replays)
    int test;                  // first layer in the solver: CG's stopping test on r_iter first
    const float *elamf, *emuf, *dgnf;  // FP32 copies of elam, emu, dgn (one-channel kernel)
    const float *jacf;                 // D^-1 in FP32 (one-channel kernel)
    // One-channel kernel, fixed-stride tile data (TileSet::hpad, lcpad; null: not built).
    const int32_t *hpad;
    const uint16_t *lcpad;
    const float *elamf_pad, *emuf_pad;
    int hstride, estride;
    const double *cmass;       // consistent mass: (4/dt^2 + 2 am/dt) times the element-mass scale
                               // per tile element (null: lumped, inside dgn)
};

__device__ __forceinline__ float mixw(const float *M, int C, int c, int k, int trans) {
    return trans ? M[k * C + c] : M[c * C + k];
}

// sigma of the input scaling: rms of r_j (from the solver's partials) or 1.
__device__ __forceinline__ double gcn_sigma(const CgParts &c, int j, int normalize, int n, double *buf) {
    if (!normalize) return 1.0;
    double v[3];
    cg_vec(c, j, buf, v);
    return v[1] > 0.0 ? sqrt(v[1] / double(n)) : 1.0;
}

// ---------------------------------------------------------------- element phase
// FP64 element routine on the staged FP32 operand of one channel.
template <int DIM, int PATH>
__device__ __forceinline__ void gcn_elem_valu(const GcnArgs &a, int e0, int ne, int nown, const float *sXc,
                                              const double *sCoord, double *sIn, int MO) {
    constexpr int NEN = Elem<DIM>::nen;
    const int tid = threadIdx.x;
    if (tid >= ne) return;
    int lc[NEN];
    const uint16_t *l = a.lconn + size_t(e0 + tid) * NEN;
#pragma unroll
    for (int c = 0; c < NEN; ++c) lc[c] = l[c];
    const double lam = a.elam[e0 + tid], mu = a.emu[e0 + tid];
    const int gid = PATH >= 2 ? int(a.egeo[e0 + tid]) : 0;
    double u[NEN][DIM], f[NEN][DIM];
#pragma unroll
    for (int c = 0; c < NEN; ++c)
#pragma unroll
        for (int d = 0; d < DIM; ++d) u[c][d] = double(sXc[lc[c] * DIM + d]);
    tile_elem_force<DIM, PATH>(a.geo, a.th, lc, sCoord, lam, mu, gid, u, f);
    if (a.cmass) tile_elem_mass<DIM, PATH>(lc, sCoord, a.cmass[e0 + tid], u, f);
#pragma unroll
    for (int c = 0; c < NEN; ++c)
        if (lc[c] < nown) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) sIn[(c * DIM + d) * MO + lc[c]] = f[c][d];
        }
}

// ---------------------------------------------------------------- layer kernel
// One node tile of a network layer: stages s * in (every channel) for owned + halo nodes,
// applies Ehat per channel, then mixes channels on the owned DOFs.
//  G_FWD: out = in M1 + (Ehat in) M2 + x0 v^T; optionally stores Ehat in.
//  G_BWD: out = in M1^T + (Ehat in) M2^T (trans = 1) and the per-tile gradient partials
//         dA[c][k] = sum hprev_c in_k, dB[c][k] = sum yl_c in_k, dc[k] = sum x0 in_k.
//  G_OUT: as G_FWD, then z = sigma s * (out w) (FP64), per-tile z.r and z.q; the last tile
//         sets the flexible-CG scalars RHO = z.r and BETA = first ? 0 : -z.q / PQ.
template <int DIM, int PATH, bool MFMA, int MODE>
__global__ __launch_bounds__(512) void k_gcn_layer(GcnArgs a) {
    constexpr int NEN = Elem<DIM>::nen;
    constexpr bool STAGE_X = PATH < 2;
    extern __shared__ double smem[];
    __shared__ double red[3 * 17];
    if constexpr (MODE != G_BWD) {
        if (a.scal && a.scal[SC_DONE] != 0.0) return;
        if (a.test && cg_stop_test(a.cg, a.jdev, a.iter, a.scal, red)) return;
    }
    double sig = 1.0;
    if constexpr (MODE == G_OUT) sig = gcn_sigma(a.cg, iter_of(a.jdev, a.iter), a.normalize, a.ndofn, red);
    const int nt = a.ntiles;
    const int bI = blockIdx.x, xcd = bI & 7, j = bI >> 3;
    const int q = nt >> 3, r = nt & 7;
    const int tile = xcd * q + min(xcd, r) + j;
    const int o0 = a.own_ptr[tile], nown = a.own_ptr[tile + 1] - o0;
    const int h0 = a.halo_ptr[tile], nh = a.halo_ptr[tile + 1] - h0;
    const int e0 = a.elem_ptr[tile], ne = a.elem_ptr[tile + 1] - e0;
    const int MS = a.max_stage, MO = a.max_own, C = a.C, N = a.ndofn;
    double *sIn = smem;                                              // [NEN*DIM][MO]
    double *sCoord = sIn + size_t(NEN * DIM) * MO;                   // [MS*DIM] if STAGE_X
    float *sX = reinterpret_cast<float *>(sCoord + (STAGE_X ? size_t(MS) * DIM : 0));  // [C][MS*DIM]
    const int tid = threadIdx.x, B = blockDim.x;
    const size_t g0 = size_t(o0) * DIM;
    // Stage s * in for every channel (and the coordinates).
    for (int i = tid; i < nown * DIM; i += B) {
        const float sv = a.sf[g0 + i];
        for (int c = 0; c < C; ++c) sX[c * MS * DIM + i] = sv * a.in[size_t(c) * N + g0 + i];
        if constexpr (STAGE_X) sCoord[i] = a.X[g0 + i];
    }
    for (int sl = tid; sl < nh; sl += B) {
        const size_t nd = size_t(a.halo_node[h0 + sl]) * DIM;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int i = (nown + sl) * DIM + d;
            const float sv = a.sf[nd + d];
            for (int c = 0; c < C; ++c) sX[c * MS * DIM + i] = sv * a.in[size_t(c) * N + nd + d];
            if constexpr (STAGE_X) sCoord[i] = a.X[nd + d];
        }
    }
    bf16x8 Ak[3];
    if constexpr (MFMA) mfma_load_kblocks(a.kmat, Ak);
    double yv[kMaxC][DIM];
    for (int c = 0; c < C; ++c) {
        for (int i = tid; i < NEN * DIM * MO; i += B) sIn[i] = 0.0;
        __syncthreads();
        if constexpr (MFMA)
            mfma_hexa_elements<float>(a.lconn, a.elam, a.emu, e0, ne, nown, sX + c * MS * DIM, sIn, MO, Ak, -1.0);
        else
            gcn_elem_valu<DIM, PATH>(a, e0, ne, nown, sX + c * MS * DIM, sCoord, sIn, MO);
        __syncthreads();
        if (tid < nown) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                double fi = 0.0;
#pragma unroll
                for (int cc = 0; cc < NEN; ++cc) fi += sIn[(cc * DIM + d) * MO + tid];
                const size_t g = g0 + size_t(tid) * DIM + d;
                const double sg = a.s[g];
                const double xh = sg * double(a.in[size_t(c) * N + g]);
                yv[c][d] = sg * (fi + a.dgn[g] * xh);
            }
        }
        __syncthreads();
    }
    // Channel mixing on the owned DOFs.
    double acc0 = 0.0, acc1 = 0.0;
    constexpr int NP = 2 * kMaxC * kMaxC + kMaxC;
    double gp[NP];
    if constexpr (MODE == G_BWD) {
#pragma unroll
        for (int i = 0; i < NP; ++i) gp[i] = 0.0;
    }
    if (tid < nown) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const size_t g = g0 + size_t(tid) * DIM + d;
            float inv[kMaxC];
            for (int c = 0; c < C; ++c) inv[c] = a.in[size_t(c) * N + g];
            const float x0 = a.x0[g];
            double zsum = 0.0;
            for (int k = 0; k < C; ++k) {
                double o = 0.0;
                for (int c = 0; c < C; ++c)
                    o += double(inv[c]) * mixw(a.M1, C, c, k, a.trans) + yv[c][d] * mixw(a.M2, C, c, k, a.trans);
                if (MODE != G_BWD && a.v) o += double(x0) * a.v[k];
                const float of = float(o);
                a.out[size_t(k) * N + g] = of;
                if constexpr (MODE == G_OUT) zsum += double(of) * a.w[k];
            }
            if constexpr (MODE == G_FWD) {
                if (a.ystore)
                    for (int c = 0; c < C; ++c) a.ystore[size_t(c) * N + g] = float(yv[c][d]);
            }
            if constexpr (MODE == G_BWD) {
                // in = G^l here; hprev = H^(l-1), yl = Ehat H^(l-1).
                for (int k = 0; k < C; ++k) {
                    const double gk = inv[k];
                    for (int c = 0; c < C; ++c) {
                        gp[c * C + k] += double(a.hprev[size_t(c) * N + g]) * gk;
                        gp[C * C + c * C + k] += double(a.yl[size_t(c) * N + g]) * gk;
                    }
                    gp[2 * C * C + k] += double(x0) * gk;
                }
            }
            if constexpr (MODE == G_OUT) {
                const double zv = sig * a.s[g] * zsum;
                a.z[g] = zv;
                acc0 += zv * a.r[g];
                acc1 += zv * a.q[g];
            }
        }
    }
    if constexpr (MODE == G_BWD) {
        const int np = 2 * C * C + C;
        for (int i = 0; i < np; ++i) {
            const double sv = block_sum(gp[i], red);
            if (tid == 0) a.gpart[size_t(tile) * np + i] = sv;
        }
    }
    if constexpr (MODE == G_OUT) {
        acc0 = block_sum(acc0, red);
        acc1 = block_sum(acc1, red);
        if (tid == 0) {
            double *t = a.cg.tp + size_t(iter_of(a.jdev, a.iter) & 1) * 3 * nt;
            t[3 * tile] = acc0;
            t[3 * tile + 1] = acc1;
            t[3 * tile + 2] = 0.0;
        }
    }
}

// Inference layer for one channel with the BF16 MFMA element step (Hexa8, single element
// geometry) on the scaled states G = s H (s = D^-1/2): the layer H' = a_l H + b_l (Ehat H) +
// c_l x0 with Ehat = s Keff s reads G' = a_l G + b_l D^-1 (Keff G) + c_l s x0, so the staged
// operand is the state itself (no scaling loads) and the output layer gives z = sigma s (w H)
// = sigma w G, with the tile partials of z.r and z.q. The element data of the workgroup's
// element blocks are prefetched before staging, the corner slots and the node-phase arithmetic
// are FP32 (preconditioner precision); z and the partials are FP64.
// 16-bit staged slot of corner 2w + h from the packed connectivity (four 32-bit words).
__device__ __forceinline__ int packed_half(uint32_t w, int h) { return int((w >> (16 * h)) & 0xffffu); }

// MOC > 0: every tile owns at most MOC nodes; the slot planes then have the compile-time
// stride MOC + 1 with column MOC as the trash column (one base address per corner, immediate
// plane offsets).
// With MOC the kernel is held to 64 VGPRs (eight waves per SIMD): four 512-thread workgroups
// then share a CU, so the default tiles of the 64^3 block all run in one round.
// PAD: the tiles' fixed-stride halo lists and element data exist (TileSet::hpad, lcpad and the
// padded FP32 coefficients): the halo indices and the connectivity are loaded with the tile
// index alone, in the same round trip as the tile bounds, and the solver's done flag is read
// while they are in flight.
template <bool OUT, int MOC = 0, bool PAD = false, bool TEST = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(MOC > 0 ? 8 : 1))) void k_gcn_fast(GcnArgs a) {
    constexpr int MAXR = 4;  // element rounds per wave: tiles hold at most 512 elements
    extern __shared__ double smem[];
    __shared__ double red[3 * 17];
    const int nt = a.ntiles;
    const int bI = blockIdx.x, xcd = bI & 7, jb = bI >> 3;
    const int qq = nt >> 3, rr = nt & 7;
    const int tile = xcd * qq + min(xcd, rr) + jb;
    const int tid = threadIdx.x, B = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = B >> 6, n16 = lane & 15, g = lane >> 4;
    // First layer in the solver: the stopping test on r_iter (with the DONE flag) before the
    // prefetch, whose registers would be held across the test's barriers.
    if constexpr (TEST)
        if (cg_stop_test(a.cg, a.jdev, a.iter, a.scal, red)) return;
    int hn[2];
    uint4 pk[MAXR];
    if constexpr (PAD) {
        const int32_t *hp = a.hpad + size_t(tile) * a.hstride;
#pragma unroll
        for (int k = 0; k < 2; ++k) hn[k] = hp[min(k * B + tid, a.hstride - 1)];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int el = min(wave * 16 + r * nw * 16 + n16, a.estride - 1);
            pk[r] = *reinterpret_cast<const uint4 *>(a.lcpad + (size_t(tile) * a.estride + el) * 8);
        }
    }
    if constexpr (!TEST)
        if (a.scal && a.scal[SC_DONE] != 0.0) return;
    double sig = 1.0;
    if constexpr (OUT) sig = gcn_sigma(a.cg, iter_of(a.jdev, a.iter), a.normalize, a.ndofn, red);
    const int o0 = a.own_ptr[tile], nown = a.own_ptr[tile + 1] - o0;
    const int h0 = a.halo_ptr[tile], nh = a.halo_ptr[tile + 1] - h0;
    const int e0 = a.elem_ptr[tile], ne = a.elem_ptr[tile + 1] - e0;
    const int MS = a.max_stage, MO = MOC > 0 ? MOC + 1 : a.max_own;
    // LDS: corner slots [24][MO] + one trash cell (every masked-out write lands there; with
    // MOC the trash column), staged operand [MS * 3] + one zero cell.
    float *sIn = reinterpret_cast<float *>(smem);
    const int trash = 24 * MO;
    float *sX = sIn + size_t(24) * MO + 1;
    const int zcell = 3 * MS;
    // Prefetch the A-operand image (four 16 x 32 BF16 blocks) and this lane's element of every
    // round. k-slot order: lane group g carries corners 2g and 2g+1 (k = 8g + jj -> element DOF
    // 6g + jj, jj < 6; jj = 6, 7 are zero). Row order: blocks 0, 1 (K_lambda) and 2, 3 (K_mu)
    // give lane group g the rows of corner 4(b%2) + g (row 4g + i -> DOF 3 corner + i, i < 3).
    const uint4 *kimg = reinterpret_cast<const uint4 *>(a.kmat);
    uint4 ak[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) ak[b] = kimg[b * 64 + lane];
    if constexpr (!PAD)
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int el = min(wave * 16 + r * nw * 16 + n16, max(ne - 1, 0));
            pk[r] = *reinterpret_cast<const uint4 *>(a.lconn + size_t(e0 + el) * 8);
        }
    const size_t g0 = size_t(o0) * 3;
    if constexpr (PAD) {
        // One pass: the host checked that two block strides cover the owned DOFs and the halo.
        float v[2], hv[2][3];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = k * B + tid;
            v[k] = i < nown * 3 ? a.in[g0 + i] : 0.f;
#pragma unroll
            for (int c = 0; c < 3; ++c) hv[k][c] = i < nh ? a.in[size_t(hn[k]) * 3 + c] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = k * B + tid;
            if (i < nown * 3) sX[i] = v[k];
            if (i < nh)
#pragma unroll
                for (int c = 0; c < 3; ++c) sX[(nown + i) * 3 + c] = hv[k][c];
        }
    } else {
        for (int i = tid; i < nown * 3; i += B) sX[i] = a.in[g0 + i];
        for (int sl = tid; sl < nh; sl += B) {
            const size_t nd = size_t(a.halo_node[h0 + sl]) * 3;
            const float v0 = a.in[nd], v1 = a.in[nd + 1], v2 = a.in[nd + 2];
            float *d = sX + (nown + sl) * 3;
            d[0] = v0;
            d[1] = v1;
            d[2] = v2;
        }
    }
    for (int i = tid; i < 24 * MO; i += B) sIn[i] = 0.f;
    if (tid == 0) sX[zcell] = 0.f;
    __syncthreads();
    bf16x8 Ak[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) Ak[b] = *reinterpret_cast<const bf16x8 *>(&ak[b]);
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
        const int eb = wave * 16 + r * nw * 16;
        if (eb >= ne) break;  // wave-uniform
        const bool valid = eb + n16 < ne;
        // The element coefficients are read per round (issued before the gathers) rather than
        // prefetched with the connectivity: eight fewer live registers.
        const int el = min(eb + n16, ne - 1);
        const float lamv = PAD ? a.elamf_pad[size_t(tile) * a.estride + el] : a.elamf[e0 + el];
        const float muv = PAD ? a.emuf_pad[size_t(tile) * a.estride + el] : a.emuf[e0 + el];
        // Gather: word g of the packed connectivity holds corners 2g, 2g+1.
        const uint32_t wg = g == 0 ? pk[r].x : g == 1 ? pk[r].y : g == 2 ? pk[r].z : pk[r].w;
        const int sa = valid ? packed_half(wg, 0) * 3 : zcell;
        const int sb = valid ? packed_half(wg, 1) * 3 : zcell;
        const int za = valid ? 1 : 0;
        bf16x8 Bv;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            Bv[jj] = __bf16(sX[sa + jj * za]);
            Bv[3 + jj] = __bf16(sX[sb + jj * za]);
        }
        Bv[6] = __bf16(0.f);
        Bv[7] = __bf16(0.f);
        const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
        const f32x4 l0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ak[0], Bv, zero, 0, 0, 0);
        const f32x4 l1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ak[1], Bv, zero, 0, 0, 0);
        const f32x4 m0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ak[2], Bv, zero, 0, 0, 0);
        const f32x4 m1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ak[3], Bv, zero, 0, 0, 0);
        // Scatter: corner g (word g/2, half g&1) from blocks 0/2, corner 4 + g from blocks 1/3.
        const uint32_t wlo = (g >> 1) ? pk[r].y : pk[r].x, whi = (g >> 1) ? pk[r].w : pk[r].z;
        const int c0 = g, c1 = 4 + g;
        const int s0 = packed_half(wlo, g & 1), s1 = packed_half(whi, g & 1);
        const bool w0 = valid && s0 < nown, w1 = valid && s1 < nown;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if constexpr (MOC > 0) {
                sIn[(c0 * 3 + i) * MO + (w0 ? s0 : MOC)] = lamv * l0[i] + muv * m0[i];
                sIn[(c1 * 3 + i) * MO + (w1 ? s1 : MOC)] = lamv * l1[i] + muv * m1[i];
            } else {
                sIn[w0 ? (c0 * 3 + i) * MO + s0 : trash] = lamv * l0[i] + muv * m0[i];
                sIn[w1 ? (c1 * 3 + i) * MO + s1 : trash] = lamv * l1[i] + muv * m1[i];
            }
        }
    }
    __syncthreads();
    const float am = a.M1[0], bm = a.M2[0], cm = a.v[0];
    double acc0 = 0.0, acc1 = 0.0;
    if (tid < nown) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float fi = 0.f;
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) fi += sIn[(cc * 3 + d) * MO + tid];
            const size_t gd = g0 + size_t(tid) * 3 + d;
            const float xh = sX[tid * 3 + d];
            const float y = a.jacf[gd] * (fi + a.dgnf[gd] * xh);
            const float o = am * xh + bm * y + cm * a.x0[gd];
            if constexpr (OUT) {
                const double zv = sig * double(o * a.w[0]);
                a.z[gd] = zv;
                acc0 += zv * a.r[gd];
                acc1 += zv * a.q[gd];
            } else {
                a.out[gd] = o;
            }
        }
    }
    if constexpr (OUT) {
        acc0 = block_sum(acc0, red);
        acc1 = block_sum(acc1, red);
        if (tid == 0) {
            double *t = a.cg.tp + size_t(iter_of(a.jdev, a.iter) & 1) * 3 * nt;
            t[3 * tile] = acc0;
            t[3 * tile + 1] = acc1;
            t[3 * tile + 2] = 0.0;
        }
    }
}

// A-operand image of the fast MFMA element step ([4][64] x 16 B): for block b and lane l
// (row m = l % 16, k-slots 8 (l / 16) .. +7) the BF16 entries K[row DOF][column DOF] in the
// k-slot and row orders described in k_gcn_fast (zero on padding slots).
__global__ void k_gcn_kimg(const float *__restrict__ kmat, uint4 *__restrict__ img) {
    const int lane = threadIdx.x;
    if (lane >= 64) return;
    const int m = lane & 15, g = lane >> 4;
    const int gr = m >> 2, ir = m & 3;
    for (int b = 0; b < 4; ++b) {
        const float *K = kmat + (b < 2 ? 0 : 576);
        const int rdof = ir < 3 ? (4 * (b & 1) + gr) * 3 + ir : -1;
        bf16x8 v;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            const int cdof = jj < 6 ? 6 * g + jj : -1;
            v[jj] = __bf16((rdof >= 0 && cdof >= 0) ? K[rdof * 24 + cdof] : 0.f);
        }
        img[b * 64 + lane] = *reinterpret_cast<const uint4 *>(&v);
    }
}

// x0 = s r / sigma, H^1 = x0 c_1^T (s = D^-1/2, or D^-1 for the one-channel kernel's scaled states).
__global__ __launch_bounds__(512) void k_gcn_first(int n, int C, int normalize, CgParts cg, int j,
                                                   const double *__restrict__ r, const double *__restrict__ s,
                                                   const float *__restrict__ c1, const double *scal,
                                                   float *__restrict__ x0, float *__restrict__ h1,
                                                   const int *jdev) {
    __shared__ double red[3 * 17];
    if (normalize && scal[SC_DONE] != 0.0) return;
    const double sig = gcn_sigma(cg, iter_of(jdev, j), normalize, n, red);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float xv = float(s[i] * r[i] / sig);
        x0[i] = xv;
        for (int k = 0; k < C; ++k) h1[size_t(k) * n + i] = xv * c1[k];
    }
}

// Training: z = s * (H^L w).
__global__ void k_gcn_out(int n, int C, const float *__restrict__ hL, const float *__restrict__ w,
                          const double *__restrict__ s, double *__restrict__ z) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int k = 0; k < C; ++k) acc += double(hL[size_t(k) * n + i]) * w[k];
        z[i] = s[i] * acc;
    }
}

// Per-block partials of (z - e).(Keff z - r) and e.r (loss of one pair).
__global__ __launch_bounds__(512) void k_gcn_lossdots(int n, const double *__restrict__ z, const double *__restrict__ e,
                                                      const double *__restrict__ kz, const double *__restrict__ r,
                                                      double *__restrict__ part) {
    __shared__ double red[17];
    double a0 = 0.0, a1 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        a0 += (z[i] - e[i]) * (kz[i] - r[i]);
        a1 += e[i] * r[i];
    }
    a0 = block_sum(a0, red);
    a1 = block_sum(a1, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a0;
        part[2 * blockIdx.x + 1] = a1;
    }
}

// Output gradient: gz = scale (Keff z - r); G^L_k = w_k s gz; partials dw_k = sum H^L_k s gz.
__global__ __launch_bounds__(512) void k_gcn_outgrad(int n, int C, double scale, const double *__restrict__ kz,
                                                     const double *__restrict__ r, const double *__restrict__ s,
                                                     const float *__restrict__ w, const float *__restrict__ hL,
                                                     float *__restrict__ G, double *__restrict__ part) {
    __shared__ double red[17];
    double dw[kMaxC] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double sg = s[i] * scale * (kz[i] - r[i]);
        for (int k = 0; k < C; ++k) {
            G[size_t(k) * n + i] = float(w[k] * sg);
            dw[k] += double(hL[size_t(k) * n + i]) * sg;
        }
    }
    for (int k = 0; k < C; ++k) {
        const double v = block_sum(dw[k], red);
        if (threadIdx.x == 0) part[size_t(blockIdx.x) * C + k] = v;
    }
}

// Layer-1 gradient partials dc_1[k] = sum x0 G^1_k.
__global__ __launch_bounds__(512) void k_gcn_first_grad(int n, int C, const float *__restrict__ x0,
                                                        const float *__restrict__ G, double *__restrict__ part) {
    __shared__ double red[17];
    double dc[kMaxC] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < C; ++k) dc[k] += double(x0[i]) * double(G[size_t(k) * n + i]);
    for (int k = 0; k < C; ++k) {
        const double v = block_sum(dc[k], red);
        if (threadIdx.x == 0) part[size_t(blockIdx.x) * C + k] = v;
    }
}

// grad[off + i] += sum over rows of part[row][i] (fixed row order), one workgroup per i.
__global__ __launch_bounds__(256) void k_gcn_accum(int rows, int width, const double *__restrict__ part, double *grad,
                                                   int off) {
    __shared__ double red[17];
    const int i = blockIdx.x;
    double acc = 0.0;
    for (int row = threadIdx.x; row < rows; row += blockDim.x) acc += part[size_t(row) * width + i];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) grad[off + i] += acc;
}

// Columns of K_lambda and K_mu of geometry class 0 (unit Lame constants) by unit probes.
__global__ void k_gcn_kmat(const double *__restrict__ geo, int diag, float *__restrict__ kmat) {
    const int col = threadIdx.x;  // 0..47
    if (col >= 48) return;
    const int which = col / 24, k = col % 24;
    double Lam[3][3];
    for (int i = 0; i < 3; ++i)
        for (int jj = 0; jj < 3; ++jj) Lam[i][jj] = geo[i * 3 + jj];
    double u[8][3], f[8][3];
    for (int c = 0; c < 8; ++c)
        for (int d = 0; d < 3; ++d) u[c][d] = (c * 3 + d == k) ? 1.0 : 0.0;
    const double lam = which == 0 ? 1.0 : 0.0, mu = which == 0 ? 0.0 : 1.0;
    if (diag)
        hexa_modal_force<true>(Lam, lam, mu, u, f);
    else
        hexa_modal_force<false>(Lam, lam, mu, u, f);
    for (int c = 0; c < 8; ++c)
        for (int d = 0; d < 3; ++d) kmat[which * 576 + (c * 3 + d) * 24 + k] = float(f[c][d]);
}

__global__ void k_gcn_scale(int n, const double *__restrict__ jac, double *__restrict__ s, float *__restrict__ sf,
                            float *__restrict__ jacf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = jac[i] > 0.0 ? sqrt(jac[i]) : 0.0;
    s[i] = v;
    sf[i] = float(v);
    jacf[i] = float(jac[i]);
}

// e = s * w with w uniform in [-1, 1] from a counter-based hash of (seed, i).
__global__ void k_gcn_white(int n, unsigned long long seed, const double *__restrict__ s, double *__restrict__ e) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + (unsigned long long)(i) * 0xBF58476D1CE4E5B9ull;
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = double(z >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
    e[i] = s[i] * (2.0 * u - 1.0);
}

__global__ void k_gcn_scale_mul(int n, const double *__restrict__ jac, double a, double *__restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = a * jac[i] * d[i];
}

__global__ void k_axpby(int n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// ---------------------------------------------------------------- host side
static constexpr int kVecGrid = 256, kVecBlock = 512;

// K_lambda, K_mu of the single element geometry of a Hexa8 model (for the MFMA element step).
int build_hexa_kmat(Model &m, DevBuf<float> &kmat) {
    SVLG_CHECK(m.ndim == 3 && m.elem_path >= 2 && m.ngeo == 1,
               "the MFMA element step needs a Hexa8 mesh with a single element geometry");
    SVLG_HIP(kmat.alloc(2 * 576));
    hipLaunchKernelGGL(k_gcn_kmat, dim3(1), dim3(64), 0, m.stream, m.geo.p, m.elem_path == 3 ? 1 : 0, kmat.p);
    SVLG_HIP(hipGetLastError());
    return 0;
}

static size_t gcn_lds(const Model &m) {
    const int nen = m.nen, dim = m.ndim;
    size_t b = size_t(nen * dim) * m.tiles.max_own * sizeof(double);
    if (m.elem_path < 2) b += size_t(m.tiles.max_stage) * dim * sizeof(double);
    b += size_t(m.gcn_C) * m.tiles.max_stage * dim * sizeof(float);
    return b;
}

// Owned-node bounds of the compile-time slot strides of the one-channel layer kernel (tiles
// of up to 168 and up to 336 owned nodes); 0: the stride read from the tile set.
static int gcn_moc(const Model &m) {
    const int mo = m.tiles.max_own;
    return mo <= 168 ? 168 : mo <= 336 ? 336 : 0;
}

// Corner slots [24][stride] + one trash cell, staged operand [max_stage * 3] + one zero cell
// (below 40 KB for the default tiles, so four workgroups fit a CU).
static size_t gcn_fast_lds(const Model &m) {
    const int moc = gcn_moc(m);
    const size_t stride = moc ? moc + 1 : m.tiles.max_own;
    return (size_t(24) * stride + 1 + size_t(m.tiles.max_stage) * 3 + 1) * sizeof(float);
}

template <int MOC, bool TEST>
static const void *gcn_fast_fn_t(bool out, bool pad) {
    if (out) return reinterpret_cast<const void *>(&k_gcn_fast<true, MOC, false, TEST>);
    if constexpr (MOC > 0)
        if (pad) return reinterpret_cast<const void *>(&k_gcn_fast<false, MOC, true, TEST>);
    return reinterpret_cast<const void *>(&k_gcn_fast<false, MOC, false, TEST>);
}

// test: the solver's first layer (CG's stopping test first).
static const void *gcn_fast_fn(const Model &m, bool out, bool test) {
    // The output layer keeps the tile-bound loads first: with the fixed-stride prefetch its
    // registers (z, the partials) would spill.
    const bool pad = !out && m.tiles.hstride > 0 && m.gcn_elamf_pad.p != nullptr;
    switch (gcn_moc(m)) {
        case 168: return test ? gcn_fast_fn_t<168, true>(out, pad) : gcn_fast_fn_t<168, false>(out, pad);
        case 336: return test ? gcn_fast_fn_t<336, true>(out, pad) : gcn_fast_fn_t<336, false>(out, pad);
        default: return test ? gcn_fast_fn_t<0, true>(out, false) : gcn_fast_fn_t<0, false>(out, false);
    }
}

template <int DIM, int PATH, bool MFMA>
static const void *gcn_fn_t(int mode) {
    switch (mode) {
        case G_BWD: return reinterpret_cast<const void *>(&k_gcn_layer<DIM, PATH, MFMA, G_BWD>);
        case G_OUT: return reinterpret_cast<const void *>(&k_gcn_layer<DIM, PATH, MFMA, G_OUT>);
        default: return reinterpret_cast<const void *>(&k_gcn_layer<DIM, PATH, MFMA, G_FWD>);
    }
}

static const void *gcn_fn(const Model &m, int mode, bool mfma) {
    if (m.ndim == 3) {
        if (mfma) return m.elem_path == 3 ? gcn_fn_t<3, 3, true>(mode) : gcn_fn_t<3, 2, true>(mode);
        switch (m.elem_path) {
            case 3: return gcn_fn_t<3, 3, false>(mode);
            case 2: return gcn_fn_t<3, 2, false>(mode);
            case 1: return gcn_fn_t<3, 1, false>(mode);
            default: return gcn_fn_t<3, 0, false>(mode);
        }
    }
    switch (m.elem_path) {
        case 3: return gcn_fn_t<2, 3, false>(mode);
        case 2: return gcn_fn_t<2, 2, false>(mode);
        case 1: return gcn_fn_t<2, 1, false>(mode);
        default: return gcn_fn_t<2, 0, false>(mode);
    }
}

static GcnArgs gcn_args(const Model &m) {
    const TileSet &t = m.tiles;
    GcnArgs a;
    std::memset(&a, 0, sizeof(a));
    a.own_ptr = t.own_ptr.p;
    a.halo_ptr = t.halo_ptr.p;
    a.halo_node = t.halo_node.p;
    a.elem_ptr = t.elem_ptr.p;
    a.lconn = t.elem_lconn.p;
    a.egeo = t.elem_geo.p;
    a.elam = m.gcn_elam.p;
    a.emu = m.gcn_emu.p;
    a.elamf = m.gcn_elamf.p;
    a.emuf = m.gcn_emuf.p;
    a.dgnf = m.gcn_dgnf.p;
    a.jacf = m.gcn_jacf.p;
    a.hpad = m.tiles.hpad.p;
    a.lcpad = m.tiles.lcpad.p;
    a.elamf_pad = m.gcn_elamf_pad.p;
    a.emuf_pad = m.gcn_emuf_pad.p;
    a.hstride = m.tiles.hstride;
    a.estride = m.tiles.estride;
    a.geo = m.geo.p;
    a.X = m.X.p;
    a.th = m.th;
    a.ntiles = t.ntiles;
    a.max_stage = t.max_stage;
    a.max_own = t.max_own;
    a.kmat = m.gcn_kmat.p;
    a.dgn = m.dgn.p;
    a.cmass = m.lumped ? nullptr : m.cmass_keff.p;
    a.s = m.gcn_s.p;
    a.sf = m.gcn_sf.p;
    a.C = m.gcn_C;
    a.ndofn = m.ndofn;
    a.x0 = m.gcn_x0.p;
    return a;
}

static int launch_gcn(Model &m, int mode, bool mfma, const GcnArgs &a) {
    void *args[] = {const_cast<GcnArgs *>(&a)};
    SVLG_HIP(hipLaunchKernel(gcn_fn(m, mode, mfma), dim3(m.tiles.ntiles), dim3(m.tiles.block), args, gcn_lds(m),
                             m.stream));
    return 0;
}

static const float *theta_A(const Model &m, int l) { return m.gcn_w.p + size_t(l - 1) * theta_layer(m.gcn_C); }
static const float *theta_B(const Model &m, int l) { return theta_A(m, l) + m.gcn_C * m.gcn_C; }
static const float *theta_c(const Model &m, int l) { return theta_A(m, l) + 2 * m.gcn_C * m.gcn_C; }
static const float *theta_w(const Model &m) { return m.gcn_w.p + size_t(m.gcn_L) * theta_layer(m.gcn_C); }

static int upload_theta(Model &m) {
    std::vector<float> f(m.gcn_theta.begin(), m.gcn_theta.end());
    SVLG_HIP(hipMemcpyAsync(m.gcn_w.p, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

// Inference: z = P(r) (node layout). In the solver (normalize = 1) r is r_j of iteration j
// with its partials in `cg`; the output layer writes the tp[j&1] partials.
// first = false: x0 and H^1 are already in gcn_x0 / gcn_h[0] (formed by the CG vector kernel).
// The one-channel MFMA inference kernel runs on scaled states G = D^-1/2 H (its input layer
// is x0 = D^-1 r); every other layer kernel on H (x0 = D^-1/2 r).
static bool gcn_scaled_states(const Model &m) { return m.gcn_mfma && m.gcn_C == 1; }
static const double *gcn_input_scale(const Model &m) { return gcn_scaled_states(m) ? m.jac.p : m.gcn_s.p; }

static int gcn_run(Model &m, const CgParts &cg, int j, const double *r, const double *q, double *z, int normalize,
                   const int *jdev = nullptr, bool first = true, bool test = false) {
    const int n = m.ndofn, C = m.gcn_C, L = m.gcn_L;
    if (first)
        hipLaunchKernelGGL(k_gcn_first, dim3(kVecGrid), dim3(kVecBlock), 0, m.stream, n, C, normalize, cg, j, r,
                           gcn_input_scale(m), theta_c(m, 1), m.scal.p, m.gcn_x0.p, m.gcn_h[0].p, jdev);
    GcnArgs a = gcn_args(m);
    int cur = 0;
    for (int l = 2; l <= L; ++l) {
        a.in = m.gcn_h[cur].p;
        a.out = m.gcn_h[1 - cur].p;
        a.M1 = theta_A(m, l);
        a.M2 = theta_B(m, l);
        a.v = theta_c(m, l);
        a.trans = 0;
        a.scal = m.scal.p;
        const int mode = l == L ? G_OUT : G_FWD;
        a.test = test && l == 2 ? 1 : 0;
        if (a.test || mode == G_OUT) {
            a.cg = cg;
            a.iter = j;
            a.jdev = jdev;
        }
        if (mode == G_OUT) {
            a.w = theta_w(m);
            a.r = r;
            a.q = q;
            a.z = z;
            a.cg = cg;
            a.iter = j;
            a.jdev = jdev;
            a.normalize = normalize;
        }
        if (gcn_scaled_states(m)) {
            a.kmat = reinterpret_cast<const float *>(m.gcn_kimg.p);
            void *args[] = {&a};
            const void *fn = gcn_fast_fn(m, mode == G_OUT, a.test != 0);
            SVLG_HIP(hipLaunchKernel(fn, dim3(m.tiles.ntiles), dim3(m.tiles.block), args, gcn_fast_lds(m), m.stream));
        } else if (int rc = launch_gcn(m, mode, m.gcn_mfma != 0, a)) {
            return rc;
        }
        cur = 1 - cur;
    }
    return 0;
}

// In the solver the network runs unnormalised (it is linear in r, and FP32 / BF16 cover the
// residual's range), so its input layer comes from the CG vector kernel with no global norm.
int gcn_precond(Model &m, const CgParts &c, int j, const double *r, const double *q, double *z, const int *jdev) {
    return gcn_run(m, c, j, r, q, z, 0, jdev, false, true);
}

GcnInput gcn_input_args(Model &m) {
    return GcnInput{gcn_input_scale(m), theta_c(m, 1), m.gcn_x0.p, m.gcn_h[0].p, m.gcn_C};
}

// Loss and gradient of a batch of pairs (node layout, device); grad is FP64 on the device,
// accumulated (+=) over the pairs. loss_sum receives sum over pairs of the pair losses.
static int gcn_loss_grad(Model &m, int np, const double *const *R, const double *const *E, double weight,
                         double *grad_dev, double *loss_sum) {
    const int n = m.ndofn, C = m.gcn_C, L = m.gcn_L, TL = theta_layer(C);
    const size_t plane = size_t(C) * n;
    hipStream_t s = m.stream;
    double *z = m.gcn_tmp[0].p, *kz = m.gcn_tmp[1].p;
    std::vector<double> hp(2 * kVecGrid);
    double total = 0.0;
    for (int p = 0; p < np; ++p) {
        // Forward with stored states: H^l at gcn_H + l * plane (l = 1..L), Ehat H^(l-1) at gcn_Y + l * plane.
        hipLaunchKernelGGL(k_gcn_first, dim3(kVecGrid), dim3(kVecBlock), 0, s, n, C, 0, CgParts{}, 0, R[p],
                           m.gcn_s.p, theta_c(m, 1), m.scal.p, m.gcn_x0.p, m.gcn_H.p + plane, nullptr);
        GcnArgs a = gcn_args(m);
        for (int l = 2; l <= L; ++l) {
            a.in = m.gcn_H.p + size_t(l - 1) * plane;
            a.out = m.gcn_H.p + size_t(l) * plane;
            a.ystore = m.gcn_Y.p + size_t(l) * plane;
            a.M1 = theta_A(m, l);
            a.M2 = theta_B(m, l);
            a.v = theta_c(m, l);
            a.trans = 0;
            if (int rc = launch_gcn(m, G_FWD, false, a)) return rc;
        }
        hipLaunchKernelGGL(k_gcn_out, dim3(kVecGrid), dim3(kVecBlock), 0, s, n, C, m.gcn_H.p + size_t(L) * plane,
                           theta_w(m), m.gcn_s.p, z);
        if (int rc = keff_apply_node(m, z, kz)) return rc;
        hipLaunchKernelGGL(k_gcn_lossdots, dim3(kVecGrid), dim3(kVecBlock), 0, s, n, z, E[p], kz, R[p], m.gcn_part.p);
        SVLG_HIP(hipMemcpyAsync(hp.data(), m.gcn_part.p, sizeof(double) * 2 * kVecGrid, hipMemcpyDeviceToHost, s));
        SVLG_HIP(hipStreamSynchronize(s));
        double num = 0.0, den = 0.0;
        for (int i = 0; i < kVecGrid; ++i) {
            num += hp[2 * i];
            den += hp[2 * i + 1];
        }
        if (!(den > 0.0)) {
            set_error("training pair with e . r <= 0");
            return 13;
        }
        total += weight * num / den;
        // Backward.
        const double scale = weight * 2.0 / den;
        int gc = L & 1;
        hipLaunchKernelGGL(k_gcn_outgrad, dim3(kVecGrid), dim3(kVecBlock), 0, s, n, C, scale, kz, R[p], m.gcn_s.p,
                           theta_w(m), m.gcn_H.p + size_t(L) * plane, m.gcn_G[gc].p, m.gcn_part.p);
        hipLaunchKernelGGL(k_gcn_accum, dim3(C), dim3(256), 0, s, kVecGrid, C, m.gcn_part.p, grad_dev, L * TL);
        for (int l = L; l >= 2; --l) {
            a = gcn_args(m);
            a.in = m.gcn_G[gc].p;
            a.out = m.gcn_G[1 - gc].p;
            a.M1 = theta_A(m, l);
            a.M2 = theta_B(m, l);
            a.v = nullptr;
            a.trans = 1;
            a.hprev = m.gcn_H.p + size_t(l - 1) * plane;
            a.yl = m.gcn_Y.p + size_t(l) * plane;
            a.gpart = m.gcn_part.p;
            if (int rc = launch_gcn(m, G_BWD, false, a)) return rc;
            hipLaunchKernelGGL(k_gcn_accum, dim3(TL), dim3(256), 0, s, m.tiles.ntiles, TL, m.gcn_part.p, grad_dev,
                               (l - 1) * TL);
            gc = 1 - gc;
        }
        hipLaunchKernelGGL(k_gcn_first_grad, dim3(kVecGrid), dim3(kVecBlock), 0, s, n, C, m.gcn_x0.p, m.gcn_G[gc].p,
                           m.gcn_part.p);
        hipLaunchKernelGGL(k_gcn_accum, dim3(C), dim3(256), 0, s, kVecGrid, C, m.gcn_part.p, grad_dev, 2 * C * C);
    }
    SVLG_HIP(hipGetLastError());
    *loss_sum = total;
    return 0;
}

// Chebyshev (Richardson) initial parameters on [lmin, lmax]; extra channels start small.
static void theta_chebyshev(int C, int L, double lmin, double lmax, std::vector<double> &th) {
    th.assign(theta_size(C, L), 0.0);
    const int TL = theta_layer(C);
    std::vector<double> om(L);
    for (int l = 0; l < L; ++l) {
        const double node = 0.5 * (lmax + lmin) + 0.5 * (lmax - lmin) * std::cos(M_PI * (2.0 * l + 1.0) / (2.0 * L));
        om[l] = 1.0 / node;
    }
    // Alternate large and small steps (a simple stable ordering of the Richardson factors).
    std::vector<double> ord;
    for (int i = 0, jx = L - 1; i <= jx; ++i, --jx) {
        ord.push_back(om[i]);
        if (i != jx) ord.push_back(om[jx]);
    }
    std::mt19937_64 rng(20260302);
    std::normal_distribution<double> nd(0.0, 0.02);
    for (int l = 1; l <= L; ++l) {
        double *A = th.data() + size_t(l - 1) * TL, *Bm = A + C * C, *c = A + 2 * C * C;
        for (int i = 0; i < C; ++i)
            for (int k = 0; k < C; ++k) {
                A[i * C + k] = (i == k) ? 1.0 : nd(rng);
                Bm[i * C + k] = (i == 0 && k == 0) ? -ord[l - 1] : nd(rng);
            }
        c[0] = ord[l - 1];
        for (int k = 1; k < C; ++k) c[k] = nd(rng);
    }
    double *w = th.data() + size_t(L) * TL;
    w[0] = 1.0;
    for (int k = 1; k < C; ++k) w[k] = nd(rng);
}

// Eigenvalues of a symmetric tridiagonal matrix (diagonal a, off-diagonal b) by cyclic
// Jacobi rotations on the dense k x k copy (k is a few tens).
static void tridiag_eigs(const std::vector<double> &a, const std::vector<double> &b, std::vector<double> &ev) {
    const int k = int(a.size());
    std::vector<double> T(size_t(k) * k, 0.0);
    for (int i = 0; i < k; ++i) {
        T[size_t(i) * k + i] = a[i];
        if (i + 1 < k) T[size_t(i) * k + i + 1] = T[size_t(i + 1) * k + i] = b[i];
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q) off += T[size_t(p) * k + q] * T[size_t(p) * k + q];
        if (off < 1e-30) break;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q) {
                const double apq = T[size_t(p) * k + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = T[size_t(p) * k + p], aqq = T[size_t(q) * k + q];
                const double th = 0.5 * (aqq - app) / apq;
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int r = 0; r < k; ++r) {
                    const double arp = T[size_t(r) * k + p], arq = T[size_t(r) * k + q];
                    T[size_t(r) * k + p] = c * arp - sn * arq;
                    T[size_t(r) * k + q] = sn * arp + c * arq;
                }
                for (int r = 0; r < k; ++r) {
                    const double apr = T[size_t(p) * k + r], aqr = T[size_t(q) * k + r];
                    T[size_t(p) * k + r] = c * apr - sn * aqr;
                    T[size_t(q) * k + r] = sn * apr + c * aqr;
                }
            }
    }
    ev.resize(k);
    for (int i = 0; i < k; ++i) ev[i] = T[size_t(i) * k + i];
    std::sort(ev.begin(), ev.end());
}

// Extreme eigenvalues of Ehat from `iters` Lanczos steps (FP64 Keff applies, no
// reorthogonalisation: the extreme Ritz values converge first and are what is needed).
// The Ritz values lie inside the spectrum: lmax from below, lmin from above.
static int ehat_bounds(Model &m, int iters, double *lmin, double *lmax) {
    const int n = m.ndofn;
    std::vector<double> hs(n), v(n, 0.0), vprev(n, 0.0), w(n), xs(n);
    SVLG_HIP(hipMemcpyAsync(hs.data(), m.gcn_s.p, sizeof(double) * n, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> ud(-1.0, 1.0);
    double nv = 0.0;
    for (int i = 0; i < n; ++i) {
        v[i] = hs[i] > 0.0 ? ud(rng) : 0.0;
        nv += v[i] * v[i];
    }
    nv = std::sqrt(nv);
    for (double &e : v) e /= nv;
    double *x = m.gcn_tmp[0].p, *y = m.gcn_tmp[1].p;
    std::vector<double> al, be;
    double bprev = 0.0;
    for (int it = 0; it < iters; ++it) {
        for (int i = 0; i < n; ++i) xs[i] = hs[i] * v[i];
        SVLG_HIP(hipMemcpyAsync(x, xs.data(), sizeof(double) * n, hipMemcpyHostToDevice, m.stream));
        if (int rc = keff_apply_node(m, x, y)) return rc;
        SVLG_HIP(hipMemcpyAsync(w.data(), y, sizeof(double) * n, hipMemcpyDeviceToHost, m.stream));
        SVLG_HIP(hipStreamSynchronize(m.stream));
        double a = 0.0;
        for (int i = 0; i < n; ++i) {
            w[i] *= hs[i];
            a += w[i] * v[i];
        }
        double b = 0.0;
        for (int i = 0; i < n; ++i) {
            w[i] -= a * v[i] + bprev * vprev[i];
            b += w[i] * w[i];
        }
        b = std::sqrt(b);
        al.push_back(a);
        if (b < 1e-14 * std::fabs(a) || it == iters - 1) break;
        be.push_back(b);
        for (int i = 0; i < n; ++i) {
            vprev[i] = v[i];
            v[i] = w[i] / b;
        }
        bprev = b;
    }
    be.resize(al.size() > 0 ? al.size() - 1 : 0);
    std::vector<double> ev;
    tridiag_eigs(al, be, ev);
    *lmin = std::max(ev.front(), 1e-6 * ev.back());
    *lmax = ev.back();
    return 0;
}

// The network approximates the inverse of the initial effective stiffness: while it is set
// up or trained, J2 elements enter every operator with their initial tangent (the Newmark
// steps that make training pairs then run on the linear surrogate).
struct InitialOperator {
    Model &m;
    bool saved;
    explicit InitialOperator(Model &mm) : m(mm), saved(mm.nm_initial) { m.nm_initial = true; }
    ~InitialOperator() { m.nm_initial = saved; }
};

}  // namespace svlg

using namespace svlg;

extern "C" {

int svlg_gnn_init(svlg_model *h, int32_t channels, int32_t layers, int32_t mfma, const double *theta) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    InitialOperator initial(m);
    SVLG_CHECK(m.nm_ready, "svlg_nm_init has not been called");
    SVLG_CHECK(channels >= 1 && channels <= kMaxC, "channels must be in [1, 4]");
    SVLG_CHECK(layers >= 2 && layers <= kMaxL, "layers must be in [2, 16]");
    const bool mfma_ok = m.ndim == 3 && m.elem_path >= 2 && m.ngeo == 1 && m.lumped;
    SVLG_CHECK(!(mfma == 1 && !mfma_ok),
               "the MFMA element step needs a Hexa8 mesh with a single element geometry and a lumped mass");
    m.gcn_ready = false;
    m.drop_cg_graphs();  // the network buffers are reallocated below
    m.gcn_C = channels;
    m.gcn_L = layers;
    m.gcn_mfma = (mfma == 1 || (mfma < 0 && mfma_ok)) ? 1 : 0;
    SVLG_CHECK(gcn_lds(m) <= 160 * 1024, "network layer tile needs more than 160 KB of LDS; use smaller tiles");
    hipStream_t s = m.stream;
    const int n = m.ndofn;
    SVLG_HIP(m.gcn_s.alloc(n));
    SVLG_HIP(m.gcn_sf.alloc(n));
    SVLG_HIP(m.gcn_jacf.alloc(n));
    hipLaunchKernelGGL(k_gcn_scale, dim3((n + 255) / 256), dim3(256), 0, s, n, m.jac.p, m.gcn_s.p, m.gcn_sf.p,
                       m.gcn_jacf.p);
    // (lambda, mu) * (1 + 2 ak/dt) per tile element.
    {
        const size_t ne = m.tile_elem_id.size();
        std::vector<double> el(ne), em(ne), ck(ne);
        SVLG_HIP(hipMemcpyAsync(el.data(), m.tiles.elem_lam.p, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
        SVLG_HIP(hipMemcpyAsync(em.data(), m.tiles.elem_mu.p, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
        SVLG_HIP(hipMemcpyAsync(ck.data(), m.coef_keff.p, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
        SVLG_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < ne; ++i) {
            el[i] *= ck[i];
            em[i] *= ck[i];
        }
        SVLG_HIP(m.gcn_elam.upload(el.data(), ne, s));
        SVLG_HIP(m.gcn_emu.upload(em.data(), ne, s));
        std::vector<float> elf(el.begin(), el.end()), emf(em.begin(), em.end());
        SVLG_HIP(m.gcn_elamf.upload(elf.data(), ne, s));
        SVLG_HIP(m.gcn_emuf.upload(emf.data(), ne, s));
        if (m.tiles.hstride > 0) {
            const int nt = m.tiles.ntiles;
            const size_t es = size_t(m.tiles.estride);
            std::vector<int32_t> ep(nt + 1);
            SVLG_HIP(hipMemcpyAsync(ep.data(), m.tiles.elem_ptr.p, sizeof(int32_t) * (nt + 1), hipMemcpyDeviceToHost,
                                    s));
            SVLG_HIP(hipStreamSynchronize(s));
            std::vector<float> lp(size_t(nt) * es, 0.f), mp(size_t(nt) * es, 0.f);
            for (int t = 0; t < nt; ++t)
                for (int i = ep[t]; i < ep[t + 1]; ++i) {
                    lp[size_t(t) * es + (i - ep[t])] = elf[i];
                    mp[size_t(t) * es + (i - ep[t])] = emf[i];
                }
            SVLG_HIP(m.gcn_elamf_pad.upload(lp.data(), lp.size(), s));
            SVLG_HIP(m.gcn_emuf_pad.upload(mp.data(), mp.size(), s));
            SVLG_HIP(hipStreamSynchronize(s));
        }
        std::vector<double> dg(n);
        SVLG_HIP(hipMemcpyAsync(dg.data(), m.dgn.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        SVLG_HIP(hipStreamSynchronize(s));
        std::vector<float> dgf(dg.begin(), dg.end());
        SVLG_HIP(m.gcn_dgnf.upload(dgf.data(), size_t(n), s));
        SVLG_HIP(hipStreamSynchronize(s));  // the host vectors above end with this block
    }
    if (m.gcn_mfma) {
        if (int rc = build_hexa_kmat(m, m.gcn_kmat)) return rc;
        SVLG_HIP(m.gcn_kimg.alloc(4 * 64));
        hipLaunchKernelGGL(k_gcn_kimg, dim3(1), dim3(64), 0, s, m.gcn_kmat.p, m.gcn_kimg.p);
    }
    SVLG_HIP(m.gcn_x0.alloc(n));
    for (int i = 0; i < 2; ++i) SVLG_HIP(m.gcn_h[i].alloc(size_t(channels) * n));
    for (int i = 0; i < 3; ++i) SVLG_HIP(m.gcn_tmp[i].alloc(n));
    const size_t np = std::max<size_t>(size_t(m.tiles.ntiles) * theta_layer(channels), 2 * kVecGrid * kMaxC);
    SVLG_HIP(m.gcn_part.alloc(np));
    SVLG_HIP(m.gcn_w.alloc(theta_size(channels, layers)));
    SVLG_HIP(hipGetLastError());
    if (theta) {
        m.gcn_theta.assign(theta, theta + theta_size(channels, layers));
    } else {
        double lmin = 0.0, lmax = 0.0;
        if (int rc = ehat_bounds(m, 40, &lmin, &lmax)) return rc;
        m.gcn_lmin = lmin;
        m.gcn_lmax = lmax;
        // Widened interval: the Ritz values lie inside the spectrum.
        theta_chebyshev(channels, layers, 0.9 * lmin, 1.05 * lmax, m.gcn_theta);
    }
    if (int rc = upload_theta(m)) return rc;
    if (gcn_fast_lds(m) > 64 * 1024)
        for (const void *fn : {gcn_fast_fn(m, true, false), gcn_fast_fn(m, false, false), gcn_fast_fn(m, true, true),
                               gcn_fast_fn(m, false, true)})
            SVLG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(gcn_fast_lds(m))));
    const size_t lds = gcn_lds(m);
    if (lds > 64 * 1024)
        for (int mode : {G_FWD, G_BWD, G_OUT})
            for (bool mf : {false, true}) {
                if (mf && !m.gcn_mfma) continue;
                SVLG_HIP(hipFuncSetAttribute(gcn_fn(m, mode, mf), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            }
    m.gcn_ready = true;
    return 0;
}

int svlg_gnn_info(svlg_model *h, int32_t *info6, double *bounds2) {
    int32_t *info4 = info6;
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    SVLG_CHECK(m.gcn_ready, "svlg_gnn_init has not been called");
    if (info4) {
        info4[0] = m.gcn_C;
        info4[1] = m.gcn_L;
        info4[2] = m.gcn_mfma;
        info4[3] = theta_size(m.gcn_C, m.gcn_L);
        info4[4] = m.gcn_val_iters[0];
        info4[5] = m.gcn_val_iters[1];
    }
    if (bounds2) {
        bounds2[0] = m.gcn_lmin;
        bounds2[1] = m.gcn_lmax;
    }
    return 0;
}

int svlg_gnn_get_params(svlg_model *h, double *theta) {
    SVLG_CHECK(h && theta, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.gcn_ready, "svlg_gnn_init has not been called");
    std::copy(m.gcn_theta.begin(), m.gcn_theta.end(), theta);
    return 0;
}

int svlg_gnn_set_params(svlg_model *h, const double *theta) {
    SVLG_CHECK(h && theta, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.gcn_ready, "svlg_gnn_init has not been called");
    m.gcn_theta.assign(theta, theta + theta_size(m.gcn_C, m.gcn_L));
    return upload_theta(m);
}

int svlg_gnn_apply(svlg_model *h, int32_t mfma, const double *rfree, double *zfree) {
    SVLG_CHECK(h && rfree && zfree, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.gcn_ready, "svlg_gnn_init has not been called");
    SVLG_CHECK(!(mfma && m.gcn_kmat.p == nullptr), "the MFMA element step is not set up for this model");
    DevBuf<double> f, rn, zn;
    SVLG_HIP(f.upload(rfree, m.nfree, m.stream));
    SVLG_HIP(rn.alloc(m.ndofn));
    SVLG_HIP(zn.alloc(m.ndofn));
    SVLG_HIP(hipMemsetAsync(rn.p, 0, sizeof(double) * m.ndofn, m.stream));
    if (int rc = scatter_free(m, f.p, rn.p)) return rc;
    const int saved = m.gcn_mfma;
    m.gcn_mfma = mfma ? 1 : 0;
    // The output layer also writes flexible-CG partials: point them at a scratch block.
    SVLG_HIP(hipMemsetAsync(m.scal.p, 0, sizeof(double) * SC_COUNT, m.stream));
    DevBuf<double> tp;
    SVLG_HIP(tp.alloc(6 * size_t(m.tiles.ntiles)));
    CgParts c{};
    c.ntiles = m.tiles.ntiles;
    c.tp = tp.p;
    const int rc = gcn_run(m, c, 0, rn.p, rn.p, zn.p, 0);
    m.gcn_mfma = saved;
    if (rc) return rc;
    if (int rc2 = gather_free(m, zn.p, f.p)) return rc2;
    SVLG_HIP(hipMemcpyAsync(zfree, f.p, sizeof(double) * m.nfree, hipMemcpyDeviceToHost, m.stream));
    SVLG_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

// Ensures the training buffers exist.
static int gcn_train_buffers(Model &m) {
    const size_t plane = size_t(m.gcn_C) * m.ndofn;
    if (m.gcn_H.n != plane * (m.gcn_L + 1)) {
        SVLG_HIP(m.gcn_H.alloc(plane * (m.gcn_L + 1)));
        SVLG_HIP(m.gcn_Y.alloc(plane * (m.gcn_L + 1)));
        for (int i = 0; i < 2; ++i) SVLG_HIP(m.gcn_G[i].alloc(plane));
    }
    return 0;
}

int svlg_gnn_loss_grad(svlg_model *h, int32_t npairs, const double *r, const double *e, double *loss,
                       double *grad) {
    SVLG_CHECK(h && r && e && loss, "null argument");
    Model &m = h->m;
    SVLG_CHECK(m.gcn_ready, "svlg_gnn_init has not been called");
    SVLG_CHECK(npairs > 0, "npairs must be positive");
    if (int rc = gcn_train_buffers(m)) return rc;
    const int nt = theta_size(m.gcn_C, m.gcn_L);
    DevBuf<double> pr, pe, f, gd;
    SVLG_HIP(pr.alloc(size_t(npairs) * m.ndofn));
    SVLG_HIP(pe.alloc(size_t(npairs) * m.ndofn));
    SVLG_HIP(f.alloc(m.nfree));
    SVLG_HIP(gd.alloc(nt));
    SVLG_HIP(hipMemsetAsync(pr.p, 0, sizeof(double) * pr.n, m.stream));
    SVLG_HIP(hipMemsetAsync(pe.p, 0, sizeof(double) * pe.n, m.stream));
    SVLG_HIP(hipMemsetAsync(gd.p, 0, sizeof(double) * nt, m.stream));
    std::vector<const double *> R(npairs), E(npairs);
    for (int p = 0; p < npairs; ++p) {
        SVLG_HIP(hipMemcpyAsync(f.p, r + size_t(p) * m.nfree, sizeof(double) * m.nfree, hipMemcpyHostToDevice,
                                m.stream));
        if (int rc = scatter_free(m, f.p, pr.p + size_t(p) * m.ndofn)) return rc;
        SVLG_HIP(hipMemcpyAsync(f.p, e + size_t(p) * m.nfree, sizeof(double) * m.nfree, hipMemcpyHostToDevice,
                                m.stream));
        if (int rc = scatter_free(m, f.p, pe.p + size_t(p) * m.ndofn)) return rc;
        SVLG_HIP(hipStreamSynchronize(m.stream));
        R[p] = pr.p + size_t(p) * m.ndofn;
        E[p] = pe.p + size_t(p) * m.ndofn;
    }
    double ls = 0.0;
    if (int rc = gcn_loss_grad(m, npairs, R.data(), E.data(), 1.0 / npairs, gd.p, &ls)) return rc;
    *loss = ls;
    if (grad) {
        SVLG_HIP(hipMemcpyAsync(grad, gd.p, sizeof(double) * nt, hipMemcpyDeviceToHost, m.stream));
        SVLG_HIP(hipStreamSynchronize(m.stream));
    }
    return 0;
}

int svlg_nm_steps(svlg_model *h, int32_t k0, int32_t nsteps, int32_t nrec, const int32_t *rec_dofs, double *hist,
                  int32_t *iters, double *relres);

// Generates residual/solution pairs and trains the network with Adam. From rest, steps
// k = 1..nsteps are solved with Jacobi-PCG; every `every`-th step contributes the pair
// (Feff, dU) of the model's own Newmark solve, `sweeps` Jacobi-smoothed pairs
// (Feff - Keff x_j, dU - x_j), x_j = x_(j-1) + D^-1 (Feff - Keff x_(j-1)) / 2, and one
// white pair (Keff e, e), e = D^-1/2 w with w uniform: the Newmark pairs carry the smooth
// part of the spectrum that the solves meet, the white pairs keep every mode in the loss.
// The trained parameters are kept only if a validation solve (the first Newmark pair's Feff,
// flexible CG at the model's tolerance) needs fewer iterations than with the parameters the
// training started from. The state is back at rest afterwards. loss_hist[epochs] (may be
// NULL) receives the mean loss of each epoch; accepted (may be NULL) 1 if kept.
int svlg_gnn_train(svlg_model *h, int32_t nsteps, int32_t every, int32_t sweeps, int32_t epochs, double lr,
                   double *loss_hist, int32_t *accepted) {
    SVLG_CHECK(h, "null model");
    Model &m = h->m;
    InitialOperator initial(m);
    SVLG_CHECK(m.gcn_ready, "svlg_gnn_init has not been called");
    SVLG_CHECK(nsteps > 0 && every > 0 && sweeps >= 0 && epochs >= 0 && lr > 0.0, "invalid training settings");
    if (int rc = gcn_train_buffers(m)) return rc;
    hipStream_t s = m.stream;
    const int n = m.ndofn;
    const int nbase = nsteps / every;
    const int npairs = nbase * (2 + sweeps);
    SVLG_CHECK(nbase > 0, "no training pairs (nsteps < every)");
    DevBuf<double> pr, pe;
    SVLG_HIP(pr.alloc(size_t(npairs) * n));
    SVLG_HIP(pe.alloc(size_t(npairs) * n));
    const int nb = (n + 255) / 256;
    int p = 0;
    const int saved_pc = m.nm_precond;
    m.nm_precond = SVLG_PRECOND_JACOBI;
    for (int k = 1; k <= nsteps; ++k) {
        if (int rc = svlg_nm_steps(h, k, 1, 0, nullptr, nullptr, nullptr, nullptr)) {
            m.nm_precond = saved_pc;
            return rc;
        }
        if (k % every) continue;
        double *rb = pr.p + size_t(p) * n, *eb = pe.p + size_t(p) * n;
        SVLG_HIP(hipMemcpyAsync(rb, m.cg_b.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        SVLG_HIP(hipMemcpyAsync(eb, m.ndU.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        ++p;
        for (int sw = 0; sw < sweeps; ++sw) {
            double *rp = pr.p + size_t(p - 1) * n, *ep = pe.p + size_t(p - 1) * n;
            double *rn = pr.p + size_t(p) * n, *en = pe.p + size_t(p) * n;
            double *d = m.gcn_tmp[0].p, *kd = m.gcn_tmp[1].p;
            SVLG_HIP(hipMemcpyAsync(d, rp, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(k_gcn_scale_mul, dim3(nb), dim3(256), 0, s, n, m.jac.p, 0.5, d);
            if (int rc = keff_apply_node(m, d, kd)) return rc;
            SVLG_HIP(hipMemcpyAsync(rn, rp, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            SVLG_HIP(hipMemcpyAsync(en, ep, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, s, n, -1.0, kd, 1.0, rn);
            hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, s, n, -1.0, d, 1.0, en);
            ++p;
        }
        double *rw = pr.p + size_t(p) * n, *ew = pe.p + size_t(p) * n;
        hipLaunchKernelGGL(k_gcn_white, dim3(nb), dim3(256), 0, s, n, (unsigned long long)(1000 + k), m.gcn_s.p, ew);
        if (int rc = keff_apply_node(m, ew, rw)) return rc;
        ++p;
    }
    m.nm_precond = saved_pc;
    for (DevBuf<double> *b : {&m.nU, &m.nV, &m.nA, &m.ndU})
        SVLG_HIP(hipMemsetAsync(b->p, 0, sizeof(double) * n, s));
    m.nm_steps_done = 0;
    // Validation: flexible-CG iterations on the first Newmark right-hand side.
    auto validate = [&](int *its) -> int {
        SVLG_HIP(hipMemsetAsync(m.gcn_tmp[2].p, 0, sizeof(double) * n, s));
        double rel = 0.0;
        const int rc = cg_solve(m, pr.p, m.gcn_tmp[2].p, m.nm_rtol, m.nm_maxit, SVLG_PRECOND_LEARNED, its, &rel);
        if (rc == 12) {  // not converged within maxit: count as maxit
            *its = m.nm_maxit;
            return 0;
        }
        return rc;
    };
    int it0 = 0;
    if (int rc = validate(&it0)) return rc;
    const std::vector<double> theta0 = m.gcn_theta;
    // Adam over full-batch gradients.
    const int nt = theta_size(m.gcn_C, m.gcn_L);
    DevBuf<double> gd;
    SVLG_HIP(gd.alloc(nt));
    std::vector<const double *> R(npairs), E(npairs);
    for (int i = 0; i < npairs; ++i) {
        R[i] = pr.p + size_t(i) * n;
        E[i] = pe.p + size_t(i) * n;
    }
    std::vector<double> g(nt), m1(nt, 0.0), m2(nt, 0.0);
    const double b1 = 0.9, b2 = 0.999, eps = 1e-12;
    for (int ep = 0; ep < epochs; ++ep) {
        SVLG_HIP(hipMemsetAsync(gd.p, 0, sizeof(double) * nt, s));
        double ls = 0.0;
        if (int rc = gcn_loss_grad(m, npairs, R.data(), E.data(), 1.0 / npairs, gd.p, &ls)) return rc;
        SVLG_HIP(hipMemcpyAsync(g.data(), gd.p, sizeof(double) * nt, hipMemcpyDeviceToHost, s));
        SVLG_HIP(hipStreamSynchronize(s));
        if (loss_hist) loss_hist[ep] = ls;
        const double c1 = 1.0 - std::pow(b1, ep + 1), c2 = 1.0 - std::pow(b2, ep + 1);
        for (int i = 0; i < nt; ++i) {
            m1[i] = b1 * m1[i] + (1.0 - b1) * g[i];
            m2[i] = b2 * m2[i] + (1.0 - b2) * g[i] * g[i];
            m.gcn_theta[i] -= lr * (m1[i] / c1) / (std::sqrt(m2[i] / c2) + eps);
        }
        if (int rc = upload_theta(m)) return rc;
    }
    int it1 = 0;
    if (int rc = validate(&it1)) return rc;
    const bool keep = it1 < it0;
    if (!keep) {
        m.gcn_theta = theta0;
        if (int rc = upload_theta(m)) return rc;
    }
    m.gcn_val_iters[0] = it0;
    m.gcn_val_iters[1] = it1;
    if (accepted) *accepted = keep ? 1 : 0;
    return 0;
}

}  // extern "C"
