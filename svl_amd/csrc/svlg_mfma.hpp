// svlg_mfma.hpp -- BF16 MFMA element step for Hexa8 meshes with a single element geometry
// (gfx950 v_mfma_f32_16x16x32_bf16, FP32 accumulation).
//
// Every element shares K_lambda and K_mu (the stiffness for unit Lame constants), so
// f_e = lambda_e K_lambda u_e + mu_e K_mu u_e. One wave handles 16 elements at a time:
// B (32 x 16) holds the 24 DOF values of each element (8 zero rows) and three 16 x 32 A blocks
// hold [K_lambda rows 0-15], [K_mu rows 0-15], [K_lambda rows 16-23; K_mu rows 16-23].
// Operand layouts: lane l holds A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16] and
// D[4(l/16)..+3][l%16]; only the agreement of the k-slot order between A and B matters.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace svlg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// kmat: [2][24][24] FP32, K_lambda then K_mu, row = output DOF (corner*3 + component).
__device__ __forceinline__ void mfma_load_kblocks(const float *kmat, bf16x8 (&A)[3]) {
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = 8 * g + j;
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        if (k < 24) {
            a0 = kmat[m * 24 + k];
            a1 = kmat[576 + m * 24 + k];
            a2 = m < 8 ? kmat[(16 + m) * 24 + k] : kmat[576 + (8 + m) * 24 + k];
        }
        A[0][j] = __bf16(a0);
        A[1][j] = __bf16(a1);
        A[2][j] = __bf16(a2);
    }
}

// Element forces of the tile elements [e0, e0 + ne) from the staged operand sU (slot-major,
// 3 values per staged node), written to the corner slots sIn[(corner*3 + d) * MO + slot] of
// owned corners (slot < nown). If ftol >= 0 entries with |f| <= ftol are stored as 0 (the
// Assembler filter). Must be called by every wave of the workgroup.
template <class T>
__device__ __forceinline__ void mfma_hexa_elements(const uint16_t *lconn, const double *elam, const double *emu,
                                                   int e0, int ne, int nown, const T *sU, double *sIn, int MO,
                                                   const bf16x8 (&A)[3], double ftol) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    for (int eb = wave * 16; eb < ne; eb += nw * 16) {  // wave-uniform bound
        const int el = eb + n;
        const bool valid = el < ne;
        int lc[8];
        double lam = 0.0, mu = 0.0;
        if (valid) {
            const uint4 pk = *reinterpret_cast<const uint4 *>(lconn + size_t(e0 + el) * 8);
            const uint32_t w4[4] = {pk.x, pk.y, pk.z, pk.w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                lc[2 * c] = int(w4[c] & 0xffffu);
                lc[2 * c + 1] = int(w4[c] >> 16);
            }
            lam = elam[e0 + el];
            mu = emu[e0 + el];
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) lc[c] = 0;
        }
        bf16x8 Bv;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * g + j;
            float v = 0.f;
            if (valid && k < 24) v = float(sU[lc[k / 3] * 3 + (k % 3)]);
            Bv[j] = __bf16(v);
        }
        const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
        const f32x4 d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], Bv, zero, 0, 0, 0);
        const f32x4 d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], Bv, zero, 0, 0, 0);
        const f32x4 d2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[2], Bv, zero, 0, 0, 0);
        float o2[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) o2[i] = __shfl_down(d2[i], 32, 64);
        if (valid) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 4 * g + i, c = row / 3;
                const double f = lam * double(d0[i]) + mu * double(d1[i]);
                if (lc[c] < nown) sIn[row * MO + lc[c]] = (ftol < 0.0 || fabs(f) > ftol) ? f : 0.0;
            }
            if (g < 2) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 16 + 4 * g + i, c = row / 3;
                    const double f = lam * double(d2[i]) + mu * double(o2[i]);
                    if (lc[c] < nown) sIn[row * MO + lc[c]] = (ftol < 0.0 || fabs(f) > ftol) ? f : 0.0;
                }
            }
        }
    }
}

}  // namespace svlg
