// potrf_diag.hip -- diagonal-block Cholesky of one 128 x 128 block, the panel TRSM below
// it, and the diagonal-block inverses used by the solve.
//
// potrf128_kernel replaces llt_inplace<double,Lower>::unblocked on the diagonal block A11
// of llt_inplace::blocked (Eigen/src/Cholesky/LLT.h:312-336, called at :363).  The block is
// held in LDS and factored in 16-column panels:
//   * panel factor: one wave holds the panel (rows c0..127, 16 columns) in registers, two
//     rows per lane, and runs the column loop of LLT.h:320-334 (pivot test, sqrt, scale,
//     rank-1 update) with lane broadcasts (readlane) instead of barriers;
//   * then, concurrently: that wave inverts the 16 x 16 diagonal leaf L_pp while the other
//     three waves apply the rank-16 update to the rest of the block with FP64 MFMA from LDS.
// The pivot test is Eigen's `x <= 0 -> fail at column k` (LLT.h:327-330), so a NaN pivot
// passes as in Eigen; the first failing column goes to `info` (LAPACK k+1).  A last block
// narrower than 128 is padded with the identity (L_pad = [L 0; 0 I]).
//
// trsm_lt_kernel replaces the right-side triangular solve of LLT.h:364
// (A11^T.triangularView<Upper>().solveInPlace<OnTheRight>(A21), TriangularSolverMatrix.h:
// 186-329): P <- P L11^{-T}, blocked by 16 columns with the leaf inverses.  Each wave owns
// whole rows, and the MFMA result layout (lane l, element r = C[l%16][4r + l/16]) is exactly
// the operand layout of the next k-step, so solved blocks feed the next MFMAs straight from
// registers: no LDS, no barriers.  Run on an identity panel it yields L11^{-T} itself.
#include "potrf_device.hpp"

namespace mi {

__global__ __launch_bounds__(256) void potrf128_kernel(double *__restrict__ A, int64_t lda, int64_t n, int64_t k0,
                                                       double *__restrict__ Wleaf, int *info)
{
    __shared__ double S[kSmem];
    __shared__ int s_fail;
    if (factor_failed(info)) return;
    const int bs = (int)((n - k0) < 128 ? (n - k0) : 128);
    double *Ab = A + k0 + k0 * lda;
    block_to_lds(S, Ab, lda, bs);
    const int f = potrf_in_lds(S, Wleaf + (size_t)(k0 / 128) * kLeafDoubles, &s_fail);
    if (f >= 0) {
        if (threadIdx.x == 0) atomicCAS(info, 0, (int)(k0 + f + 1));
        return;
    }
    lds_to_block(S, Ab, lda, bs);
}

// Standalone panel solve (rows k0+128.. of block column k0), or, IDENTITY, the inverse
// L_bb^{-T} / L_bb^{-1} of every diagonal block b = blockIdx.x into Wout.
template <bool IDENTITY>
__global__ __launch_bounds__(256) void trsm_lt_kernel(double *__restrict__ P, int64_t ldp, int m,
                                                      const double *__restrict__ Lbase, int64_t ldl, int64_t n,
                                                      const double *__restrict__ Wleaf, double *__restrict__ Wout,
                                                      const int *info)
{
    __shared__ double smem[kSmem];
    if (factor_failed(info)) return;
    const int b = IDENTITY ? (int)blockIdx.x : 0;
    const int64_t kb = (int64_t)b * 128;
    const int bs = IDENTITY ? (int)((n - kb) < 128 ? (n - kb) : 128) : 128;
    const int r0 = IDENTITY ? 0 : (int)blockIdx.x * 128;
    trsm_rows<IDENTITY>(smem, IDENTITY ? P : P + r0, ldp, m - r0, IDENTITY ? Lbase + kb + kb * ldl : Lbase, ldl, bs,
                        Wleaf + (size_t)b * kLeafDoubles, IDENTITY ? Wout + (size_t)b * kInvDoubles : nullptr);
}

// One block step of the factorization, fused: the update of block column k0 by the previous
// panel (columns k0-128..k0-1), the diagonal-block factor, and the panel solve below it.
// Workgroups take roles from a ticket counter (so the diagonal role, ticket 0, is always
// running before any other role waits for it):
//   role 0   : C = A_kk - P_0 P_0^T straight into LDS, Cholesky in LDS, L11 and the leaf
//              inverses to memory, then publish `flag` = epoch (release);
//   role g>0 : rows k0+128g..+127: C = A_gk - P_g P_0^T written back, wait for the flag,
//              then rows <- rows * L11^{-T}.
// The diagonal role publishes even on a failing pivot, so that no waiter is left behind;
// waiters then see `info` and stop.  (LLT.h:350-366: the update of block column k by all
// earlier panels precedes its unblocked factor and the right-side solve.)
__global__ __launch_bounds__(256) void panel_fused_kernel(double *__restrict__ A, int64_t lda, int64_t n, int64_t k0,
                                                          double *__restrict__ Wleaf, int *info, uint32_t *tickets,
                                                          uint32_t ticket_base, uint32_t *flag, uint32_t epoch,
                                                          int *status)
{
    __shared__ double smem[kSmem];
    __shared__ int s_role, s_fail, s_ok;
    // every workgroup draws its ticket, even when it then stops: the host counts them
    if (threadIdx.x == 0) s_role = (int)(atomicAdd(tickets, 1u) - ticket_base);
    __syncthreads();
    if (factor_failed(info)) return;
    panel_role(s_role, smem, &s_fail, &s_ok, A, lda, n, k0, Wleaf, info, flag, epoch, status);
}

// Leaf inverses of the diagonal blocks of an already factored matrix: workgroup b, wave w
// inverts leaves 2w and 2w+1 of block b (lane i < 16 holds row i of the leaf).
__global__ __launch_bounds__(256) void leaf_inverse_kernel(const double *__restrict__ L, int64_t lda, int64_t n,
                                                           double *__restrict__ Wleaf)
{
    const int l = lane_id(), w = threadIdx.x >> 6;
    const int64_t kb = (int64_t)blockIdx.x * 128;
    const int bs = (int)((n - kb) < 128 ? (n - kb) : 128);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = 2 * w + h;
        const int c0 = 16 * p;
        double v[16], x[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int r = c0 + l, cc = c0 + c;
            double val = 0.0;
            const bool in = l < 16 && r < bs && cc < bs;
            const double lraw = L[kb + (in ? r : 0) + (kb + (in ? cc : 0)) * lda];
            if (l < 16 && c <= l) val = in ? lraw : (r == cc ? 1.0 : 0.0);
            v[c] = val;
            x[c] = (c == l) ? 1.0 : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const double lkk = readlane_d(v[k], k);
            if (l == k) {
#pragma unroll
                for (int c = 0; c <= k; ++c) x[c] = x[c] / lkk;
            }
#pragma unroll
            for (int c = 0; c <= k; ++c) {
                const double xkc = readlane_d(x[c], k);
                if (l > k) x[c] -= v[k] * xkc;
            }
        }
        if (l < 16) {
#pragma unroll
            for (int c = 0; c < 16; ++c)
                Wleaf[(size_t)blockIdx.x * 8 * 256 + p * 256 + l + 16 * c] = (c <= l) ? x[c] : 0.0;
        }
    }
}


void launch_panel_fused(double *A, int64_t lda, int64_t n, int64_t k0, double *Wleaf, int *info, uint32_t *tickets,
                        uint32_t ticket_base, uint32_t *flag, uint32_t epoch, int *status, hipStream_t s)
{
    const unsigned G = (unsigned)((n - k0 + 127) / 128);
    hipLaunchKernelGGL(panel_fused_kernel, dim3(G), dim3(256), 0, s, A, lda, n, k0, Wleaf, info, tickets, ticket_base,
                       flag, epoch, status);
}

void launch_potrf_diag(double *A, int64_t lda, int64_t n, int64_t k0, double *Wleaf, int *info, hipStream_t s)
{
    hipLaunchKernelGGL(potrf128_kernel, dim3(1), dim3(256), 0, s, A, lda, n, k0, Wleaf, info);
}

void launch_trsm_panel(double *A, int64_t lda, int64_t n, int64_t k0, const double *Wleaf, const int *info,
                       hipStream_t s)
{
    const int64_t rs = n - k0 - 128;
    if (rs <= 0) return;
    hipLaunchKernelGGL(trsm_lt_kernel<false>, dim3((unsigned)((rs + 127) / 128)), dim3(256), 0, s,
                       A + (k0 + 128) + k0 * lda, lda, (int)rs, (const double *)(A + k0 + k0 * lda), lda, n,
                       Wleaf + (size_t)(k0 / 128) * 8 * 256, (double *)nullptr, info);
}

void launch_block_inverses(const double *L, int64_t lda, int64_t n, const double *Wleaf, double *Winv,
                           const int *info, hipStream_t s)
{
    const unsigned nblk = (unsigned)((n + 127) / 128);
    hipLaunchKernelGGL(trsm_lt_kernel<true>, dim3(nblk), dim3(256), 0, s, (double *)nullptr, (int64_t)0, 0, L, lda,
                       n, Wleaf, Winv, info);
}

void launch_leaf_inverses(const double *L, int64_t lda, int64_t n, double *Wleaf, hipStream_t s)
{
    const unsigned nblk = (unsigned)((n + 127) / 128);
    hipLaunchKernelGGL(leaf_inverse_kernel, dim3(nblk), dim3(256), 0, s, L, lda, n, Wleaf);
}

}  // namespace mi
