// potrf_diag.hip -- inverses of the diagonal blocks of a factor, used by the solve.
//
// The diagonal-block Cholesky itself (llt_inplace<double,Lower>::unblocked on A11,
// Eigen/src/Cholesky/LLT.h:312-336, called at :363) and the panel solve (LLT.h:364) run
// inside the fused factorization step (factor_step.hip, potrf_device.hpp):
//   * the 128 x 128 block is held in LDS and factored in 16-column panels; one wave holds
//     a panel in registers, two rows per lane, and runs the column loop of LLT.h:320-334
//     (pivot test, sqrt, scale, rank-1 update) with lane broadcasts (readlane) instead of
//     barriers; then that wave inverts the 16 x 16 leaf while the other three apply the
//     rank-16 update to the rest of the block with FP64 MFMA from LDS;
//   * the pivot test is Eigen's `x <= 0 -> fail at column k` (LLT.h:327-330), so a NaN
//     pivot passes as in Eigen; the first failing column goes to `info` (LAPACK k+1); a last
//     block narrower than 128 is padded with the identity (L_pad = [L 0; 0 I]);
//   * the panel solve A21 <- A21 L11^{-T} (TriangularSolverMatrix.h:186-329) is blocked by
//     16 columns with the leaf inverses; the MFMA result layout is the operand layout of the
//     next k-step, so solved blocks feed the next MFMAs straight from registers.
// Run on an identity panel, the same solve yields L_bb^{-T}: block_inverse_kernel.
#include "potrf_device.hpp"

namespace mi {

// Workgroup b: X_b^T = L_bb^{-T} to Wout + b*kInvDoubles (ld 128) and X_b right after it.
template <typename T>
__global__ __launch_bounds__(256) void block_inverse_kernel(const T *__restrict__ L, int64_t ldl, int64_t n,
                                                            const double *__restrict__ Wleaf,
                                                            double *__restrict__ Wout, const int *info)
{
    __shared__ double smem[kSmem];
    if (factor_failed(info)) return;
    const int b = (int)blockIdx.x;
    const int64_t kb = (int64_t)b * 128;
    const int bs = (int)((n - kb) < 128 ? (n - kb) : 128);
    trsm_rows<true, T>(smem, (T *)nullptr, 0, 128, L + kb + kb * ldl, ldl, bs,
                       Wleaf + (size_t)b * kPanelDoubles + kLeafOffset, Wout + (size_t)b * kInvDoubles);
}

// Leaf inverses of the diagonal blocks of an already factored matrix: workgroup b, wave w
// inverts leaves 2w and 2w+1 of block b (lane i < 16 holds row i of the leaf).
template <typename T>
__global__ __launch_bounds__(256) void leaf_inverse_kernel(const T *__restrict__ L, int64_t lda, int64_t n,
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
            const double lraw = (double)L[kb + (in ? r : 0) + (kb + (in ? cc : 0)) * lda];
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
                Wleaf[(size_t)blockIdx.x * kPanelDoubles + kLeafOffset + p * 256 + l + 16 * c] = (c <= l) ? x[c] : 0.0;
        }
    }
}

template <typename T>
void launch_block_inverses(const T *L, int64_t lda, int64_t n, const double *Wleaf, double *Winv, const int *info,
                           hipStream_t s)
{
    const unsigned nblk = (unsigned)((n + 127) / 128);
    hipLaunchKernelGGL(block_inverse_kernel<T>, dim3(nblk), dim3(256), 0, s, L, lda, n, Wleaf, Winv, info);
}

template <typename T>
void launch_leaf_inverses(const T *L, int64_t lda, int64_t n, double *Wleaf, hipStream_t s)
{
    const unsigned nblk = (unsigned)((n + 127) / 128);
    hipLaunchKernelGGL(leaf_inverse_kernel<T>, dim3(nblk), dim3(256), 0, s, L, lda, n, Wleaf);
}

template void launch_block_inverses<double>(const double *, int64_t, int64_t, const double *, double *, const int *,
                                            hipStream_t);
template void launch_block_inverses<float>(const float *, int64_t, int64_t, const double *, double *, const int *,
                                           hipStream_t);
template void launch_leaf_inverses<double>(const double *, int64_t, int64_t, double *, hipStream_t);
template void launch_leaf_inverses<float>(const float *, int64_t, int64_t, double *, hipStream_t);

}  // namespace mi
