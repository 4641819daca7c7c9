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

// Workgroup b: X_b^T = L_bb^{-T} to Wout + b*kInvDoubles (ld 128) and X_b at + kInvX.
// Batched launches (every kernel of this file): system z = blockIdx.z, its pointers advanced
// by z times their strides (s*), its info at info + z.
template <typename T>
__global__ __launch_bounds__(256) void block_inverse_kernel(const T *__restrict__ L, int64_t ldl, int64_t n,
                                                            const double *__restrict__ Wleaf,
                                                            double *__restrict__ Wout, const int *info, int64_t sL,
                                                            int64_t sWl, int64_t sWo)
{
    __shared__ double smem[kSmem];
    const int64_t z = blockIdx.z;
    L += z * sL;
    Wleaf += z * sWl;
    Wout += z * sWo;
    if (info) info += z;
    if (factor_failed(info)) return;
    const int b = (int)blockIdx.x;
    const int64_t kb = (int64_t)b * 128;
    const int bs = (int)((n - kb) < 128 ? (n - kb) : 128);
    trsm_rows<true, T>(smem, (T *)nullptr, 0, 128, L + kb + kb * ldl, ldl, bs,
                       Wleaf + (size_t)b * kPanelDoubles + kLeafOffset, Wout + (size_t)b * kInvDoubles);
}

// Couplings of the block substitution: with them a block of the solution depends on its three
// predecessors through one product each,
//   Y_i = X_i (B_i - sum_{j<i-3} L_ij Y_j) - M3_i Y_{i-3} - M2_i Y_{i-2} - M_i Y_{i-1},
// so everything else is done before they are published (solve_couplings: none below 8 blocks,
// the third only from 16 blocks on).  Workgroup q nblk + b, distance d = 1 + q / 2: q even, M_b / M2_b / M3_b = X_b L_{b,b-d} (b >= d); q odd, N_b / N2_b / N3_b
// = X_b^T L_{b+d,b}^T (b + d < nblk).  128 x 128 each at Winv[b] + kInvM / kInvN / kInvM2 /
// kInvN2 / kInvM3 / kInvN3 (ld 128); entries of L outside the matrix read as 0.  4 waves of
// 64 x 64 (4 x 4 FP64 MFMA tiles), depth 128.
template <typename T>
__global__ __launch_bounds__(256) void block_coupling_kernel(const T *__restrict__ L, int64_t ldl, int64_t n,
                                                             double *__restrict__ Winv, const int *info, int64_t sL,
                                                             int64_t sWo)
{
    const int64_t z = blockIdx.z;
    L += z * sL;
    Winv += z * sWo;
    if (info) info += z;
    if (factor_failed(info)) return;
    const int nblk = (int)((n + 127) / 128);
    const int q = (int)blockIdx.x / nblk, b = (int)blockIdx.x % nblk;
    const bool fwd = (q & 1) == 0;
    const int d = 1 + q / 2;
    if (fwd ? b < d : b + d > nblk - 1) return;
    double *Wb = Winv + (size_t)b * kInvDoubles;
    const double *P = Wb + (fwd ? kInvX : 0);  // P[i][k] = X_b[i][k] or X_b^T[i][k] at P + i + 128 k
    // Q[j][k] = L[128 b + k][128 (b-d) + j] (forward; rows k < lim exist) or
    //           L[128 (b+d) + j][128 b + k] (backward; rows j < lim exist)
    const int64_t r0 = fwd ? 128 * (int64_t)b : 128 * (int64_t)(b + d);
    const int64_t c0 = fwd ? 128 * (int64_t)(b - d) : 128 * (int64_t)b;
    const int lim = (int)((n - r0) < 128 ? (n - r0) : 128);
    const int l = lane_id(), w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1, kl = l >> 4;
    dvec4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[a][q] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
    for (int s = 0; s < 32; ++s) {
        const int k = 4 * s + kl;
        double pv[4], qv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) pv[a] = P[64 * wr + 16 * a + (l & 15) + 128 * k];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = 64 * wc + 16 * q + (l & 15);
            const int rr = fwd ? k : j;  // the row of L read, relative to r0
            const bool in = rr < lim;
            const int64_t off = fwd ? (r0 + (in ? k : 0)) + (c0 + j) * ldl : (r0 + (in ? j : 0)) + (c0 + k) * ldl;
            const double v = (double)L[off];
            qv[q] = in ? v : 0.0;
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[a][q] = mfma_f64(qv[q], pv[a], acc[a][q]);
    }
    double *C = Wb + (fwd ? (d == 1 ? kInvM : d == 2 ? kInvM2 : kInvM3) : (d == 1 ? kInvN : d == 2 ? kInvN2 : kInvN3));
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                C[(64 * wr + 16 * a + (l & 15)) + 128 * (64 * wc + 16 * q + 4 * r + kl)] = acc[a][q][r];
}

// Leaf inverses of the diagonal blocks of an already factored matrix: workgroup b, wave w
// inverts leaves 2w and 2w+1 of block b (lane i < 16 holds row i of the leaf).
template <typename T>
__global__ __launch_bounds__(256) void leaf_inverse_kernel(const T *__restrict__ L, int64_t lda, int64_t n,
                                                           double *__restrict__ Wleaf, int64_t sL, int64_t sWl)
{
    L += (int64_t)blockIdx.z * sL;
    Wleaf += (int64_t)blockIdx.z * sWl;
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
                           hipStream_t s, bool couplings, const Batch &bt)
{
    const unsigned nblk = (unsigned)((n + 127) / 128);
    hipLaunchKernelGGL(block_inverse_kernel<T>, dim3(nblk, 1, bt.count), dim3(256), 0, s, L, lda, n, Wleaf, Winv,
                       info, bt.a, bt.leaf, bt.inv);
    if (couplings && solve_couplings(nblk) > 0)
        hipLaunchKernelGGL(block_coupling_kernel<T>, dim3(2 * solve_couplings(nblk) * nblk, 1, bt.count), dim3(256), 0,
                           s, L, lda, n, Winv, info, bt.a, bt.inv);
}

template <typename T>
void launch_leaf_inverses(const T *L, int64_t lda, int64_t n, double *Wleaf, hipStream_t s, const Batch &bt)
{
    const unsigned nblk = (unsigned)((n + 127) / 128);
    hipLaunchKernelGGL(leaf_inverse_kernel<T>, dim3(nblk, 1, bt.count), dim3(256), 0, s, L, lda, n, Wleaf, bt.a,
                       bt.leaf);
}

template void launch_block_inverses<double>(const double *, int64_t, int64_t, const double *, double *, const int *,
                                            hipStream_t, bool, const Batch &);
template void launch_block_inverses<float>(const float *, int64_t, int64_t, const double *, double *, const int *,
                                           hipStream_t, bool, const Batch &);
template void launch_leaf_inverses<double>(const double *, int64_t, int64_t, double *, hipStream_t, const Batch &);
template void launch_leaf_inverses<float>(const float *, int64_t, int64_t, double *, hipStream_t, const Batch &);

}  // namespace mi
