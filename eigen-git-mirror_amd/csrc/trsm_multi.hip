// trsm_multi.hip -- Y <- B L^{-1} for many rows (the triangular solves of the Cholesky
// adjoint, where B is n x n).
//
// Right-side block substitution by the factorization's 128-column blocks, Y L = B with L
// lower (TriangularSolverMatrix.h:186-329, OnTheRight):
//   Y_i = (B_i - sum_{k>i} Y_k L_ki) X_ii,   X_ii = L_ii^{-1},   Y_i = Y[:, block i]
// organised for many rows: one launch per column block i, from the last, with a one-block
// lookahead, so that nothing inside a launch waits for anything else in it.  Launch i:
//   panel tiles (row tile r, block i): B[r, i] -= Y[r, i+1] L_{i+1,i}, then
//            Y[r, i] = B[r, i] X_ii (X_ii lower: half its k-steps are skipped);
//   trailing tiles (row tile r, block j < i): B[r, j] -= Y[r, i+1] L_{i+1,j}.
// Y_{i+1} was finished by the previous launch, and every block j receives Y_k's update in the
// launch that finishes Y_{k-1}.  Both MFMA operands are row-contiguous (element (row, k) at
// base + row + k * ld): Y's rows, and L_kj through the transposed copy Lt = L^T (the product
// runs over L's row index, which is contiguous only in L^T's columns).
//
// UPPER: only the block-upper tiles r <= j are computed.  For S = Z L^{-1} with
// Z = (M L^{-1})^T (M symmetric), S = L^{-T} M L^{-1} is symmetric and its block-upper part
// depends only on block-upper tiles of the later column blocks, so half the tiles suffice.
#include "common.hpp"
#include "kernels.hpp"
#include "mfma_tile.hpp"

namespace mi {

// Launch for column block i (grid: the items below).  B (nrow x n, ld ldb) holds B, becoming
// Y; Lt = L^T (ld ldlt; only Lt[j][k] for k > j, i.e. L's lower part, is read); Winv: per
// block b, X_b^T (ld 128) at Winv + b * kInvDoubles.
//
// Panel items are 32 rows x 128 columns (rows of a right-side solve are independent), each
// wave 32 x 32: the panel's two products are on the critical path of every launch, and a
// quarter of a 128-row tile's work per workgroup shortens it about fourfold.  Trailing items
// are 128 x 128 tiles (4 waves of 64 x 64).
template <bool UPPER>
__global__ __launch_bounds__(256, 2) void trsm_xl_step_kernel(const double *__restrict__ Lt, int64_t ldlt,
                                                              int64_t n, double *__restrict__ B, int64_t ldb,
                                                              int64_t nrow, const double *__restrict__ Winv, int i,
                                                              int64_t sLt, int64_t sB, int64_t sWinv)
{
    {  // batched launches: system z = blockIdx.z
        const int64_t z = blockIdx.z;
        Lt += z * sLt;
        B += z * sB;
        Winv += z * sWinv;
    }
    const int nblk = (int)((n + 127) / 128);
    const int nrt = (int)((nrow + 127) / 128);
    const int item = (int)blockIdx.x;
    // panel rows: all rows, or (UPPER) the rows of blocks 0..i
    const int prow = UPPER ? ((128 * (i + 1)) < nrow ? 128 * (i + 1) : (int)nrow) : (int)nrow;
    const int npanel = (prow + 31) / 32;
    const int w = threadIdx.x >> 6;
    const int k = i + 1;  // the block whose update is applied
    const int k0 = 128 * k;
    const int bk = k < nblk ? (int)((n - k0) < 128 ? (n - k0) : 128) : 0;
    if (item < npanel) {
        const int r0 = 32 * item, j0 = 128 * i;
        const int br = (prow - r0) < 32 ? prow - r0 : 32;
        const int bj = (int)((n - j0) < 128 ? (n - j0) : 128);
        double *Ct = B + r0 + (int64_t)j0 * ldb;  // rows r0..r0+31, the 128 columns of block i
        RowTile<2, 2> t;
        if (k < nblk) {
            // B[r, i] -= Y[r, k] L_ki: P[row][m] = Y[r0+row][k0+m], Q[c][m] = Lt[j0+c][k0+m]
            t.zero();
            t.run(B + r0 + (int64_t)k0 * ldb, ldb, 0, br, Lt + j0 + (int64_t)k0 * ldlt, ldlt, 32 * w, bj, 0, bk);
            t.template store<false>(Ct, ldb, 0, 32 * w, br, bj);
        }
        // Y[r, i] = T X_ii, T = the updated rows: every wave's stores reach L2 (and this CU's
        // view) before any wave reads them back.  P[row][m] = T[row][m], Q[c][m] = X_ii[m][c] =
        // X_ii^T[c][m] (ld 128; X_ii lower: columns 32 w.. need m >= 32 w), m < bj.
        global_barrier();
        t.zero();
        t.run(Ct, ldb, 0, br, Winv + (size_t)i * kInvDoubles, 128, 32 * w, 128, 32 * w, bj);
        __syncthreads();  // every wave has read T before any wave overwrites it
        t.template store<true>(Ct, ldb, 0, 32 * w, br, bj);
        return;
    }
    if (k >= nblk) return;  // no trailing items in the first launch
    int rt, jb;
    const int t2 = item - npanel;
    if (UPPER) {
        lower_tile(t2, jb, rt);  // blocks j < i, row tiles r <= j
    } else {
        jb = t2 / nrt;
        rt = t2 % nrt;
    }
    const int wr = w >> 1, wc = w & 1;
    const int r0 = 128 * rt, j0 = 128 * jb;
    const int br = (int)((nrow - r0) < 128 ? (nrow - r0) : 128);
    const int bj = (int)((n - j0) < 128 ? (n - j0) : 128);
    double *Ct = B + r0 + (int64_t)j0 * ldb;
    // B[r, j] -= Y[r, k] L_kj: P[row][m] = Y[r0+row][k0+m], Q[c][m] = L[k0+m][j0+c] = Lt[j0+c][k0+m]
    if (bk == 128) {  // the common case: WaveTile's scalar-base loads, fixed depth
        WaveTile<4, 4> t;
        t.zero();
        t.setup(B + r0 + (int64_t)k0 * ldb, ldb, 64 * wr, br, Lt + j0 + (int64_t)k0 * ldlt, ldlt, 64 * wc, bj, 128);
        t.template run_fixed<32>();
        t.template store<false, false>(Ct, ldb, 64 * wr, 64 * wc, br, bj);
    } else {  // a last block shorter than 128
        RowTile64 t;
        t.zero();
        t.run(B + r0 + (int64_t)k0 * ldb, ldb, 64 * wr, br, Lt + j0 + (int64_t)k0 * ldlt, ldlt, 64 * wc, bj, 0, bk);
        t.template store<false>(Ct, ldb, 64 * wr, 64 * wc, br, bj);
    }
}

// B (nrow x n, ld ldb) <- B L^{-1} with Lt = L^T; upper: only the block-upper tiles (nrow == n).
void launch_trsm_xl_multi(const double *Lt, int64_t ldlt, int64_t n, double *B, int64_t ldb, int64_t nrow,
                          const double *Winv, bool upper, hipStream_t s, unsigned count, int64_t sLt, int64_t sB,
                          int64_t sWinv)
{
    const int nblk = (int)((n + 127) / 128);
    const int nrt = (int)((nrow + 127) / 128);
    for (int i = nblk - 1; i >= 0; --i) {
        const bool first = i == nblk - 1;  // nothing to apply yet: panel items only
        if (upper) {
            const int64_t prow = (128 * (int64_t)(i + 1)) < nrow ? 128 * (int64_t)(i + 1) : nrow;
            const unsigned items = (unsigned)((prow + 31) / 32 + (first ? 0 : i * (i + 1) / 2));
            hipLaunchKernelGGL(trsm_xl_step_kernel<true>, dim3(items, 1, count), dim3(256), 0, s, Lt, ldlt, n, B, ldb,
                               nrow, Winv, i, sLt, sB, sWinv);
        } else {
            const unsigned items = (unsigned)((nrow + 31) / 32 + (first ? 0 : (int64_t)i * nrt));
            hipLaunchKernelGGL(trsm_xl_step_kernel<false>, dim3(items, 1, count), dim3(256), 0, s, Lt, ldlt, n, B, ldb,
                               nrow, Winv, i, sLt, sB, sWinv);
        }
    }
}

}  // namespace mi
