// potrf_diag.hip -- diagonal-block Cholesky (POTRF) and triangular inverse (TRTRI) of one
// NB x NB block, held in registers of one 256-thread workgroup, columns broadcast through
// LDS.
//
// Replaces llt_inplace<double,Lower>::unblocked on the diagonal block A11 of
// llt_inplace::blocked (Eigen/src/Cholesky/LLT.h:312-336, called at :363).  Eigen runs a
// left-looking column loop with a GEMV per column; here the block is factored right-looking
// (each column j: pivot test, sqrt, scale, rank-1 update of the trailing part), which is
// the same arithmetic in a different summation order.  The pivot test is Eigen's
// `x <= 0 -> fail at column k` (LLT.h:327-330), so a NaN pivot passes as in Eigen.
//
// The kernel also forms X = L11^{-1} (lower triangular), used by the panel TRSM
// (A21 <- A21 * X^T, replacing the right-side triangular_solve_matrix of LLT.h:364) and by
// the solve (X applied to a block of right-hand sides).
//
// Register layout: thread (ti, tj) = (tid % 16, tid / 16) owns rows ti + 16a and columns
// tj + 16b, a, b in [0, NB/16): a cyclic distribution, so the triangle is balanced across
// threads.  The 16 owners of one column are 16 lanes of one wave (tid = ti + 16 tj), so the
// pivot of column j reaches them with one lane shuffle; the column itself is broadcast to
// all threads through a double-buffered LDS vector, one barrier per column.
//
// A final block smaller than NB is padded with the identity (L_pad = [L 0; 0 I]).
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

// Packed lower triangle of an R x R grid of register slots: slot (a, b), b <= a.
template <int R>
struct Tri {
    double s[R * (R + 1) / 2];
    __device__ __forceinline__ double &operator()(int a, int b) { return s[a * (a + 1) / 2 + b]; }
};

// One block-column JB (16 columns) of the right-looking factorization.  Returns the failing
// column (block-local) or -1.  Template recursion keeps every register index static.
template <int NB, int JB>
__device__ __forceinline__ int potrf_columns(Tri<NB / 16> &v, double (*colbuf)[NB], double *diag, int *s_fail,
                                             int ti, int tj)
{
    constexpr int R = NB / 16;
    if constexpr (JB == R) {
        return -1;
    } else {
        for (int jj = 0; jj < 16; ++jj) {
            const int j = JB * 16 + jj;
            double *cb = colbuf[j & 1];
            if (tj == jj) {
                // pivot a_jj lives in thread (ti=jj, tj=jj): lane jj + 16*(jj%4) of this wave
                const double x = __shfl(v(JB, JB), jj + 16 * (jj & 3));
                if (x <= 0.0 && ti == 0) *s_fail = j;  // Eigen's test, LLT.h:327
                const double d = sqrt(x);
                if (ti == jj) diag[j] = d;
#pragma unroll
                for (int a = 0; a < R; ++a) cb[ti + 16 * a] = 0.0;
#pragma unroll
                for (int a = JB; a < R; ++a) {
                    const int row = ti + 16 * a;
                    double val = v(a, JB);
                    if (row == j)
                        val = d;
                    else if (row > j)
                        val = val / d;
                    if (row >= j) v(a, JB) = val;
                    if (row > j) cb[row] = val;
                }
            }
            __syncthreads();
            if (*s_fail >= 0) return *s_fail;
            double lr[R], lc[R];
#pragma unroll
            for (int a = JB; a < R; ++a) lr[a] = cb[ti + 16 * a];
#pragma unroll
            for (int b = JB; b < R; ++b) lc[b] = cb[tj + 16 * b];
#pragma unroll
            for (int a = JB; a < R; ++a)
#pragma unroll
                for (int b = JB; b <= a; ++b) v(a, b) -= lr[a] * lc[b];
        }
        return potrf_columns<NB, JB + 1>(v, colbuf, diag, s_fail, ti, tj);
    }
}

// One block-row KB of the inverse X = L^{-1} (right-looking over rows k of X).
template <int NB, int KB>
__device__ __forceinline__ void trtri_rows(Tri<NB / 16> &v, Tri<NB / 16> &x, double (*colbuf)[NB],
                                           double (*rowbuf)[NB], const double *diag, int ti, int tj)
{
    constexpr int R = NB / 16;
    if constexpr (KB < R) {
        for (int kk = 0; kk < 16; ++kk) {
            const int k = KB * 16 + kk;
            double *rb = rowbuf[k & 1];
            double *cb = colbuf[k & 1];
            if (ti == kk) {  // owners of row k of X: scale by 1/L_kk and publish
                const double dk = diag[k];
#pragma unroll
                for (int b = 0; b < R; ++b) {
                    const int col = tj + 16 * b;
                    double out = 0.0;
                    if (b <= KB) {
                        if (col <= k) {
                            x(KB, b) = x(KB, b) / dk;
                            out = x(KB, b);
                        }
                    }
                    rb[col] = out;
                }
            }
            if (tj == kk) {  // owners of column k of L: publish L_ik, i > k
#pragma unroll
                for (int a = 0; a < R; ++a) {
                    const int row = ti + 16 * a;
                    double out = 0.0;
                    if (a >= KB) {
                        if (row > k) out = v(a, KB);
                    }
                    cb[row] = out;
                }
            }
            __syncthreads();
            double lr[R], xr[R];
#pragma unroll
            for (int a = KB; a < R; ++a) lr[a] = cb[ti + 16 * a];
#pragma unroll
            for (int b = 0; b <= KB; ++b) xr[b] = rb[tj + 16 * b];
#pragma unroll
            for (int a = KB; a < R; ++a)
#pragma unroll
                for (int b = 0; b <= KB; ++b) x(a, b) -= lr[a] * xr[b];
        }
        trtri_rows<NB, KB + 1>(v, x, colbuf, rowbuf, diag, ti, tj);
    }
}

template <int NB, bool FACTOR>
__global__ __launch_bounds__(256) void potrf_diag_kernel(double *__restrict__ A, int64_t lda, int64_t n,
                                                         int64_t k0_first, double *__restrict__ Winv,
                                                         int *info)
{
    constexpr int R = NB / 16;
    const int tid = threadIdx.x;
    const int ti = tid & 15, tj = tid >> 4;
    const int64_t k0 = k0_first + (int64_t)blockIdx.x * NB;
    if (k0 >= n) return;
    const int bs = (int)((n - k0) < NB ? (n - k0) : NB);
    double *W = Winv + (size_t)(k0 / NB) * NB * NB;
    if (FACTOR && factor_failed(info)) return;

    __shared__ double colbuf[2][NB];
    __shared__ double rowbuf[2][NB];
    __shared__ double diag[NB];
    __shared__ int s_fail;

    Tri<R> v;
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b <= a; ++b) {
            const int row = ti + 16 * a, col = tj + 16 * b;
            double x = 0.0;
            if (row >= col) {
                if (row < bs && col < bs)
                    x = A[(k0 + row) + (k0 + col) * lda];
                else
                    x = (row == col) ? 1.0 : 0.0;
            }
            v(a, b) = x;
        }

    if (FACTOR) {
        if (tid == 0) s_fail = -1;
        __syncthreads();
        const int fail = potrf_columns<NB, 0>(v, colbuf, diag, &s_fail, ti, tj);
        if (fail >= 0) {
            if (tid == 0) atomicCAS(info, 0, (int)(k0 + fail + 1));
            return;
        }
        // L11 back to A (lower triangle of the block only)
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
            for (int b = 0; b <= a; ++b) {
                const int row = ti + 16 * a, col = tj + 16 * b;
                if (row >= col && row < bs) A[(k0 + row) + (k0 + col) * lda] = v(a, b);
            }
    } else {
        if (ti == tj) {
#pragma unroll
            for (int a = 0; a < R; ++a) diag[ti + 16 * a] = v(a, a);
        }
        __syncthreads();
    }

    // ---- TRTRI: X = L^{-1}
    Tri<R> x;
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b <= a; ++b) x(a, b) = ((ti + 16 * a) == (tj + 16 * b)) ? 1.0 : 0.0;
    trtri_rows<NB, 0>(v, x, colbuf, rowbuf, diag, ti, tj);

    // full NB x NB tile of X (zeros above the diagonal), column-major, ld = NB
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const int row = ti + 16 * a, col = tj + 16 * b;
            double val = 0.0;
            if (b <= a) {
                if (row >= col) val = x(a, b);
            }
            W[row + (size_t)col * NB] = val;
        }
}

void launch_potrf_diag(double *A, int64_t lda, int64_t n, int64_t k0, double *Winv, int *info,
                       hipStream_t s)
{
    hipLaunchKernelGGL((potrf_diag_kernel<kNB, true>), dim3(1), dim3(256), 0, s, A, lda, n, k0, Winv, info);
}

void launch_trtri_blocks(const double *L, int64_t lda, int64_t n, double *Winv, hipStream_t s)
{
    const int nblk = (int)((n + kNB - 1) / kNB);
    hipLaunchKernelGGL((potrf_diag_kernel<kNB, false>), dim3(nblk), dim3(256), 0, s,
                       const_cast<double *>(L), lda, n, (int64_t)0, Winv, (int *)nullptr);
}

}  // namespace mi
