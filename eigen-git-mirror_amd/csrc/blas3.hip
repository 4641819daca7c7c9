// blas3.hip -- the level-3 pieces that Eigen's EIGEN_USE_BLAS route binds for this path:
// a triangular solve with many right-hand sides (dtrsm_) and a symmetric rank-k update
// (dsyrk_), on the device.
//
// Reference interfaces: triangular_solve_matrix specialised to call ?trsm_
// (Eigen/src/Core/products/TriangularSolverMatrix_BLAS.h:41-161: LLT::solve's two sweeps and
// the panel solve of llt_inplace::blocked, LLT.h:364) and general_matrix_matrix_rankupdate
// calling ?syrk_ (Eigen/src/Core/products/GeneralMatrixMatrixTriangular_BLAS.h:47-134: the
// trailing update of LLT.h:365); their semantics as Eigen's own BLAS implements them
// (blas/level3_impl.h:78-178 trsm, :357-420 syrk).
//
// TRSM reduces every (side, uplo, transa, diag) to one of the two sweeps of the LLT solve on a
// lower-triangular operand (solve.hip: forward L Y = C or backward L^T Y = C):
//   the triangle of A goes to a lower workspace (transposed when A is upper, diagonal set to 1
//   when it is unit), and a right-side solve X op(A) = B runs as op(A)^T X^T = B^T on a
//   transposed copy of B.  The diagonal-block inverses the sweeps use are built from the
//   triangle like after a factorization (potrf_diag.hip).
// SYRK runs 128 x 128 tiles of the requested triangle of C on FP64 MFMA (mfma_tile.hpp), each
// workgroup four 64 x 64 wave tiles over the whole depth, C = alpha op(A) op(A)^T + beta C.
#include "common.hpp"
#include "kernels.hpp"
#include "mfma_tile.hpp"

namespace mi {

// L (lower, ld ldl) from the `upper` or lower triangle of A (k x k, ld lda); the diagonal is 1
// when `unit`.  32 x 32 tiles through LDS, so that the transposed reads are coalesced.
__global__ __launch_bounds__(256) void tri_stage_kernel(const double *__restrict__ A, int64_t lda, int64_t k,
                                                        bool upper, bool unit, double *__restrict__ L, int64_t ldl)
{
    __shared__ double t[32][33];
    const int bi = blockIdx.x, bj = blockIdx.y;  // tile (rows bi, columns bj) of L, bi >= bj
    if (bj > bi) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t i0 = (int64_t)bi * 32, j0 = (int64_t)bj * 32;
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int y = ty + 8 * q;
        // upper: source rows j0.., columns i0.. (L(i, j) = A(j, i)); lower: rows i0.., columns j0..
        const int64_t rr = (upper ? j0 : i0) + tx, cc = (upper ? i0 : j0) + y;
        v[q] = (rr < k && cc < k) ? A[rr + cc * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) t[ty + 8 * q][tx] = v[q];
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int64_t rr = i0 + tx, cc = j0 + y;  // element (rr, cc) of L
        if (rr < k && cc < k && rr >= cc) {
            const double s = upper ? t[tx][y] : t[y][tx];
            L[rr + cc * ldl] = (unit && rr == cc) ? 1.0 : s;
        }
    }
}

// D (n x m, ld ldd) = alpha * S^T, S m x n (ld lds); alpha == 0 writes zeros without reading S.
__global__ __launch_bounds__(256) void transpose_scale_kernel(const double *__restrict__ S, int64_t lds_, int64_t m,
                                                              int64_t n, double alpha, double *__restrict__ D,
                                                              int64_t ldd)
{
    __shared__ double t[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t r0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;  // tile of S
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t rr = r0 + tx, cc = c0 + ty + 8 * q;
        v[q] = (alpha != 0.0 && rr < m && cc < n) ? alpha * S[rr + cc * lds_] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) t[ty + 8 * q][tx] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t dr = c0 + tx, dc = r0 + ty + 8 * q;  // D(dr, dc) = S(dc, dr)
        if (dr < n && dc < m) D[dr + dc * ldd] = t[tx][ty + 8 * q];
    }
}

// B (m x n) *= alpha; alpha == 0 writes zeros without reading B.
__global__ __launch_bounds__(256) void scale_kernel(double *__restrict__ B, int64_t ldb, int64_t m, int64_t n,
                                                    double alpha)
{
    const int64_t total = m * n;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e % m, j = e / m;
        double *p = B + i + j * ldb;
        *p = alpha == 0.0 ? 0.0 : alpha * *p;
    }
}

void launch_tri_stage(const double *A, int64_t lda, int64_t k, bool upper, bool unit, double *L, int64_t ldl,
                      hipStream_t s)
{
    const unsigned T = (unsigned)((k + 31) / 32);
    hipLaunchKernelGGL(tri_stage_kernel, dim3(T, T), dim3(256), 0, s, A, lda, k, upper, unit, L, ldl);
}

void launch_transpose_scale(const double *S, int64_t lds_, int64_t m, int64_t n, double alpha, double *D,
                            int64_t ldd, hipStream_t s)
{
    const dim3 g((unsigned)((m + 31) / 32), (unsigned)((n + 31) / 32));
    hipLaunchKernelGGL(transpose_scale_kernel, g, dim3(256), 0, s, S, lds_, m, n, alpha, D, ldd);
}

void launch_scale(double *B, int64_t ldb, int64_t m, int64_t n, double alpha, hipStream_t s)
{
    const int64_t total = m * n;
    const unsigned g = (unsigned)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
    hipLaunchKernelGGL(scale_kernel, dim3(g), dim3(256), 0, s, B, ldb, m, n, alpha);
}

// ---------------------------------------------------------------------------------- SYRK
// W = op(A) as an n x kp column-major operand (k columns of data, zeros to kp, a multiple of
// 16 so the wave tiles need no tail): trans = false copies A (n x k), true transposes A (k x n).
__global__ __launch_bounds__(256) void syrk_stage_kernel(const double *__restrict__ A, int64_t lda, int64_t n,
                                                         int64_t k, int64_t kp, bool trans, double *__restrict__ W,
                                                         int64_t ldw)
{
    __shared__ double t[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t r0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;  // tile of W
    if (!trans) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t rr = r0 + tx, cc = c0 + ty + 8 * q;
            if (rr < n && cc < kp) W[rr + cc * ldw] = cc < k ? A[rr + cc * lda] : 0.0;
        }
        return;
    }
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // A(c, r) for W(r, c): source rows c0.., columns r0..
        const int64_t ar = c0 + tx, ac = r0 + ty + 8 * q;
        v[q] = (ar < k && ac < n) ? A[ar + ac * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) t[ty + 8 * q][tx] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t rr = r0 + tx, cc = c0 + ty + 8 * q;
        if (rr < n && cc < kp) W[rr + cc * ldw] = t[tx][ty + 8 * q];
    }
}

// One 128 x 128 tile of the `upper` or lower triangle of C per workgroup (tile u of the lower
// tile enumeration; for upper, the transposed tile), C = alpha W W^T + beta C over the
// elements of that triangle; beta == 0 does not read C.  4 waves of 64 x 64.
__global__ __launch_bounds__(256) void syrk_kernel(int64_t n, int64_t kp, double alpha, const double *__restrict__ W,
                                                   int64_t ldw, double beta, double *__restrict__ C, int64_t ldc,
                                                   bool upper)
{
    int ta, tb;
    lower_tile((int)blockIdx.x, ta, tb);  // ta >= tb
    const int ti = upper ? tb : ta, tj = upper ? ta : tb;
    const int w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    // the wave tile strictly on the other side of a diagonal tile has nothing to store
    if (ti == tj && (upper ? (wr == 1 && wc == 0) : (wr == 0 && wc == 1))) return;
    const int i0 = ti * 128 + wr * 64, j0 = tj * 128 + wc * 64;
    if (i0 >= n || j0 >= n) return;
    WaveTile<4, 4, 3> t;
    t.zero();
    t.setup(W, ldw, i0, (int)n, W, ldw, j0, (int)n, (int)kp);
    t.run();
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = i0 + t.row_of(a), j = j0 + t.col_of(b, r);
                if (i < n && j < n && (upper ? i <= j : j <= i)) {
                    double *p = C + i + j * ldc;
                    const double v = alpha * t.acc[a][b][r];
                    *p = beta == 0.0 ? v : v + beta * *p;
                }
            }
}

// beta * C on the triangle (alpha == 0 or k == 0: no product), beta == 0 writes zeros.
__global__ __launch_bounds__(256) void tri_scale_kernel(double *__restrict__ C, int64_t ldc, int64_t n, bool upper,
                                                        double beta)
{
    const int64_t j = blockIdx.x;
    const int64_t lo = upper ? 0 : j, hi = upper ? j + 1 : n;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
        double *p = C + i + j * ldc;
        *p = beta == 0.0 ? 0.0 : beta * *p;
    }
}

int64_t syrk_depth_padded(int64_t k) { return (k + 15) / 16 * 16; }

void launch_syrk(bool upper, bool trans, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                 double beta, double *C, int64_t ldc, double *W, hipStream_t s)
{
    if (n == 0) return;
    if (k == 0 || alpha == 0.0) {
        if (beta != 1.0) hipLaunchKernelGGL(tri_scale_kernel, dim3((unsigned)n), dim3(256), 0, s, C, ldc, n, upper, beta);
        return;
    }
    const int64_t kp = syrk_depth_padded(k);
    const double *op = A;
    int64_t ldo = lda;
    if (trans || kp != k) {  // op(A) as a zero-padded n x kp operand
        const dim3 g((unsigned)((n + 31) / 32), (unsigned)((kp + 31) / 32));
        hipLaunchKernelGGL(syrk_stage_kernel, g, dim3(256), 0, s, A, lda, n, k, kp, trans, W, n);
        op = W;
        ldo = n;
    }
    const int64_t T = (n + 127) / 128;
    hipLaunchKernelGGL(syrk_kernel, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, n, kp, alpha, op, ldo, beta,
                       C, ldc, upper);
}

}  // namespace mi
