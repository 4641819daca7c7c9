// mixed.hip -- the FP64 side of the mixed-precision solve (mi_dposv_mixed): the FP32 working
// copy of A with its Frobenius norm, the FP64 residual R = B - A X of the symmetric matrix
// held in one triangle, and the normwise backward error of each right-hand side.
//
// There is no reference counterpart: LLT works in one precision (LLT.h:430-460, 483-498).
// The residual and the error measure are those of the FP64 acceptance test,
// ||A x - b||_2 / (||A||_F ||x||_2) per right-hand side.
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

// W (T, lower, ld ldw) = the lower triangle of symmetric A (lower or upper storage, FP64),
// and part[t] = the contribution of lower 64 x 64 tile t to ||A||_F^2 (off-diagonal
// elements twice).  One workgroup per lower tile; upper storage is transposed through LDS.
template <typename T>
__global__ __launch_bounds__(256) void sym_copy_lower_kernel(const double *__restrict__ A, int64_t lda, int64_t n,
                                                             int upper, T *__restrict__ W, int64_t ldw,
                                                             double *__restrict__ part)
{
    __shared__ double tile[64][65];
    __shared__ double red[256];
    int ti, tj;
    lower_tile((int)blockIdx.x, ti, tj);
    const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    double s = 0.0;
    if (upper) {
        // source block A[j0.., i0..] (upper), element (i, j) of the lower triangle is A[j][i]
        for (int y = ty; y < 64; y += 4) {
            const int64_t rr = j0 + tx, cc = i0 + y;
            tile[y][tx] = (rr < n && cc < n) ? A[rr + cc * lda] : 0.0;
        }
        __syncthreads();
    }
    // all 16 elements of a thread first (one memory round trip), then the stores
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int y = ty + 4 * q;
        const int64_t i = i0 + tx, j = j0 + y;
        const bool in = i < n && j < n && i >= j;
        v[q] = upper ? tile[tx][y] : (in ? A[i + j * lda] : 0.0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int y = ty + 4 * q;
        const int64_t i = i0 + tx, j = j0 + y;
        if (i < n && j < n && i >= j) {
            W[i + j * ldw] = (T)v[q];
            s += (i == j ? 1.0 : 2.0) * v[q] * v[q];
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out = sqrt(sum of v[0..count)) in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void sqrt_sum_kernel(const double *__restrict__ v, int64_t count,
                                                       double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < count; i += 256) s += v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sqrt(red[0]);
}

// Partial products of A X, A symmetric in one triangle of FP64 storage: workgroup
// (row tile of 64, group of 16 right-hand sides, k-chunk) -> part[chunk] (n x nrhs, ld n).
// Each wave owns 16 rows; lane (i = l%16, k = l/16) supplies A_sym(i, k), X is the MFMA's other
// operand; 32 k (8 MFMA steps) per group, all loads of a group issued before its MFMAs.
// Where the group's k lie on the stored side of the row tile (k <= i for the lower triangle),
// A(i, k) is read in place -- lanes over consecutive i, coalesced.  Where they all lie on the
// other side, A_sym(i, k) = A(k, i) is stored with k contiguous: the workgroup reads the 32 x 64
// piece with lanes over k (coalesced) into LDS and the waves take their operand from there
// (reading it in place would stride by lda across the lanes).  The groups that straddle the
// diagonal pick per element.
__global__ __launch_bounds__(256) void symm_part_kernel(const double *__restrict__ A, int64_t lda, int64_t n,
                                                        int upper, const double *__restrict__ X, int64_t ldx,
                                                        int64_t nrhs, int64_t kchunk, double *__restrict__ part)
{
    __shared__ double Ts[33 * 64];  // Ts[kk + 33 * ii]: 32 k (padded to 33) x 64 rows of the tile
    const int l = lane_id(), w = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int64_t i = r0 + 16 * w + (l & 15);
    const int64_t c0 = (int64_t)blockIdx.y * 16;
    const int64_t kb = (int64_t)blockIdx.z * kchunk;
    const int64_t ke = (kb + kchunk) < n ? (kb + kchunk) : n;
    const int64_t ic = i < n ? i : n - 1;
    const int64_t cj = c0 + (l & 15);
    const int64_t cjc = cj < nrhs ? cj : nrhs - 1;
    const int64_t rlast = (r0 + 63) < n ? r0 + 63 : n - 1;  // last valid row of the tile
    dvec4 acc = {0.0, 0.0, 0.0, 0.0};
    constexpr int U = 8;
    for (int64_t k0 = kb; k0 < ke; k0 += 4 * U) {
        const int64_t kl = (k0 + 4 * U - 1) < ke ? k0 + 4 * U - 1 : ke - 1;  // last k of the group
        // the group lies entirely on the unstored side of the whole row tile: transposed read
        const bool transposed = upper ? (kl < r0) : (k0 > rlast);
        double av[U], xv[U];
        bool ok[U];
        if (transposed) {  // uniform over the workgroup
            // Ts[kk + 33 * ii] = A_sym(r0 + ii, k0 + kk) = A(k0 + kk, r0 + ii), stored (for either
            // triangle) at A[k + r * lda]: lanes over kk, coalesced; all eight loads of a thread
            // are issued before its first LDS store (one memory round trip, not eight)
            double tv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = threadIdx.x + 256 * q, kk = e & 31, ii = e >> 5;
                const int64_t k = k0 + kk, r = r0 + ii;
                const int64_t kc2 = k < ke ? k : ke - 1, rc2 = r < n ? r : n - 1;
                tv[q] = A[kc2 + rc2 * lda];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = threadIdx.x + 256 * q;
                Ts[(e & 31) + 33 * (e >> 5)] = tv[q];
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = 4 * u + (l >> 4);
                const int64_t k = k0 + kk;
                ok[u] = k < ke;
                av[u] = Ts[kk + 33 * (16 * w + (l & 15))];
                xv[u] = X[(ok[u] ? k : kb) + cjc * ldx];
            }
            __syncthreads();  // Ts is rewritten by the next transposed group
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t k = k0 + 4 * u + (l >> 4);
                ok[u] = k < ke;
                const int64_t kc = ok[u] ? k : kb;
                const bool low = upper ? (ic <= kc) : (ic >= kc);  // (ic, kc) is in the stored triangle
                av[u] = low ? A[ic + kc * lda] : A[kc + ic * lda];
                xv[u] = X[kc + cjc * ldx];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc = mfma_f64(ok[u] ? xv[u] : 0.0, av[u], acc);
    }
    double *P = part + (size_t)blockIdx.z * (size_t)(n * nrhs);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t j = c0 + 4 * r + (l >> 4);
        if (i < n && j < nrhs) P[i + j * n] = acc[r];
    }
}

// R = B - sum_c part[c]  (element-wise, n x nrhs)
__global__ __launch_bounds__(256) void residual_finish_kernel(const double *__restrict__ B, int64_t ldb,
                                                              const double *__restrict__ part, int chunks,
                                                              int64_t n, int64_t nrhs, double *__restrict__ R,
                                                              int64_t ldr)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * nrhs) return;
    const int64_t i = e % n, j = e / n;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += part[(size_t)c * (size_t)(n * nrhs) + e];
    R[i + j * ldr] = B[i + j * ldb] - s;
}

// out[j] = ||V(:, j)||_2^2, one workgroup per column, fixed order.
__global__ __launch_bounds__(256) void col_sumsq_kernel(const double *__restrict__ V, int64_t ldv, int64_t n,
                                                        double *__restrict__ out)
{
    __shared__ double red[256];
    const int64_t j = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double v = V[i + j * ldv];
        s += v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

// berr = max_j sqrt(rr_j) / (anorm sqrt(xx_j)); a zero residual counts as 0, a zero x with a
// nonzero residual as +inf.
__global__ void berr_kernel(const double *rr, const double *xx, int64_t nrhs, const double *anorm, double *berr)
{
    if (threadIdx.x != 0) return;
    double m = 0.0;
    for (int64_t j = 0; j < nrhs; ++j) {
        const double r = sqrt(rr[j]), d = *anorm * sqrt(xx[j]);
        const double b = r == 0.0 ? 0.0 : (d > 0.0 ? r / d : INFINITY);
        m = b > m || b != b ? b : m;
    }
    *berr = m;
}

// X += D (n x nrhs)
__global__ __launch_bounds__(256) void add_in_place_kernel(double *__restrict__ X, int64_t ldx,
                                                           const double *__restrict__ D, int64_t ldd, int64_t n,
                                                           int64_t nrhs)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * nrhs) return;
    const int64_t i = e % n, j = e / n;
    X[i + j * ldx] += D[i + j * ldd];
}

int64_t sym_tiles64(int64_t n)
{
    const int64_t T = (n + 63) / 64;
    return T * (T + 1) / 2;
}

template <typename T>
void launch_sym_copy_lower(const double *A, int64_t lda, int64_t n, bool upper, T *W, int64_t ldw, double *part,
                           double *anorm, hipStream_t s)
{
    const int64_t tiles = sym_tiles64(n);
    hipLaunchKernelGGL(sym_copy_lower_kernel<T>, dim3((unsigned)tiles), dim3(256), 0, s, A, lda, n, upper ? 1 : 0, W,
                       ldw, part);
    hipLaunchKernelGGL(sqrt_sum_kernel, dim3(1), dim3(256), 0, s, part, tiles, anorm);
}
template void launch_sym_copy_lower<float>(const double *, int64_t, int64_t, bool, float *, int64_t, double *,
                                           double *, hipStream_t);
template void launch_sym_copy_lower<double>(const double *, int64_t, int64_t, bool, double *, int64_t, double *,
                                            double *, hipStream_t);

int residual_chunks(int64_t n, int64_t nrhs)
{
    const int64_t wg = ((n + 63) / 64) * ((nrhs + 15) / 16);
    int64_t c = (2560 + wg - 1) / wg;  // two full rounds of 5 workgroups on each of 256 CUs
    c = c < 1 ? 1 : (c > 16 ? 16 : c);
    return (int)c;
}

void launch_residual(const double *A, int64_t lda, int64_t n, bool upper, const double *X, int64_t ldx,
                     const double *B, int64_t ldb, int64_t nrhs, double *R, int64_t ldr, double *part, double *rr,
                     double *xx, const double *anorm, double *berr, hipStream_t s)
{
    const int chunks = residual_chunks(n, nrhs);
    int64_t kchunk = (n + chunks - 1) / chunks;
    kchunk = (kchunk + 31) / 32 * 32;
    const int used = (int)((n + kchunk - 1) / kchunk);
    const dim3 g((unsigned)((n + 63) / 64), (unsigned)((nrhs + 15) / 16), (unsigned)used);
    hipLaunchKernelGGL(symm_part_kernel, g, dim3(256), 0, s, A, lda, n, upper ? 1 : 0, X, ldx, nrhs, kchunk, part);
    const unsigned ge = (unsigned)((n * nrhs + 255) / 256);
    hipLaunchKernelGGL(residual_finish_kernel, dim3(ge), dim3(256), 0, s, B, ldb, part, used, n, nrhs, R, ldr);
    hipLaunchKernelGGL(col_sumsq_kernel, dim3((unsigned)nrhs), dim3(256), 0, s, R, ldr, n, rr);
    hipLaunchKernelGGL(col_sumsq_kernel, dim3((unsigned)nrhs), dim3(256), 0, s, X, ldx, n, xx);
    hipLaunchKernelGGL(berr_kernel, dim3(1), dim3(64), 0, s, rr, xx, nrhs, anorm, berr);
}

void launch_add_in_place(double *X, int64_t ldx, const double *D, int64_t ldd, int64_t n, int64_t nrhs, hipStream_t s)
{
    const unsigned ge = (unsigned)((n * nrhs + 255) / 256);
    hipLaunchKernelGGL(add_in_place_kernel, dim3(ge), dim3(256), 0, s, X, ldx, D, ldd, n, nrhs);
}

}  // namespace mi
