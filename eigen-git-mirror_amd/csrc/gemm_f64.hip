// gemm_f64.hip -- symmetric-storage helpers: triangle transposes for uplo = 'U' and the
// symmetric L1 norm of LLT::compute.
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

// ------------------------------------------------------------------------------ helpers
// W (lower) = A (upper)^T through 32x32 LDS tiles; tiles strictly above the diagonal of W
// are skipped.  Used for uplo = 'U' (LLT.h:377-399 factors the transpose).
// D may be S itself (lower -> its own upper triangle): a tile's reads and writes are separated
// by the workgroup barrier, and different tiles touch disjoint elements.
template <bool TO_LOWER, typename T>
__global__ __launch_bounds__(256) void tri_transpose_kernel(const T *S, int64_t lds_, int64_t n, T *D, int64_t ldd)
{
    // TO_LOWER: D[i][j] = S[j][i] for i >= j (D lower from S upper)
    // else    : D[j][i] = S[i][j] for i >= j (D upper from S lower)
    __shared__ T t[32][33];
    const int bi = blockIdx.x, bj = blockIdx.y;  // tile of the lower triangle (rows bi, cols bj)
    if (bj > bi) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int64_t i0 = (int64_t)bi * 32, j0 = (int64_t)bj * 32;
    // read source block: TO_LOWER reads S[j0+.., i0+..] (upper), else S[i0+.., j0+..] (lower);
    // a thread's four loads are issued before its LDS stores (one memory round trip)
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int y = ty + 8 * q;
        const int64_t rr = (TO_LOWER ? j0 : i0) + tx, cc = (TO_LOWER ? i0 : j0) + y;
        v[q] = (rr < n && cc < n) ? S[rr + cc * lds_] : T(0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) t[ty + 8 * q][tx] = v[q];
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int64_t rr = (TO_LOWER ? i0 : j0) + tx, cc = (TO_LOWER ? j0 : i0) + y;
        // element (rr, cc) of D takes source (cc, rr); keep only D's target triangle
        const bool keep = TO_LOWER ? (rr >= cc) : (rr <= cc);
        if (rr < n && cc < n && keep) D[rr + cc * ldd] = t[tx][y];
    }
}

void launch_upper_to_lower(const double *A, int64_t lda, int64_t n, double *W, int64_t ldw, hipStream_t s)
{
    const unsigned T = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL((tri_transpose_kernel<true, double>), dim3(T, T), dim3(256), 0, s, A, lda, n, W, ldw);
}

template <typename T>
void launch_upper_to_lower_t(const T *A, int64_t lda, int64_t n, T *W, int64_t ldw, hipStream_t s)
{
    const unsigned G = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL((tri_transpose_kernel<true, T>), dim3(G, G), dim3(256), 0, s, A, lda, n, W, ldw);
}
template void launch_upper_to_lower_t<float>(const float *, int64_t, int64_t, float *, int64_t, hipStream_t);

// dst (m x n, ld ldd) = src converted element by element.
template <typename S, typename D>
__global__ __launch_bounds__(256) void convert_kernel(const S *__restrict__ src, int64_t lds_, int64_t m, int64_t n,
                                                      D *__restrict__ dst, int64_t ldd)
{
    const int64_t total = m * n;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e % m, j = e / m;
        dst[i + j * ldd] = (D)src[i + j * lds_];
    }
}

template <typename S, typename D>
void launch_convert(const S *src, int64_t lds_, int64_t m, int64_t n, D *dst, int64_t ldd, hipStream_t s)
{
    const int64_t total = m * n;
    if (total == 0) return;
    const unsigned g = (unsigned)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
    hipLaunchKernelGGL((convert_kernel<S, D>), dim3(g), dim3(256), 0, s, src, lds_, m, n, dst, ldd);
}
template void launch_convert<float, double>(const float *, int64_t, int64_t, int64_t, double *, int64_t, hipStream_t);
template void launch_convert<double, float>(const double *, int64_t, int64_t, int64_t, float *, int64_t, hipStream_t);

template <typename T>
void launch_lower_to_upper(const T *W, int64_t ldw, int64_t n, T *A, int64_t lda, hipStream_t s)
{
    const unsigned G = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL((tri_transpose_kernel<false, T>), dim3(G, G), dim3(256), 0, s, W, ldw, n, A, lda);
}
template void launch_lower_to_upper<double>(const double *, int64_t, int64_t, double *, int64_t, hipStream_t);
template void launch_lower_to_upper<float>(const float *, int64_t, int64_t, float *, int64_t, hipStream_t);

// Symmetric L1 norm (LLT.h:442-453).  For lower storage column c sums |A(i,c)|, i >= c
// (col(c).tail) plus |A(c,j)|, j < c (row(c).head); upper storage mirrors it.  The column
// part is a workgroup reduction per column (coalesced), the row part one thread per c
// (coalesced across threads); both in a fixed order, so the result is deterministic.
__global__ __launch_bounds__(256) void sy_colpart_kernel(const double *__restrict__ A, int64_t lda, int64_t n,
                                                         bool upper, double *__restrict__ colsum)
{
    __shared__ double red[256];
    const int64_t c = blockIdx.x;
    const int64_t lo = upper ? 0 : c, hi = upper ? c : n;
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) s += fabs(A[i + c * lda]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) colsum[c] = red[0];
}

__global__ __launch_bounds__(256) void sy_rowpart_kernel(const double *__restrict__ A, int64_t lda, int64_t n,
                                                         bool upper, double *__restrict__ colsum)
{
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int64_t lo = upper ? c : 0, hi = upper ? n : c;
    double s = 0.0;
    for (int64_t j = lo; j < hi; ++j) s += fabs(A[c + j * lda]);
    colsum[c] += s;
}

__global__ __launch_bounds__(256) void max_reduce_kernel(const double *__restrict__ v, int64_t n, double *out)
{
    __shared__ double red[256];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) m = fmax(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

void launch_sy_l1norm(const double *A, int64_t lda, int64_t n, bool upper, double *colsum, double *out,
                      hipStream_t s)
{
    hipLaunchKernelGGL(sy_colpart_kernel, dim3((unsigned)n), dim3(256), 0, s, A, lda, n, upper, colsum);
    hipLaunchKernelGGL(sy_rowpart_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A, lda, n, upper,
                       colsum);
    hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(256), 0, s, colsum, n, out);
}

}  // namespace mi
