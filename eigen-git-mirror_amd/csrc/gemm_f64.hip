// gemm_f64.hip -- FP64 MFMA tile kernels of the blocked factorization and of the solve.
//
//   syrk_lower_kernel  A22 (lower) -= A21 * A21^T
//       replaces the trailing rank update of llt_inplace::blocked, LLT.h:365
//       (SelfadjointProduct.h:82-117 -> GeneralMatrixMatrixTriangular.h:57-194, whose cycles
//       are spent in gebp_kernel, GeneralBlockPanelKernel.h:1548-2217).
//   trsm_panel_kernel  A21 <- A21 * X^T with X = L11^{-1}
//       replaces A11^T.triangularView<Upper>().solveInPlace<OnTheRight>(A21), LLT.h:364
//       (triangular_solve_matrix<..,OnTheRight,..>, TriangularSolverMatrix.h:186-329).
//   solve_diag_kernel / solve_update_kernel
//       block forward/backward substitution of LLT::_solve_impl (LLT.h:483-498,
//       TriangularSolverMatrix.h:36-184): Y_b = X_b * B_b, then B_rest -= L_rest,b * Y_b.
//
// All products have the form C[i][j] (+)= sum_k P[i][k] * Q[j][k] and run on
// v_mfma_f64_16x16x4_f64.  Operand tiles are staged global -> registers -> LDS in chunks
// of KC = 16 along k; the next chunk is fetched into registers while the current one feeds
// the MFMAs.
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

constexpr int KC = 16;   // k-chunk staged through LDS
constexpr int PAD = 16;  // LDS row padding (doubles): rows of a chunk land in alternate bank halves

// KC x ROWS operand tile.  Source element (r, k) is at g[r * rs + k * ks]; elements with
// r >= rmax or k >= kmax read as 0.  KCONTIG selects the lane mapping that keeps global
// reads contiguous when ks == 1 (else rs == 1 is assumed).
template <int ROWS, bool KCONTIG>
struct Tile {
    static constexpr int PER = ROWS * KC / 256;
    double v[PER];

    __device__ __forceinline__ static void coords(int q, int &r, int &k)
    {
        const int e = q * 256 + (int)threadIdx.x;
        if (KCONTIG) {
            k = e % KC;
            r = e / KC;
        } else {
            r = e % ROWS;
            k = e / ROWS;
        }
    }
    __device__ __forceinline__ void fetch(const double *g, int64_t rs, int64_t ks, int rmax, int kmax)
    {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int r, k;
            coords(q, r, k);
            v[q] = (r < rmax && k < kmax) ? g[r * rs + k * ks] : 0.0;
        }
    }
    __device__ __forceinline__ void stash(double *s) const
    {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int r, k;
            coords(q, r, k);
            s[k * (ROWS + PAD) + r] = v[q];
        }
    }
};

// ------------------------------------------------------------------------------ SYRK
// One 128x128 tile of the lower triangle of A22 per workgroup; 4 waves in a 2x2 grid of
// 64x64 sub-tiles (4x4 MFMA tiles each).  On diagonal tiles the strictly-upper 64x64
// quadrant is skipped.
__global__ __launch_bounds__(256) void syrk_lower_kernel(double *__restrict__ A, int64_t lda, int64_t n,
                                                         int64_t k0, const int *info)
{
    if (factor_failed(info)) return;
    const int64_t r0 = k0 + kNB;
    const int rs = (int)(n - r0);
    int ti, tj;
    lower_tile((int)blockIdx.x, ti, tj);
    const int i0 = ti * 128, j0 = tj * 128;
    const double *P = A + r0 + k0 * lda;

    __shared__ double Ps[KC * (128 + PAD)];
    __shared__ double Qs[KC * (128 + PAD)];
    const int w = threadIdx.x >> 6, l = lane_id();
    const int wr = w >> 1, wc = w & 1;
    const bool active = !(ti == tj && wr == 0 && wc == 1);

    dvec4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};

    Tile<128, false> tp, tq;
    tp.fetch(P + i0, 1, lda, rs - i0, KC);
    tq.fetch(P + j0, 1, lda, rs - j0, KC);
    for (int kc = 0; kc < kNB; kc += KC) {
        __syncthreads();
        tp.stash(Ps);
        tq.stash(Qs);
        __syncthreads();
        if (kc + KC < kNB) {
            tp.fetch(P + i0 + (kc + KC) * lda, 1, lda, rs - i0, KC);
            tq.fetch(P + j0 + (kc + KC) * lda, 1, lda, rs - j0, KC);
        }
        if (active) {
#pragma unroll
            for (int s = 0; s < KC / 4; ++s) {
                const int kr = (s * 4 + (l >> 4)) * (128 + PAD);
                double pf[4], qf[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) pf[a] = Ps[kr + wr * 64 + a * 16 + (l & 15)];
#pragma unroll
                for (int b = 0; b < 4; ++b) qf[b] = Qs[kr + wc * 64 + b * 16 + (l & 15)];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = mfma_f64(qf[b], pf[a], acc[a][b]);
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wr * 64 + a * 16 + (l & 15);
                const int j = j0 + wc * 64 + b * 16 + 4 * r + (l >> 4);
                if (i < rs && j <= i) A[(r0 + i) + (r0 + j) * lda] -= acc[a][b][r];
            }
}

void launch_syrk_update(double *A, int64_t lda, int64_t n, int64_t k0, const int *info, hipStream_t s)
{
    const int64_t rs = n - k0 - kNB;
    if (rs <= 0) return;
    const int T = (int)((rs + 127) / 128);
    hipLaunchKernelGGL(syrk_lower_kernel, dim3(T * (T + 1) / 2), dim3(256), 0, s, A, lda, n, k0, info);
}

// ------------------------------------------------------------------------------ panel TRSM
// 64 panel rows x kNB columns per workgroup; wave w owns columns 32w..32w+31.  The
// workgroup reads all of its rows before writing them, so the update is in place.
__global__ __launch_bounds__(256) void trsm_panel_kernel(double *__restrict__ A, int64_t lda, int64_t n,
                                                         int64_t k0, const double *__restrict__ Winv,
                                                         const int *info)
{
    if (factor_failed(info)) return;
    const int64_t r0 = k0 + kNB;
    const int rs = (int)(n - r0);
    const int i0 = blockIdx.x * 64;
    double *P = A + r0 + k0 * lda;
    const double *X = Winv + (size_t)(k0 / kNB) * kNB * kNB;

    __shared__ double Ps[KC * (64 + PAD)];
    __shared__ double Qs[KC * (128 + PAD)];
    const int w = threadIdx.x >> 6, l = lane_id();

    dvec4 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};

    Tile<64, false> tp;
    Tile<128, false> tq;
    tp.fetch(P + i0, 1, lda, rs - i0, KC);
    tq.fetch(X, 1, kNB, kNB, KC);
    for (int kc = 0; kc < kNB; kc += KC) {
        __syncthreads();
        tp.stash(Ps);
        tq.stash(Qs);
        __syncthreads();
        if (kc + KC < kNB) {
            tp.fetch(P + i0 + (kc + KC) * lda, 1, lda, rs - i0, KC);
            tq.fetch(X + (kc + KC) * kNB, 1, kNB, kNB, KC);
        }
#pragma unroll
        for (int s = 0; s < KC / 4; ++s) {
            const int kr = s * 4 + (l >> 4);
            double pf[4], qf[2];
#pragma unroll
            for (int a = 0; a < 4; ++a) pf[a] = Ps[kr * (64 + PAD) + a * 16 + (l & 15)];
#pragma unroll
            for (int b = 0; b < 2; ++b) qf[b] = Qs[kr * (128 + PAD) + w * 32 + b * 16 + (l & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = mfma_f64(qf[b], pf[a], acc[a][b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + a * 16 + (l & 15);
                const int j = w * 32 + b * 16 + 4 * r + (l >> 4);
                if (i < rs) P[i + j * lda] = acc[a][b][r];
            }
}

void launch_trsm_panel(double *A, int64_t lda, int64_t n, int64_t k0, const double *Winv, const int *info,
                       hipStream_t s)
{
    const int64_t rs = n - k0 - kNB;
    if (rs <= 0) return;
    hipLaunchKernelGGL(trsm_panel_kernel, dim3((unsigned)((rs + 63) / 64)), dim3(256), 0, s, A, lda, n, k0,
                       Winv, info);
}

// ------------------------------------------------------------------------------ solve
// Y_b = op(X_b) * B_b for the rows of block b (op = X for the forward sweep with L,
// X^T for the backward sweep with L^T).  kNB rows x 64 right-hand sides per workgroup;
// wave w owns rows 32w..32w+31.  In place on B.
template <bool TRANS>
__global__ __launch_bounds__(256) void solve_diag_kernel(double *__restrict__ B, int64_t ldb, int64_t n,
                                                         int64_t nrhs, int64_t k0,
                                                         const double *__restrict__ Winv, const int *info)
{
    if (factor_failed(info)) return;
    const int bs = (int)((n - k0) < kNB ? (n - k0) : kNB);
    const int c0 = blockIdx.x * 64;
    const int nc = (int)(nrhs - c0);
    const double *X = Winv + (size_t)(k0 / kNB) * kNB * kNB;
    double *Bb = B + k0 + (int64_t)c0 * ldb;

    __shared__ double Ps[KC * (128 + PAD)];
    __shared__ double Qs[KC * (64 + PAD)];
    const int w = threadIdx.x >> 6, l = lane_id();

    dvec4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};

    // P[i][k] = X[i][k] (forward) or X[k][i] (backward); Q[c][k] = B[k0 + k][c0 + c]
    Tile<128, TRANS> tp;
    Tile<64, true> tq;
    const int64_t prs = TRANS ? kNB : 1, pks = TRANS ? 1 : kNB;
    tp.fetch(X, prs, pks, kNB, KC);
    tq.fetch(Bb, ldb, 1, nc, bs);
    for (int kc = 0; kc < kNB; kc += KC) {
        __syncthreads();
        tp.stash(Ps);
        tq.stash(Qs);
        __syncthreads();
        if (kc + KC < kNB) {
            tp.fetch(X + (kc + KC) * pks, prs, pks, kNB, KC);
            tq.fetch(Bb + (kc + KC), ldb, 1, nc, bs - (kc + KC));
        }
#pragma unroll
        for (int s = 0; s < KC / 4; ++s) {
            const int kr = s * 4 + (l >> 4);
            double pf[2], qf[4];
#pragma unroll
            for (int a = 0; a < 2; ++a) pf[a] = Ps[kr * (128 + PAD) + w * 32 + a * 16 + (l & 15)];
#pragma unroll
            for (int b = 0; b < 4; ++b) qf[b] = Qs[kr * (64 + PAD) + b * 16 + (l & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = mfma_f64(qf[b], pf[a], acc[a][b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = w * 32 + a * 16 + (l & 15);
                const int c = b * 16 + 4 * r + (l >> 4);
                if (i < bs && c < nc) Bb[i + (int64_t)c * ldb] = acc[a][b][r];
            }
}

// Forward (TRANS=false): B[rows below block] -= L[rows, block] * Y_b.
// Backward (TRANS=true): B[rows above block] -= L[block, rows]^T * Y_b.
// 64 rows x 64 right-hand sides per workgroup, 2x2 waves of 32x32.
template <bool TRANS>
__global__ __launch_bounds__(256) void solve_update_kernel(const double *__restrict__ L, int64_t lda,
                                                           int64_t n, double *__restrict__ B, int64_t ldb,
                                                           int64_t nrhs, int64_t k0, const int *info)
{
    if (factor_failed(info)) return;
    const int bs = (int)((n - k0) < kNB ? (n - k0) : kNB);
    const int64_t rbase = TRANS ? 0 : k0 + kNB;
    const int m = (int)(TRANS ? k0 : n - rbase);
    const int i0 = blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int nc = (int)(nrhs - c0);

    // P[i][k] = L[rbase + i0 + i][k0 + k] (forward) or L[k0 + k][i0 + i] (backward)
    const double *Pg = TRANS ? (L + k0 + (i0 + rbase) * lda) : (L + (rbase + i0) + k0 * lda);
    const int64_t prs = TRANS ? lda : 1, pks = TRANS ? 1 : lda;
    const double *Qg = B + k0 + (int64_t)c0 * ldb;  // Q[c][k] = Y[k][c]

    __shared__ double Ps[KC * (64 + PAD)];
    __shared__ double Qs[KC * (64 + PAD)];
    const int w = threadIdx.x >> 6, l = lane_id();
    const int wr = w >> 1, wc = w & 1;

    dvec4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};

    Tile<64, TRANS> tp;
    Tile<64, true> tq;
    tp.fetch(Pg, prs, pks, m - i0, bs);
    tq.fetch(Qg, ldb, 1, nc, bs);
    for (int kc = 0; kc < bs; kc += KC) {
        __syncthreads();
        tp.stash(Ps);
        tq.stash(Qs);
        __syncthreads();
        if (kc + KC < bs) {
            tp.fetch(Pg + (kc + KC) * pks, prs, pks, m - i0, bs - (kc + KC));
            tq.fetch(Qg + (kc + KC), ldb, 1, nc, bs - (kc + KC));
        }
#pragma unroll
        for (int s = 0; s < KC / 4; ++s) {
            const int kr = (s * 4 + (l >> 4)) * (64 + PAD);
            double pf[2], qf[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) pf[a] = Ps[kr + wr * 32 + a * 16 + (l & 15)];
#pragma unroll
            for (int b = 0; b < 2; ++b) qf[b] = Qs[kr + wc * 32 + b * 16 + (l & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = mfma_f64(qf[b], pf[a], acc[a][b]);
        }
    }
    double *Bo = B + rbase + i0 + (int64_t)c0 * ldb;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wr * 32 + a * 16 + (l & 15);
                const int c = wc * 32 + b * 16 + 4 * r + (l >> 4);
                if (i0 + i < m && c < nc) Bo[i + (int64_t)c * ldb] -= acc[a][b][r];
            }
}

void launch_solve(const double *L, int64_t lda, int64_t n, double *B, int64_t ldb, int64_t nrhs,
                  const double *Winv, const int *info, hipStream_t s)
{
    const int64_t nblk = (n + kNB - 1) / kNB;
    const unsigned gc = (unsigned)((nrhs + 63) / 64);
    for (int64_t b = 0; b < nblk; ++b) {  // forward: L Y = B
        const int64_t k0 = b * kNB;
        hipLaunchKernelGGL(solve_diag_kernel<false>, dim3(gc), dim3(256), 0, s, B, ldb, n, nrhs, k0, Winv, info);
        const int64_t m = n - k0 - kNB;
        if (m > 0)
            hipLaunchKernelGGL(solve_update_kernel<false>, dim3((unsigned)((m + 63) / 64), gc), dim3(256), 0, s, L,
                               lda, n, B, ldb, nrhs, k0, info);
    }
    for (int64_t b = nblk - 1; b >= 0; --b) {  // backward: L^T X = Y
        const int64_t k0 = b * kNB;
        hipLaunchKernelGGL(solve_diag_kernel<true>, dim3(gc), dim3(256), 0, s, B, ldb, n, nrhs, k0, Winv, info);
        if (k0 > 0)
            hipLaunchKernelGGL(solve_update_kernel<true>, dim3((unsigned)((k0 + 63) / 64), gc), dim3(256), 0, s, L,
                               lda, n, B, ldb, nrhs, k0, info);
    }
}

// ------------------------------------------------------------------------------ helpers
// W (lower) = A (upper)^T through 32x32 LDS tiles; tiles strictly above the diagonal of W
// are skipped.  Used for uplo = 'U' (LLT.h:377-399 factors the transpose).
template <bool TO_LOWER>
__global__ __launch_bounds__(256) void tri_transpose_kernel(const double *__restrict__ S, int64_t lds_,
                                                            int64_t n, double *__restrict__ D, int64_t ldd)
{
    // TO_LOWER: D[i][j] = S[j][i] for i >= j (D lower from S upper)
    // else    : D[j][i] = S[i][j] for i >= j (D upper from S lower)
    __shared__ double t[32][33];
    const int bi = blockIdx.x, bj = blockIdx.y;  // tile of the lower triangle (rows bi, cols bj)
    if (bj > bi) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int64_t i0 = (int64_t)bi * 32, j0 = (int64_t)bj * 32;
    // read source block: TO_LOWER reads S[j0+.., i0+..] (upper), else S[i0+.., j0+..] (lower)
    for (int y = ty; y < 32; y += 8) {
        const int64_t rr = (TO_LOWER ? j0 : i0) + tx, cc = (TO_LOWER ? i0 : j0) + y;
        if (rr < n && cc < n) t[y][tx] = S[rr + cc * lds_];
    }
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
    hipLaunchKernelGGL(tri_transpose_kernel<true>, dim3(T, T), dim3(256), 0, s, A, lda, n, W, ldw);
}

void launch_lower_to_upper(const double *W, int64_t ldw, int64_t n, double *A, int64_t lda, hipStream_t s)
{
    const unsigned T = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL(tri_transpose_kernel<false>, dim3(T, T), dim3(256), 0, s, W, ldw, n, A, lda);
}

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
