// adjoint.hip -- reverse mode of the Cholesky factorization and of the solve with its factor.
//
// No reference counterpart (Eigen has no derivatives of LLT).  For A = L L^T (L lower) and a
// scalar loss with gradient Lbar = dloss/dL (lower), the gradient with respect to the
// symmetric A (blocked reverse mode, e.g. I. Murray, "Differentiation of the Cholesky
// decomposition", 2016, with the symmetric convention of torch.linalg.cholesky):
//   P    = tril(L^T Lbar)                                  (triangular-times-triangular)
//   M    = (tril(P) + tril(P, -1)^T) / 2                   (symmetric completion, halved)
//   Abar = sym(L^{-T} M L^{-1}):  Z = L^{-T} M,  S = L^{-T} Z^T,  Abar = (S + S^T) / 2
// The two triangular solves reuse the persistent backward sweep (n right-hand sides) with
// the diagonal-block inverses recomputed from L (nothing of the forward pass is stored).
// For X = A^{-1} B (solve with the factor) and Xbar: Bbar = A^{-1} Xbar and
//   Lbar -= tril(Bbar (L^T X)^T + X (L^T Bbar)^T)          (O(n^2 r), no n^3 product).
#include "common.hpp"
#include "kernels.hpp"
#include "mfma_tile.hpp"

namespace mi {

// P[i][j] = sum_{k >= i} L[k][i] G[k][j] for i >= j (tril(L^T G), L and G lower; their
// upper triangles are never read).  128 x 128 tiles of P's lower triangle, 4 waves of 64 x 64
// (4 x 4 FP64 MFMA tiles); operands k-major: element (row r, k) at X[k + r * ld].
__global__ __launch_bounds__(256) void trmm_tn_lower_kernel(const double *__restrict__ L, int64_t ldl,
                                                            const double *__restrict__ G, int64_t ldg, int64_t n,
                                                            double *__restrict__ P, int64_t ldp)
{
    const int T = (int)((n + 127) / 128);
    int ti, tj;
    lower_tile(xcd_remap((int)blockIdx.x, T * (T + 1) / 2), ti, tj);
    const int w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1, l = lane_id();
    if (ti == tj && wr == 0 && wc == 1) return;
    const int64_t i0 = (int64_t)ti * 128 + wr * 64, j0 = (int64_t)tj * 128 + wc * 64;
    if (i0 >= n || j0 >= n) return;
    int64_t prow[4], qrow[4];
    const double *pa[4], *qb[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        prow[a] = i0 + 16 * a + (l & 15);
        pa[a] = L + (prow[a] < n ? prow[a] : n - 1) * ldl;
        qrow[a] = j0 + 16 * a + (l & 15);
        qb[a] = G + (qrow[a] < n ? qrow[a] : n - 1) * ldg;
    }
    dvec4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
    const int64_t kb = i0;  // rows i >= i0 have L[k][i] = 0 for k < i0
    const int64_t S = (n - kb + 3) / 4;
    constexpr int U = 4;
    double pbuf[U][4], qbuf[U][4];
    auto load = [&](int u, int64_t s) {
        int64_t k = kb + 4 * (s < S ? s : S - 1) + (l >> 4);
        k = k < n ? k : n - 1;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            pbuf[u][a] = pa[a][k];
            qbuf[u][a] = qb[a][k];
        }
    };
    auto mma = [&](int u, int64_t s) {
        const int64_t k = kb + 4 * s + (l >> 4);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double pv = (k < n && k >= prow[a]) ? pbuf[u][a] : 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double qv = (k < n && k >= qrow[b]) ? qbuf[u][b] : 0.0;
                acc[a][b] = mfma_f64(qv, pv, acc[a][b]);
            }
        }
    };
    const int64_t Sp = (S + U - 1) / U * U;
#pragma unroll
    for (int u = 0; u < U - 1; ++u) load(u, u);
    for (int64_t s0 = 0; s0 < Sp; s0 += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            load((u + U - 1) % U, s0 + u + U - 1);
            mma(u, s0 + u);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = prow[a], j = j0 + 16 * b + 4 * r + (l >> 4);
                if (i < n && j < n && j <= i) P[i + j * ldp] = acc[a][b][r];
            }
}

// Full n x n outputs from a square input through 32 x 32 LDS tiles:
//   MODE 0: out[i][j] = in[max(i,j)][min(i,j)] / 2   (symmetric completion of tril(in), halved)
//   MODE 1: out = in^T
//   MODE 2: out = (in + in^T) / 2
template <int MODE>
__global__ __launch_bounds__(256) void sym_tile_kernel(const double *__restrict__ in, int64_t ldi, int64_t n,
                                                       double *__restrict__ out, int64_t ldo)
{
    __shared__ double t[32][33];
    const int64_t bi = blockIdx.x, bj = blockIdx.y;  // output tile rows bi, cols bj
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // mirrored tile in[bj.., bi..] -> t[x][y] = in[bj*32 + x][bi*32 + y]
    for (int y = ty; y < 32; y += 8) {
        const int64_t r = bj * 32 + tx, c = bi * 32 + y;
        t[tx][y] = (r < n && c < n) ? in[r + c * ldi] : 0.0;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int64_t i = bi * 32 + tx, j = bj * 32 + y;
        if (i >= n || j >= n) continue;
        const double tr = t[y][tx];  // in[j][i]
        double v;
        if (MODE == 0) v = 0.5 * (i >= j ? in[i + j * ldi] : tr);
        else if (MODE == 1) v = tr;
        else v = 0.5 * (in[i + j * ldi] + tr);
        out[i + j * ldo] = v;
    }
}

// U[i][c] = sum_{k >= i} L[k][i] X[k][c] (U = L^T X, L lower): one workgroup per row i.
__global__ __launch_bounds__(256) void trmv_t_kernel(const double *__restrict__ L, int64_t ldl, int64_t n,
                                                     const double *__restrict__ X, int64_t ldx, int64_t r,
                                                     double *__restrict__ U, int64_t ldu)
{
    __shared__ double red[256];
    const int64_t i = blockIdx.x;
    for (int64_t c = 0; c < r; ++c) {
        double s = 0.0;
        for (int64_t k = i + threadIdx.x; k < n; k += 256) s += L[k + i * ldl] * X[k + c * ldx];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) U[i + c * ldu] = red[0];
        __syncthreads();
    }
}

// Lbar[i][j] -= sum_c (Bb[i][c] U[j][c] + X[i][c] V[j][c]) for i >= j.
__global__ __launch_bounds__(256) void rank2_lower_kernel(double *__restrict__ Lb, int64_t ldlb, int64_t n,
                                                          const double *__restrict__ Bb, int64_t ldbb,
                                                          const double *__restrict__ U, int64_t ldu,
                                                          const double *__restrict__ X, int64_t ldx,
                                                          const double *__restrict__ V, int64_t ldv, int64_t r)
{
    const int64_t j = blockIdx.y;
    const int64_t i = j + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t c = 0; c < r; ++c) s += Bb[i + c * ldbb] * U[j + c * ldu] + X[i + c * ldx] * V[j + c * ldv];
    Lb[i + j * ldlb] -= s;
}

void launch_trmm_tn_lower(const double *L, int64_t ldl, const double *G, int64_t ldg, int64_t n, double *P,
                          int64_t ldp, hipStream_t s)
{
    const int T = (int)((n + 127) / 128);
    hipLaunchKernelGGL(trmm_tn_lower_kernel, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, L, ldl, G, ldg, n,
                       P, ldp);
}

void launch_sym_tile(int mode, const double *in, int64_t ldi, int64_t n, double *out, int64_t ldo, hipStream_t s)
{
    const dim3 g((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
    if (mode == 0)
        hipLaunchKernelGGL(sym_tile_kernel<0>, g, dim3(256), 0, s, in, ldi, n, out, ldo);
    else if (mode == 1)
        hipLaunchKernelGGL(sym_tile_kernel<1>, g, dim3(256), 0, s, in, ldi, n, out, ldo);
    else
        hipLaunchKernelGGL(sym_tile_kernel<2>, g, dim3(256), 0, s, in, ldi, n, out, ldo);
}

void launch_solve_adjoint_update(const double *L, int64_t ldl, int64_t n, int64_t r, const double *X, int64_t ldx,
                                 const double *Bb, int64_t ldbb, double *Lb, int64_t ldlb, double *U, double *V,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(trmv_t_kernel, dim3((unsigned)n), dim3(256), 0, s, L, ldl, n, X, ldx, r, U, n);
    hipLaunchKernelGGL(trmv_t_kernel, dim3((unsigned)n), dim3(256), 0, s, L, ldl, n, Bb, ldbb, r, V, n);
    const dim3 g((unsigned)((n + 255) / 256), (unsigned)n);
    hipLaunchKernelGGL(rank2_lower_kernel, g, dim3(256), 0, s, Lb, ldlb, n, Bb, ldbb, U, n, X, ldx, V, n, r);
}

}  // namespace mi
