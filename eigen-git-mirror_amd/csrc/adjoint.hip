// adjoint.hip -- reverse mode of the Cholesky factorization and of the solve with its factor.
//
// No reference counterpart (Eigen has no derivatives of LLT).  For A = L L^T (L lower) and a
// scalar loss with gradient Lbar = dloss/dL (lower), the gradient with respect to the
// symmetric A (blocked reverse mode, e.g. I. Murray, "Differentiation of the Cholesky
// decomposition", 2016, with the symmetric convention of torch.linalg.cholesky):
//   P    = tril(L^T Lbar)                                  (triangular-times-triangular)
//   M    = (tril(P) + tril(P, -1)^T) / 2                   (symmetric completion, halved)
//   Abar = sym(L^{-T} M L^{-1}):  Y = M L^{-1} (= (L^{-T} M)^T),  S = Y^T L^{-1},  Abar = sym(S)
// The two triangular solves run as blocked right-side substitutions over n rows
// (trsm_multi.hip, operands row-contiguous through a transposed copy of L) with the
// diagonal-block inverses recomputed from L (nothing of the forward pass is stored); the
// second computes only the block-upper half of the symmetric S.
// For X = A^{-1} B (solve with the factor) and Xbar: Bbar = A^{-1} Xbar and
//   Lbar -= tril(Bbar (L^T X)^T + X (L^T Bbar)^T)          (O(n^2 r), no n^3 product).
#include "common.hpp"
#include "kernels.hpp"
#include "mfma_tile.hpp"

namespace mi {

// P[i][j] = sum_{k >= i} L[k][i] G[k][j] for i >= j (tril(L^T G), L and G lower), from the
// transposed lower parts Lt = tril(L)^T and Gt = tril(G)^T (zero above their diagonals), so that
// P = Lt Gt^T is a plain product with row-contiguous operands.  128 x 128 tiles of P's lower
// triangle, 4 waves of 64 x 64 (RowTile64).
//
// Tile row ti needs k in [128 ti, n): the first tiles' depth is n, the last ones' 128.  So each
// tile's k range is split into chunks of ck (a multiple of 128, about n/3), one workgroup per
// (tile, chunk) writing its 128 x 128 partial (ld 128) to Wp[item]; trmm_reduce_kernel then
// adds each tile's partials in chunk order (deterministic).  The launcher grows ck until the
// partials fit in the n^2 doubles of scratch; when even one chunk per tile would not (n below
// about 300), each tile is written straight to its place in an n x n output (its entries above
// the diagonal too: nothing reads them).  Items are numbered tile row by tile row, (tj, chunk)
// within a row.
__device__ __forceinline__ int trmm_chunks(int64_t n, int ti, int64_t ck)
{
    return (int)((n - 128 * (int64_t)ti + ck - 1) / ck);
}

__global__ __launch_bounds__(256, 2) void trmm_part_kernel(const double *__restrict__ Lt, int64_t ldlt,
                                                           const double *__restrict__ Gt, int64_t ldgt, int64_t n,
                                                           int64_t ck, double *__restrict__ Wp, double *__restrict__ P,
                                                           int64_t ldp, int direct, int64_t sLt, int64_t sGt,
                                                           int64_t sWp, int64_t sP)
{
    {  // batched launches (every kernel of this file): system z = blockIdx.z
        const int64_t z = blockIdx.z;
        Lt += z * sLt;
        Gt += z * sGt;
        if (Wp) Wp += z * sWp;
        if (P) P += z * sP;
    }
    int u = (int)blockIdx.x, ti = 0, nch = trmm_chunks(n, 0, ck);
    while (u >= (ti + 1) * nch) {  // uniform: item -> tile row
        u -= (ti + 1) * nch;
        ++ti;
        nch = trmm_chunks(n, ti, ck);
    }
    const int tj = u / nch, ch = u % nch;
    const int w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    const int64_t klo = (int64_t)ti * 128 + ch * ck, khi = (klo + ck) < n ? (klo + ck) : n;
    const int64_t i0w = (int64_t)ti * 128 + wr * 64;  // Lt rows i >= i0w are zero for k < i0w
    const int kstart = (int)(klo > i0w ? klo : i0w);
    RowTile64 t;
    t.zero();
    if (khi > kstart && (khi - kstart) % 16 == 0) {  // WaveTile's scalar-base loads (depth a multiple of 16)
        WaveTile<4, 4> wt;
        wt.zero();
        wt.setup(Lt + (int64_t)kstart * ldlt, ldlt, ti * 128 + wr * 64, (int)n, Gt + (int64_t)kstart * ldgt, ldgt,
                 tj * 128 + wc * 64, (int)n, (int)(khi - kstart));
        wt.run();
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) t.acc[a][b] = wt.acc[a][b];
    } else {
        t.run(Lt, ldlt, ti * 128 + wr * 64, (int)n, Gt, ldgt, tj * 128 + wc * 64, (int)n, kstart, (int)khi);
    }
    if (direct)
        t.template store<true>(P + (int64_t)ti * 128 + (int64_t)tj * 128 * ldp, ldp, wr * 64, wc * 64,
                               (int)(n - ti * 128), (int)(n - tj * 128));
    else
        t.template store<true>(Wp + (size_t)blockIdx.x * 128 * 128, 128, wr * 64, wc * 64, 128, 128);
}

// P tile (ti, tj) = sum of its partials in chunk order; lower part only on diagonal tiles.
__global__ __launch_bounds__(256) void trmm_reduce_kernel(const double *__restrict__ Wp, int64_t n, int64_t ck,
                                                          double *__restrict__ P, int64_t ldp, int64_t sWp, int64_t sP)
{
    Wp += (int64_t)blockIdx.z * sWp;
    P += (int64_t)blockIdx.z * sP;
    int ti, tj;
    lower_tile((int)blockIdx.x, ti, tj);
    int64_t base = 0;
    for (int t = 0; t < ti; ++t) base += (int64_t)(t + 1) * trmm_chunks(n, t, ck);
    const int nch = trmm_chunks(n, ti, ck);
    const double *src = Wp + (size_t)(base + (int64_t)tj * nch) * 128 * 128;
    for (int e = threadIdx.x; e < 128 * 128; e += 256) {
        const int r = e & 127, c = e >> 7;
        const int64_t i = (int64_t)ti * 128 + r, j = (int64_t)tj * 128 + c;
        if (i >= n || j >= n || j > i) continue;
        double v = 0.0;
        for (int ch = 0; ch < nch; ++ch) v += src[(size_t)ch * 128 * 128 + e];
        P[i + j * ldp] = v;
    }
}

// Full n x n outputs from a square input through 32 x 32 LDS tiles:
//   MODE 0: out[i][j] = in[max(i,j)][min(i,j)] / 2   (symmetric completion of tril(in), halved)
//   MODE 1: out = in^T
//   MODE 2: out = (in + in^T) / 2
//   MODE 3: as MODE 2 inside the 128 x 128 diagonal blocks, in[min(i,j)][max(i,j)] outside
//           (in is symmetric and only its block-upper part was computed)
//   MODE 4: out = tril(in)^T  (out[i][j] = in[j][i] for j >= i, else 0)
template <int MODE>
__global__ __launch_bounds__(256) void sym_tile_kernel(const double *__restrict__ in, int64_t ldi, int64_t n,
                                                       double *__restrict__ out, int64_t ldo, int64_t sIn,
                                                       int64_t sOut)
{
    __shared__ double t[32][33];
    in += (int64_t)blockIdx.z * sIn;
    out += (int64_t)blockIdx.z * sOut;
    const int64_t bi = blockIdx.x, bj = blockIdx.y;  // output tile rows bi, cols bj
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // mirrored tile in[bj.., bi..] -> t[x][y] = in[bj*32 + x][bi*32 + y]; each thread issues all
    // its loads (the mirrored ones and, except for modes 1 and 4, its own) before using them
    double m[4], d[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int y = ty + 8 * q;
        const int64_t r = bj * 32 + tx, c = bi * 32 + y;
        m[q] = (r < n && c < n) ? in[r + c * ldi] : 0.0;
        const int64_t i = bi * 32 + tx, j = bj * 32 + y;
        d[q] = (MODE != 1 && MODE != 4 && i < n && j < n) ? in[i + j * ldi] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) t[tx][ty + 8 * q] = m[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int y = ty + 8 * q;
        const int64_t i = bi * 32 + tx, j = bj * 32 + y;
        if (i >= n || j >= n) continue;
        const double tr = t[y][tx];  // in[j][i]
        double v;
        if (MODE == 0) v = 0.5 * (i >= j ? d[q] : tr);
        else if (MODE == 1) v = tr;
        else if (MODE == 3 && (i >> 7) != (j >> 7)) v = i < j ? d[q] : tr;
        else if (MODE == 4) v = j >= i ? tr : 0.0;
        else v = 0.5 * (d[q] + tr);
        out[i + j * ldo] = v;
    }
}

// U[i][c] = sum_{k >= i} L[k][i] X[k][c] (U = L^T X, L lower): one workgroup per row i.
__global__ __launch_bounds__(256) void trmv_t_kernel(const double *__restrict__ L, int64_t ldl, int64_t n,
                                                     const double *__restrict__ X, int64_t ldx, int64_t r,
                                                     double *__restrict__ U, int64_t ldu, int64_t sL, int64_t sX,
                                                     int64_t sU)
{
    __shared__ double red[256];
    L += (int64_t)blockIdx.z * sL;
    X += (int64_t)blockIdx.z * sX;
    U += (int64_t)blockIdx.z * sU;
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
                                                          const double *__restrict__ V, int64_t ldv, int64_t r,
                                                          int64_t sLb, int64_t sBb, int64_t sU, int64_t sX, int64_t sV)
{
    const int64_t z = blockIdx.z;
    Lb += z * sLb;
    Bb += z * sBb;
    U += z * sU;
    X += z * sX;
    V += z * sV;
    const int64_t j = blockIdx.y;
    const int64_t i = j + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t c = 0; c < r; ++c) s += Bb[i + c * ldbb] * U[j + c * ldu] + X[i + c * ldx] * V[j + c * ldv];
    Lb[i + j * ldlb] -= s;
}

// out (n x n, ld ldo) = in (ld ldi), one column per workgroup, system z = blockIdx.z.
__global__ __launch_bounds__(256) void copy_square_kernel(const double *__restrict__ in, int64_t ldi, int64_t n,
                                                          double *__restrict__ out, int64_t ldo, int64_t sIn,
                                                          int64_t sOut)
{
    const int64_t z = blockIdx.z, j = blockIdx.x;
    for (int64_t i = threadIdx.x; i < n; i += 256) out[z * sOut + i + j * ldo] = in[z * sIn + i + j * ldi];
}

// Lt = tril(L)^T, Gt = tril(G)^T; Wp: n*n doubles of scratch for the partials.  count systems,
// each operand of system z at its base + z times its stride.
void launch_trmm_tn_lower(const double *Lt, int64_t ldlt, const double *Gt, int64_t ldgt, int64_t n, double *P,
                          int64_t ldp, double *Wp, hipStream_t s, unsigned count, int64_t sLt, int64_t sGt,
                          int64_t sP, int64_t sWp)
{
    const int T = (int)((n + 127) / 128);
    auto items_of = [&](int64_t ck) {
        int64_t items = 0;
        for (int t = 0; t < T; ++t) items += (int64_t)(t + 1) * ((n - 128 * (int64_t)t + ck - 1) / ck);
        return items;
    };
    int C = (T + 2) / 3 > 1 ? (T + 2) / 3 : 1;  // chunk of 128 C >= n/3
    int64_t items = items_of(128 * (int64_t)C);
    while (items * 128 * 128 > n * n && C < T) items = items_of(128 * (int64_t)(++C));
    const int direct = items * 128 * 128 > n * n ? 1 : 0;  // then C == T: one chunk per tile
    const int64_t ck = 128 * (int64_t)C;
    if (direct) {
        // P may be Gt's buffer, still read by other workgroups of the launch: the tiles go to
        // Wp (as an n x n matrix) and are copied over afterwards (small n only)
        hipLaunchKernelGGL(trmm_part_kernel, dim3((unsigned)items, 1, count), dim3(256), 0, s, Lt, ldlt, Gt, ldgt, n,
                           ck, (double *)nullptr, Wp, n, 1, sLt, sGt, (int64_t)0, sWp);
        hipLaunchKernelGGL(copy_square_kernel, dim3((unsigned)n, 1, count), dim3(256), 0, s, Wp, n, n, P, ldp, sWp,
                           sP);
    } else {
        hipLaunchKernelGGL(trmm_part_kernel, dim3((unsigned)items, 1, count), dim3(256), 0, s, Lt, ldlt, Gt, ldgt, n,
                           ck, Wp, (double *)nullptr, (int64_t)0, 0, sLt, sGt, sWp, (int64_t)0);
        hipLaunchKernelGGL(trmm_reduce_kernel, dim3((unsigned)(T * (T + 1) / 2), 1, count), dim3(256), 0, s, Wp, n,
                           ck, P, ldp, sWp, sP);
    }
}

void launch_sym_tile(int mode, const double *in, int64_t ldi, int64_t n, double *out, int64_t ldo, hipStream_t s,
                     unsigned count, int64_t sIn, int64_t sOut)
{
    const dim3 g((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32), count);
    if (mode == 0)
        hipLaunchKernelGGL(sym_tile_kernel<0>, g, dim3(256), 0, s, in, ldi, n, out, ldo, sIn, sOut);
    else if (mode == 1)
        hipLaunchKernelGGL(sym_tile_kernel<1>, g, dim3(256), 0, s, in, ldi, n, out, ldo, sIn, sOut);
    else if (mode == 2)
        hipLaunchKernelGGL(sym_tile_kernel<2>, g, dim3(256), 0, s, in, ldi, n, out, ldo, sIn, sOut);
    else if (mode == 3)
        hipLaunchKernelGGL(sym_tile_kernel<3>, g, dim3(256), 0, s, in, ldi, n, out, ldo, sIn, sOut);
    else
        hipLaunchKernelGGL(sym_tile_kernel<4>, g, dim3(256), 0, s, in, ldi, n, out, ldo, sIn, sOut);
}

void launch_solve_adjoint_update(const double *L, int64_t ldl, int64_t n, int64_t r, const double *X, int64_t ldx,
                                 const double *Bb, int64_t ldbb, double *Lb, int64_t ldlb, double *U, double *V,
                                 hipStream_t s, unsigned count, int64_t sL, int64_t sX, int64_t sBb, int64_t sLb,
                                 int64_t sUV)
{
    hipLaunchKernelGGL(trmv_t_kernel, dim3((unsigned)n, 1, count), dim3(256), 0, s, L, ldl, n, X, ldx, r, U, n, sL, sX,
                       sUV);
    hipLaunchKernelGGL(trmv_t_kernel, dim3((unsigned)n, 1, count), dim3(256), 0, s, L, ldl, n, Bb, ldbb, r, V, n, sL,
                       sBb, sUV);
    const dim3 g((unsigned)((n + 255) / 256), (unsigned)n, count);
    hipLaunchKernelGGL(rank2_lower_kernel, g, dim3(256), 0, s, Lb, ldlb, n, Bb, ldbb, U, n, X, ldx, V, n, r, sLb, sBb,
                       sUV, sX, sUV);
}

}  // namespace mi
