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
// The solve's block inverses L_bb^{-1} are built after the factor from its leaf inverses
// (block_inverse_kernel), and the couplings from them (block_coupling_kernel).
#include "potrf_device.hpp"

namespace mi {

// Packed lower 16 x 16 tiles of a 128 x 128 block in LDS, the diagonal included: tile (ti, tj),
// ti >= tj, at 256 (ti (ti + 1) / 2 + tj), column-major.
__device__ __forceinline__ int tile_at(int ti, int tj) { return 256 * ((ti * (ti + 1)) / 2 + tj); }

// acc += A B for 16 x 16 tiles at LDS offsets a and bb (column-major): the MFMA operands are
// P[i][k] = A[i][k] and Q[j][k] = B[k][j], k = 4s + lane/16; acc[r] is (lane%16, 4r + lane/16).
__device__ __forceinline__ void tile_mma(const double *S, int a, int bb, dvec4 &acc)
{
    const int l = lane_id(), i = l & 15, kl = l >> 4;
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = mfma_f64(S[bb + 4 * s + kl + 16 * i], S[a + i + 16 * (4 * s + kl)], acc);
}

// Workgroup b: X_b^T = L_bb^{-T} to Wout + b*kInvDoubles (ld 128) and X_b at + kInvX.
// A triangular inverse by a tree over the diagonal 16 x 16 leaves (the recursive form of
// LAPACK's dtrtri): from the 8 leaf inverses X_pp of the factorization, three levels double
// the inverted diagonal blocks, X_BA = -X_BB (L_BA X_AA) for [A; B] of 16 + 16, 32 + 32 and
// 64 + 64 rows, each level's products over 16 x 16 tiles on FP64 MFMA, in place in LDS (L_BA
// is replaced by the product T = L_BA X_AA, then by X_BA).  Three levels of two products each
// instead of eight dependent 16-column solve steps: the kernel is bound by one workgroup's
// latency, not by its flops.  Beyond bs (a last block narrower than 128) L reads as 0 and the
// leaf inverses are the identity's, so X is padded with the identity.
// Batched launches (every kernel of this file): system z = blockIdx.z, its pointers advanced
// by z times their strides (s*), its info at info + z.
template <typename T>
__global__ __launch_bounds__(256) void block_inverse_kernel(const T *__restrict__ L, int64_t ldl, int64_t n,
                                                            const double *__restrict__ Wleaf,
                                                            double *__restrict__ Wout, const int *info, int64_t sL,
                                                            int64_t sWl, int64_t sWo)
{
    __shared__ double S[kSmem];  // the 36 packed lower tiles
    const int64_t z = blockIdx.z;
    L += z * sL;
    Wleaf += z * sWl;
    Wout += z * sWo;
    if (info) info += z;
    if (factor_failed(info)) return;
    const int b = (int)blockIdx.x;
    const int64_t kb = (int64_t)b * 128;
    const int bs = (int)((n - kb) < 128 ? (n - kb) : 128);
    const T *Lb = L + kb + kb * ldl;
    const double *Xl = Wleaf + (size_t)b * kPanelDoubles + kLeafOffset;
    const int tid = threadIdx.x, l = lane_id();
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);

    // the strictly lower tiles of L_bb (0 beyond bs) and the leaf inverses on the diagonal: a
    // thread takes row r, columns of parity tid / 128, its loads issued 16 at a time (from
    // clamped, in-bounds addresses) before their LDS stores
    {
        const int r = tid & 127, ch = tid >> 7, rt = r >> 4;
        double x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = Xl[tid + 256 * q];
#pragma unroll
        for (int q0 = 0; q0 < 64; q0 += 16) {
            double v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int c = 2 * (q0 + q) + ch;
                const bool in = r < bs && c < bs;
                v[q] = (double)Lb[(in ? r : 0) + (int64_t)(in ? c : 0) * ldl];
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int c = 2 * (q0 + q) + ch;
                const bool in = r < bs && c < bs;
                if ((c >> 4) < rt) S[tile_at(rt, c >> 4) + (r & 15) + 16 * (c & 15)] = in ? v[q] : 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {  // element tid of leaf q: (tid % 16, tid / 16)
            const int i = tid & 15, j = tid >> 4;
            S[tile_at(q, q) + tid] = j <= i ? x[q] : 0.0;
        }
    }
    __syncthreads();

    // level v: pairs of sz-tile blocks A (tiles A0 ..) and B (tiles A0 + sz ..); 4 sz output
    // tiles (B0 + ti, A0 + tj) per level, dealt over the 4 waves
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const int sz = 1 << v, nout = 4 * sz;
        dvec4 acc[4];
        // T = L_BA X_AA (X_AA lower: k >= tj)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc[q] = dvec4{0.0, 0.0, 0.0, 0.0};
            const int o = w + 4 * q;
            if (o < nout) {
                const int A0 = (o / (sz * sz)) * 2 * sz, B0 = A0 + sz, ti = (o % (sz * sz)) / sz, tj = o % sz;
                for (int k = tj; k < sz; ++k) tile_mma(S, tile_at(B0 + ti, A0 + k), tile_at(A0 + k, A0 + tj), acc[q]);
            }
        }
        __syncthreads();  // every wave has read L_BA
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int o = w + 4 * q;
            if (o < nout) {
                const int A0 = (o / (sz * sz)) * 2 * sz, B0 = A0 + sz, ti = (o % (sz * sz)) / sz, tj = o % sz;
#pragma unroll
                for (int r = 0; r < 4; ++r) S[tile_at(B0 + ti, A0 + tj) + (l & 15) + 16 * (4 * r + (l >> 4))] = acc[q][r];
            }
        }
        __syncthreads();
        // X_BA = -X_BB T (X_BB lower: k <= ti)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc[q] = dvec4{0.0, 0.0, 0.0, 0.0};
            const int o = w + 4 * q;
            if (o < nout) {
                const int A0 = (o / (sz * sz)) * 2 * sz, B0 = A0 + sz, ti = (o % (sz * sz)) / sz, tj = o % sz;
                for (int k = 0; k <= ti; ++k) tile_mma(S, tile_at(B0 + ti, B0 + k), tile_at(B0 + k, A0 + tj), acc[q]);
            }
        }
        __syncthreads();  // every wave has read T
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int o = w + 4 * q;
            if (o < nout) {
                const int A0 = (o / (sz * sz)) * 2 * sz, B0 = A0 + sz, ti = (o % (sz * sz)) / sz, tj = o % sz;
#pragma unroll
                for (int r = 0; r < 4; ++r) S[tile_at(B0 + ti, A0 + tj) + (l & 15) + 16 * (4 * r + (l >> 4))] = -acc[q][r];
            }
        }
        __syncthreads();
    }

    // X to Wout + kInvX and X^T to Wout (ld 128), zeros above the diagonal
    double *Wt = Wout + (size_t)b * kInvDoubles, *Wx = Wt + kInvX;
    for (int e = tid; e < 128 * 128; e += 256) {
        const int r = e & 127, c = e >> 7;
        const double x = (r >> 4) >= (c >> 4) ? S[tile_at(r >> 4, c >> 4) + (r & 15) + 16 * (c & 15)] : 0.0;
        Wx[r + 128 * c] = x;
        Wt[c + 128 * r] = x;
    }
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
