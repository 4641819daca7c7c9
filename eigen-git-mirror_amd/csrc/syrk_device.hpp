// syrk_device.hpp -- one 128 x 128 tile of the trailing update C (lower) -= P P^T.
//
// Replaces the trailing rank update of llt_inplace::blocked, LLT.h:365
// (SelfadjointProduct.h:82-117 -> general_matrix_matrix_triangular_product<..,Lower>,
// GeneralMatrixMatrixTriangular.h:57-194, whose cycles go to gebp_kernel,
// GeneralBlockPanelKernel.h:1548-2217).  4 waves in a 2x2 grid of 64x64 wave tiles (4x4
// FP64 MFMA tiles each), operands from L2 (mfma_tile.hpp); the strictly-upper wave tile of a
// diagonal tile returns at once.
#pragma once
#include "common.hpp"
#include "mfma_tile.hpp"

namespace mi {

// Tile (ti, tj) of C (m x m, lower), columns restricted to j < ncol; depth K (compile time).
// T = double: FP64 MFMA; T = float: FP32 storage with the low-precision MFMA of PREC.
template <int K, typename T = double, int PREC = kPrecF64>
__device__ __forceinline__ void syrk_tile(T *__restrict__ C, int64_t ldc, const T *__restrict__ P, int64_t ldp,
                                          int m, int ncol, int ti, int tj)
{
    const int w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    if (ti == tj && wr == 0 && wc == 1) return;
    const int i0 = ti * 128 + wr * 64, j0 = tj * 128 + wc * 64;
    if constexpr (sizeof(T) == 8) {
        // C is read into the accumulators first, so its loads overlap the first operand
        // prefetch, and the negated product accumulates into it: the tile ends with stores only
        // (about 2% off the factorization at n = 8192 against a read of C after the last MFMA)
        WaveTile<4, 4, 3, true> t;
        const int rows[4] = {i0, i0 + 16, i0 + 32, i0 + 48};
        t.load_acc(C, ldc, rows, j0, m, ncol);
        t.setup(P, ldp, i0, m, P, ldp, j0, m, K);
        t.template run_fixed<K / 4>();
        t.template store<true, true>(C, ldc, i0, j0, m, ncol);
    } else {
        using Tile = typename TileSel<T, PREC>::type;
        Tile t;
        t.zero();
        t.setup(P, ldp, i0, m, P, ldp, j0, m, K);
        t.template run_fixed<K / Tile::KS>();
        t.template store<false, true>(C, ldc, i0, j0, m, ncol);
    }
}

// The split-BF16 form (FP32 storage): the K-deep operand comes from split-panel slots, S0 for
// depth columns 0..127 and S1 for 128..255; C's row 0 is row r0 of the slots.
template <int K>
__device__ __forceinline__ void syrk_tile_split(float *__restrict__ C, int64_t ldc, const __bf16 *S0, const __bf16 *S1,
                                                int64_t lo_off, int r0, int m, int ncol, int ti, int tj)
{
    const int w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    if (ti == tj && wr == 0 && wc == 1) return;
    const int i0 = ti * 128 + wr * 64, j0 = tj * 128 + wc * 64;
    WaveTileSplit<4, 4, 1> t;
    t.zero();
    t.setup(S0, S1, lo_off, r0, i0, m, j0, m);
    t.template run_fixed<K / 32>();
    t.template store<false, true>(C, ldc, i0, j0, m, ncol);
}

}  // namespace mi
