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
    using Tile = typename TileSel<T, PREC>::type;
    const int w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    if (ti == tj && wr == 0 && wc == 1) return;
    const int i0 = ti * 128 + wr * 64, j0 = tj * 128 + wc * 64;
    Tile t;
    t.zero();
    t.setup(P, ldp, i0, m, P, ldp, j0, m, K);
    t.template run_fixed<K / Tile::KS>();
    t.template store<false, true>(C, ldc, i0, j0, m, ncol);
}

}  // namespace mi
