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
#include "kernels.hpp"

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

// The split forms (FP32 storage) read from LDS: PARTS = 2 (hi, lo: ~16 significant bits; three
// products hi.hi, hi.lo, lo.hi) or 3 (hi, mid, lo: all 24 bits; six products hi.hi, hi.mid,
// mid.hi, hi.lo, lo.hi, mid.mid -- the dropped terms are below 2^-24 of the product).  The
// depth is nks k-steps of 32 columns; columns 128 q .. 128 q + 127 are those of the panel in
// ring slot (sp + q) % kSplitRing (at ring + slot * slot_elems), part t at + t * part_off; C's
// row 0 is row r0 of the slots.  Each k-step's operands (128 P rows and 128 Q rows x PARTS,
// 16 KiB per part) are staged once per workgroup in LDS (`lds` >= 32 KiB * PARTS; two waves
// read every operand row), the next k-step's loads held in registers meanwhile (a raw 16-byte
// vector type: HIP's uint4 struct would be kept in scratch memory), and the waves read their
// parts from LDS per half wave tile and a Q part at a time (fewer operand registers).  All
// four waves take part in the staging, so the strictly upper wave tile of a diagonal tile is
// computed too (and not stored).
template <int PARTS>
__device__ __forceinline__ void syrk_tile_split_lds(float *__restrict__ C, int64_t ldc, const __bf16 *ring, int64_t sp,
                                                    int64_t slot_elems, int64_t part_off, int nks, int r0, int m,
                                                    int ncol, int ti, int tj, char *lds)
{
    constexpr int NC = 4 * PARTS;  // chunks per thread per k-step
    const int S = nks;
    const int tid = threadIdx.x, l = lane_id();
    const int w = tid >> 6, wr = w >> 1, wc = w & 1;
    // 16-byte chunks (operand P / Q, part, group of 16 rows, lane chunk); thread t moves, for
    // each operand and part, the chunks (group t / 64 and t / 64 + 4, lane chunk t % 64): global
    // byte offset base(operand, group) + part * part_off * 2 + 16 (t % 64), LDS byte offset
    // 8192 (PARTS operand + part) + 1024 group + 16 (t % 64).
    const uint32_t gp = (uint32_t)(((r0 + ti * 128) >> 4) + (tid >> 6)) * 4096u + 16u * (uint32_t)(tid & 63);
    const uint32_t gq = (uint32_t)(((r0 + tj * 128) >> 4) + (tid >> 6)) * 4096u + 16u * (uint32_t)(tid & 63);
    const uint32_t pbytes = (uint32_t)(part_off * 2);
    // chunk i of a thread: operand i / (2 PARTS), group half (i / PARTS) % 2, part i % PARTS
#define MI_SL_GLOAD(x, s_)                                                                                   \
    do {                                                                                                     \
        const char *b_ = (const char *)(ring + ((sp + (s_) / 4) & (kSplitRing - 1)) * slot_elems) + 1024 * ((s_)&3); \
        _Pragma("unroll") for (int i_ = 0; i_ < NC; ++i_) x[i_] =                                            \
            *(const u32x4 *)(b_ + ((i_ / (2 * PARTS)) ? gq : gp) + 16384u * ((i_ / PARTS) & 1) + pbytes * (i_ % PARTS)); \
    } while (0)
#define MI_SL_LSTORE(x)                                                                                      \
    do {                                                                                                     \
        _Pragma("unroll") for (int i_ = 0; i_ < NC; ++i_)                                                    \
            *(u32x4 *)(lds + 8192 * (PARTS * (i_ / (2 * PARTS)) + i_ % PARTS) +                               \
                       1024 * ((tid >> 6) + 4 * ((i_ / PARTS) & 1)) + 16 * (tid & 63)) = x[i_];              \
    } while (0)
    auto part = [&](int op, int t, int g) -> bf16x8 {
        return *(const bf16x8 *)(lds + 8192 * (PARTS * op + t) + 1024 * g + 16 * l);
    };
    WaveTileF<4, 4, kPrecBF16x3> acc;  // FP32 accumulators and store
    acc.zero();
    u32x4 v[NC];
    MI_SL_GLOAD(v, 0);
    MI_SL_LSTORE(v);
    lds_barrier();
    for (int s = 0; s < S; ++s) {
        if (s + 1 < S) MI_SL_GLOAD(v, s + 1);  // in flight during this k-step
#pragma unroll
        for (int h = 0; h < 2; ++h) {  // row tiles 2h, 2h + 1
            bf16x8 ph[2], pm[2], pl[2], q[4];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                ph[a] = part(0, 0, 4 * wr + 2 * h + a);
                if (PARTS == 3) pm[a] = part(0, 1, 4 * wr + 2 * h + a);
                pl[a] = part(0, PARTS - 1, 4 * wr + 2 * h + a);
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) q[b] = part(1, PARTS - 1, 4 * wc + b);  // Q lo
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc.acc[2 * h + a][b] = mfma_bf16(q[b], ph[a], acc.acc[2 * h + a][b]);
            if (PARTS == 3) {
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int b = 0; b < 4; ++b) q[b] = part(1, 1, 4 * wc + b);  // Q mid
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        acc.acc[2 * h + a][b] = mfma_bf16(q[b], pm[a], acc.acc[2 * h + a][b]);
                        acc.acc[2 * h + a][b] = mfma_bf16(q[b], ph[a], acc.acc[2 * h + a][b]);
                    }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int b = 0; b < 4; ++b) q[b] = part(1, 0, 4 * wc + b);  // Q hi
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    acc.acc[2 * h + a][b] = mfma_bf16(q[b], pl[a], acc.acc[2 * h + a][b]);
                    if (PARTS == 3) acc.acc[2 * h + a][b] = mfma_bf16(q[b], pm[a], acc.acc[2 * h + a][b]);
                    acc.acc[2 * h + a][b] = mfma_bf16(q[b], ph[a], acc.acc[2 * h + a][b]);
                }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (s + 1 < S) {
            lds_barrier();  // every wave has read this k-step's parts
            MI_SL_LSTORE(v);
            lds_barrier();
        }
    }
    acc.template store<false, true>(C, ldc, ti * 128 + wr * 64, tj * 128 + wc * 64, m, ncol);
}
#undef MI_SL_GLOAD
#undef MI_SL_LSTORE

}  // namespace mi
