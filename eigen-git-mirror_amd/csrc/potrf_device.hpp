// potrf_device.hpp -- device-side pieces of the diagonal-block factor and the panel solve,
// shared by the stand-alone kernels (potrf_diag.hip) and the fused factorization step
// (factor_step.hip).  See potrf_diag.hip for the algorithms and their reference citations.
#pragma once
#include "common.hpp"
#include "kernels.hpp"
#include "mfma_tile.hpp"

namespace mi {

// The 128 x 128 diagonal block lives in LDS as its 36 lower 16 x 16 tiles (72 KiB instead of a
// full square), each column-major, tile (ti, tj) at 256 * (ti(ti+1)/2 + tj): so two
// workgroups fit on a CU.  sl(r, c) addresses element (r, c) for r >= c, or any element of a
// diagonal tile; an MFMA operand (16 rows of 4 consecutive columns) is then 64 consecutive
// doubles.
__device__ __forceinline__ int sl(int r, int c)
{
    const int ti = r >> 4, tj = c >> 4;
    return 256 * ((ti * (ti + 1)) / 2 + tj) + (r & 15) + 16 * (c & 15);
}
// The strictly lower tiles only (the panel solve keeps L11 without its diagonal tiles, whose
// inverses it uses instead): r >> 4 > c >> 4.
__device__ __forceinline__ int sls(int r, int c)
{
    const int ti = r >> 4, tj = c >> 4;
    return 256 * ((ti * (ti - 1)) / 2 + tj) + (r & 15) + 16 * (c & 15);
}

__device__ __forceinline__ double readlane_d(double x, int lane)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Panel p (columns c0..c0+15 of the 128 x 128 block in S, tiled: sl), waves 0..2 (wave 3 is
// meanwhile busy with the previous panel's deferred update, see potrf_in_lds).  Each of the
// three waves holds the 16 x 16 leaf in lanes 0..15 (row c0+l) and runs the column loop of
// LLT.h:320-334 on it (pivot test, sqrt, scale, rank-1 update) redundantly, so its lane
// broadcasts (readlane) need no barrier.  Lanes 16..63 hold the rows c0+16+48w+(l-16) below
// the leaf (up to 112 rows over the three waves) and receive the same column operations, which
// is the panel solve A_ip L_pp^{-T}; lanes 48..63 of wave 2 (never a row below) hold the rows
// of the 16 x 16 identity, which the same operations turn into L_pp^{-T}, i.e. column i of
// X_pp = L_pp^{-1} in lane 48+i, written to xleaf column-major.  1/sqrt from the hardware
// estimate and one Newton step, d = x/sqrt(x).
//
// The column operations run on every lane without lane predicates: at step j, v[j] *= 1/sqrt(x)
// gives sqrt(x) on lane j (where v[j] is the pivot x) and the scaled column below it, and the
// rank-1 update v[c] -= v[j] * L(c0+c, c0+j), c > j, leaves every lane's entries at or below
// its diagonal exact.  Entries above the diagonal of a leaf lane l (c > l) do change, but they
// are never broadcast (readlane reads lane c of column j < c) nor stored.  Without per-lane
// selects the loop holds no predicate masks, so the broadcasts get scalar registers of their
// own and column j+1 -- the next pivot -- is updated first, off the remaining updates.
// (Broadcasting the multipliers through LDS instead -- the leaf lanes store column j, every
// lane reads it back with uniform addresses -- measured slower: 23 vs 21.5 us per block.)
//
// Returns the failing panel column or -1, the same in the three waves.  The factored leaf is
// left in v (lanes 0..15): it is written back to S only after a barrier, since the other waves
// read the unfactored leaf from S here and may still be doing so when one wave is done.
__device__ __forceinline__ int panel_factor3(double *S, int c0, double *__restrict__ xleaf, double (&v)[16])
{
    const int l = lane_id(), w = threadIdx.x >> 6;
    const bool leaf = l < 16;
    const int rb = c0 + 16 + 48 * w + (l - 16);
    const bool ident = w == 2 && l >= 48;
    const bool below = !leaf && !ident && rb < 128;
    const int r = leaf ? c0 + l : (below ? rb : c0);
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = S[sl(r, c0 + c)];
    if (ident) {
#pragma unroll
        for (int c = 0; c < 16; ++c) v[c] = (c == l - 48) ? 1.0 : 0.0;
    }
    int fail = -1;
    double x = readlane_d(v[0], 0);  // pivot: leaf row j lives in lane j
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (fail < 0 && x <= 0.0) fail = j;  // LLT.h:327-330; a NaN pivot passes
        double rd = __builtin_amdgcn_rsq(x);
        rd = rd * (1.5 - 0.5 * x * rd * rd);
        v[j] *= rd;
        if (j + 1 < 16) {  // the next pivot first: it is the loop's serial chain
            v[j + 1] -= v[j] * readlane_d(v[j], j + 1);
            x = readlane_d(v[j + 1], j + 1);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int c = j + 2; c < 16; ++c) v[c] -= v[j] * readlane_d(v[j], c);  // L(c0+c, c0+j)
    }
    if (fail >= 0) return fail;
    if (below) {  // rows no other wave reads before the next barrier
#pragma unroll
        for (int c = 0; c < 16; ++c) S[sl(r, c0 + c)] = v[c];
    }
    if (ident) {
#pragma unroll
        for (int c = 0; c < 16; ++c) xleaf[c + 16 * (l - 48)] = v[c];
    }
    return -1;
}

// LDS of one workgroup, shared by the two roles: the 36 lower tiles of the block for the
// factor, or the 28 strictly lower tiles of L11 plus the 8 leaf inverses for the panel solve.
constexpr int kSmem = 36 * 256;

// A (global, ld lda, valid order bs) -> S (lower part, identity beyond bs); 32 loads in
// flight per thread, from clamped in-bounds addresses.  T = float widens to FP64.
template <typename T>
__device__ __forceinline__ void block_to_lds(double *S, const T *__restrict__ Ab, int64_t lda, int bs)
{
    const int tid = threadIdx.x, r = tid & 127, ch = tid >> 7;
#pragma unroll
    for (int q0 = 0; q0 < 64; q0 += 32) {
        T a[32];
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int c = 2 * (q0 + q) + ch;
            const bool in = r < bs && c < bs;
            a[q] = Ab[(in ? r : 0) + (int64_t)(in ? c : 0) * lda];
        }
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int c = 2 * (q0 + q) + ch;
            const bool in = r < bs && c < bs;
            if (r >= c) S[sl(r, c)] = in ? (double)a[q] : (r == c ? 1.0 : 0.0);
        }
    }
}

template <typename T>
__device__ __forceinline__ void lds_to_block(const double *S, T *__restrict__ Ab, int64_t lda, int bs)
{
    const int tid = threadIdx.x, r = tid & 127, ch = tid >> 7;
#pragma unroll 8
    for (int q = 0; q < 64; ++q) {
        const int c = 2 * q + ch;
        if (r >= c && r < bs) Ab[r + (int64_t)c * lda] = (T)S[sl(r, c)];
    }
}

// One 16 x 16 tile of the rank-16 update in LDS: C -= Pi Pj^T (operands in registers, lane l
// k-step s: row l%16, column 4s + l/16), C at tile offset + l + 64 r.  A diagonal tile is
// updated whole: the entries above its diagonal are never read as data.
__device__ __forceinline__ void lds_tile_update(double *C, const double (&Pi)[4], const double (&Pj)[4])
{
    double old[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) old[r] = C[64 * r];
    dvec4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = mfma_f64(Pj[s], Pi[s], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) C[64 * r] = old[r] - acc[r];
}

// Cholesky of the 128 x 128 lower block in S (all 256 threads), in eight 16-column panels with
// a one-panel lookahead inside the block.  After panel p is factored, with T = 7-p tile rows
// below it and their operands P_x (tile (p+1+x, p), x < T) in every wave's registers:
//   A(p): the leaf of p goes back to S; the T tiles of column p+1 get panel p's update (the
//         waves round-robin), barrier;
//   B(p): waves 0..2 factor panel p+1 (panel_factor3: VALU and lane broadcasts) while wave 3
//         applies panel p to the tiles right of column p+1 on the matrix core -- up to kCap of
//         them, any more go to waves 0..2 after their column loop -- barrier.
// Column p+1 thus has every earlier panel's update before its factor, the tiles beyond it get
// panel p's before phase A(p+1) touches them, and the rank-16 updates mostly run under the
// column loop.  Returns the first failing column or -1 (uniform); writes the 8 leaf inverses
// to Xl.  s_fail: an LDS int.
__device__ __forceinline__ int potrf_in_lds(double *S, double *__restrict__ Xl, int *s_fail)
{
    constexpr int kCap = 12;  // deferred tiles taken by wave 3 (about one column loop's worth)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = lane_id();
    double lv[16];
    __syncthreads();
    if (w < 3) {
        const int f = panel_factor3(S, 0, Xl, lv);
        if (w == 0 && l == 0) *s_fail = f;
    }
    __syncthreads();
    {
        const int f = __builtin_amdgcn_readfirstlane(*s_fail);
        if (f >= 0) return f;
    }
    for (int p = 0; p < 7; ++p) {
        const int T = 7 - p;
        const int tb = p + 1;  // first trailing tile row / column
        // ---- A(p)
        if (w == 0 && l < 16) {  // the leaf of p: read by nobody until the final copy out
#pragma unroll
            for (int c = 0; c < 16; ++c) S[sl(16 * p + l, 16 * p + c)] = lv[c];
        }
        double P[7][4];
#pragma unroll
        for (int x = 0; x < 7; ++x) {
            if (x < T) {
                const double *src = S + 256 * (((tb + x) * (tb + x + 1)) / 2 + p) + l;
#pragma unroll
                for (int s = 0; s < 4; ++s) P[x][s] = src[64 * s];
            }
        }
#pragma unroll
        for (int x = 0; x < 7; ++x)
            if (x < T && (x & 3) == w) lds_tile_update(S + 256 * (((tb + x) * (tb + x + 1)) / 2 + tb) + l, P[x], P[0]);
        __syncthreads();
        // ---- B(p)
        if (w < 3) {
            const int f = panel_factor3(S, 16 * tb, Xl + 256 * tb, lv);
            if (w == 0 && l == 0) *s_fail = f;
        }
#pragma unroll
        for (int ti = 1; ti < 7; ++ti) {
#pragma unroll
            for (int tj = 1; tj <= ti; ++tj) {
                const int t = ((ti - 1) * ti) / 2 + (tj - 1);  // enumerates the deferred tiles
                const int owner = t < kCap ? 3 : (t - kCap) % 3;
                if (ti < T && w == owner)
                    lds_tile_update(S + 256 * (((tb + ti) * (tb + ti + 1)) / 2 + tb + tj) + l, P[ti], P[tj]);
            }
        }
        __syncthreads();
        const int f = __builtin_amdgcn_readfirstlane(*s_fail);
        if (f >= 0) return 16 * tb + f;
    }
    if (w == 0 && l < 16) {  // the last leaf
#pragma unroll
        for (int c = 0; c < 16; ++c) S[sl(112 + l, 112 + c)] = lv[c];
    }
    __syncthreads();
    return -1;
}

// The panel solve works on a wave's 32 rows x 128 columns in the FP64 MFMA result layout:
// X[t][p][r] is row 16t + lane%16, column 16p + 4r + lane/16 (t = 0..1 row tiles, p = 0..7
// column strips).  It needs, in LDS, the 28 strictly lower tiles of L (Ls, sls order) and the
// 8 leaf inverses (Xs, 256 doubles each, column-major): the panel image of kernels.hpp.
//
// X <- X L^{-T} in place, strip by strip (TriangularSolverMatrix.h:186-329, blocked by 16):
//   X_p <- (X_p - sum_{q<p} X_q L_pq^T) X_pp^T,   X_pp = L_pp^{-1}
// with the accumulator of a solved strip used directly as the operand of the next ones (its
// layout is that of an MFMA P operand with k = 4s + lane/16).
__device__ __forceinline__ void trsm_in_registers(const double *smem, dvec4 (&X)[2][8])
{
    const int l = lane_id();
    const double *Ls = smem;
    const double *Xs = smem + kLeafOffset;
#pragma unroll
    for (int p = 0; p < 8; ++p) {
#pragma unroll
        for (int q = 0; q < p; ++q) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const double lv = Ls[256 * ((p * (p - 1)) / 2 + q) + l + 64 * s];  // L[16p + l%16][16q + 4s + l/16]
#pragma unroll
                for (int t = 0; t < 2; ++t) X[t][p] = mfma_f64(-lv, X[t][q][s], X[t][p]);
            }
        }
        dvec4 z[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const double xv = Xs[256 * p + l + 64 * s];  // X_pp[l%16][4s + l/16]
#pragma unroll
            for (int t = 0; t < 2; ++t) z[t] = mfma_f64(xv, X[t][p][s], z[t]);
        }
        X[0][p] = z[0];
        X[1][p] = z[1];
    }
}

// L (128 x 128 lower, ld ldl, valid order bs, identity beyond; T = float widens) and the leaf
// inverses Xl -> the panel image in LDS.
template <typename T>
__device__ __forceinline__ void stage_factor_image(double *smem, const T *__restrict__ L, int64_t ldl, int bs,
                                                   const double *__restrict__ Xl)
{
    for (int e = threadIdx.x; e < 128 * 128; e += 256) {  // strictly lower tiles of L
        const int r = e & 127, c = e >> 7;
        if ((r >> 4) > (c >> 4)) {
            const bool in = r < bs && c < bs;
            const double v = (double)L[(in ? r : 0) + (int64_t)(in ? c : 0) * ldl];
            smem[sls(r, c)] = in ? v : 0.0;
        }
    }
    for (int e = threadIdx.x; e < 8 * 256; e += 256) smem[kLeafOffset + e] = Xl[e];
}

// A published panel image (kPanelDoubles, 16-byte aligned) -> LDS, 16 bytes per load.
__device__ __forceinline__ void stage_panel_image(double *smem, const double *__restrict__ W)
{
    const double2 *src = reinterpret_cast<const double2 *>(W);
    double2 *dst = reinterpret_cast<double2 *>(smem);
#pragma unroll
    for (int i = 0; i < (int)(kPanelDoubles / 2 / 256); ++i) dst[threadIdx.x + 256 * i] = src[threadIdx.x + 256 * i];
}

// Rows 0..127 of P (valid m) <- P * L^{-T} with L and its 8 leaf inverses Xl (read from
// global into the LDS panel image).  Each wave owns 32 rows.  T = float: P and L are read as
// FP32 and widened, the arithmetic is FP64, the result is rounded back to FP32.  If IMAGE is given, the panel image is copied from there instead of built from
// L and Xl.  If SPLIT (a split-panel slot) is given (T = float), the rounded rows are also
// written split into split_parts BF16 parts (2 or 3), row i as row split_row0 + i of the slot
// (split_store_row; part t at split + t * split_lo).
template <typename T>
__device__ __forceinline__ void trsm_rows(double *smem, T *__restrict__ P, int64_t ldp, int m,
                                          const T *__restrict__ L, int64_t ldl, int bs,
                                          const double *__restrict__ Xl,
                                          const double *__restrict__ image = nullptr, __bf16 *split = nullptr,
                                          int64_t split_lo = 0, int64_t split_row0 = 0, int split_parts = 2)
{
    const int l = lane_id(), w = threadIdx.x >> 6;
    const int row0 = 32 * w;
    dvec4 X[2][8];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int i = row0 + 16 * t + (l & 15);
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = 16 * p + 4 * r + (l >> 4);
                const double pv = (double)P[(i < m ? i : m - 1) + (int64_t)j * ldp];
                X[t][p][r] = (i < m) ? pv : 0.0;
            }
    }
    if (image)
        stage_panel_image(smem, image);
    else
        stage_factor_image<T>(smem, L, ldl, bs, Xl);
    __syncthreads();
    trsm_in_registers(smem, X);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int i = row0 + 16 * t + (l & 15);
        if (i < m) {
#pragma unroll
            for (int p = 0; p < 8; ++p)
#pragma unroll
                for (int r = 0; r < 4; ++r) P[i + (int64_t)(16 * p + 4 * r + (l >> 4)) * ldp] = (T)X[t][p][r];
            if constexpr (sizeof(T) == 4) {
                if (split) {
                    float v[8][4];
#pragma unroll
                    for (int p = 0; p < 8; ++p)
#pragma unroll
                        for (int r = 0; r < 4; ++r) v[p][r] = (float)X[t][p][r];
                    if (split_parts == 3)
                        split_store_row<3>(split, split_lo, split_row0 + i, v);
                    else
                        split_store_row<2>(split, split_lo, split_row0 + i, v);
                }
            }
        }
    }
}

// FP64 update of the diagonal block straight into LDS: S = A_kk - P_0 P_0^T for tile rows W
// and 7-W, their tiles at and left of the diagonal -- 9 tiles per wave, the 36 lower tiles
// over the 4 waves -- with acc starting from A_kk; identity beyond bs.  A diagonal tile is
// written whole (its upper entries are never read as data).
template <int W>
__device__ __forceinline__ void diag_update_f64(double *S, const double *__restrict__ Akk,
                                                const double *__restrict__ P0, int64_t lda, int bs)
{
    constexpr uint32_t kActive = ((1u << (W + 1)) - 1u) | (((1u << (8 - W)) - 1u) << 8);
    WaveTile<2, 8, 2, true, kActive> t;
    const int rows[2] = {16 * W, 16 * (7 - W)};
    t.zero();
    t.load_acc(Akk, lda, rows, 0, bs, bs);
    t.setup_rows(P0, lda, rows, bs, P0, lda, 0, bs, 128);
    t.template run_fixed<32>();
    const int l = lane_id();
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int ti = a == 0 ? W : 7 - W;
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (t.on(a, b)) {
                double *C = S + 256 * ((ti * (ti + 1)) / 2 + b) + l;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * ti + (l & 15), j = 16 * b + 4 * r + (l >> 4);
                    C[64 * r] = (i < bs && j < bs) ? t.acc[a][b][r] : (i == j ? 1.0 : 0.0);
                }
            }
    }
}

// The same for FP32 storage with low-precision MFMA (the mixed-precision factorization): 64 x
// 64 wave tiles over the three lower quadrants, the old values read after the MFMAs, in two
// halves of tile rows.
template <typename T, int PREC>
__device__ __forceinline__ void diag_update_lowp(double *S, const T *__restrict__ A, int64_t lda, int64_t k0, int bs)
{
    const int w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    if (wr == 0 && wc == 1) return;  // strictly upper quadrant
    const T *Ck = A + k0 + k0 * lda;
    using Tile = typename TileSel<T, PREC>::type;
    Tile t;
    t.zero();
    t.setup(A + k0 + (k0 - 128) * lda, lda, wr * 64, bs, A + k0 + (k0 - 128) * lda, lda, wc * 64, bs, 128);
    t.template run_fixed<128 / Tile::KS>();
#pragma unroll
    for (int a0 = 0; a0 < 4; a0 += 2) {
        T old[2][4][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int i = wr * 64 + t.row_of(a0 + a), j = wc * 64 + t.col_of(b, r);
                    i = i < bs ? i : bs - 1;
                    j = j < bs ? j : bs - 1;
                    old[a][b][r] = Ck[i + (int64_t)j * lda];
                }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = wr * 64 + t.row_of(a0 + a), j = wc * 64 + t.col_of(b, r);
                    if (i >= j) {
                        const bool in = i < bs && j < bs;
                        S[sl(i, j)] = in ? (double)old[a][b][r] - (double)t.acc[a0 + a][b][r] : (i == j ? 1.0 : 0.0);
                    }
                }
    }
}

// The diagonal block's strictly lower tiles in S (sl order) -> the panel image W (sls order).
__device__ __forceinline__ void publish_lower_tiles(const double *S, double *__restrict__ W)
{
#pragma unroll
    for (int q = 0; q < 28; ++q) {
        const int ti = 1 + (q >= 1) + (q >= 3) + (q >= 6) + (q >= 10) + (q >= 15) + (q >= 21);
        W[256 * q + threadIdx.x] = S[256 * (q + ti) + threadIdx.x];  // sl tile index = sls index + ti
    }
}

// Waits (thread 0) until *flag >= epoch, then every wave acquires: each invalidates its own
// view of the caches, so that the loads it issues next see what the flag's release published.
// Returns false (uniformly) on a timeout, recorded as MI_PANEL_TIMEOUT.
__device__ __forceinline__ bool wait_panel_flag(int *s_ok, const uint32_t *flag, uint32_t epoch, int *status)
{
    if (threadIdx.x == 0) {
        uint32_t spins = 0;
        int ok = 1;
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1u << 22)) {
                atomicOr(status, MI_PANEL_TIMEOUT);
                ok = 0;
                break;
            }
        }
        *s_ok = ok;
    }
    __syncthreads();
    const bool ok = *s_ok != 0;
    if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok;
}

// The panel role of one block step; g is the role from the ticket:
//   role 0   : C = A_kk - P_0 P_0^T straight into LDS, Cholesky in LDS, the panel image
//              (L11's strictly lower tiles and the leaf inverses) to the workspace, publish
//              `flag` = epoch (release), then L11 into A;
//   role g>0 : rows k0+128g..+127: C = A_gk - P_g P_0^T, wait for the flag, then rows <-
//              rows * L11^{-T} with the published image, stored once.
// The diagonal role publishes even on a failing pivot, so that no waiter is left behind;
// waiters then see `info` and stop.  (LLT.h:350-366: the update of block column k by all
// earlier panels precedes its unblocked factor and the right-side solve.)
// smem: kSmem doubles of LDS; s_ok: an LDS int.  T/PREC: storage and update arithmetic
// (the diagonal factor and the panel solve are FP64 for every T).  With T = double the row
// role keeps its rows in registers from the update to the store; with T = float the update
// is rounded to FP32 storage first (as the mixed-precision trailing updates are) and read back.
template <typename T, int PREC>
__device__ __forceinline__ void panel_role(int g, double *smem, int *s_ok, T *__restrict__ A,
                                           int64_t lda, int64_t n, int64_t k0, double *__restrict__ Wleaf, int *info,
                                           uint32_t *flag, uint32_t epoch, int *status, __bf16 *split = nullptr,
                                           int64_t split_lo = 0)
{
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t R0 = k0 + 128 * (int64_t)g;
    const int rows = (int)((n - R0) < 128 ? (n - R0) : 128);
    const int bs = (int)((n - k0) < 128 ? (n - k0) : 128);
    T *Ck = A + R0 + k0 * lda;  // this role's 128 rows of block column k0
    double *Wp = Wleaf + (size_t)(k0 / 128) * kPanelDoubles;  // the panel image
    double *Xl = Wp + kLeafOffset;

    if (g == 0) {
        if (k0 > 0) {
            if constexpr (sizeof(T) == 8) {
                const double *P0 = A + k0 + (k0 - 128) * lda;
                switch (w) {
                case 0: diag_update_f64<0>(smem, Ck, P0, lda, bs); break;
                case 1: diag_update_f64<1>(smem, Ck, P0, lda, bs); break;
                case 2: diag_update_f64<2>(smem, Ck, P0, lda, bs); break;
                default: diag_update_f64<3>(smem, Ck, P0, lda, bs); break;
                }
            } else {
                diag_update_lowp<T, PREC>(smem, A, lda, k0, bs);
            }
        } else {
            block_to_lds(smem, Ck, lda, bs);
        }
        const int f = potrf_in_lds(smem, Xl, s_ok);  // begins with a barrier
        if (f >= 0) {
            if (tid == 0) atomicCAS(info, 0, (int)(k0 + f + 1));
        } else {
            publish_lower_tiles(smem, Wp);
        }
        global_barrier();  // every wave's stores of the image have reached L2
        if (tid == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (f < 0) lds_to_block(smem, Ck, lda, bs);  // read by later launches only
        return;
    }

    if constexpr (sizeof(T) == 8) {
        WaveTile<2, 8, 2, true, 0xffffffffu, true> t;  // P_0 has all 128 rows (k0 + 128 < n)
        const int r2[2] = {32 * w, 32 * w + 16};
        t.zero();
        t.load_acc(Ck, lda, r2, 0, rows, 128);
        if (k0 > 0) {  // C -= P_g P_0^T, P = previous panel (columns k0-128..k0-1): rows R0.. and k0..
            t.setup_rows(A + R0 + (k0 - 128) * lda, lda, r2, rows, A + k0 + (k0 - 128) * lda, lda, 0, 128, 128);
            t.template run_fixed<32>();
        }
        if (!wait_panel_flag(s_ok, flag, epoch, status) || factor_failed(info)) return;
        stage_panel_image(smem, Wp);
        __syncthreads();
        trsm_in_registers(smem, t.acc);
        const int l = lane_id();
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int i = r2[a] + (l & 15);
            if (i < rows) {
#pragma unroll
                for (int b = 0; b < 8; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) Ck[i + (int64_t)(16 * b + 4 * r + (l >> 4)) * lda] = t.acc[a][b][r];
            }
        }
    } else {
        if (k0 > 0) {
            const int wr = w >> 1, wc = w & 1;
            using Tile = typename TileSel<T, PREC>::type;
            Tile t;
            t.zero();
            t.setup(A + R0 + (k0 - 128) * lda, lda, wr * 64, rows, A + k0 + (k0 - 128) * lda, lda, wc * 64, bs, 128);
            t.template run_fixed<128 / Tile::KS>();
            t.template store<false, false>(Ck, lda, wr * 64, wc * 64, rows, 128);
        }
        global_barrier();  // this role's updated rows are in L2 before other waves read them back
        if (!wait_panel_flag(s_ok, flag, epoch, status) || factor_failed(info)) return;
        trsm_rows<T>(smem, Ck, lda, rows, (const T *)nullptr, 0, 128, nullptr, Wp, split, split_lo,
                            R0, PREC == kPrecBF16x6 ? 3 : 2);
    }
}

}  // namespace mi
