// mfma_tile.hpp -- FP64 MFMA wave-tile engine with operands read straight from global
// memory (L2) into registers.
//
// An FP64 MFMA 16x16x4 occupies the matrix core for 64 cycles, so a wave tile of
// TA x TB MFMA tiles consumes 4 k-values of (16*TA + 16*TB) rows every 64*TA*TB cycles:
// a few bytes per cycle per CU, far below L2 bandwidth.  Staging through LDS would only add
// barriers; instead each wave loads its own fragments and keeps D k-steps of loads in
// flight in a rolling register buffer, so a workgroup's waves never synchronise.
//
// Operand convention (see mfma_f64 in common.hpp): C[i][j] += sum_k P[i][k] * Q[j][k],
// both operands "rows x k", column-major: element (r, k) at X[r + k * ldx].  Rows past the
// operand's end are clamped (their products land in C rows/cols the caller never stores);
// k past kmax reads as 0.
#pragma once
#include "common.hpp"

namespace mi {

// Addressing: each operand has a wave-uniform base pointer, advanced by 4 columns per k-step
// (scalar registers), and a per-lane 32-bit byte offset of (its row, column lane/16), fixed
// for the whole run -- global loads in the scalar-base + vector-offset form, no per-load
// address arithmetic.  Requires (row + 3 * ld) * 8 < 4 GiB for both operands.
// NEG: accumulate -P Q^T (acc starts from C for C - P Q^T, see load_acc).  ACTIVE: bit
// a*TB+b set for the MFMA tiles computed; the others are never issued (an irregular tile set,
// e.g. the lower-triangle tiles of two tile rows).  QDENSE: the 16*TB rows of Q are all valid
// (no clamping), so one offset plus constants serves them (fewer registers).
template <int TA, int TB, int D = 3, bool NEG = false, uint32_t ACTIVE = 0xffffffffu, bool QDENSE = false>
struct WaveTile {
    static_assert(TA * TB <= 32, "ACTIVE has one bit per MFMA tile");
    static constexpr int U = D + 1;  // rolling buffer depth
    static constexpr int KS = 4;     // k per MFMA step
    __device__ __forceinline__ static constexpr bool on(int a, int b) { return (ACTIVE >> (a * TB + b)) & 1u; }
    dvec4 acc[TA][TB];
    const double *pbase, *qbase;  // wave-uniform
    uint32_t po[TA], qo[TB];      // per-lane byte offsets
    int64_t ldp, ldq;
    int kmax;  // number of valid k (a multiple of 4)

    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int a = 0; a < TA; ++a)
#pragma unroll
            for (int b = 0; b < TB; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
    }

    // P rows p0[a] .. p0[a]+15 for tile row a (valid rows < pmax; the rows of different tile
    // rows need not be adjacent), Q rows q0 .. q0+16*TB-1 (valid < qmax).
    __device__ __forceinline__ void setup_rows(const double *P, int64_t ldp_, const int (&p0)[TA], int pmax,
                                               const double *Q, int64_t ldq_, int q0, int qmax, int kmax_)
    {
        const int l = lane_id();
        pbase = P;
        qbase = Q;
        ldp = ldp_;
        ldq = ldq_;
        kmax = kmax_;
        const int kl = l >> 4;
#pragma unroll
        for (int a = 0; a < TA; ++a) {
            int r = p0[a] + (l & 15);
            r = r < pmax ? r : pmax - 1;
            po[a] = (uint32_t)((r + kl * ldp_) * 8);
        }
        if (QDENSE) {
            qo[0] = (uint32_t)((q0 + (l & 15) + kl * ldq_) * 8);
        } else {
#pragma unroll
            for (int b = 0; b < TB; ++b) {
                int r = q0 + 16 * b + (l & 15);
                r = r < qmax ? r : qmax - 1;
                qo[b] = (uint32_t)((r + kl * ldq_) * 8);
            }
        }
    }

    // P rows p0 .. p0+16*TA-1 (valid rows < pmax), Q rows q0 .. (valid < qmax).
    __device__ __forceinline__ void setup(const double *P, int64_t ldp_, int p0, int pmax, const double *Q,
                                          int64_t ldq_, int q0, int qmax, int kmax_)
    {
        int rows[TA];
#pragma unroll
        for (int a = 0; a < TA; ++a) rows[a] = p0 + 16 * a;
        setup_rows(P, ldp_, rows, pmax, Q, ldq_, q0, qmax, kmax_);
    }

    // acc <- C for the active tiles: tile (a, b) is rows p0[a].., columns col0 + 16b..;
    // loads from clamped (in-bounds) addresses, i < imax, j < jmax.
    __device__ __forceinline__ void load_acc(const double *C, int64_t ldc, const int (&p0)[TA], int col0, int imax,
                                             int jmax)
    {
#pragma unroll
        for (int a = 0; a < TA; ++a)
#pragma unroll
            for (int b = 0; b < TB; ++b)
                if (on(a, b)) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        int i = p0[a] + (lane_id() & 15), j = col0 + col_of(b, r);
                        i = i < imax ? i : imax - 1;
                        j = j < jmax ? j : jmax - 1;
                        acc[a][b][r] = C[i + (int64_t)j * ldc];
                    }
                }
    }

    struct Frag {
        double p[TA], q[TB];
    };

    __device__ __forceinline__ static double at(const double *base, uint32_t off)
    {
        return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + off);
    }

    // Loads of k-step s (clamped to the last one: prefetches past the end re-read it).
    __device__ __forceinline__ void load(Frag &f, int s, int S) const
    {
        s = s < S ? s : S - 1;
        const double *pb = pbase + (int64_t)(4 * s) * ldp;
        const double *qb = qbase + (int64_t)(4 * s) * ldq;
#pragma unroll
        for (int a = 0; a < TA; ++a) f.p[a] = at(pb, po[a]);
#pragma unroll
        for (int b = 0; b < TB; ++b) f.q[b] = QDENSE ? at(qb + 16 * b, qo[0]) : at(qb, qo[b]);
    }

    __device__ __forceinline__ void mma(const Frag &f, int /*s*/)
    {
#pragma unroll
        for (int a = 0; a < TA; ++a) {
            const double p = NEG ? -f.p[a] : f.p[a];
#pragma unroll
            for (int b = 0; b < TB; ++b)
                if (on(a, b)) acc[a][b] = mfma_f64(f.q[b], p, acc[a][b]);
        }
    }

    // Accumulate over k in [0, kmax), 4 k per MFMA step, D steps of loads in flight.  The
    // loop body has no branches, so the compiler's wait counts stay exact: the step count
    // is padded to a multiple of U, which requires kmax % (4U) == 0.
    __device__ __forceinline__ void run()
    {
        const int S = (kmax + 3) >> 2;
        const int Sp = (S + U - 1) / U * U;
        Frag buf[U];
#pragma unroll
        for (int u = 0; u < D; ++u) load(buf[u], u, S);
        for (int s0 = 0; s0 < Sp; s0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int s = s0 + u;
                load(buf[(u + D) % U], s + D, S);
                mma(buf[u], s);
                // keep each step's loads where they are: hoisting them reorders the wait
                // counts so a step ends up waiting for the newest prefetches
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }

    // Same for a compile-time number of steps S (kmax == 4S), fully unrolled: without a loop
    // back edge the wait counts are exact for every step.
    template <int S>
    __device__ __forceinline__ void run_fixed()
    {
        Frag buf[U];
#pragma unroll
        for (int u = 0; u < D; ++u) load(buf[u], u, S);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (s + D < S) load(buf[(s + D) % U], s + D, S);
            mma(buf[s % U], s);
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // C[i][j] -= acc (or = acc when ASSIGN) for the elements with i < imax, j < jmax and,
    // when LOWER, j <= i; i = row0 + row_of(a), j = col0 + col_of(b, r).  In groups of HA
    // tile rows, all loads of C are issued before any store, so they overlap (two groups keep
    // the old values within the register budget of two waves per SIMD).
    template <bool ASSIGN, bool LOWER>
    __device__ __forceinline__ void store(double *C, int64_t ldc, int row0, int col0, int imax, int jmax) const
    {
        constexpr int HA = TA > 1 ? TA / 2 : 1;
#pragma unroll
        for (int a0 = 0; a0 < TA; a0 += HA) {
            double old[HA][TB][4];
            if (!ASSIGN) {  // unconditional loads from clamped (in-bounds) addresses
#pragma unroll
                for (int a = 0; a < HA; ++a)
#pragma unroll
                    for (int b = 0; b < TB; ++b)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            int i = row0 + row_of(a0 + a), j = col0 + col_of(b, r);
                            i = i < imax ? i : imax - 1;
                            j = j < jmax ? j : jmax - 1;
                            old[a][b][r] = C[i + (int64_t)j * ldc];
                        }
            }
#pragma unroll
            for (int a = 0; a < HA; ++a)
#pragma unroll
                for (int b = 0; b < TB; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = row0 + row_of(a0 + a), j = col0 + col_of(b, r);
                        const bool ok = i < imax && j < jmax && (!LOWER || j <= i);
                        if (ok)
                            C[i + (int64_t)j * ldc] =
                                ASSIGN ? acc[a0 + a][b][r] : old[a][b][r] - acc[a0 + a][b][r];
                    }
        }
    }

    // Element r (0..3) of tile (a, b) held by this lane is C[i][j] with
    // i = row0 + 16a + (lane % 16), j = col0 + 16b + 4r + lane / 16.
    __device__ __forceinline__ static int row_of(int a) { return 16 * a + (lane_id() & 15); }
    __device__ __forceinline__ static int col_of(int b, int r) { return 16 * b + 4 * r + (lane_id() >> 4); }
};

// The same engine for FP32 storage with low-precision MFMA (FP32 accumulation), used by the
// mixed-precision factorization: PREC = kPrecF32 (v_mfma_f32_16x16x4_f32, one k per lane
// per step), kPrecBF16 (v_mfma_f32_16x16x32_bf16, 8 consecutive k per lane, operands
// rounded to BF16 when issued) or kPrecBF16x3 (each operand split x = hi + lo into two BF16
// parts; hi*hi + lo*hi + hi*lo on three MFMAs, ~16 significant bits per product).  The
// depth is a multiple of KS (no masking); result element r of tile (a, b) is
// C[row0 + 16a + lane%16][col0 + 16b + 4*(lane/16) + r].
template <int TA, int TB, int PREC>
struct WaveTileF {
    static constexpr int E = PREC == kPrecF32 ? 1 : 8;  // k per lane per MFMA
    static constexpr int KS = 4 * E;                     // k per MFMA step
    static constexpr int D = E == 1 ? 3 : 1;             // steps of loads in flight
    static constexpr int U = D + 1;
    fvec4 acc[TA][TB];
    const float *pa[TA];
    const float *qb[TB];
    int64_t ldp, ldq;

    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int a = 0; a < TA; ++a)
#pragma unroll
            for (int b = 0; b < TB; ++b) acc[a][b] = fvec4{0.f, 0.f, 0.f, 0.f};
    }

    __device__ __forceinline__ void setup(const float *P, int64_t ldp_, int p0, int pmax, const float *Q,
                                          int64_t ldq_, int q0, int qmax, int /*kmax*/)
    {
        const int l = lane_id();
        ldp = ldp_;
        ldq = ldq_;
#pragma unroll
        for (int a = 0; a < TA; ++a) {
            int r = p0 + 16 * a + (l & 15);
            r = r < pmax ? r : pmax - 1;
            pa[a] = P + r;
        }
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            int r = q0 + 16 * b + (l & 15);
            r = r < qmax ? r : qmax - 1;
            qb[b] = Q + r;
        }
    }

    struct Frag {
        float p[TA][E], q[TB][E];
    };

    __device__ __forceinline__ void load(Frag &f, int s, int S) const
    {
        s = s < S ? s : S - 1;
        const int kb = KS * s + E * (lane_id() >> 4);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int64_t op = (int64_t)(kb + e) * ldp, oq = (int64_t)(kb + e) * ldq;
#pragma unroll
            for (int a = 0; a < TA; ++a) f.p[a][e] = pa[a][op];
#pragma unroll
            for (int b = 0; b < TB; ++b) f.q[b][e] = qb[b][oq];
        }
    }

    __device__ __forceinline__ void mma(const Frag &f, int)
    {
        if constexpr (PREC == kPrecF32) {
#pragma unroll
            for (int a = 0; a < TA; ++a)
#pragma unroll
                for (int b = 0; b < TB; ++b) acc[a][b] = mfma_f32(f.q[b][0], f.p[a][0], acc[a][b]);
        } else {
            bf16x8 ph[TA], qh[TB], pl[TA], ql[TB];
#pragma unroll
            for (int e = 0; e < E; ++e) {
#pragma unroll
                for (int a = 0; a < TA; ++a) {
                    ph[a][e] = (__bf16)f.p[a][e];
                    if (PREC == kPrecBF16x3) pl[a][e] = (__bf16)(f.p[a][e] - (float)ph[a][e]);
                }
#pragma unroll
                for (int b = 0; b < TB; ++b) {
                    qh[b][e] = (__bf16)f.q[b][e];
                    if (PREC == kPrecBF16x3) ql[b][e] = (__bf16)(f.q[b][e] - (float)qh[b][e]);
                }
            }
#pragma unroll
            for (int a = 0; a < TA; ++a)
#pragma unroll
                for (int b = 0; b < TB; ++b) {
                    if (PREC == kPrecBF16x3) {
                        acc[a][b] = mfma_bf16(ql[b], ph[a], acc[a][b]);
                        acc[a][b] = mfma_bf16(qh[b], pl[a], acc[a][b]);
                    }
                    acc[a][b] = mfma_bf16(qh[b], ph[a], acc[a][b]);
                }
        }
    }

    template <int S>
    __device__ __forceinline__ void run_fixed()
    {
        Frag buf[U];
#pragma unroll
        for (int u = 0; u < D; ++u) load(buf[u], u, S);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (s + D < S) load(buf[(s + D) % U], s + D, S);
            mma(buf[s % U], s);
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    template <bool ASSIGN, bool LOWER>
    __device__ __forceinline__ void store(float *C, int64_t ldc, int row0, int col0, int imax, int jmax) const
    {
        float old[TA][TB][4];
        if (!ASSIGN) {
#pragma unroll
            for (int a = 0; a < TA; ++a)
#pragma unroll
                for (int b = 0; b < TB; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        int i = row0 + row_of(a), j = col0 + col_of(b, r);
                        i = i < imax ? i : imax - 1;
                        j = j < jmax ? j : jmax - 1;
                        old[a][b][r] = C[i + (int64_t)j * ldc];
                    }
        }
#pragma unroll
        for (int a = 0; a < TA; ++a)
#pragma unroll
            for (int b = 0; b < TB; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = row0 + row_of(a), j = col0 + col_of(b, r);
                    const bool ok = i < imax && j < jmax && (!LOWER || j <= i);
                    if (ok) C[i + (int64_t)j * ldc] = ASSIGN ? acc[a][b][r] : old[a][b][r] - acc[a][b][r];
                }
    }

    __device__ __forceinline__ static int row_of(int a) { return 16 * a + (lane_id() & 15); }
    __device__ __forceinline__ static int col_of(int b, int r) { return 16 * b + 4 * (lane_id() >> 4) + r; }
};

// Split BF16 operands, written once per panel when it is finished: a panel slot holds each
// part of every row in the MFMA operand order -- per group of 16 rows, per k-step s (32
// columns), per lane quad q, per row, 8 values: element (16 g + rr, s, q, e) at 2048 g + 512 s
// + 128 q + 8 rr + e -- so a wave's 16-byte operand loads of a 16-row tile read 1 KiB
// contiguously.  The columns are permuted (the same way for every row, so a product over k is
// unchanged).  Readers: syrk_tile_split_lds.
// Writes row R's values v[p][r] (column j = 16 p + 4 r + q of a 128-column panel, q = lane / 16:
// the layout of trsm_rows' result) to a split slot, part t at slot + t * part_off: column j
// goes to k-step q, lane quad p / 2, position 4 (p % 2) + r.  PARTS = 2:
// hi = bf16(x), lo = bf16(x - hi) (16 significant bits); PARTS = 3: hi, mid = bf16(x - hi),
// lo = bf16(x - hi - mid) (all 24 bits of x; both differences are exact in FP32).
template <int PARTS = 2>
__device__ __forceinline__ void split_store_row(__bf16 *slot, int64_t part_off, int64_t R, const float (&v)[8][4])
{
    const int q = lane_id() >> 4;
    __bf16 *row = slot + (R >> 4) * 2048 + 512 * q + 8 * (R & 15);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        bf16x8 part[PARTS];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float x = v[2 * u + (e >> 2)][e & 3];
#pragma unroll
            for (int t = 0; t < PARTS; ++t) {
                const __bf16 h = (__bf16)x;
                part[t][e] = h;
                x -= (float)h;
            }
        }
#pragma unroll
        for (int t = 0; t < PARTS; ++t) *(bf16x8 *)(row + 128 * u + t * part_off) = part[t];
    }
}

// Tile engine for storage type T and trailing-update arithmetic PREC.
template <typename T, int PREC>
struct TileSel {
    using type = WaveTile<4, 4>;
};
template <int PREC>
struct TileSel<float, PREC> {
    // the updates inside a panel step of the 3-part split arithmetic use FP32 MFMAs (the same
    // 24-bit accuracy; the trailing update reads the split panels instead, syrk_tile_split_lds)
    using type = WaveTileF<4, 4, PREC == kPrecBF16x6 ? kPrecF32 : PREC>;
};

// A second FP64 wave tile, for irregular k ranges: TA x TB MFMA tiles, acc[a][b] = tile
// (16a.., 16b..); row-contiguous operands (element (row r, k) at base + r + k * ld).  Rows >=
// rmax are clamped (their results are never stored); k outside [klo, kmax) reads as 0 (klo a
// multiple of 4), so any depth works (WaveTile needs a multiple of 4 k-steps per buffer
// round).  U k-steps of loads in flight in a rolling register buffer.
template <int TA, int TB>
struct RowTile {
    dvec4 acc[TA][TB];
    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int a = 0; a < TA; ++a)
#pragma unroll
            for (int b = 0; b < TB; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
    }
    // acc += P Q^T over k in [klo, kmax), P rows p0.., Q rows q0..
    __device__ __forceinline__ void run(const double *P, int64_t ldp, int p0, int pmax, const double *Q, int64_t ldq,
                                        int q0, int qmax, int klo, int kmax)
    {
        const int l = lane_id(), kl = l >> 4;
        const double *pa[TA], *qb[TB];
#pragma unroll
        for (int a = 0; a < TA; ++a) {
            int r = p0 + 16 * a + (l & 15);
            r = r < pmax ? r : pmax - 1;
            pa[a] = P + r;
        }
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            int q = q0 + 16 * b + (l & 15);
            q = q < qmax ? q : qmax - 1;
            qb[b] = Q + q;
        }
        const int s0 = klo >> 2, S = (kmax + 3) >> 2;  // k-steps [s0, S)
        if (S <= s0) return;
        constexpr int U = 4;
        double pbuf[U][TA], qbuf[U][TB];
        auto load = [&](int u, int s) {
            int k = 4 * (s < S ? s : S - 1) + kl;
            k = k < kmax ? k : kmax - 1;
#pragma unroll
            for (int a = 0; a < TA; ++a) pbuf[u][a] = pa[a][(int64_t)k * ldp];
#pragma unroll
            for (int b = 0; b < TB; ++b) qbuf[u][b] = qb[b][(int64_t)k * ldq];
        };
        auto mma = [&](int u, int s) {
            const bool ok = 4 * s + kl < kmax;
#pragma unroll
            for (int a = 0; a < TA; ++a) {
                const double pv = ok ? pbuf[u][a] : 0.0;
#pragma unroll
                for (int b = 0; b < TB; ++b) acc[a][b] = mfma_f64(qbuf[u][b], pv, acc[a][b]);
            }
        };
        const int Sp = s0 + (S - s0 + U - 1) / U * U;
#pragma unroll
        for (int u = 0; u < U - 1; ++u) load(u, s0 + u);
        for (int sb = s0; sb < Sp; sb += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                load((u + U - 1) % U, sb + u + U - 1);
                if (sb + u < S) mma(u, sb + u);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // C[i][j] (column-major, rows i0.., cols j0..) = acc (ASSIGN) or -= acc, i < imax, j < jmax.
    template <bool ASSIGN>
    __device__ __forceinline__ void store(double *C, int64_t ldc, int i0, int j0, int imax, int jmax) const
    {
        const int l = lane_id();
#pragma unroll
        for (int a = 0; a < TA; ++a) {
            const int i = i0 + 16 * a + (l & 15);
            double old[TB][4];
            if (!ASSIGN) {
#pragma unroll
                for (int b = 0; b < TB; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = j0 + 16 * b + 4 * r + (l >> 4);
                        old[b][r] = C[(int64_t)(i < imax ? i : imax - 1) + (int64_t)(j < jmax ? j : jmax - 1) * ldc];
                    }
            }
            if (i < imax) {
#pragma unroll
                for (int b = 0; b < TB; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = j0 + 16 * b + 4 * r + (l >> 4);
                        if (j < jmax) C[i + (int64_t)j * ldc] = ASSIGN ? acc[a][b][r] : old[b][r] - acc[a][b][r];
                    }
            }
        }
    }
};
using RowTile64 = RowTile<4, 4>;

// Workgroup -> tile remap that gives each XCD a contiguous range of the tile sequence
// (workgroups are dispatched round-robin over the 8 XCDs; each XCD has its own L2).
__device__ __forceinline__ int xcd_remap(int w, int ntiles)
{
    constexpr int X = 8;
    if (ntiles < 2 * X) return w;
    const int x = w % X, idx = w / X;
    const int q = ntiles / X, rem = ntiles % X;
    const int start = x * q + (x < rem ? x : rem);
    return start + idx;
}

}  // namespace mi
