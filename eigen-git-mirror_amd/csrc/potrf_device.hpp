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

// Panel p (columns c0..c0+15 of the 128 x 128 block in S, tiled: sl), all four waves at once.
// Every wave holds the 16 x 16 leaf in lanes 0..15 (row c0+l) and runs the column loop of
// LLT.h:320-334 on it (pivot test, sqrt, scale, rank-1 update) redundantly, so its lane
// broadcasts (readlane) need no barrier.  Lanes 16..63 of waves 0..2 hold the 48 rows
// c0+16+48w+(l-16) below the leaf and receive the same column operations, which is the panel
// solve A_ip L_pp^{-T}; lanes 16..31 of wave 3 hold the rows of the 16 x 16 identity, which the
// same operations turn into L_pp^{-T}, i.e. column i of X_pp = L_pp^{-1} in lane 16+i, written
// to xleaf column-major.  1/sqrt from the hardware estimate and one Newton step, d = x/sqrt(x).
// Returns the failing panel column or -1, the same in every wave.  The factored leaf is left
// in v (lanes 0..15): the caller writes it back to S only after a barrier, since the other
// waves read the unfactored leaf from S here and may still be doing so when this wave is done.
__device__ __forceinline__ int panel_factor4(double *S, int c0, double *__restrict__ xleaf, double (&v)[16])
{
    const int l = lane_id(), w = threadIdx.x >> 6;
    const bool leaf = l < 16;
    const int rb = c0 + 16 + 48 * w + (l - 16);
    const bool below = !leaf && w < 3 && rb < 128;
    const bool ident = !leaf && w == 3 && l < 32;
    const int r = leaf ? c0 + l : (below ? rb : c0);
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const double a = S[sl(r, c0 + c)];
        v[c] = leaf ? (c <= l ? a : 0.0) : (below ? a : ((ident && c == l - 16) ? 1.0 : 0.0));
    }
    int fail = -1;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const double x = readlane_d(v[j], j);  // pivot: leaf row j lives in lane j
        if (fail < 0 && x <= 0.0) fail = j;    // LLT.h:327-330; a NaN pivot passes
        double rd = __builtin_amdgcn_rsq(x);
        rd = rd * (1.5 - 0.5 * x * rd * rd);
        const double d = x * rd;
        v[j] = l == j ? d : (l > j ? v[j] * rd : v[j]);
#pragma unroll
        for (int c = j + 1; c < 16; ++c) {
            const double lcj = readlane_d(v[j], c);  // L(c0+c, c0+j)
            v[c] -= v[j] * lcj;
        }
    }
    if (fail >= 0) return fail;
    if (below) {  // rows no other wave reads before the next barrier
#pragma unroll
        for (int c = 0; c < 16; ++c) S[sl(r, c0 + c)] = v[c];
    }
    if (ident) {
#pragma unroll
        for (int c = 0; c < 16; ++c) xleaf[c + 16 * (l - 16)] = v[c];
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

// Cholesky of the 128 x 128 lower block in S (all 256 threads), in eight 16-column panels:
// panel factor and solve by all four waves (panel_factor4), then the rank-16 update of the
// rest of the block by the panel on FP64 MFMA from LDS, the 28..1 lower 16 x 16 tiles dealt
// round-robin to the waves, all MFMAs before the read-modify-writes.  Returns the first
// failing column or -1 (uniform); writes the 8 leaf inverses to Xl.
__device__ __forceinline__ int potrf_in_lds(double *S, double *__restrict__ Xl)
{
    const int w = threadIdx.x >> 6, l = lane_id();
    __syncthreads();
    for (int p = 0; p < 8; ++p) {
        const int c0 = 16 * p;
        double lv[16];
        const int f = panel_factor4(S, c0, Xl + p * 256, lv);
        if (f >= 0) return c0 + f;
        __syncthreads();  // every wave has read the leaf; the rows below it are in S
        if (w == 0 && l < 16) {  // the leaf itself: read by nobody until the final copy out
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c <= l) S[sl(c0 + l, c0 + c)] = lv[c];
        }
        const int T = 7 - p;
        const int nt = T * (T + 1) / 2;
        dvec4 acc[7];
        int tr[7], tc[7];
#pragma unroll
        for (int h = 0; h < 7; ++h) {
            const int t = w + 4 * h;
            int ti = 0, tj = 0;
            if (t < nt) lower_tile(t, ti, tj);
            tr[h] = c0 + 16 + 16 * ti;
            tc[h] = c0 + 16 + 16 * tj;
            acc[h] = dvec4{0.0, 0.0, 0.0, 0.0};
            if (t < nt) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int kc = c0 + 4 * s + (l >> 4);
                    acc[h] = mfma_f64(S[sl(tc[h] + (l & 15), kc)], S[sl(tr[h] + (l & 15), kc)], acc[h]);
                }
            }
        }
#pragma unroll
        for (int h = 0; h < 7; ++h) {
            if (w + 4 * h < nt) {
                double old[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) old[r] = S[sl(tr[h] + (l & 15), tc[h] + 4 * r + (l >> 4))];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = tr[h] + (l & 15), j = tc[h] + 4 * r + (l >> 4);
                    if (i >= j) S[sl(i, j)] = old[r] - acc[h][r];
                }
            }
        }
        __syncthreads();
    }
    return -1;
}

// Rows 0..127 of P (valid m) <- P * L^{-T} with L (128 x 128 lower, ld ldl, valid order bs,
// identity beyond) and its 8 leaf inverses Xl, using smem for L and the leaves.  Each wave
// owns 2 row tiles of 16 rows.  IDENTITY: P is the identity (not read); the result L^{-T}
// goes to O (ld 128) and its transpose L^{-1} to O + 128*128.  T = float: P and L are read
// as FP32 and widened, the arithmetic is FP64, the result is rounded back to FP32.
template <bool IDENTITY, typename T>
__device__ __forceinline__ void trsm_rows(double *smem, T *__restrict__ P, int64_t ldp, int m,
                                          const T *__restrict__ L, int64_t ldl, int bs,
                                          const double *__restrict__ Xl, double *__restrict__ O)
{
    const int l = lane_id(), w = threadIdx.x >> 6;
    const int row0 = 32 * w;
    double *Ls = smem;
    double *Xs = smem + 28 * 256;
    // L11 and its leaf inverses were published by another workgroup: coherent loads
    for (int e = threadIdx.x; e < 128 * 128; e += 256) {  // strictly lower tiles of L
        const int r = e & 127, c = e >> 7;
        if ((r >> 4) > (c >> 4)) {
            const bool in = r < bs && c < bs;
            const double v = (double)coherent_load(L + (in ? r : 0) + (int64_t)(in ? c : 0) * ldl);
            Ls[sls(r, c)] = in ? v : 0.0;
        }
    }
    for (int e = threadIdx.x; e < 8 * 256; e += 256) Xs[e] = coherent_load(Xl + e);
    __syncthreads();

    dvec4 sol[2][8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        dvec4 acc[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int i = row0 + 16 * t + (l & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = 16 * p + 4 * r + (l >> 4);
                if (IDENTITY) {
                    acc[t][r] = (i == j) ? 1.0 : 0.0;
                } else {
                    const double pv = (double)P[(i < m ? i : m - 1) + (int64_t)j * ldp];
                    acc[t][r] = (i < m) ? pv : 0.0;
                }
            }
        }
        // acc -= sum_{q<p} sol_q * L[16p.., 16q..]^T
#pragma unroll
        for (int q = 0; q < p; ++q) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int li = 16 * p + (l & 15), lk = 16 * q + 4 * s + (l >> 4);
                const double lv = Ls[sls(li, lk)];
#pragma unroll
                for (int t = 0; t < 2; ++t) acc[t] = mfma_f64(-lv, sol[t][q][s], acc[t]);
            }
        }
        // sol_p = acc * X_pp^T
        dvec4 z[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const double xv = Xs[p * 256 + (l & 15) + 16 * (4 * s + (l >> 4))];
#pragma unroll
            for (int t = 0; t < 2; ++t) z[t] = mfma_f64(xv, acc[t][s], z[t]);
        }
        sol[0][p] = z[0];
        sol[1][p] = z[1];
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int i = row0 + 16 * t + (l & 15);
        if (IDENTITY || i < m) {
#pragma unroll
            for (int p = 0; p < 8; ++p)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = 16 * p + 4 * r + (l >> 4);
                    if (IDENTITY) {
                        O[i + (int64_t)j * 128] = sol[t][p][r];
                        O[128 * 128 + j + 128 * i] = sol[t][p][r];
                    } else {
                        P[i + (int64_t)j * ldp] = (T)sol[t][p][r];
                    }
                }
        }
    }
}

// The panel role of one block step; g is the role from the ticket:
//   role 0   : C = A_kk - P_0 P_0^T straight into LDS, Cholesky in LDS, L11 and the leaf
//              inverses to memory, then publish `flag` = epoch (release);
//   role g>0 : rows k0+128g..+127: C = A_gk - P_g P_0^T written back, wait for the flag,
//              then rows <- rows * L11^{-T}.
// The diagonal role publishes even on a failing pivot, so that no waiter is left behind;
// waiters then see `info` and stop.  (LLT.h:350-366: the update of block column k by all
// earlier panels precedes its unblocked factor and the right-side solve.)
// smem: kSmem doubles of LDS; s_ok: an LDS int.  T/PREC: storage and update arithmetic
// (the diagonal factor and the panel solve are FP64 for every T).
template <typename T, int PREC>
__device__ __forceinline__ void panel_role(int g, double *smem, int *s_ok, T *__restrict__ A,
                                           int64_t lda, int64_t n, int64_t k0, double *__restrict__ Wleaf, int *info,
                                           uint32_t *flag, uint32_t epoch, int *status)
{
    const int tid = threadIdx.x, w = tid >> 6;
    const int64_t R0 = k0 + 128 * (int64_t)g;
    const int rows = (int)((n - R0) < 128 ? (n - R0) : 128);
    const int bs = (int)((n - k0) < 128 ? (n - k0) : 128);
    T *Ck = A + R0 + k0 * lda;  // this role's 128 rows of block column k0
    double *Xl = Wleaf + (size_t)(k0 / 128) * kLeafDoubles;

    if (k0 > 0) {
        // C -= P_g P_0^T, P = previous panel (columns k0-128..k0-1): rows R0.. and k0..
        const int wr = w >> 1, wc = w & 1;
        const bool active = !(g == 0 && wr == 0 && wc == 1);  // strictly upper quadrant of the diagonal tile
        using Tile = typename TileSel<T, PREC>::type;
        Tile t;
        t.zero();
        if (active) {
            t.setup(A + R0 + (k0 - 128) * lda, lda, wr * 64, rows, A + k0 + (k0 - 128) * lda, lda, wc * 64, bs, 128);
            t.template run_fixed<128 / Tile::KS>();
        }
        if (g == 0) {
            double *S = smem;
            if (active) {  // S = A_kk - acc on the lower part, identity beyond bs; two halves
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
                                    S[sl(i, j)] = in ? (double)old[a][b][r] - (double)t.acc[a0 + a][b][r]
                                                     : (i == j ? 1.0 : 0.0);
                                }
                            }
                }
            }
        } else {
            t.template store<false, false>(Ck, lda, wr * 64, wc * 64, rows, 128);
        }
    } else if (g == 0) {
        block_to_lds(smem, Ck, lda, bs);
    }

    if (g == 0) {
        const int f = potrf_in_lds(smem, Xl);  // begins with a barrier
        if (f >= 0) {
            if (tid == 0) atomicCAS(info, 0, (int)(k0 + f + 1));
        } else {
            lds_to_block(smem, Ck, lda, bs);
        }
        global_barrier();  // every wave's stores of L11 and the leaves have reached L2
        if (tid == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    global_barrier();  // this role's updated rows are in L2 before other waves read them back
    if (tid == 0) {
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
        if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_ok = ok;
    }
    __syncthreads();
    if (!*s_ok || factor_failed(info)) return;
    trsm_rows<false, T>(smem, Ck, lda, rows, A + k0 + k0 * lda, lda, 128, Xl, nullptr);
}

}  // namespace mi
