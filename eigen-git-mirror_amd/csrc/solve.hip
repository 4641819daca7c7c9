// solve.hip -- forward and backward block substitution of LLT::solve in one persistent
// launch per sweep.
//
// Replaces LLT::_solve_impl (Eigen/src/Cholesky/LLT.h:483-498): dst = rhs; L Y = dst
// (TriangularSolverMatrix.h:36-184, OnTheLeft, Lower) then L^T X = Y (same template on the
// transposed view).  Blocked by the factorization's 128-row blocks:
//   forward : Y_i = X_i (B_i - sum_{j<i} L_ij Y_j)
//   backward: X_i = X_i^T (Y_i - sum_{j>i} L_ji^T X_j),      X_i = L_ii^{-1}
// with the diagonal-block inverses built after the factorization (and the couplings to the
// predecessor block, which take X_i's product off the chain from block to block).
//
// Work items are (row block i, group of 16 right-hand sides, half h of the block's 128 rows).
// A workgroup takes the next item from an atomic ticket counter, in dependency order, so every
// block it waits for belongs to a workgroup that started earlier and is resident: the waits
// always end.
//
// A finished half block is published into a packed copy of the block [c + 16 * row] (the LDS
// layout of its consumers, zero-padded to 128 x 16) written with agent-scope atomic stores,
// and the data is its own flag: before each sweep the copies are filled with a signalling
// NaN, which no arithmetic result can be (results of arithmetic are quiet NaNs), and a
// consumer polls one element, then loads all 2048 and re-polls any still unwritten.  So a
// producer never waits for its stores to drain and no separate flag is written or read.  A
// poll that exceeds a large bound records MI_SOLVE_TIMEOUT in the status word and the item is
// abandoned instead of spinning forever.
//
// Contributions are accumulated as soon as their block is published (left-looking per row
// block): 4 waves, one 16-row MFMA tile each, K = 128 per contributing block, the block
// staged through LDS (it is the shared operand of all 4 waves), so that the 64 lanes of an
// MFMA operand read 64 consecutive doubles (no bank conflicts).  Barriers that only protect
// LDS wait for LDS alone, so the prefetch of the next operand stays in flight across them.
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

constexpr int SB = 128;          // row block (= factorization block)
constexpr int SH = 64;           // rows of one work item (half a block)
constexpr int SC = 16;           // right-hand sides per work item
constexpr int SP = SB * SC;      // doubles per packed block
constexpr int SPH = SH * SC;     // doubles per packed half block
constexpr int ST = 256;          // threads per work item (two resident per CU)
constexpr uint32_t kSpinLimit = 1u << 22;  // ~0.2 s of polling
// Unwritten marker of the packed copies: a signalling NaN (quiet bit 51 clear).
constexpr unsigned long long kUnwritten = 0x7ff40000000005eeull;
static_assert(solve_pack_doubles(1, 1) == SP + SPH, "per block and group: the block's copy, then a half block");

__device__ __forceinline__ bool unwritten(double v) { return (unsigned long long)__double_as_longlong(v) == kUnwritten; }

// Until every element of v (loaded from p + e * stride) reads as written, reloads all those
// that do not, together: one round trip per round however many are stale (false: timed out,
// recorded in the status word).
template <int M>
__device__ __forceinline__ bool await_written(double (&v)[M], const double *p, int stride, int *status)
{
    for (uint32_t spins = 0;;) {
        bool stale = false;
#pragma unroll
        for (int e = 0; e < M; ++e) stale |= unwritten(v[e]);
        if (!stale) return true;
        if (++spins > kSpinLimit) {
            atomicOr(status, MI_SOLVE_TIMEOUT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int e = 0; e < M; ++e)
            if (unwritten(v[e])) v[e] = coherent_load(p + e * stride);
    }
}

// Waits until the packed copy src[0, N) is written, then hands each element to put(index,
// value) (false: given up -- the status word is set, or the poll timed out).  Every thread
// loads its N / ST at once; the one holding the last element, if that is not yet written,
// polls it alone while the others wait, and then every thread re-polls its own elements not
// yet visible, all of them in each round.  A copy that is already written (the usual case on the chain: the consumer
// finishes its previous product after the producer publishes) costs one round trip.
template <int N, class Put>
__device__ __forceinline__ bool fetch_published(const double *src, int *status, int *s_ok, int *s_fail, Put put)
{
    const int tid = threadIdx.x;
    double v[N / ST];
#pragma unroll
    for (int e = 0; e < N / ST; ++e) v[e] = coherent_load(src + e * ST + tid);
    if (tid == ST - 1) {  // holds element N - 1 in v[N / ST - 1]
        uint32_t spins = 0;
        int ok = 1;
        while (unwritten(v[N / ST - 1])) {
            __builtin_amdgcn_s_sleep(1);
            if ((++spins & 255) == 0 && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (spins > kSpinLimit) {
                atomicOr(status, MI_SOLVE_TIMEOUT);
                ok = 0;
                break;
            }
            v[N / ST - 1] = coherent_load(src + N - 1);
        }
        *s_ok = ok;
    }
    lds_barrier();
    if (!*s_ok) return false;
    if (!await_written(v, src + tid, ST, status)) *s_fail = 1;
#pragma unroll
    for (int e = 0; e < N / ST; ++e) put(e * ST + tid, v[e]);
    lds_barrier();
    return *s_fail == 0;
}

// BACKWARD = false: forward sweep with L; true: backward sweep with L^T.
// Winv holds, per block b (at Winv + b * kInvDoubles, each ld 128): X_b^T, X_b, and the
// couplings M_b, M2_b, M3_b = X_b L_{b,b-d} and N_b, N2_b, N3_b = X_b^T L_{b+d,b}^T (d = 1, 2,
// 3).  Ypk: per group of right-hand sides, nblk packed blocks, then per group nblk packed half
// blocks (the first half's share Q of the other half's rows, below), all filled with kUnwritten
// before the launch.
// T: storage of L (float for a mixed-precision factor, widened on load; FP64 arithmetic).
//
// Block i depends on its three predecessors (i-1, i-2, i-3 forward; i+1, i+2, i+3 backward)
// through one product each:
//   forward : Y_i = Z_i - M3_i Y_{i-3} - M2_i Y_{i-2} - M_i Y_{i-1},
//             Z_i = X_i (B_i - sum_{j<i-3} L_ij Y_j)
//   backward: X_i = Z_i - N3_i X_{i+3} - N2_i X_{i+2} - N_i X_{i+1},
//             Z_i = X_i^T (Y_i - sum_{j>i+3} L_ji^T X_j)
// (with fewer than 16 blocks two, no M3_i, and Z_i sums to j < i-2; with fewer than 8 none, and
// Y_i = Z_i sums over all j < i), so a workgroup finishes
// Z_i (its triangular X_i product included) three blocks ahead, the i-3 and i-2 products two
// and one block ahead, and the step from one block to the next is:
// see i-1, stage it, 32 MFMA steps (on four independent chains), store.  Two workgroups share
// a block, 64 rows each, so that step's MFMAs (one CU's FP64 rate) are half those of a whole
// block.
//
// X_i is lower triangular (X_i^T upper), so the rows of one half of Z_i need only that half's
// residual rows T: the first half forward, the second backward.  That half (taken first from
// the tickets) also multiplies the other half's rows of op(X_i) over its own k, and publishes
// that 64 x 16 partial product Q; the other half multiplies only its own triangle and adds Q
// just before the step on the chain, by which time Q has long been written.
//
// UC (backward only): L's strictly upper triangle holds L^T, so the operand L_ji^T of block i
// is read as the upper block (i, j) -- lanes over consecutive rows, like the forward sweep --
// instead of across 16 columns of L per load.
template <bool BACKWARD, typename T, bool UC = false>
__global__ __launch_bounds__(ST, 2) void solve_sweep_kernel(const T *__restrict__ L, int64_t lda, int64_t n,
                                                         double *__restrict__ B, int64_t ldb, int64_t nrhs,
                                                         const double *__restrict__ Winv, double *Ypk, int *ticket,
                                                         const int *info, int *status, int64_t sL, int64_t sB,
                                                         int64_t sWinv, int64_t sY, int64_t sT)
{
    // batched launches: system z = blockIdx.z, its pointers advanced by z times their strides
    const int64_t z = blockIdx.z;
    L += z * sL;
    B += z * sB;
    Winv += z * sWinv;
    Ypk += z * sY;
    ticket += z * sT;
    if (info) info += z;
    // uniform: info and the status word are final before this kernel starts
    if (factor_failed(info) || device_failed(status)) return;
    // Ys[c + SC*k + 16 (k / 32)]: the published block being consumed, one padding row per
    // 32 k, so that lanes reading k = 32 q + s (q = lane / 16, the permuted order below) hit
    // distinct banks
    __shared__ double Ys[SP + 4 * 16];
    __shared__ double Ts[SP];         // Ts[c + SC*r]: the block's residual before the inverse
    __shared__ int s_item, s_ok, s_fail;
    const int nblk = (int)((n + SB - 1) / SB);
    const int ngrp = (int)((nrhs + SC - 1) / SC);
    const int tid = threadIdx.x, l = lane_id();
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);

    if (tid == 0) {
        s_item = atomicAdd(ticket, 1);
        s_fail = 0;
    }
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_item);
    if (t >= 2 * nblk * ngrp) return;
    const int tb = t / (2 * ngrp);             // position of the block in the sweep
    const int ib = BACKWARD ? nblk - 1 - tb : tb;
    const int g = (t % (2 * ngrp)) >> 1;
    const int hind = BACKWARD ? 1 : 0;         // the half whose Z rows need only its own T rows
    const bool indep = (t & 1) == 0;           // taken first: the other half waits for its Q
    const int h = indep ? hind : 1 - hind;
    const int64_t i0 = (int64_t)ib * SB, c0 = (int64_t)g * SC;
    const int bi = (int)((n - i0) < SB ? (n - i0) : SB);
    const int nc = (int)((nrhs - c0) < SC ? (nrhs - c0) : SC);
    double *gpk = Ypk + (size_t)g * nblk * SP;  // block jb of this group at gpk + jb * SP
    // Q, the first half's share of the other half's Z rows, packed [c + 16 * (row - 64 (1 - hind))]
    double *qpk = Ypk + (size_t)ngrp * nblk * SP + ((size_t)g * nblk + ib) * SPH;

    // this wave's 16 rows of the block
    const int rt = 4 * h + w;
    const int r = 16 * rt + (l & 15);
    const int rc = r < bi ? r : bi - 1;  // clamped row (in bounds)
    const int kl = l >> 4;
    const double *Wi = Winv + (size_t)ib * kInvDoubles;

    // B_i (this wave's rows; lane l, element e: row r, column 4e + l/16): read before any wait
    double bv[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = 4 * e + kl;
        bv[e] = B[(i0 + rc) + (c0 + (c < nc ? c : 0)) * ldb];
    }
    // the predecessors and the couplings to them (P[row][k] = M_i[row][k], M2_i[row][k], ...)
    const int jstep = BACKWARD ? -1 : 1;
    const int npred = BACKWARD ? nblk - 1 - ib : ib;  // blocks before this one in the sweep
    const int ncp = solve_couplings(nblk);  // coupled predecessors: 3, 2, or 0 for few blocks
    const bool crit1 = npred >= 1 && ncp >= 1, crit2 = npred >= 2 && ncp >= 2, crit3 = npred >= 3 && ncp >= 3;
    const int jend = ib - jstep * (npred < ncp ? npred : ncp);  // plain contributions: [start, jend)
    const double *M1p = Wi + (BACKWARD ? kInvN : kInvM) + r + SB * kl;
    const double *M2p = Wi + (BACKWARD ? kInvN2 : kInvM2) + r + SB * kl;
    const double *M3p = Wi + (BACKWARD ? kInvN3 : kInvM3) + r + SB * kl;

    // Stages packed block jb into Ys once it is written (false: given up).
    auto stage = [&](int jb) -> bool {
        return fetch_published<SP>(gpk + (size_t)jb * SP, status, &s_ok, &s_fail,
                                   [&](int e, double v) { Ys[e + 16 * (e >> 9)] = v; });
    };

    // P operand of a contributing block, k-step s: P[row][k] = L_ij[row][k] (forward) or
    // L_ji[k][row] (backward); k clamped in bounds.  `aim` points it at block jb.
    // PERM (FP64 backward without the upper copy): there the operand is a column of L, and a
    // lane takes k = 32 q + s at step s (q = lane / 16) instead of 4 s + q -- the 32 k of a lane
    // are contiguous in memory and load as 16-byte pairs (when lda is even and the block is
    // full), not as 32 loads that each touch 16 columns.  The sum over k is the same.
    constexpr bool PERM = BACKWARD && !UC && sizeof(T) == 8;
    const T *lb = nullptr;  // this lane's row (forward) or column (backward) of the block
    int bjm = SB - 1;       // last valid k of the block
    auto aim = [&](int jb) {
        const int64_t j0 = (int64_t)jb * SB;
        bjm = (int)((n - j0) < SB ? (n - j0) : SB) - 1;
        lb = (BACKWARD && !UC) ? L + j0 + (i0 + rc) * lda : L + (i0 + rc) + j0 * lda;
    };
    auto kof = [&](int s) -> int { return PERM ? 32 * kl + s : 4 * s + kl; };
    auto load_operand = [&](double (&dst)[SB / 4]) {
        if (PERM && (lda & 1) == 0 && bjm == SB - 1) {  // 16-byte aligned: lda even, j0 % 128 == 0
#pragma unroll
            for (int u = 0; u < SB / 8; ++u) {
                typedef double d2 __attribute__((ext_vector_type(2)));
                const d2 d = *(const d2 *)((const double *)lb + 32 * kl + 2 * u);
                dst[2 * u] = d[0];
                dst[2 * u + 1] = d[1];
            }
        } else {
#pragma unroll
            for (int s = 0; s < SB / 4; ++s) {
                const int k = kof(s) < bjm ? kof(s) : bjm;
                dst[s] = (double)lb[(BACKWARD && !UC) ? (int64_t)k : (int64_t)k * lda];
            }
        }
    };
    // Ys element (lane's column, k)
    auto ys = [&](int k) -> double { return Ys[(l & 15) + SC * k + 16 * (k >> 5)]; };
    // op(X_i): column-major P[row][k] = X_i[row][k] (forward, lower: k < 16(rt+1)) or X_i[k][row]
    // (backward, upper: k >= 16 rt); only the nonzero k-steps of this half's own k (16 h <= s <
    // 16 h + 16) are loaded and multiplied.  The first half's Q takes the other half's rows
    // (r ^ 64) over the same k, all nonzero, in the operand registers s ^ 16 that its own
    // product leaves free.
    const double *Xp = Wi + (BACKWARD ? 0 : kInvX) + r + SB * kl;
    // register s < 16 loads from Xlo + 4 SB s, s >= 16 from Xhi + 4 SB s: Xp, except for the
    // first half's Q registers, whose k-step s ^ 16 is 16 steps away (rows r ^ 64)
    const double *Xq = Wi + (BACKWARD ? 0 : kInvX) + (r ^ SH) + SB * kl;
    const double *Xlo = (indep && h == 1) ? Xq + 4 * SB * 16 : Xp;
    const double *Xhi = (indep && h == 0) ? Xq - 4 * SB * 16 : Xp;
    const int s_lo = (BACKWARD ? 4 * rt : 0) > 16 * h ? (BACKWARD ? 4 * rt : 0) : 16 * h;
    const int s_hi = (BACKWARD ? SB / 4 : 4 * (rt + 1)) < 16 * h + 16 ? (BACKWARD ? SB / 4 : 4 * (rt + 1)) : 16 * h + 16;

    double pv[SB / 4];
    // operand register s of the X_i products, loaded whether or not it is used (no load sits in a
    // branch, so the compiler's counts of loads in flight stay exact and no MFMA waits for a
    // later load)
    auto x_slot = [&](int s) { pv[s] = (s < 16 ? Xlo : Xhi)[(size_t)4 * SB * s]; };
    dvec4 acc = {0.0, 0.0, 0.0, 0.0};
    int jb = BACKWARD ? nblk - 1 : 0;
    const bool contrib = jb != jend;  // any plain contributions (before the three coupled blocks)
    if (contrib) {
        aim(jb);
        load_operand(pv);
    }
    for (; jb != jend; jb += jstep) {
        if (!stage(jb)) return;
        // acc[row][c] += sum_k P[row][k] * Ys[k][c]; rows k >= bj of Ys are zero, so the
        // clamped (finite) P values past a partial block contribute nothing.  In the last
        // contribution each k-step's operand register is refilled with X_i's right after its
        // MFMA issues, so X_i is in flight during these MFMAs.
        if (jb + jstep != jend) {
#pragma unroll
            for (int s = 0; s < SB / 4; ++s) acc = mfma_f64(ys(kof(s)), pv[s], acc);
            aim(jb + jstep);
            load_operand(pv);  // in flight during the next wait
        } else {
#pragma unroll
            for (int s = 0; s < SB / 4; ++s) {
                acc = mfma_f64(ys(kof(s)), pv[s], acc);
                x_slot(s);
            }
        }
        lds_barrier();  // Ys is rewritten by the next contribution
    }

    if (!contrib) {
#pragma unroll
        for (int s = 0; s < SB / 4; ++s) x_slot(s);
    }
    // T = B_i - acc  (lane l, element e: row r, column 4e + l/16), this half's rows
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = 4 * e + kl;
        const bool in = r < bi && c < nc;
        Ts[c + SC * r] = in ? bv[e] - acc[e] : 0.0;
    }
    lds_barrier();
    // Z_i = op(X_i) T over this half's k (and the first half's Q), in groups of four k-steps
    // (s_lo, s_hi and the first half's range are multiples of four), the first half's range
    // first: then every group is the last use of its own four operand registers in both halves
    // (the first half's Q registers of a group in the second range were used by the group 16
    // steps away, before it), and they take the first coupling's operands at once, which are
    // in flight while that predecessor is awaited and staged.  The loads are unconditional (read
    // only with crit1): a load in a branch would make the compiler's count of loads in flight
    // inexact, and every later MFMA would wait for all of them.
    constexpr int HQ = BACKWARD ? 1 : 0;  // the first half (h of an item with indep)
    const double *Cp = crit3 ? M3p : crit2 ? M2p : M1p;
    dvec4 y = {0.0, 0.0, 0.0, 0.0}, yq = {0.0, 0.0, 0.0, 0.0};
    __builtin_amdgcn_s_setprio(1);  // the rest of this item leads to the chain
#pragma unroll
    for (int gi = 0; gi < 8; ++gi) {
        const int sb = 4 * ((gi + 4 * HQ) & 7);
        const bool own = sb >= s_lo && sb < s_hi;
        const bool q = indep && (sb >> 4) == HQ;  // Q's k-steps, operands in registers s ^ 16
        if (own || q) {
            double tv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) tv[u] = Ts[(l & 15) + SC * (4 * (sb + u) + kl)];
            if (own) {
#pragma unroll
                for (int u = 0; u < 4; ++u) y = mfma_f64(tv[u], pv[sb + u], y);
            }
            if (q) {
#pragma unroll
                for (int u = 0; u < 4; ++u) yq = mfma_f64(tv[u], pv[(sb + u) ^ 16], yq);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) pv[sb + u] = Cp[(size_t)4 * SB * (sb + u)];
    }
    if (indep) {  // publish Q: rows r ^ 64 of Z_i, this half's share
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = 4 * e + kl;
            const bool in = (r ^ SH) < bi && c < nc;
            double v = in ? yq[e] : 0.0;
            if (unwritten(v)) v = __builtin_nan("");  // never: arithmetic yields quiet NaNs
            __hip_atomic_store(qpk + c + SC * ((r ^ SH) - SH * (1 - h)), v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    // the second half's Q (lane l, element e: row r, column 4e + l/16), loaded now so that it is
    // in flight during the wait for the predecessor (the first half loads and ignores its own)
    const double *qsrc = qpk + kl + SC * (r - SH * h);
    double qv[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) qv[e] = coherent_load(qsrc + 4 * e);

    // the coupling products, y -= (M3_i or N3_i) Y_{i-3}, y -= (M2_i or N2_i) Y_{i-2}, then
    // y -= (M_i or N_i) Y_{i-1}, each on four independent MFMA chains; each one's operand loads
    // follow the previous one's MFMAs.  The i-3 and i-2 products come before the chain needs
    // them (their waves keep the raised issue priority over the other item resident on this CU,
    // whose contributions are further ahead).
    dvec4 cq[4];
    if (crit3) {
        if (!stage(ib - 3 * jstep)) return;
#pragma unroll
        for (int q = 0; q < 4; ++q) cq[q] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < SB / 4; ++s) {
            cq[s & 3] = mfma_f64(ys(4 * s + kl), pv[s], cq[s & 3]);
            pv[s] = M2p[(size_t)4 * SB * s];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) y[e] -= (cq[0][e] + cq[1][e]) + (cq[2][e] + cq[3][e]);
    }
    if (crit2) {
        if (!stage(ib - 2 * jstep)) return;
#pragma unroll
        for (int q = 0; q < 4; ++q) cq[q] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < SB / 4; ++s) {
            cq[s & 3] = mfma_f64(ys(4 * s + kl), pv[s], cq[s & 3]);
            pv[s] = M1p[(size_t)4 * SB * s];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) y[e] -= (cq[0][e] + cq[1][e]) + (cq[2][e] + cq[3][e]);
    }
    if (!indep) {  // Q, the first half's share of these rows (a timeout fails the next stage)
        if (!await_written(qv, qsrc, 4, status)) {  // timed out (recorded)
            if (!crit1) return;                     // no stage follows to stop the item
            s_fail = 1;                             // the next stage gives up
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) y[e] += qv[e];
    }
    // From here the item is on the chain from block to block: its waves take the highest issue
    // priority.
    __builtin_amdgcn_s_setprio(2);
    if (crit1) {  // the step that waits
        if (!stage(ib - jstep)) return;
#pragma unroll
        for (int q = 0; q < 4; ++q) cq[q] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < SB / 4; ++s) cq[s & 3] = mfma_f64(ys(4 * s + kl), pv[s], cq[s & 3]);
#pragma unroll
        for (int e = 0; e < 4; ++e) y[e] -= (cq[0][e] + cq[1][e]) + (cq[2][e] + cq[3][e]);
    }
    // publish: this half of the packed copy (atomic stores, so they reach every XCD without a
    // flush), then the result itself
    double *dst = gpk + (size_t)ib * SP;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = 4 * e + kl;
        const bool in = r < bi && c < nc;
        double v = in ? y[e] : 0.0;
        if (unwritten(v)) v = __builtin_nan("");  // never: arithmetic yields quiet NaNs
        __hip_atomic_store(dst + c + SC * r, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = 4 * e + kl;
        if (r < bi && c < nc) B[(i0 + r) + (c0 + c) * ldb] = y[e];
    }
}

// Before a sweep: the ticket counter to 0 and the sweep's packed copies to kUnwritten.
__global__ __launch_bounds__(256) void solve_reset_kernel(int *ticket, double *Ypk, int64_t count, int64_t sT,
                                                          int64_t sY)
{
    ticket += (int64_t)blockIdx.z * sT;
    Ypk += (int64_t)blockIdx.z * sY;
    const int64_t i0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i0 == 0) *ticket = 0;
    const double u = __longlong_as_double((long long)kUnwritten);
    for (int64_t i = i0; i < count; i += (int64_t)gridDim.x * 256) Ypk[i] = u;
}

// tickets: 2 ints; Ypk: solve_pack_doubles(n, nrhs) doubles, refilled before each sweep.
template <typename T>
void launch_solve_persistent(const T *L, int64_t lda, int64_t n, double *B, int64_t ldb, int64_t nrhs,
                             const double *Winv, double *Ypk, int *tickets, const int *info, int *status, hipStream_t s,
                             int sweeps, bool upper_copy, const Batch &bt)
{
    const unsigned nb = bt.count;
    const int nblk = (int)((n + SB - 1) / SB);
    const int ngrp = (int)((nrhs + SC - 1) / SC);
    const unsigned items = (unsigned)(2 * nblk * ngrp);  // two halves per block and group
    const int64_t pk = solve_pack_doubles(n, nrhs);
    const unsigned fill = (unsigned)((pk + 255) / 256 < 1024 ? (pk + 255) / 256 : 1024);
    if (sweeps & 1) {
        hipLaunchKernelGGL(solve_reset_kernel, dim3(fill, 1, nb), dim3(256), 0, s, tickets, Ypk, pk, bt.sync, bt.pack);
        hipLaunchKernelGGL((solve_sweep_kernel<false, T>), dim3(items, 1, nb), dim3(ST), 0, s, L, lda, n, B, ldb,
                           nrhs, Winv, Ypk, tickets, info, status, bt.a, bt.b, bt.inv, bt.pack, bt.sync);
    }
    if (sweeps & 2) {
        hipLaunchKernelGGL(solve_reset_kernel, dim3(fill, 1, nb), dim3(256), 0, s, tickets + 1, Ypk, pk, bt.sync,
                           bt.pack);
        if (upper_copy)
            hipLaunchKernelGGL((solve_sweep_kernel<true, T, true>), dim3(items, 1, nb), dim3(ST), 0, s, L, lda, n, B,
                               ldb, nrhs, Winv, Ypk, tickets + 1, info, status, bt.a, bt.b, bt.inv, bt.pack,
                               bt.sync);
        else
            hipLaunchKernelGGL((solve_sweep_kernel<true, T>), dim3(items, 1, nb), dim3(ST), 0, s, L, lda, n, B, ldb,
                               nrhs, Winv, Ypk, tickets + 1, info, status, bt.a, bt.b, bt.inv, bt.pack, bt.sync);
    }
}

template void launch_solve_persistent<double>(const double *, int64_t, int64_t, double *, int64_t, int64_t,
                                              const double *, double *, int *, const int *, int *, hipStream_t, int,
                                              bool, const Batch &);
template void launch_solve_persistent<float>(const float *, int64_t, int64_t, double *, int64_t, int64_t,
                                             const double *, double *, int *, const int *, int *, hipStream_t, int,
                                             bool, const Batch &);

}  // namespace mi
