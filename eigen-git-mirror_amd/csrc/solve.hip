// solve.hip -- forward and backward block substitution of LLT::solve in one persistent
// launch per sweep.
//
// Replaces LLT::_solve_impl (Eigen/src/Cholesky/LLT.h:483-498): dst = rhs; L Y = dst
// (TriangularSolverMatrix.h:36-184, OnTheLeft, Lower) then L^T X = Y (same template on the
// transposed view).  Blocked by the factorization's 128-row blocks:
//   forward : Y_i = X_i (B_i - sum_{j<i} L_ij Y_j)
//   backward: X_i = X_i^T (Y_i - sum_{j>i} L_ji^T X_j),      X_i = L_ii^{-1}
// with the diagonal-block inverses built after the factorization.
//
// Work items are (row block i, group of 16 right-hand sides).  A workgroup takes the next
// item from an atomic ticket counter, in dependency order, so every block it waits for
// belongs to a workgroup that started earlier and is resident: the waits always end.  Each
// finished block is published with a release store of the call's epoch into its flag;
// consumers poll with relaxed loads and issue one acquire fence once the flag is set.  Flags
// carry an increasing epoch, so they never need clearing; only the ticket counter is reset
// (by a one-thread kernel) before each sweep.
// A poll that exceeds a large bound records MI_SOLVE_TIMEOUT in the status word and the
// item is abandoned instead of spinning forever.
//
// Contributions are accumulated as soon as their block is published (left-looking per row
// block): 8 waves, one 16-row MFMA tile each, K = 128 per contributing block, Y_j staged
// through LDS (it is the shared operand of all 8 waves) as [rhs + 16 * row], so that the
// 64 lanes of an MFMA operand read 64 consecutive doubles (no bank conflicts).  Barriers
// that only protect LDS wait for LDS alone, so the prefetch of the next operand stays in
// flight across them.
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

constexpr int SB = 128;          // row block (= factorization block)
constexpr int SC = 16;           // right-hand sides per work item
constexpr uint32_t kSpinLimit = 1u << 22;  // ~0.2 s of polling

// Workgroup barrier ordering LDS accesses only (outstanding global loads stay in flight).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Waits until *flag >= epoch (called by one thread; the caller's barrier then orders the
// workgroup's reads after it).  Polls are relaxed agent-scope loads, which read the
// coherent value without invalidating caches; a single acquire fence follows success.  (An
// acquire on every poll invalidates the caches each time, and with hundreds of polling
// workgroups that starves every other memory access.)  Gives up, recording
// MI_SOLVE_TIMEOUT, after kSpinLimit polls, or at once when another workgroup gave up.
__device__ __forceinline__ bool wait_flag(const uint32_t *flag, uint32_t epoch, int *status)
{
    uint32_t spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
        __builtin_amdgcn_s_sleep(2);
        if ((++spins & 255) == 0 && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return false;
        if (spins > kSpinLimit) {
            atomicOr(status, MI_SOLVE_TIMEOUT);
            return false;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// BACKWARD = false: forward sweep with L; true: backward sweep with L^T.
// Winv holds, per block b, X_b^T (ld 128) at Winv + b*2*16384 and X_b at +16384.
// T: storage of L (float for a mixed-precision factor, widened on load; FP64 arithmetic).
//
// Each wave owns one 16-row tile of the block, rt = w for waves 0..3 and 11 - w for waves
// 4..7: waves w and w+4 share a SIMD, and the triangular X_i product costs row tile rt
// 4(rt+1) MFMA steps forward (4(8-rt) backward), so every SIMD issues 36 of them.
template <bool BACKWARD, typename T>
__global__ __launch_bounds__(512) void solve_sweep_kernel(const T *__restrict__ L, int64_t lda, int64_t n,
                                                          double *__restrict__ B, int64_t ldb, int64_t nrhs,
                                                          const double *__restrict__ Winv, uint32_t *flags,
                                                          int *ticket, uint32_t epoch, const int *info, int *status)
{
    if (factor_failed(info)) return;  // uniform: info is final before this kernel starts
    __shared__ double Ys[SB * SC];    // Ys[c + SC*k]: the published block being consumed
    __shared__ double Ts[SB * SC];    // Ts[c + SC*r]: the block's residual before the inverse
    __shared__ int s_item;
    __shared__ int s_ok;
    const int nblk = (int)((n + SB - 1) / SB);
    const int ngrp = (int)((nrhs + SC - 1) / SC);
    const int tid = threadIdx.x, l = lane_id();
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);

    if (tid == 0) s_item = atomicAdd(ticket, 1);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_item);
    if (t >= nblk * ngrp) return;
    const int ib = BACKWARD ? nblk - 1 - t / ngrp : t / ngrp;
    const int g = t % ngrp;
    const int64_t i0 = (int64_t)ib * SB, c0 = (int64_t)g * SC;
    const int bi = (int)((n - i0) < SB ? (n - i0) : SB);
    const int nc = (int)((nrhs - c0) < SC ? (nrhs - c0) : SC);
    uint32_t *gflags = flags + (size_t)g * nblk;

    // this wave's 16 rows of the block
    const int rt = w < 4 ? w : 11 - w;
    const int r = 16 * rt + (l & 15);
    const int rc = r < bi ? r : bi - 1;  // clamped row (in bounds)
    const int kl = l >> 4;

    // B_i (this wave's rows; lane l, element e: row r, column 4e + l/16): read before any wait
    double bv[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = 4 * e + kl;
        bv[e] = B[(i0 + rc) + (c0 + (c < nc ? c : 0)) * ldb];
    }

    // P operand of a contributing block jb, all 32 k-steps of this lane at once:
    // P[row][k] = L_ij[row][k] (forward) or L_ji[k][row] (backward); k clamped in bounds.
    const int jstep = BACKWARD ? -1 : 1;
    double pv[SB / 4];
#define MI_LOAD_OPERAND(jb_)                                                                       \
    do {                                                                                           \
        const int64_t j0_ = (int64_t)(jb_)*SB;                                                     \
        const int bj_ = (int)((n - j0_) < SB ? (n - j0_) : SB);                                    \
        const T *colbase_ = BACKWARD ? L + j0_ + (i0 + rc) * lda : L + (i0 + rc) + j0_ * lda;       \
        _Pragma("unroll") for (int s = 0; s < SB / 4; ++s)                                         \
        {                                                                                          \
            const int k_ = (4 * s + kl) < bj_ ? (4 * s + kl) : bj_ - 1;                             \
            pv[s] = (double)colbase_[BACKWARD ? (int64_t)k_ : (int64_t)k_ * lda];                  \
        }                                                                                          \
    } while (0)

    dvec4 acc = {0.0, 0.0, 0.0, 0.0};
    int jb = BACKWARD ? nblk - 1 : 0;
    if (jb != ib) MI_LOAD_OPERAND(jb);
    for (; jb != ib; jb += jstep) {
        if (tid == 0) s_ok = wait_flag(gflags + jb, epoch, status) ? 1 : 0;
        __syncthreads();
        if (!s_ok) return;
        const int64_t j0 = (int64_t)jb * SB;
        const int bj = (int)((n - j0) < SB ? (n - j0) : SB);
        for (int e = tid; e < SB * SC; e += 512) {  // Y_j (bj x nc, zero-padded to 128 x 16)
            const int c = e & (SC - 1), k = e >> 4;
            const bool in = k < bj && c < nc;
            const double v = coherent_load(B + (j0 + (in ? k : 0)) + (c0 + (in ? c : 0)) * ldb);
            Ys[e] = in ? v : 0.0;
        }
        lds_barrier();
        // acc[row][c] += sum_k P[row][k] * Ys[k][c]; rows k >= bj of Ys are zero, so the
        // clamped (finite) P values past a partial block contribute nothing
#pragma unroll
        for (int s = 0; s < SB / 4; ++s) acc = mfma_f64(Ys[(l & 15) + SC * (4 * s + kl)], pv[s], acc);
        if (jb + jstep != ib) MI_LOAD_OPERAND(jb + jstep);  // next block: in flight during the next wait
        lds_barrier();  // Ys is rewritten by the next contribution
    }
#undef MI_LOAD_OPERAND

    // op(X_i): column-major P[row][k] = X_i[row][k] (forward, lower: k < 16(rt+1)) or X_i[k][row]
    // (backward, upper: k >= 16 rt); only the nonzero k-steps are loaded and multiplied
    const double *Xp = Winv + (size_t)ib * 2 * SB * SB + (BACKWARD ? 0 : SB * SB) + r + SB * kl;
    const int s_lo = BACKWARD ? 4 * rt : 0;
    const int s_hi = BACKWARD ? SB / 4 : 4 * (rt + 1);
#pragma unroll
    for (int s = 0; s < SB / 4; ++s)
        if (s >= s_lo && s < s_hi) pv[s] = Xp[(size_t)4 * SB * s];

    // T = B_i - acc  (lane l, element e: row r, column 4e + l/16)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = 4 * e + kl;
        const bool in = r < bi && c < nc;
        Ts[c + SC * r] = in ? bv[e] - acc[e] : 0.0;
    }
    lds_barrier();
    // Y_i = op(X_i) T
    dvec4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < SB / 4; ++s)
        if (s >= s_lo && s < s_hi) y = mfma_f64(Ts[(l & 15) + SC * (4 * s + kl)], pv[s], y);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = 4 * e + kl;
        if (r < bi && c < nc) B[(i0 + r) + (c0 + c) * ldb] = y[e];
    }
    // every wave's stores reach L2 before the barrier; one agent-scope release store then
    // publishes them all
    global_barrier();
    if (tid == 0) __hip_atomic_store(gflags + ib, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void reset_ticket_kernel(int *ticket) { *ticket = 0; }

// flags: ngrp * nblk uint32 (persistent across calls, epochs increase); tickets: 2 ints.
template <typename T>
void launch_solve_persistent(const T *L, int64_t lda, int64_t n, double *B, int64_t ldb, int64_t nrhs,
                             const double *Winv, uint32_t *flags, int *tickets, uint32_t epoch_fwd,
                             uint32_t epoch_bwd, const int *info, int *status, hipStream_t s, int sweeps)
{
    const int nblk = (int)((n + SB - 1) / SB);
    const int ngrp = (int)((nrhs + SC - 1) / SC);
    const unsigned items = (unsigned)(nblk * ngrp);
    if (sweeps & 1) {
        hipLaunchKernelGGL(reset_ticket_kernel, dim3(1), dim3(1), 0, s, tickets);
        hipLaunchKernelGGL((solve_sweep_kernel<false, T>), dim3(items), dim3(512), 0, s, L, lda, n, B, ldb, nrhs, Winv,
                           flags, tickets, epoch_fwd, info, status);
    }
    if (sweeps & 2) {
        hipLaunchKernelGGL(reset_ticket_kernel, dim3(1), dim3(1), 0, s, tickets + 1);
        hipLaunchKernelGGL((solve_sweep_kernel<true, T>), dim3(items), dim3(512), 0, s, L, lda, n, B, ldb, nrhs, Winv,
                           flags, tickets + 1, epoch_bwd, info, status);
    }
}

template void launch_solve_persistent<double>(const double *, int64_t, int64_t, double *, int64_t, int64_t,
                                              const double *, uint32_t *, int *, uint32_t, uint32_t, const int *,
                                              int *, hipStream_t, int);
template void launch_solve_persistent<float>(const float *, int64_t, int64_t, double *, int64_t, int64_t,
                                             const double *, uint32_t *, int *, uint32_t, uint32_t, const int *,
                                             int *, hipStream_t, int);

}  // namespace mi
