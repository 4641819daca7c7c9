// factor_step.hip -- one launch per block step of the factorization: the panel of block
// column k+1 (its update by panel k, diagonal factor, panel solve) together with the
// trailing update by panel k of the columns from k+2 on.
//
// This is llt_inplace<double,Lower>::blocked (Eigen/src/Cholesky/LLT.h:338-368) with a
// one-panel lookahead: the two halves of a step touch disjoint data, and everything each
// depends on was finished by an earlier launch on the same stream, so a single stream and a
// single launch per step carry all the ordering (no events, no second stream).  Workgroups
// take roles from a ticket counter, in the order: the diagonal role of the panel (the
// critical path, always running before anyone waits for it), the trailing-update tiles,
// then the panel's row roles -- which wait for the diagonal block, so they should not hold
// CUs while it is being factored.
#include "potrf_device.hpp"
#include "syrk_device.hpp"

namespace mi {

// kp: block column (first column) of the panel to factor, or -1 when none; Gp its roles.
// ks: first column of the panel whose trailing update runs here, or -1 when none; the
//     update covers rows/columns from r2 = ks + 256 on, T x T tiles (lower triangle).
// S: storage type (double, or float for the mixed-precision factorization), PREC: the
// trailing-update arithmetic (kPrecF64 with double; FP32/BF16/split BF16 MFMA with float).
template <typename S, int PREC>
__global__ __launch_bounds__(256) void factor_step_kernel(S *__restrict__ A, int64_t lda, int64_t n, int64_t kp,
                                                          int Gp, int64_t ks, int T, double *__restrict__ Wleaf,
                                                          int *info, uint32_t *tickets, uint32_t ticket_base,
                                                          uint32_t *flag, uint32_t epoch, int *status)
{
    __shared__ double smem[kSmem];
    __shared__ int s_role, s_ok;
    // every workgroup draws its ticket, even when it then stops: the host counts them
    if (threadIdx.x == 0) s_role = (int)(atomicAdd(tickets, 1u) - ticket_base);
    __syncthreads();
    if (factor_failed(info)) return;
    const int t = s_role;
    const int nsyrk = T * (T + 1) / 2;
    if (Gp > 0 && (t == 0 || t > nsyrk)) {  // ticket 0: diagonal role; last Gp-1: row roles
        panel_role<S, PREC>(t == 0 ? 0 : t - nsyrk, smem, &s_ok, A, lda, n, kp, Wleaf, info, flag, epoch,
                            status);
        return;
    }
    const int u = Gp > 0 ? t - 1 : t;
    const int64_t r2 = ks + 256;
    int ti, tj;
    lower_tile(xcd_remap(u, nsyrk), ti, tj);
    syrk_tile<kNB, S, PREC>(A + r2 + r2 * lda, lda, A + r2 + ks * lda, lda, (int)(n - r2), (int)(n - r2), ti, tj);
}

template <typename S>
void launch_factor_step(S *A, int64_t lda, int64_t n, int64_t kp, int64_t ks, double *Wleaf, int *info,
                        uint32_t *tickets, uint32_t ticket_base, uint32_t *flag, uint32_t epoch, int *status,
                        hipStream_t s, uint32_t *grid_out, int prec)
{
    const int Gp = kp >= 0 ? (int)((n - kp + 127) / 128) : 0;
    const int64_t m2 = ks >= 0 ? n - ks - 256 : 0;
    const int T = m2 > 0 ? (int)((m2 + 127) / 128) : 0;
    const unsigned grid = (unsigned)(Gp + T * (T + 1) / 2);
    *grid_out = grid;
    if (grid == 0) return;
    const int64_t ksx = m2 > 0 ? ks : (int64_t)-1;
    if constexpr (sizeof(S) == 8) {
        hipLaunchKernelGGL((factor_step_kernel<S, kPrecF64>), dim3(grid), dim3(256), 0, s, A, lda, n, kp, Gp, ksx, T,
                           Wleaf, info, tickets, ticket_base, flag, epoch, status);
    } else if (prec == kPrecBF16) {
        hipLaunchKernelGGL((factor_step_kernel<S, kPrecBF16>), dim3(grid), dim3(256), 0, s, A, lda, n, kp, Gp, ksx, T,
                           Wleaf, info, tickets, ticket_base, flag, epoch, status);
    } else if (prec == kPrecBF16x3) {
        hipLaunchKernelGGL((factor_step_kernel<S, kPrecBF16x3>), dim3(grid), dim3(256), 0, s, A, lda, n, kp, Gp, ksx,
                           T, Wleaf, info, tickets, ticket_base, flag, epoch, status);
    } else {
        hipLaunchKernelGGL((factor_step_kernel<S, kPrecF32>), dim3(grid), dim3(256), 0, s, A, lda, n, kp, Gp, ksx, T,
                           Wleaf, info, tickets, ticket_base, flag, epoch, status);
    }
}

template void launch_factor_step<double>(double *, int64_t, int64_t, int64_t, int64_t, double *, int *, uint32_t *,
                                         uint32_t, uint32_t *, uint32_t, int *, hipStream_t, uint32_t *, int);
template void launch_factor_step<float>(float *, int64_t, int64_t, int64_t, int64_t, double *, int *, uint32_t *,
                                        uint32_t, uint32_t *, uint32_t, int *, hipStream_t, uint32_t *, int);

}  // namespace mi
