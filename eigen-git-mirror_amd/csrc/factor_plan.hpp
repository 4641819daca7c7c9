// factor_plan.hpp -- the host-side sequence of launches of one blocked factorization, shared
// by the handle entry points (api.hip) and the allocation-free one (noalloc.hip).
//
// Right-looking blocked Cholesky (llt_inplace::blocked, Eigen/src/Cholesky/LLT.h:338-368) with
// a one-panel lookahead, one launch per step on one stream:
//   launch -1 : panel 0 (diagonal factor + panel solve of block column 0)
//   launch k  : panel k+1 (its update by panel k, diagonal factor, panel solve)
//               + trailing update by panels k-1 and k (depth 256) of every other block
//               column from k+2 on, or (FP32 storage) by panels k-1, k of column k+2 and
//               panels k-3 .. k of every fourth column from k+4 (factor_step.hip explains
//               both schedules).
// The two halves of a launch touch disjoint data; everything they read was finished by an
// earlier launch, so stream order is the only synchronisation the host needs.
#pragma once
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

// Zeroes the ticket counter and the panel flag (sync[kSyncFactorTickets], sync[kSyncFlag]) on
// the stream: a kernel, not a memset, so that a captured factorization replays it
// (factor_step.hip).
void launch_factor_reset(uint32_t *sync, hipStream_t s, const Batch &bt = Batch());

// The trailing update of launch k in the four-panel schedule of the FP32-storage
// factorizations (factor_step.hip): npair tiles of block column c0 = k+2 at depth K from
// column slab, then nquad tiles of the block columns cq0 = k+4, k+8, ... at depth KQ from
// column slabq.  k = -1: none.
// oct (the 2-part split-BF16 arithmetic): block column k+4 alone takes the four panels, and
// the block columns co0 = k+8, k+16, ... take the eight panels k-7 .. k (depth Ko from column
// slabo, nocts tiles).
struct QuadPlan {
    int c0 = 0, cq0 = 0, npair = 0, nquad = 0, K = 128, KQ = 128;
    int64_t slab = 0, slabq = 0;
    bool oct = false;
    int co0 = 0, nocts = 0, Ko = 128;
    int64_t slabo = 0;
};
QuadPlan quad_plan(int64_t n, int64_t k, bool oct);
inline bool oct_schedule(int prec) { return prec == kPrecBF16x3; }

// Algorithmic flop counts (LAPACK convention): POTRF of b: b^3/3 + b^2/2 + b/6; TRSM m x b
// panel: m b^2; panel update by the previous panel: 2 (n - k0) b NB minus its upper part.
inline double panel_flops(int64_t n, int64_t k0, int64_t NB)
{
    const double b = (double)((n - k0) < NB ? (n - k0) : NB);
    const double m = (double)(n - k0 - NB > 0 ? n - k0 - NB : 0);
    const double mc = (double)(n - k0);
    return (k0 > 0 ? 2.0 * mc * b * NB - b * (b - 1) * NB : 0.0) + b * b * b / 3 + b * b / 2 + b / 6 + m * b * b;
}

// Queues the factorization of the lower triangle of A (n x n, ld lda) on stream s.
// sync: the synchronisation words whose ticket counter and panel flag the factorization uses
// (at kSyncFactorTickets, kSyncFlag); status: the status word; wleaf: the panel images
// (kPanelDoubles per 128-block).  each(flops, launch) runs every step's launch (the tracing
// hook of the handle API); the ticket bases and flag epochs count from the reset, so every
// call and every replay of a captured call has the same accounting.  bt: a batch of systems of
// the same order (each launch covers all of them; every system has its own block of sync
// words, bt.sync apart, and its info at d_info + z).
template <typename T, class Each>
void enqueue_factorization(hipStream_t s, T *A, int64_t lda, int64_t n, double *wleaf, int *d_info, uint32_t *sync,
                           int *status, int prec, __bf16 *wsplit, Each &&each, const Batch &bt = Batch())
{
    constexpr int64_t NB = kNB;
    const int64_t nblk = (n + NB - 1) / NB;
    launch_factor_reset(sync, s, bt);
    uint32_t tickets = 0, epoch_now = 0;
    for (int64_t k = -1; k < nblk; ++k) {
        const int64_t kp = (k + 1 < nblk) ? (k + 1) * NB : -1;  // panel of this launch
        // trailing update of launch k: panels k-1, k (k >= 1; panel 0 alone for k = 0) to the
        // block columns c = k+2, k+4, ... (all rows from c down); with FP32 storage panels
        // k-1, k to column k+2 and panels k-3 .. k to the columns k+4, k+8, ...
        double fl_upd = 0.0;
        auto column = [&](int64_t c, double K) {
            const double m = (double)(n - c * NB), w = m < NB ? m : (double)NB;
            fl_upd += 2.0 * K * (m * w - w * (w - 1) / 2);
        };
        if constexpr (sizeof(T) == 8) {
            const double K = k >= 1 ? 2.0 * NB : (double)NB;
            for (int64_t c = k + 2; k >= 0 && c < nblk; c += 2) column(c, K);
        } else if (k >= 0) {
            const QuadPlan q = quad_plan(n, k, oct_schedule(prec));
            if (q.npair > 0) column(q.c0, (double)q.K);
            for (int64_t c = q.cq0; c < nblk; c += q.oct ? nblk : 4) column(c, (double)q.KQ);
            for (int64_t c = q.co0; q.oct && c < nblk; c += 8) column(c, (double)q.Ko);
        }
        if (kp < 0 && fl_upd == 0.0) break;
        const double fl = (kp >= 0 ? panel_flops(n, kp, NB) : 0.0) + fl_upd;
        const uint32_t epoch = (kp >= 0) ? ++epoch_now : epoch_now;
        uint32_t grid = 0;
        each(fl, [&] {
            launch_factor_step<T>(A, lda, n, kp, k, wleaf, d_info, sync + kSyncFactorTickets, tickets, sync + kSyncFlag,
                                  epoch, status, s, &grid, prec, wsplit, bt);
        });
        tickets += grid;
    }
}

}  // namespace mi
