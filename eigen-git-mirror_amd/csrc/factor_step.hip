// factor_step.hip -- one launch per block step of the factorization: the panel of block
// column k+1 (its update by panel k, diagonal factor, panel solve) together with a trailing
// update by the two most recent panels.
//
// This is llt_inplace<double,Lower>::blocked (Eigen/src/Cholesky/LLT.h:338-368) with a
// one-panel lookahead: the two halves of a step touch disjoint data, and everything each
// depends on was finished by an earlier launch on the same stream, so a single stream and a
// single launch per step carry all the ordering (no events, no second stream).
//
// Trailing updates two panels deep.  Launch k (panel k just finished) applies the pair of
// panels k-1, k -- a contiguous 256-column slab -- with depth K = 256 to every other block
// column c >= k+2, those with c - k even; launch k+1 applies the pair k, k+1 to the others.
// So every panel reaches every later block column exactly once (panel q through launch q
// when c - q is even, through launch q+1 when it is odd, and block column q+1 through the
// panel roles of launch q), each launch does the flops of a one-panel update, and the
// trailing matrix is read and written half as often as with K = 128 -- at K = 128 that
// traffic, not the MFMA, bounds the update.  Launch 0 applies panel 0 alone (K = 128).
//
// The FP32-storage factorizations of the mixed-precision solve use four-panel slabs instead:
// there the trailing update moves data rather than MFMA work (C read and written in FP32, the
// panels' BF16 parts staged through LDS), and a deeper slab reads and writes C once per four
// panels.  Launch k applies the pair of panels k-1, k to block column k+2 alone and the four
// panels k-3 .. k (depth K = 512) to the block columns c = k+4, k+8, ...  Block column c then
// receives panel c-1 from the panel role of launch c-1, panels c-3, c-2 from launch c-2, and
// panels 0 .. c-4 from launches c-4, c-8, ..., each exactly once and after it was factored;
// the first launches use the panels that exist (depth 128 (k + 1) for k < 3).  With FP64
// storage a launch's update is MFMA work, and the deeper tiles would leave the last round of
// each launch emptier: the FP64 factorization keeps the two-panel alternation.
//
// Workgroups take roles from a ticket counter, in the order: the diagonal role of the panel
// (the critical path, always running before anyone waits for it), the trailing-update tiles,
// then the panel's row roles -- which wait for the diagonal block, so they should not hold
// CUs while it is being factored.
#include "factor_plan.hpp"
#include "potrf_device.hpp"
#include "syrk_device.hpp"

namespace mi {

// Tile u of an update of the block columns c0, c0 + step, ... (rows c .. nblk-1 of each) ->
// (block column c = c0 + step m, block row i >= c).
__device__ __forceinline__ void strided_tile(int u, int c0, int step, int nblk, int &i, int &c)
{
    int m = 0;
    while (u >= nblk - c0 - step * m) {
        u -= nblk - c0 - step * m;
        ++m;
    }
    c = c0 + step * m;
    i = c + u;
}

// kp: block column (first column) of the panel to factor, or -1 when none; Gp its roles.
// c0: first block column of the trailing update (every other one from there), nsyrk its
//     tiles (0: none); slab: first column of the K-wide operand (the last one or two panels).
// KQ > 0 (the four-panel schedule): c0 is the one block column of the pair update (depth K,
//     slab), and the nsyrk tiles continue over the block columns cq0, cq0 + 4, ... with
//     depth KQ from column slabq; with the 2-part split arithmetic (the eight-panel schedule)
//     block column cq0 alone takes depth KQ, and the tiles continue over the block columns
//     co0, co0 + 8, ... with nko k-steps of 32 from column slabo.
// S: storage type (double, or float for the mixed-precision factorization), PREC: the
// trailing-update arithmetic (kPrecF64 with double; FP32/BF16/split BF16 MFMA with float).
// Wsplit (split BF16 arithmetic only, else null): a ring of 4 panel slots of slot_elems BF16
// each (2 or 3 parts of split_part_elems(n): hi, lo or hi, mid, lo); the row roles of panel q
// write slot q % 4 and the trailing update reads its depth from the slots of the slab's panels.
// Workgroups per CU the kernel is compiled for: 2 fits two 72 KiB LDS blocks and 256 VGPRs
// per lane on a CU (the second workgroup hides the first one's barrier and memory waits).
#ifndef MI_FACTOR_WG_PER_CU
#define MI_FACTOR_WG_PER_CU 2
#endif

template <typename S, int PREC, int K, int KQ>
__global__ __launch_bounds__(256, MI_FACTOR_WG_PER_CU) void factor_step_kernel(S *__restrict__ A, int64_t lda, int64_t n, int64_t kp,
                                                          int Gp, int c0, int nsyrk, int64_t slab, int cq0,
                                                          int64_t slabq, int co0, int64_t slabo, int nko,
                                                          double *__restrict__ Wleaf, int *info, uint32_t *tickets,
                                                          uint32_t ticket_base, uint32_t *flag, uint32_t epoch,
                                                          int *status, __bf16 *Wsplit, int64_t slot_elems,
                                                          int64_t sA, int64_t sW, int64_t sSync)
{
    __shared__ __align__(16) double smem[kSmem];
    {  // batched launches: system z = blockIdx.z (its matrix, panel images, tickets, flag, info)
        const int64_t z = blockIdx.z;
        A += z * sA;
        Wleaf += z * sW;
        tickets += z * sSync;
        flag += z * sSync;
        info += z;
    }
    __shared__ int s_role, s_ok, s_stop;
    // Every workgroup draws its ticket, even when it then stops: the host counts them.  The
    // failure words (info, status) are read in the same round trip as the ticket, so a
    // workgroup starts after one memory latency, not three.
    if (threadIdx.x == 0) {
        const uint32_t tk = atomicAdd(tickets, 1u);
        const int iv = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int sv = __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_role = (int)(tk - ticket_base);
        s_stop = iv | sv;
    }
    __syncthreads();
    if (s_stop) return;
    const int t = __builtin_amdgcn_readfirstlane(s_role);  // uniform: every derived address is scalar
    // A role outside this launch means the counter was not reset with the host's accounting
    // (never in a correct call sequence): record it and touch nothing, rather than turn the
    // ticket into an address.
    if (t < 0 || t >= Gp + nsyrk) {
        if (threadIdx.x == 0) atomicOr(status, MI_FACTOR_TICKET);
        return;
    }
    if (Gp > 0 && (t == 0 || t > nsyrk)) {  // ticket 0: diagonal role; last Gp-1: row roles
        __bf16 *split = Wsplit ? Wsplit + ((kp / kNB) & (kSplitRing - 1)) * slot_elems : nullptr;
        panel_role<S, PREC>(t == 0 ? 0 : t - nsyrk, smem, &s_ok, A, lda, n, kp, Wleaf, info, flag, epoch,
                            status, split, split_part_elems(n));
        return;
    }
    const int u = xcd_remap(Gp > 0 ? t - 1 : t, nsyrk);
    const int nblk = (int)((n + kNB - 1) / kNB);
    constexpr bool OCT = KQ > 0 && PREC == kPrecBF16x3;
    int i, c;
    bool quad = false, oct = false;
    if constexpr (KQ == 0) {
        strided_tile(u, c0, 2, nblk, i, c);
    } else {
        const int npair = c0 < nblk ? nblk - c0 : 0;
        const int nq = cq0 < nblk ? nblk - cq0 : 0;
        if (u < npair) {
            c = c0;
            i = c0 + u;
        } else if (!OCT) {
            strided_tile(u - npair, cq0, 4, nblk, i, c);
            quad = true;
        } else if (u - npair < nq) {
            c = cq0;
            i = cq0 + (u - npair);
            quad = true;
        } else {
            strided_tile(u - npair - nq, co0, 8, nblk, i, c);
            oct = true;
        }
    }
    const int64_t r = (int64_t)c * kNB;
    const int64_t m = n - r;
    const int ncol = (int)(m < kNB ? m : kNB);
    if constexpr (sizeof(S) == 4 && (PREC == kPrecBF16x3 || PREC == kPrecBF16x6)) {
        if (Wsplit) {
            constexpr int PARTS = PREC == kPrecBF16x3 ? 2 : 3;
            if constexpr (KQ > 0) {
                if (quad || oct) {
                    syrk_tile_split_lds<PARTS>(A + r + r * lda, lda, Wsplit, (oct ? slabo : slabq) / kNB, slot_elems,
                                               split_part_elems(n), oct ? nko : KQ / 32, (int)r, (int)m, ncol, i - c,
                                               0, (char *)smem);
                    return;
                }
            }
            syrk_tile_split_lds<PARTS>(A + r + r * lda, lda, Wsplit, slab / kNB, slot_elems, split_part_elems(n), K / 32,
                                       (int)r, (int)m, ncol, i - c, 0, (char *)smem);
            return;
        }
    }
    if constexpr (KQ > 0) {
        if (quad) {
            syrk_tile<KQ, S, PREC>(A + r + r * lda, lda, A + r + slabq * lda, lda, (int)m, ncol, i - c, 0);
            return;
        }
    }
    syrk_tile<K, S, PREC>(A + r + r * lda, lda, A + r + slab * lda, lda, (int)m, ncol, i - c, 0);
}

__global__ void factor_reset_kernel(uint32_t *sync, int64_t sSync)
{
    sync += (int64_t)blockIdx.z * sSync;
    if (threadIdx.x == 0) sync[kSyncFactorTickets] = 0u;
    if (threadIdx.x == 1) sync[kSyncFlag] = 0u;
}

void launch_factor_reset(uint32_t *sync, hipStream_t s, const Batch &bt)
{
    hipLaunchKernelGGL(factor_reset_kernel, dim3(1, 1, bt.count), dim3(64), 0, s, sync, bt.sync);
}

// Trailing-update tiles of launch k (panel k just finished; k = -1: none) in the two-panel
// alternation: the block columns k+2, k+4, ...
int pair_tiles(int64_t n, int64_t k, int *c0_out)
{
    const int nblk = (int)((n + kNB - 1) / kNB);
    const int c0 = (int)k + 2;
    *c0_out = c0;
    if (k < 0) return 0;
    int tiles = 0;
    for (int c = c0; c < nblk; c += 2) tiles += nblk - c;
    return tiles;
}

QuadPlan quad_plan(int64_t n, int64_t k, bool oct)
{
    QuadPlan q;
    const int nblk = (int)((n + kNB - 1) / kNB);
    q.oct = oct;
    q.c0 = (int)k + 2;
    q.cq0 = (int)k + 4;
    q.co0 = (int)k + 8;
    if (k < 0) return q;
    q.npair = q.c0 < nblk ? nblk - q.c0 : 0;
    if (oct) {
        q.nquad = q.cq0 < nblk ? nblk - q.cq0 : 0;
        for (int c = q.co0; c < nblk; c += 8) q.nocts += nblk - c;
    } else {
        for (int c = q.cq0; c < nblk; c += 4) q.nquad += nblk - c;
    }
    q.K = k >= 1 ? 256 : 128;
    q.slab = (k >= 1 ? k - 1 : 0) * kNB;
    const int64_t first = k >= 3 ? k - 3 : 0;
    q.KQ = (int)(k - first + 1) * kNB;
    q.slabq = first * kNB;
    const int64_t firsto = k >= 7 ? k - 7 : 0;
    q.Ko = (int)(k - firsto + 1) * kNB;
    q.slabo = firsto * kNB;
    return q;
}

template <typename S, int PREC, int K, int KQ>
static void launch_one(dim3 grid, hipStream_t s, S *A, int64_t lda, int64_t n, int64_t kp, int Gp, int c0, int nsyrk,
                       int64_t slab, int cq0, int64_t slabq, int co0, int64_t slabo, int nko, double *Wleaf, int *info, uint32_t *tickets,
                       uint32_t ticket_base, uint32_t *flag, uint32_t epoch, int *status, __bf16 *Wsplit,
                       int64_t slot_elems, const Batch &bt)
{
    hipLaunchKernelGGL((factor_step_kernel<S, PREC, K, KQ>), grid, dim3(256), 0, s, A, lda, n, kp, Gp, c0, nsyrk, slab,
                       cq0, slabq, co0, slabo, nko, Wleaf, info, tickets, ticket_base, flag, epoch, status, Wsplit, slot_elems, bt.a,
                       bt.leaf, bt.sync);
}

template <typename S, int PREC>
static void launch_prec(dim3 grid, hipStream_t s, int K, int KQ, S *A, int64_t lda, int64_t n, int64_t kp, int Gp,
                        int c0, int nsyrk, int64_t slab, int cq0, int64_t slabq, int co0, int64_t slabo, int nko,
                        double *Wleaf, int *info,
                        uint32_t *tickets, uint32_t ticket_base, uint32_t *flag, uint32_t epoch, int *status,
                        __bf16 *Wsplit, int64_t slot_elems, const Batch &bt)
{
#define MI_FS_ARGS grid, s, A, lda, n, kp, Gp, c0, nsyrk, slab, cq0, slabq, co0, slabo, nko, Wleaf, info, tickets, \
                   ticket_base, flag, \
                   epoch, status, Wsplit, slot_elems, bt
    if constexpr (sizeof(S) == 8) {
        if (K == 256)
            launch_one<S, PREC, 256, 0>(MI_FS_ARGS);
        else
            launch_one<S, PREC, 128, 0>(MI_FS_ARGS);
    } else {
        if (K == 128)
            launch_one<S, PREC, 128, 128>(MI_FS_ARGS);
        else if (KQ == 256)
            launch_one<S, PREC, 256, 256>(MI_FS_ARGS);
        else if (KQ == 384)
            launch_one<S, PREC, 256, 384>(MI_FS_ARGS);
        else
            launch_one<S, PREC, 256, 512>(MI_FS_ARGS);
    }
#undef MI_FS_ARGS
}

template <typename S>
void launch_factor_step(S *A, int64_t lda, int64_t n, int64_t kp, int64_t k, double *Wleaf, int *info,
                        uint32_t *tickets, uint32_t ticket_base, uint32_t *flag, uint32_t epoch, int *status,
                        hipStream_t s, uint32_t *grid_out, int prec, __bf16 *Wsplit, const Batch &bt)
{
    const int Gp = kp >= 0 ? (int)((n - kp + 127) / 128) : 0;
    int c0 = 0, nsyrk = 0, K = k >= 1 ? 256 : 128, KQ = 0, cq0 = 0, co0 = 0, nko = 0;
    int64_t slab = k >= 1 ? (k - 1) * kNB : 0, slabq = 0, slabo = 0;
    if constexpr (sizeof(S) == 8) {
        nsyrk = pair_tiles(n, k, &c0);
    } else {
        const QuadPlan q = quad_plan(n, k, oct_schedule(prec));
        c0 = q.c0;
        cq0 = q.cq0;
        co0 = q.co0;
        nsyrk = q.npair + q.nquad + q.nocts;
        K = q.K;
        KQ = q.KQ;
        nko = q.Ko / 32;
        slab = q.slab;
        slabq = q.slabq;
        slabo = q.slabo;
    }
    const unsigned grid = (unsigned)(Gp + nsyrk);
    *grid_out = grid;
    if (grid == 0) return;
    const dim3 g(grid, 1, bt.count);
    const int64_t slot = split_slot_elems(n, prec == kPrecBF16x6 ? 3 : 2);
#define MI_FS_ARGS g, s, K, KQ, A, lda, n, kp, Gp, c0, nsyrk, slab, cq0, slabq, co0, slabo, nko, Wleaf, info, \
                   tickets, ticket_base, flag, epoch, status
    if constexpr (sizeof(S) == 8) {
        launch_prec<S, kPrecF64>(MI_FS_ARGS, nullptr, slot, bt);
    } else if (prec == kPrecBF16) {
        launch_prec<S, kPrecBF16>(MI_FS_ARGS, nullptr, slot, bt);
    } else if (prec == kPrecBF16x6) {
        launch_prec<S, kPrecBF16x6>(MI_FS_ARGS, Wsplit, slot, bt);
    } else if (prec == kPrecBF16x3) {
        launch_prec<S, kPrecBF16x3>(MI_FS_ARGS, Wsplit, slot, bt);
    } else {
        launch_prec<S, kPrecF32>(MI_FS_ARGS, nullptr, slot, bt);
    }
#undef MI_FS_ARGS
}

template void launch_factor_step<double>(double *, int64_t, int64_t, int64_t, int64_t, double *, int *, uint32_t *,
                                         uint32_t, uint32_t *, uint32_t, int *, hipStream_t, uint32_t *, int,
                                         __bf16 *, const Batch &);
template void launch_factor_step<float>(float *, int64_t, int64_t, int64_t, int64_t, double *, int *, uint32_t *,
                                        uint32_t, uint32_t *, uint32_t, int *, hipStream_t, uint32_t *, int,
                                        __bf16 *, const Batch &);

}  // namespace mi
