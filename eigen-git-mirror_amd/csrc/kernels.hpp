// kernels.hpp -- host-side launchers of the LLT kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

// Block size of the blocked factorization: the diagonal block factored in one workgroup,
// the depth of each trailing update, and the granularity of the stored inverses.
constexpr int kNB = 128;

// Workspace of one factorization, per 128-block b:
//   Wleaf[b]: the panel image of L_bb, kPanelDoubles: its 28 strictly lower 16 x 16 tiles
//             (tile (ti, tj), ti > tj, at 256 * (ti(ti-1)/2 + tj), each column-major), then at
//             kLeafOffset the 8 inverses of its 16 x 16 diagonal leaves (column-major).  The
//             factorization's diagonal role publishes it for the panel solve, which copies it
//             to LDS as it is; the solve-only path fills the leaf part alone.
//   Winv[b] : L_bb^{-T}, 128 x 128, ld 128 (built for the solve)
constexpr int64_t kPanelDoubles = 36 * 256;
constexpr int64_t kLeafOffset = 28 * 256;
// Per diagonal block b, six 128 x 128 matrices (ld 128): X_b^T, X_b (X_b = L_bb^{-1}), then the
// couplings of the solve M_b = X_b L_{b,b-1}, N_b = X_b^T L_{b+1,b}^T, M2_b = X_b L_{b,b-2},
// N2_b = X_b^T L_{b+2,b}^T (zero where undefined).
constexpr int64_t kInvDoubles = 8 * 128 * 128;
constexpr int64_t kInvX = 128 * 128, kInvM = 2 * 128 * 128, kInvN = 3 * 128 * 128;  // offsets
constexpr int64_t kInvM2 = 4 * 128 * 128, kInvN2 = 5 * 128 * 128;
constexpr int64_t kInvM3 = 6 * 128 * 128, kInvN3 = 7 * 128 * 128;
// Couplings of the block substitution per direction: M, M2 from 8 blocks on, and M3 from 16.
// Each takes a block's triangular product one predecessor further off the chain from block
// to block, which shortens long chains; for a few blocks (a batch of small systems, whose
// chains are short and whose launch is bound by throughput) they cost more to build than they
// save, and the substitution is plain.
__host__ __device__ constexpr int solve_couplings(int64_t nblk) { return nblk >= 16 ? 3 : nblk >= 8 ? 2 : 0; }

// A batched launch: `count` systems of one order, system z (blockIdx.z) at every base pointer
// plus z times that buffer's stride in elements -- a: the matrix or factor, b: right-hand
// sides, leaf: panel images, inv: block inverses, pack: the solve's packed copies, sync: the
// system's block of synchronisation words, g / aux / aux2: the adjoint's operands and scratch;
// an info pointer advances by z.  The default is one system (the strides are then unused).
struct Batch {
    unsigned count = 1;
    int64_t a = 0, b = 0, leaf = 0, inv = 0, pack = 0, sync = 0, g = 0, aux = 0, aux2 = 0;
};

// Synchronisation words of a handle (32-bit words in device memory).  Each group sits on its
// own 128-byte line: every starting workgroup of a factorization step takes an atomic on the
// ticket counter and the waiting ones poll the flag, and a read of another word of that line
// (the status, read by every workgroup as it starts) waits behind those atomics.
constexpr int kSyncStatus = 0;          // status bits (MI_SOLVE_TIMEOUT, ...)
constexpr int kSyncSolveTickets = 32;   // 2 words: the forward and the backward sweep
constexpr int kSyncRankTicket = 34;     // the rank update's ticket
constexpr int kSyncFactorTickets = 64;  // the factorization's ticket counter
constexpr int kSyncFlag = 96;           // the factorization's panel flag
constexpr int kSyncWords = 128;

// Device status bits (a word in the handle's persistent device memory).
constexpr int MI_SOLVE_TIMEOUT = 1;
constexpr int MI_PANEL_TIMEOUT = 2;
constexpr int MI_FACTOR_TICKET = 4;  // a factorization workgroup drew a ticket outside its launch

// Launch k of the factorization: panel of block column kp (or none, kp < 0) + the trailing
// update by panels k-1, k of the block columns k+2, k+4, ... (none for k < 0).  Writes the grid
// size (= tickets drawn) to *grid_out.  tickets/flag: persistent words of the handle,
// ticket_base = number of tickets drawn before this launch, epoch = this step's flag value.
// S = double: FP64 throughout (prec ignored); S = float: FP32 storage, FP64 diagonal factor
// and panel solve, trailing updates in the MFMA arithmetic `prec` (kPrecF32/BF16/BF16x3);
// with kPrecBF16x3 / kPrecBF16x6 and Wsplit given, the trailing update reads panels split once
// into 2 / 3 BF16 parts by the panel solve (4 * split_slot_elems(n, parts) elements, no other
// contents needed); without Wsplit, kPrecBF16x6 runs its trailing update on FP32 MFMAs.
template <typename S>
void launch_factor_step(S *A, int64_t lda, int64_t n, int64_t kp, int64_t k, double *Wleaf, int *info,
                        uint32_t *tickets, uint32_t ticket_base, uint32_t *flag, uint32_t epoch, int *status,
                        hipStream_t s, uint32_t *grid_out, int prec, __bf16 *Wsplit = nullptr,
                        const Batch &bt = Batch());
// BF16 elements of one part (n rows rounded up to 128, x 128 columns) and of one slot (2 parts
// for kPrecBF16x3, 3 for kPrecBF16x6) of the split-panel ring (Wsplit, 4 slots).
__host__ __device__ inline int64_t split_part_elems(int64_t n) { return ((n + 127) / 128 * 128) * 128; }
// Slots of the ring of split panels: launch k writes panel k+1's slot while its trailing update
// reads panels k-7..k (the 2-part split schedule), so 9 are in use at once.
constexpr int kSplitRing = 16;
__host__ __device__ inline int64_t split_slot_elems(int64_t n, int parts) { return parts * split_part_elems(n); }
// Winv[b] = {X_b^T, X_b, M_b, N_b, M2_b, N2_b}, X_b = L_bb^{-1}, for every diagonal block (needs Wleaf).
// couplings = false: X_b^T and X_b only (for users of Winv other than the solve).
template <typename T>
void launch_block_inverses(const T *L, int64_t lda, int64_t n, const double *Wleaf, double *Winv, const int *info,
                           hipStream_t s, bool couplings = true, const Batch &bt = Batch());
// Wleaf from an already factored matrix (standalone solve).
template <typename T>
void launch_leaf_inverses(const T *L, int64_t lda, int64_t n, double *Wleaf, hipStream_t s,
                          const Batch &bt = Batch());
// Solve: forward then backward sweep in place on B (one persistent launch each, each after a
// small reset launch).  tickets: 2 ints; Ypk: solve_pack_doubles(n, nrhs) doubles of scratch
// (each finished block's packed copy for its consumers), used by the two sweeps in turn: each
// sweep's reset runs after everything queued before it on the stream.
// sweeps: 1 forward only (L Y = B), 2 backward only (L^T X = B), 3 both.
template <typename T>
// upper_copy: L's strictly upper triangle holds L^T (launch_lower_to_upper), which the backward
// sweep then reads with the forward sweep's coalesced access.
void launch_solve_persistent(const T *L, int64_t lda, int64_t n, double *B, int64_t ldb, int64_t nrhs,
                             const double *Winv, double *Ypk, int *tickets, const int *info, int *status, hipStream_t s,
                             int sweeps = 3, bool upper_copy = false, const Batch &bt = Batch());
// Doubles of one sweep's packed copies for n rows and nrhs right-hand sides: per 128-row
// block and group of 16 right-hand sides, the finished block (128 x 16) and the residual rows
// that one half of the block hands to the other (64 x 16).
constexpr int64_t solve_pack_doubles(int64_t n, int64_t nrhs)
{
    return ((n + 127) / 128) * ((nrhs + 15) / 16) * (2048 + 1024);
}
// Symmetric-storage helpers.
void launch_upper_to_lower(const double *A, int64_t lda, int64_t n, double *W, int64_t ldw, hipStream_t s);
// A's upper triangle = W's lower triangle transposed (A may be W: the copy into its own upper
// triangle that the backward solve reads with upper_copy).
template <typename T>
void launch_lower_to_upper(const T *W, int64_t ldw, int64_t n, T *A, int64_t lda, hipStream_t s);
void launch_sy_l1norm(const double *A, int64_t lda, int64_t n, bool upper, double *colsum, double *out,
                      hipStream_t s);

// Mixed-precision solve helpers (mixed.hip).
int64_t sym_tiles64(int64_t n);  // lower 64 x 64 tiles: size of the norm partials
int residual_chunks(int64_t n, int64_t nrhs);  // k-chunks of the residual (partials: chunks*n*nrhs)
// W (lower) = lower triangle of symmetric A (lower or upper storage), anorm = ||A||_F.
template <typename T>
void launch_sym_copy_lower(const double *A, int64_t lda, int64_t n, bool upper, T *W, int64_t ldw, double *part,
                           double *anorm, hipStream_t s);
// R = B - A X (A symmetric, FP64, one triangle), rr/xx = column sums of squares of R and X,
// berr = max_j ||R_j|| / (anorm ||X_j||).
void launch_residual(const double *A, int64_t lda, int64_t n, bool upper, const double *X, int64_t ldx,
                     const double *B, int64_t ldb, int64_t nrhs, double *R, int64_t ldr, double *part, double *rr,
                     double *xx, const double *anorm, double *berr, hipStream_t s);
void launch_add_in_place(double *X, int64_t ldx, const double *D, int64_t ldd, int64_t n, int64_t nrhs,
                         hipStream_t s);

// Batched uneven factor + solve (vbatched.hip): plan (argument checks, longest-first order,
// counter reset) + one persistent launch of `grid` workgroups of 4 independent waves.
// order: batch ints; counter: 1 int; scratch: vb_scratch_elems(grid) elements.
int64_t vb_scratch_elems(int grid);
template <typename T>
void launch_vbatched(int64_t batch, const int *d_n, T *const *d_A, const int *d_lda, const int *d_nrhs,
                     T *const *d_B, const int *d_ldb, int *d_info, int *order, int *counter, T *scratch, int grid,
                     int group, int lpt, hipStream_t s);

// Level-3 pieces of the BLAS route (blas3.hip).
// L (lower, ld ldl) = the lower triangle of A, or the upper one transposed; unit: diagonal 1.
void launch_tri_stage(const double *A, int64_t lda, int64_t k, bool upper, bool unit, double *L, int64_t ldl,
                      hipStream_t s);
// D (n x m) = alpha S^T for S m x n; alpha == 0 writes zeros without reading S.
void launch_transpose_scale(const double *S, int64_t lds_, int64_t m, int64_t n, double alpha, double *D,
                            int64_t ldd, hipStream_t s);
// B (m x n) *= alpha; alpha == 0 writes zeros without reading B.
void launch_scale(double *B, int64_t ldb, int64_t m, int64_t n, double alpha, hipStream_t s);
// C = alpha op(A) op(A)^T + beta C on the `upper` or lower triangle (op(A) = A, n x k, or
// A^T, A k x n, when trans); W: n * syrk_depth_padded(k) doubles of scratch (used when trans
// or when k is not a multiple of 16).
int64_t syrk_depth_padded(int64_t k);
void launch_syrk(bool upper, bool trans, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                 double beta, double *C, int64_t ldc, double *W, hipStream_t s);
// Precision conversion of an m x n column-major block (S -> D).
template <typename S, typename D>
void launch_convert(const S *src, int64_t lds_, int64_t m, int64_t n, D *dst, int64_t ldd, hipStream_t s);
template <typename T>
void launch_upper_to_lower_t(const T *A, int64_t lda, int64_t n, T *W, int64_t ldw, hipStream_t s);

// Rank-one update / downdate of a lower factor (rank_update.hip).  scratch:
// rank_update_scratch_bytes(n) bytes; ticket, info: device ints (reset by the launch).
int64_t rank_update_scratch_bytes(int64_t n);
void launch_rank_update(double *L, int64_t ldl, int64_t n, const double *v, int64_t incv, double sigma, void *scratch,
                        int *ticket, int *info, int *status, hipStream_t s);

// Reverse mode (adjoint.hip).
// P (lower) = tril(L^T G) for lower L, G, from Lt = tril(L)^T and Gt = tril(G)^T; Wp: n*n
// doubles of scratch (the k-chunk partials).  P may be Gt's buffer.
void launch_trmm_tn_lower(const double *Lt, int64_t ldlt, const double *Gt, int64_t ldgt, int64_t n, double *P,
                          int64_t ldp, double *Wp, hipStream_t s, unsigned count = 1, int64_t sLt = 0, int64_t sGt = 0,
                          int64_t sP = 0, int64_t sWp = 0);
// mode 0: out = symmetric completion of tril(in), halved; 1: out = in^T; 2: out = (in + in^T)/2.
void launch_sym_tile(int mode, const double *in, int64_t ldi, int64_t n, double *out, int64_t ldo, hipStream_t s,
                     unsigned count = 1, int64_t sIn = 0, int64_t sOut = 0);
// B (nrow x n) <- B L^{-1} by blocked right-side substitution, one launch per 128-column
// block, with Lt = L^T; upper: only the block-upper tiles (nrow == n, the result symmetric).
void launch_trsm_xl_multi(const double *Lt, int64_t ldlt, int64_t n, double *B, int64_t ldb, int64_t nrow,
                          const double *Winv, bool upper, hipStream_t s, unsigned count = 1, int64_t sLt = 0,
                          int64_t sB = 0, int64_t sWinv = 0);
// Lbar (lower) -= tril(Bb (L^T X)^T + X (L^T Bb)^T); U, V: n x r scratch.
void launch_solve_adjoint_update(const double *L, int64_t ldl, int64_t n, int64_t r, const double *X, int64_t ldx,
                                 const double *Bb, int64_t ldbb, double *Lb, int64_t ldlb, double *U, double *V,
                                 hipStream_t s, unsigned count = 1, int64_t sL = 0, int64_t sX = 0, int64_t sBb = 0,
                                 int64_t sLb = 0, int64_t sUV = 0);

}  // namespace mi
