// kernels.hpp -- host-side launchers of the LLT kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

// Block size of the blocked factorization: the diagonal block factored in one workgroup,
// the depth of each trailing update, and the granularity of the stored inverses.
constexpr int kNB = 128;

// Workspace of one factorization, per 128-block b:
//   Wleaf[b]: the 8 inverses of the 16x16 diagonal leaves of L_bb (8 x 256 doubles)
//   Winv[b] : L_bb^{-T}, 128 x 128, ld 128 (built for the solve)
constexpr int64_t kLeafDoubles = 8 * 256;
constexpr int64_t kInvDoubles = 2 * 128 * 128;  // X_b^T then X_b, each 128 x 128, ld 128

// Device status bits (a word in the handle's persistent device memory).
constexpr int MI_SOLVE_TIMEOUT = 1;
constexpr int MI_PANEL_TIMEOUT = 2;

// One fused block step (update of block column k0 by the previous panel, diagonal factor,
// panel solve), ceil((n-k0)/128) workgroups; tickets/flag: persistent words of the handle,
// ticket_base = number of tickets drawn before this launch, epoch = this step's flag value.
void launch_panel_fused(double *A, int64_t lda, int64_t n, int64_t k0, double *Wleaf, int *info, uint32_t *tickets,
                        uint32_t ticket_base, uint32_t *flag, uint32_t epoch, int *status, hipStream_t s);
// One factorization step: panel of block column kp (or none, kp < 0) + trailing update by
// the panel at column ks of the columns from ks+256 on (or none, ks < 0).  Writes the grid
// size (= tickets drawn) to *grid_out.
void launch_factor_step(double *A, int64_t lda, int64_t n, int64_t kp, int64_t ks, double *Wleaf, int *info,
                        uint32_t *tickets, uint32_t ticket_base, uint32_t *flag, uint32_t epoch, int *status,
                        hipStream_t s, uint32_t *grid_out);
// Diagonal block at column k0: POTRF in place (lower) + its leaf inverses.
void launch_potrf_diag(double *A, int64_t lda, int64_t n, int64_t k0, double *Wleaf, int *info, hipStream_t s);
// Panel: A21 <- A21 * L11^{-T}, rows k0+128..n-1, columns k0..k0+127.
void launch_trsm_panel(double *A, int64_t lda, int64_t n, int64_t k0, const double *Wleaf, const int *info,
                       hipStream_t s);
// Winv[b] = {X_b^T, X_b}, X_b = L_bb^{-1}, for every diagonal block (needs Wleaf).
void launch_block_inverses(const double *L, int64_t lda, int64_t n, const double *Wleaf, double *Winv,
                           const int *info, hipStream_t s);
// Wleaf from an already factored matrix (standalone solve).
void launch_leaf_inverses(const double *L, int64_t lda, int64_t n, double *Wleaf, hipStream_t s);
// C (m x m lower) -= P P^T with P m x K; and its restriction to the first ncol columns.
void launch_syrk(double *C, int64_t ldc, const double *P, int64_t ldp, int64_t m, int64_t K, const int *info,
                 hipStream_t s);
void launch_syrk_first_col(double *C, int64_t ldc, const double *P, int64_t ldp, int64_t m, int64_t ncol,
                           int64_t K, const int *info, hipStream_t s);
// Trailing update: A22 (lower) -= A21 * A21^T.
void launch_syrk_update(double *A, int64_t lda, int64_t n, int64_t k0, const int *info, hipStream_t s);
// Solve: forward then backward sweep in place on B (one persistent launch each).
// flags: ceil(nrhs/16) * nblk words that persist across calls; tickets: 2 ints.
void launch_solve_persistent(const double *L, int64_t lda, int64_t n, double *B, int64_t ldb, int64_t nrhs,
                             const double *Winv, uint32_t *flags, int *tickets, uint32_t epoch_fwd,
                             uint32_t epoch_bwd, const int *info, int *status, hipStream_t s);
// Symmetric-storage helpers.
void launch_upper_to_lower(const double *A, int64_t lda, int64_t n, double *W, int64_t ldw, hipStream_t s);
void launch_lower_to_upper(const double *W, int64_t ldw, int64_t n, double *A, int64_t lda, hipStream_t s);
void launch_sy_l1norm(const double *A, int64_t lda, int64_t n, bool upper, double *colsum, double *out,
                      hipStream_t s);

}  // namespace mi
