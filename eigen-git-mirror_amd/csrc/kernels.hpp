// kernels.hpp -- host-side launchers of the LLT kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

// Block size of the blocked factorization: the diagonal block factored in one workgroup,
// the depth of each trailing update, and the granularity of the stored inverses.
constexpr int kNB = 128;

// Diagonal block at column k0: POTRF in place + X = L11^{-1} into Winv[k0/kNB].
void launch_potrf_diag(double *A, int64_t lda, int64_t n, int64_t k0, double *Winv, int *info,
                       hipStream_t s);
// X_b = L_bb^{-1} for every diagonal block of an already factored matrix.
void launch_trtri_blocks(const double *L, int64_t lda, int64_t n, double *Winv, hipStream_t s);
// Panel: A21 <- A21 * X^T, rows k0+kNB..n-1, columns k0..k0+kNB-1.
void launch_trsm_panel(double *A, int64_t lda, int64_t n, int64_t k0, const double *Winv, const int *info,
                       hipStream_t s);
// Trailing update: A22 (lower) -= A21 * A21^T.
void launch_syrk_update(double *A, int64_t lda, int64_t n, int64_t k0, const int *info, hipStream_t s);
// Solve: forward (trans=false) or backward (trans=true) block substitution, in place on B.
void launch_solve(const double *L, int64_t lda, int64_t n, double *B, int64_t ldb, int64_t nrhs,
                  const double *Winv, const int *info, hipStream_t s);
// Symmetric-storage helpers.
void launch_upper_to_lower(const double *A, int64_t lda, int64_t n, double *W, int64_t ldw, hipStream_t s);
void launch_lower_to_upper(const double *W, int64_t ldw, int64_t n, double *A, int64_t lda, hipStream_t s);
void launch_sy_l1norm(const double *A, int64_t lda, int64_t n, bool upper, double *colsum, double *out,
                      hipStream_t s);

}  // namespace mi
