// lapacke_shim.cpp -- host-pointer LAPACKE / LAPACK entry points over the device C ABI
// (declarations and reference citations in include/mi_lapacke.h).
//
// Row-major storage is handled by the identity  row-major(uplo) == column-major(flipped
// uplo)  for the factor; for the solve a row-major B is transposed on the host.
#include <stdio.h>
#include <string.h>

#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/mi_lapacke.h"
#include "../../include/mi_llt.h"

namespace {

constexpr int kRowMajor = 101;
constexpr int kColMajor = 102;
constexpr int kDeviceError = -1010;

struct ThreadState {
    mi_llt_handle h = nullptr;
    double *buf = nullptr;
    size_t bytes = 0;
    ~ThreadState()
    {
        if (buf) (void)hipFree(buf);
        if (h) mi_llt_destroy(h);
    }
};

thread_local ThreadState tls;

bool ensure(size_t bytes)
{
    if (!tls.h && mi_llt_create(&tls.h) != 0) return false;
    if (bytes <= tls.bytes) return true;
    if (tls.buf) (void)hipFree(tls.buf);
    tls.buf = nullptr;
    tls.bytes = 0;
    if (hipMalloc(&tls.buf, bytes) != hipSuccess) return false;
    tls.bytes = bytes;
    return true;
}

int report(const char *fn, int rc)
{
    fprintf(stderr, "%s: device failure (%d): %s\n", fn, rc, tls.h ? mi_llt_last_error(tls.h) : "no handle");
    return kDeviceError;
}

bool uplo_ok(char u) { return u == 'L' || u == 'l' || u == 'U' || u == 'u'; }

char flip(char u) { return (u == 'L' || u == 'l') ? 'U' : 'L'; }

// Column-major factor of the n x n matrix at host pointer a (ld lda).  Returns info >= 0
// or kDeviceError.
int potrf_colmajor(char uplo, int n, double *a, int lda, const char *fn)
{
    if (n == 0) return 0;
    const size_t bytes = (size_t)n * n * sizeof(double);
    if (!ensure(bytes + 256)) return report(fn, MI_LLT_E_ALLOC);
    double *d = tls.buf;
    if (hipMemcpy2D(d, (size_t)n * 8, a, (size_t)lda * 8, (size_t)n * 8, n, hipMemcpyHostToDevice) != hipSuccess)
        return report(fn, MI_LLT_E_HIP);
    int info = 0;
    int rc = mi_dpotrf(tls.h, uplo, n, d, n, &info);
    if (rc) return report(fn, rc);
    if (hipMemcpy2D(a, (size_t)lda * 8, d, (size_t)n * 8, (size_t)n * 8, n, hipMemcpyDeviceToHost) != hipSuccess)
        return report(fn, MI_LLT_E_HIP);
    return info;
}

// Column-major solve with the factor at a; b is n x nrhs (ld ldb).
int potrs_colmajor(char uplo, int n, int nrhs, const double *a, int lda, double *b, int ldb, const char *fn)
{
    if (n == 0 || nrhs == 0) return 0;
    const size_t abytes = (size_t)n * n * sizeof(double);
    const size_t bbytes = (size_t)n * nrhs * sizeof(double);
    if (!ensure(abytes + bbytes + 256)) return report(fn, MI_LLT_E_ALLOC);
    double *da = tls.buf, *db = tls.buf + (size_t)n * n;
    if (hipMemcpy2D(da, (size_t)n * 8, a, (size_t)lda * 8, (size_t)n * 8, n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy2D(db, (size_t)n * 8, b, (size_t)ldb * 8, (size_t)n * 8, nrhs, hipMemcpyHostToDevice) != hipSuccess)
        return report(fn, MI_LLT_E_HIP);
    int rc = mi_dpotrs(tls.h, uplo, n, nrhs, da, n, db, n);
    if (rc) return report(fn, rc);
    if (hipMemcpy2D(b, (size_t)ldb * 8, db, (size_t)n * 8, (size_t)n * 8, nrhs, hipMemcpyDeviceToHost) != hipSuccess)
        return report(fn, MI_LLT_E_HIP);
    return 0;
}

}  // namespace

extern "C" {

int LAPACKE_dpotrf(int matrix_layout, char uplo, int n, double *a, int lda)
{
    if (matrix_layout != kRowMajor && matrix_layout != kColMajor) return -1;
    if (!uplo_ok(uplo)) return -2;
    if (n < 0) return -3;
    if (a == nullptr && n > 0) return -4;
    if (lda < (n > 1 ? n : 1)) return -5;
    // row-major lower (upper) storage is column-major upper (lower) storage of the same bytes
    const char u = matrix_layout == kColMajor ? uplo : flip(uplo);
    return potrf_colmajor(u, n, a, lda, "LAPACKE_dpotrf");
}

int LAPACKE_dpotrs(int matrix_layout, char uplo, int n, int nrhs, const double *a, int lda, double *b, int ldb)
{
    if (matrix_layout != kRowMajor && matrix_layout != kColMajor) return -1;
    if (!uplo_ok(uplo)) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (lda < (n > 1 ? n : 1)) return -6;
    if (matrix_layout == kColMajor) {
        if (ldb < (n > 1 ? n : 1)) return -8;
        return potrs_colmajor(uplo, n, nrhs, a, lda, b, ldb, "LAPACKE_dpotrs");
    }
    if (ldb < (nrhs > 1 ? nrhs : 1)) return -8;
    // row-major B (n x nrhs, ld ldb >= nrhs): transpose to column-major on the host
    std::vector<double> bt((size_t)n * nrhs);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < nrhs; ++c) bt[i + (size_t)c * n] = b[(size_t)i * ldb + c];
    int rc = potrs_colmajor(flip(uplo), n, nrhs, a, lda, bt.data(), n, "LAPACKE_dpotrs");
    if (rc) return rc;
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < nrhs; ++c) b[(size_t)i * ldb + c] = bt[i + (size_t)c * n];
    return 0;
}

// lapack/cholesky.cpp:14-36: argument errors give info = -i (and xerbla); a failing pivot
// at column k gives info = k+1.
void dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info)
{
    *info = 0;
    if (!uplo_ok(*uplo))
        *info = -1;
    else if (*n < 0)
        *info = -2;
    else if (*lda < (*n > 1 ? *n : 1))
        *info = -4;
    if (*info != 0) {
        fprintf(stderr, " ** On entry to DPOTRF parameter number %d had an illegal value\n", -*info);
        return;
    }
    int r = potrf_colmajor(*uplo, *n, a, *lda, "dpotrf_");
    *info = r;
}

// lapack/cholesky.cpp:41-72
void dpotrs_(const char *uplo, const int *n, const int *nrhs, const double *a, const int *lda, double *b,
             const int *ldb, int *info)
{
    *info = 0;
    if (!uplo_ok(*uplo))
        *info = -1;
    else if (*n < 0)
        *info = -2;
    else if (*nrhs < 0)
        *info = -3;
    else if (*lda < (*n > 1 ? *n : 1))
        *info = -5;
    else if (*ldb < (*n > 1 ? *n : 1))
        *info = -7;
    if (*info != 0) {
        fprintf(stderr, " ** On entry to DPOTRS parameter number %d had an illegal value\n", -*info);
        return;
    }
    *info = potrs_colmajor(*uplo, *n, *nrhs, a, *lda, b, *ldb, "dpotrs_");
}

}  // extern "C"
