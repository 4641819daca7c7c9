// lapacke_shim.cpp -- host-pointer LAPACKE / LAPACK entry points over the device C ABI
// (declarations and reference citations in include/mi_lapacke.h).
//
// Row-major storage is handled by the identity  row-major(uplo) == column-major(flipped
// uplo)  for the factor; for the solve a row-major B is transposed on the host.
#include <stdio.h>
#include <string.h>

#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/mi_blas.h"
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

// Device entry points by precision.
int dev_potrf(char uplo, int64_t n, double *d, int64_t ld, int *info) { return mi_dpotrf(tls.h, uplo, n, d, ld, info); }
int dev_potrf(char uplo, int64_t n, float *d, int64_t ld, int *info) { return mi_spotrf(tls.h, uplo, n, d, ld, info); }
int dev_potrs(char uplo, int64_t n, int64_t r, const double *a, int64_t lda, double *b, int64_t ldb)
{
    return mi_dpotrs(tls.h, uplo, n, r, a, lda, b, ldb);
}
int dev_potrs(char uplo, int64_t n, int64_t r, const float *a, int64_t lda, float *b, int64_t ldb)
{
    return mi_spotrs(tls.h, uplo, n, r, a, lda, b, ldb);
}

// Host (ld ldh) <-> device (ld rows) copies of a rows x cols column-major block.
template <typename T>
bool to_device(T *d, const T *hst, int rows, int cols, int ldh)
{
    if (rows == 0 || cols == 0) return true;
    return hipMemcpy2D(d, (size_t)rows * sizeof(T), hst, (size_t)ldh * sizeof(T), (size_t)rows * sizeof(T), cols,
                       hipMemcpyHostToDevice) == hipSuccess;
}
template <typename T>
bool to_host(T *hst, int ldh, const T *d, int rows, int cols)
{
    if (rows == 0 || cols == 0) return true;
    return hipMemcpy2D(hst, (size_t)ldh * sizeof(T), d, (size_t)rows * sizeof(T), (size_t)rows * sizeof(T), cols,
                       hipMemcpyDeviceToHost) == hipSuccess;
}

// The asynchronous work of the thread's handle reported a device failure (status word).
int device_status(const char *fn)
{
    int st = 0, rc = mi_llt_status(tls.h, &st);
    if (rc) return report(fn, rc);
    if (st) {
        fprintf(stderr, "%s: device status %d\n", fn, st);
        return kDeviceError;
    }
    return 0;
}

// Column-major factor of the n x n matrix at host pointer a (ld lda).  Returns info >= 0
// or kDeviceError.
template <typename T>
int potrf_colmajor(char uplo, int n, T *a, int lda, const char *fn)
{
    if (n == 0) return 0;
    const size_t bytes = (size_t)n * n * sizeof(T);
    if (!ensure(bytes + 256)) return report(fn, MI_LLT_E_ALLOC);
    T *d = (T *)tls.buf;
    if (!to_device(d, a, n, n, lda)) return report(fn, MI_LLT_E_HIP);
    int info = 0;
    int rc = dev_potrf(uplo, n, d, n, &info);
    if (rc) return report(fn, rc);
    if (!to_host(a, lda, d, n, n)) return report(fn, MI_LLT_E_HIP);
    return info;
}

// Column-major solve with the factor at a; b is n x nrhs (ld ldb).
template <typename T>
int potrs_colmajor(char uplo, int n, int nrhs, const T *a, int lda, T *b, int ldb, const char *fn)
{
    if (n == 0 || nrhs == 0) return 0;
    const size_t abytes = (size_t)n * n * sizeof(T);
    const size_t bbytes = (size_t)n * nrhs * sizeof(T);
    if (!ensure(abytes + bbytes + 256)) return report(fn, MI_LLT_E_ALLOC);
    T *da = (T *)tls.buf, *db = (T *)tls.buf + (size_t)n * n;
    if (!to_device(da, a, n, n, lda) || !to_device(db, b, n, nrhs, ldb)) return report(fn, MI_LLT_E_HIP);
    int rc = dev_potrs(uplo, n, nrhs, da, n, db, n);
    if (rc) return report(fn, rc);
    if ((rc = device_status(fn)) != 0) return rc;  // the solve is asynchronous
    if (!to_host(b, ldb, db, n, nrhs)) return report(fn, MI_LLT_E_HIP);
    return 0;
}

template <typename T>
int lapacke_potrf(int matrix_layout, char uplo, int n, T *a, int lda, const char *fn)
{
    if (matrix_layout != kRowMajor && matrix_layout != kColMajor) return -1;
    if (!uplo_ok(uplo)) return -2;
    if (n < 0) return -3;
    if (a == nullptr && n > 0) return -4;
    if (lda < (n > 1 ? n : 1)) return -5;
    // row-major lower (upper) storage is column-major upper (lower) storage of the same bytes
    const char u = matrix_layout == kColMajor ? uplo : flip(uplo);
    return potrf_colmajor(u, n, a, lda, fn);
}

template <typename T>
int lapacke_potrs(int matrix_layout, char uplo, int n, int nrhs, const T *a, int lda, T *b, int ldb, const char *fn)
{
    if (matrix_layout != kRowMajor && matrix_layout != kColMajor) return -1;
    if (!uplo_ok(uplo)) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (lda < (n > 1 ? n : 1)) return -6;
    if (matrix_layout == kColMajor) {
        if (ldb < (n > 1 ? n : 1)) return -8;
        return potrs_colmajor(uplo, n, nrhs, a, lda, b, ldb, fn);
    }
    if (ldb < (nrhs > 1 ? nrhs : 1)) return -8;
    // row-major B (n x nrhs, ld ldb >= nrhs): transpose to column-major on the host
    std::vector<T> bt((size_t)n * nrhs);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < nrhs; ++c) bt[i + (size_t)c * n] = b[(size_t)i * ldb + c];
    int rc = potrs_colmajor(flip(uplo), n, nrhs, a, lda, bt.data(), n, fn);
    if (rc) return rc;
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < nrhs; ++c) b[(size_t)i * ldb + c] = bt[i + (size_t)c * n];
    return 0;
}

// lapack/cholesky.cpp:14-36: argument errors give info = -i (and xerbla); a failing pivot
// at column k gives info = k+1.
template <typename T>
void fortran_potrf(const char *uplo, const int *n, T *a, const int *lda, int *info, const char *name)
{
    *info = 0;
    if (!uplo_ok(*uplo))
        *info = -1;
    else if (*n < 0)
        *info = -2;
    else if (*lda < (*n > 1 ? *n : 1))
        *info = -4;
    if (*info != 0) {
        fprintf(stderr, " ** On entry to %s parameter number %d had an illegal value\n", name, -*info);
        return;
    }
    *info = potrf_colmajor(*uplo, *n, a, *lda, name);
}

// lapack/cholesky.cpp:41-72
template <typename T>
void fortran_potrs(const char *uplo, const int *n, const int *nrhs, const T *a, const int *lda, T *b, const int *ldb,
                   int *info, const char *name)
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
        fprintf(stderr, " ** On entry to %s parameter number %d had an illegal value\n", name, -*info);
        return;
    }
    *info = potrs_colmajor(*uplo, *n, *nrhs, a, *lda, b, *ldb, name);
}

// The BLAS argument checks of blas/level3_impl.h report the illegal argument's position and
// return 0 (Eigen's xerbla_ prints and returns, blas/xerbla.cpp:15-19).
int blas_arg_error(const char *name, int pos)
{
    fprintf(stderr, " ** On entry to %s parameter number %d had an illegal value\n", name, pos);
    return 0;
}

bool op_ok(char c) { return c == 'N' || c == 'n' || c == 'T' || c == 't' || c == 'C' || c == 'c'; }
bool side_ok(char c) { return c == 'L' || c == 'l' || c == 'R' || c == 'r'; }
bool diag_ok(char c) { return c == 'N' || c == 'n' || c == 'U' || c == 'u'; }
bool is_notrans(char c) { return c == 'N' || c == 'n'; }

}  // namespace

extern "C" {

int LAPACKE_dpotrf(int matrix_layout, char uplo, int n, double *a, int lda)
{
    return lapacke_potrf(matrix_layout, uplo, n, a, lda, "LAPACKE_dpotrf");
}

int LAPACKE_spotrf(int matrix_layout, char uplo, int n, float *a, int lda)
{
    return lapacke_potrf(matrix_layout, uplo, n, a, lda, "LAPACKE_spotrf");
}

int LAPACKE_dpotrs(int matrix_layout, char uplo, int n, int nrhs, const double *a, int lda, double *b, int ldb)
{
    return lapacke_potrs(matrix_layout, uplo, n, nrhs, a, lda, b, ldb, "LAPACKE_dpotrs");
}

int LAPACKE_spotrs(int matrix_layout, char uplo, int n, int nrhs, const float *a, int lda, float *b, int ldb)
{
    return lapacke_potrs(matrix_layout, uplo, n, nrhs, a, lda, b, ldb, "LAPACKE_spotrs");
}

void dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info)
{
    fortran_potrf(uplo, n, a, lda, info, "DPOTRF");
}

void spotrf_(const char *uplo, const int *n, float *a, const int *lda, int *info)
{
    fortran_potrf(uplo, n, a, lda, info, "SPOTRF");
}

void dpotrs_(const char *uplo, const int *n, const int *nrhs, const double *a, const int *lda, double *b,
             const int *ldb, int *info)
{
    fortran_potrs(uplo, n, nrhs, a, lda, b, ldb, info, "DPOTRS");
}

void spotrs_(const char *uplo, const int *n, const int *nrhs, const float *a, const int *lda, float *b, const int *ldb,
             int *info)
{
    fortran_potrs(uplo, n, nrhs, a, lda, b, ldb, info, "SPOTRS");
}

// blas/level3_impl.h:78-178 (argument positions 1..11 as there); B is staged, A's triangle too.
int dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m, const int *n,
           const double *alpha, const double *a, const int *lda, double *b, const int *ldb)
{
    const bool left = *side == 'L' || *side == 'l';
    if (!side_ok(*side)) return blas_arg_error("DTRSM ", 1);
    if (!uplo_ok(*uplo)) return blas_arg_error("DTRSM ", 2);
    if (!op_ok(*transa)) return blas_arg_error("DTRSM ", 3);
    if (!diag_ok(*diag)) return blas_arg_error("DTRSM ", 4);
    if (*m < 0) return blas_arg_error("DTRSM ", 5);
    if (*n < 0) return blas_arg_error("DTRSM ", 6);
    const int k = left ? *m : *n;
    if (*lda < (k > 1 ? k : 1)) return blas_arg_error("DTRSM ", 9);
    if (*ldb < (*m > 1 ? *m : 1)) return blas_arg_error("DTRSM ", 11);
    if (*m == 0 || *n == 0) return 0;
    const size_t abytes = (size_t)k * k * sizeof(double), bbytes = (size_t)*m * *n * sizeof(double);
    if (!ensure(abytes + bbytes + 256)) return report("dtrsm_", MI_LLT_E_ALLOC), 0;
    double *da = tls.buf, *db = tls.buf + (size_t)k * k;
    // alpha == 0: A is not referenced (B becomes zero)
    if ((*alpha != 0.0 && !to_device(da, a, k, k, *lda)) || !to_device(db, b, *m, *n, *ldb))
        return report("dtrsm_", MI_LLT_E_HIP), 0;
    int rc = mi_dtrsm(tls.h, *side, *uplo, *transa, *diag, *m, *n, *alpha, da, k, db, *m);
    if (rc) return report("dtrsm_", rc), 0;
    if (device_status("dtrsm_") != 0) return 0;  // B left as it was on entry
    if (!to_host(b, *ldb, db, *m, *n)) report("dtrsm_", MI_LLT_E_HIP);
    return 0;
}

// blas/level3_impl.h:357-420 (argument positions 1..10 as there); C's triangle is staged with
// the whole n x n block (the other triangle is copied back unchanged).
int dsyrk_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha, const double *a,
           const int *lda, const double *beta, double *c, const int *ldc)
{
    if (!uplo_ok(*uplo)) return blas_arg_error("DSYRK ", 1);
    if (!op_ok(*trans)) return blas_arg_error("DSYRK ", 2);
    if (*n < 0) return blas_arg_error("DSYRK ", 3);
    if (*k < 0) return blas_arg_error("DSYRK ", 4);
    const int arows = is_notrans(*trans) ? *n : *k, acols = is_notrans(*trans) ? *k : *n;
    if (*lda < (arows > 1 ? arows : 1)) return blas_arg_error("DSYRK ", 7);
    if (*ldc < (*n > 1 ? *n : 1)) return blas_arg_error("DSYRK ", 10);
    if (*n == 0) return 0;
    const size_t abytes = (size_t)arows * acols * sizeof(double), cbytes = (size_t)*n * *n * sizeof(double);
    if (!ensure(abytes + cbytes + 256)) return report("dsyrk_", MI_LLT_E_ALLOC), 0;
    double *da = tls.buf, *dc = tls.buf + (size_t)arows * acols;
    // alpha == 0 or k == 0: A is not referenced
    const bool use_a = *alpha != 0.0 && *k > 0;
    if ((use_a && !to_device(da, a, arows, acols, *lda)) || !to_device(dc, c, *n, *n, *ldc))
        return report("dsyrk_", MI_LLT_E_HIP), 0;
    int rc = mi_dsyrk(tls.h, *uplo, *trans, *n, *k, *alpha, da, arows > 1 ? arows : 1, *beta, dc, *n);
    if (rc) return report("dsyrk_", rc), 0;
    if (device_status("dsyrk_") != 0) return 0;
    if (!to_host(c, *ldc, dc, *n, *n)) report("dsyrk_", MI_LLT_E_HIP);
    return 0;
}

}  // extern "C"
