// api.hip -- the C ABI of include/mi_llt.h: handle, workspace, argument checks and the
// orchestration of the blocked factorization and solve on the handle's stream.
//
// Factor (mi_dpotrf*): right-looking blocked Cholesky with block kNB, the order of
// llt_inplace<double,Lower>::blocked (Eigen/src/Cholesky/LLT.h:338-368):
//   for k0 in 0, kNB, ...:  POTRF+TRTRI of A11  ->  A21 <- A21 X11^T  ->  A22 -= A21 A21^T
// Every kernel first reads the device `info`; after a failing pivot the remaining kernels
// exit at once, so the stream never needs a host round trip to stop.
// Upper storage (LLT.h:377-399: the upper factor is the lower factor of the transpose) is
// transposed into a lower-triangle workspace, factored there and transposed back.
#include <stdio.h>
#include <string.h>

#include <hip/hip_runtime.h>

#include "../../include/mi_llt.h"
#include "kernels.hpp"

struct mi_llt_context {
    hipStream_t stream = nullptr;
    int device = 0;
    void *ws = nullptr;       // grow-only device workspace
    size_t ws_bytes = 0;
    int *d_info = nullptr;    // device int for the synchronous entry points
    char err[256] = {0};
    uint32_t magic = 0x4c4c5431u;
};

namespace {

__global__ void set_int_kernel(int *p, int v) { *p = v; }

int fail_hip(mi_llt_handle h, hipError_t e, const char *what)
{
    snprintf(h->err, sizeof h->err, "%s: %s", what, hipGetErrorString(e));
    return MI_LLT_E_HIP;
}

bool valid(mi_llt_handle h) { return h != nullptr && h->magic == 0x4c4c5431u; }

int uplo_code(char uplo)
{
    if (uplo == 'L' || uplo == 'l') return 0;
    if (uplo == 'U' || uplo == 'u') return 1;
    return -1;
}

size_t round_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Ensure the workspace holds `bytes`; returns 0 or an error code.
int reserve(mi_llt_handle h, size_t bytes)
{
    if (bytes <= h->ws_bytes) return 0;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(h, e, "hipStreamSynchronize");
    if (h->ws) hipFree(h->ws);
    h->ws = nullptr;
    h->ws_bytes = 0;
    e = hipMalloc(&h->ws, bytes);
    if (e != hipSuccess) {
        h->ws = nullptr;
        snprintf(h->err, sizeof h->err, "workspace of %zu bytes: %s", bytes, hipGetErrorString(e));
        return MI_LLT_E_ALLOC;
    }
    h->ws_bytes = bytes;
    return 0;
}

int check_launch(mi_llt_handle h, const char *what)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail_hip(h, e, what);
}

size_t inv_bytes(int64_t n)
{
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    return round_up((size_t)nblk * mi::kNB * mi::kNB * sizeof(double));
}

// Workspace layout: [ Winv | lower copy for uplo='U' (n*n) ]
struct Layout {
    double *winv;
    double *tri;
};

int prepare(mi_llt_handle h, int64_t n, bool upper, Layout &lay)
{
    size_t need = inv_bytes(n);
    size_t tri_off = need;
    if (upper) need += round_up((size_t)n * n * sizeof(double));
    int rc = reserve(h, need);
    if (rc) return rc;
    lay.winv = (double *)h->ws;
    lay.tri = upper ? (double *)((char *)h->ws + tri_off) : nullptr;
    return 0;
}

void factor_lower(double *A, int64_t lda, int64_t n, double *winv, int *d_info, hipStream_t s)
{
    for (int64_t k0 = 0; k0 < n; k0 += mi::kNB) {
        mi::launch_potrf_diag(A, lda, n, k0, winv, d_info, s);
        mi::launch_trsm_panel(A, lda, n, k0, winv, d_info, s);
        mi::launch_syrk_update(A, lda, n, k0, d_info, s);
    }
}

}  // namespace

extern "C" {

int mi_llt_create(mi_llt_handle *h)
{
    if (!h) return -1;
    mi_llt_context *c = new (std::nothrow) mi_llt_context();
    if (!c) return MI_LLT_E_ALLOC;
    hipGetDevice(&c->device);
    if (hipMalloc(&c->d_info, 256) != hipSuccess) {
        delete c;
        return MI_LLT_E_ALLOC;
    }
    *h = c;
    return 0;
}

int mi_llt_destroy(mi_llt_handle h)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    hipStreamSynchronize(h->stream);
    if (h->ws) hipFree(h->ws);
    if (h->d_info) hipFree(h->d_info);
    h->magic = 0;
    delete h;
    return 0;
}

int mi_llt_set_stream(mi_llt_handle h, void *stream)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    h->stream = (hipStream_t)stream;
    return 0;
}

const char *mi_llt_last_error(mi_llt_handle h) { return valid(h) ? h->err : "invalid handle"; }

int64_t mi_llt_workspace_bytes(mi_llt_handle h) { return valid(h) ? (int64_t)h->ws_bytes : -1; }

int mi_dpotrf_async(mi_llt_handle h, char uplo, int64_t n, double *dA, int64_t lda, int *d_info)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (n < 0) return -3;
    if (dA == nullptr && n > 0) return -4;
    if (lda < (n > 1 ? n : 1)) return -5;
    if (d_info == nullptr) return -6;
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, d_info, 0);
    if (n == 0) return check_launch(h, "set_info");
    Layout lay;
    int rc = prepare(h, n, up == 1, lay);
    if (rc) return rc;
    if (up == 0) {
        factor_lower(dA, lda, n, lay.winv, d_info, s);
    } else {
        mi::launch_upper_to_lower(dA, lda, n, lay.tri, n, s);
        factor_lower(lay.tri, n, n, lay.winv, d_info, s);
        mi::launch_lower_to_upper(lay.tri, n, n, dA, lda, s);
    }
    return check_launch(h, "factor");
}

int mi_dpotrf(mi_llt_handle h, char uplo, int64_t n, double *dA, int64_t lda, int *info)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    if (info == nullptr) return -6;
    int rc = mi_dpotrf_async(h, uplo, n, dA, lda, h->d_info);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? 0 : fail_hip(h, e, "info readback");
}

int mi_dpotrs(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, const double *dA, int64_t lda, double *dB,
              int64_t ldb)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (dA == nullptr && n > 0) return -5;
    if (lda < (n > 1 ? n : 1)) return -6;
    if (dB == nullptr && n > 0 && nrhs > 0) return -7;
    if (ldb < (n > 1 ? n : 1)) return -8;
    if (n == 0 || nrhs == 0) return 0;
    Layout lay;
    int rc = prepare(h, n, up == 1, lay);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const double *L = dA;
    int64_t ldl = lda;
    if (up == 1) {
        mi::launch_upper_to_lower(dA, lda, n, lay.tri, n, s);
        L = lay.tri;
        ldl = n;
    }
    mi::launch_trtri_blocks(L, ldl, n, lay.winv, s);
    mi::launch_solve(L, ldl, n, dB, ldb, nrhs, lay.winv, nullptr, s);
    return check_launch(h, "solve");
}

int mi_dposv_async(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, double *dA, int64_t lda, double *dB,
                   int64_t ldb, int *d_info)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (dA == nullptr && n > 0) return -5;
    if (lda < (n > 1 ? n : 1)) return -6;
    if (dB == nullptr && n > 0 && nrhs > 0) return -7;
    if (ldb < (n > 1 ? n : 1)) return -8;
    if (d_info == nullptr) return -9;
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, d_info, 0);
    if (n == 0) return check_launch(h, "set_info");
    Layout lay;
    int rc = prepare(h, n, up == 1, lay);
    if (rc) return rc;
    double *L = dA;
    int64_t ldl = lda;
    if (up == 1) {
        mi::launch_upper_to_lower(dA, lda, n, lay.tri, n, s);
        L = lay.tri;
        ldl = n;
    }
    factor_lower(L, ldl, n, lay.winv, d_info, s);
    if (nrhs > 0) mi::launch_solve(L, ldl, n, dB, ldb, nrhs, lay.winv, d_info, s);
    if (up == 1) mi::launch_lower_to_upper(lay.tri, n, n, dA, lda, s);
    return check_launch(h, "posv");
}

int mi_dsy_l1norm(mi_llt_handle h, char uplo, int64_t n, const double *dA, int64_t lda, double *d_norm)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (n < 0) return -3;
    if (dA == nullptr && n > 0) return -4;
    if (lda < (n > 1 ? n : 1)) return -5;
    if (d_norm == nullptr) return -6;
    if (n == 0) {
        hipError_t e = hipMemsetAsync(d_norm, 0, sizeof(double), h->stream);
        return e == hipSuccess ? 0 : fail_hip(h, e, "l1norm");
    }
    int rc = reserve(h, round_up((size_t)n * sizeof(double)) > h->ws_bytes ? round_up((size_t)n * sizeof(double))
                                                                            : h->ws_bytes);
    if (rc) return rc;
    mi::launch_sy_l1norm(dA, lda, n, up == 1, (double *)h->ws, d_norm, h->stream);
    return check_launch(h, "l1norm");
}

}  // extern "C"
