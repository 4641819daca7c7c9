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

#include <math.h>

#include <limits>
#include <new>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/mi_llt.h"
#include "factor_plan.hpp"
#include "kernels.hpp"

struct mi_llt_context {
    hipStream_t stream = nullptr;       // caller's stream: all work is ordered on it
    int device = 0;
    void *ws = nullptr;       // grow-only device workspace
    size_t ws_bytes = 0;
    int *d_info = nullptr;    // device int for the synchronous entry points
    // synchronisation words (status, solve and rank-update tickets, factor tickets, panel flag
    // at the mi::kSync* offsets, each group on its own cache line); zeroed on the handle's stream when allocated; the ticket and flag words of a solve or a
    // factorization are reset on the stream at its start, so every call (and every replay of a
    // captured call) counts from zero
    uint32_t *sync = nullptr;
    size_t sync_words = 0;
    double *ypk = nullptr;       // the solve's packed copies of finished blocks (grow-only)
    size_t ypk_bytes = 0;
    char err[256] = {0};
    uint32_t magic = 0x4c4c5431u;
    // optional per-kernel timing (mi_llt_trace_*): event pairs around each launch
    bool trace = false;
    struct Rec {
        int kind;
        double flops;
        hipEvent_t e0, e1;
    };
    std::vector<Rec> recs;
    double t_ms[MI_TRACE_KINDS] = {0};
    // batched path: its own grow-only workspace (stable across calls of the same or smaller
    // size, so a captured graph keeps valid pointers) and the schedule parameters
    void *vb_ws = nullptr;
    size_t vb_bytes = 0;
    int vb_wg_per_cu = 2, vb_group = 0, vb_lpt = 1;  // group 0: by precision (FP64 2, FP32 4)
    int cus = 0;
    double t_flops[MI_TRACE_KINDS] = {0};
    int64_t t_calls[MI_TRACE_KINDS] = {0};
};

namespace {

__global__ void set_int_kernel(int *p, int v) { *p = v; }
__global__ void set_double_kernel(double *p, double v) { *p = v; }
__global__ void or_int_kernel(int *p, int v) { atomicOr(p, v); }
__global__ void zero_words_kernel(uint32_t *p, int count)
{
    for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0u;
}

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

// Run `launch` (which enqueues on stream s), bracketed by events when tracing is on.
template <class F>
void traced(mi_llt_handle h, hipStream_t s, int kind, double flops, F &&launch)
{
    if (!h->trace) {
        launch();
        return;
    }
    mi_llt_context::Rec r{kind, flops, nullptr, nullptr};
    if (hipEventCreate(&r.e0) != hipSuccess || hipEventCreate(&r.e1) != hipSuccess) {
        if (r.e0) (void)hipEventDestroy(r.e0);
        launch();  // untimed: tracing is best effort
        return;
    }
    const bool rec = hipEventRecord(r.e0, s) == hipSuccess;
    launch();
    if (rec && hipEventRecord(r.e1, s) == hipSuccess) {
        h->recs.push_back(r);
    } else {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
}

size_t round_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Ensure the workspace holds `bytes`; returns 0 or an error code.
int reserve(mi_llt_handle h, size_t bytes)
{
    if (bytes <= h->ws_bytes) return 0;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(h, e, "hipStreamSynchronize");
    if (h->ws) (void)hipFree(h->ws);
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

// Workspace layout: [ Wleaf | Winv | lower copy for uplo='U' (n*n) ]
struct Layout {
    double *wleaf;
    double *winv;
    double *tri;
};

int prepare(mi_llt_handle h, int64_t n, bool upper, Layout &lay)
{
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    const size_t leaf = round_up((size_t)nblk * mi::kPanelDoubles * sizeof(double));
    const size_t inv = round_up((size_t)nblk * mi::kInvDoubles * sizeof(double));
    size_t need = leaf + inv;
    if (upper) need += round_up((size_t)n * n * sizeof(double));
    int rc = reserve(h, need);
    if (rc) return rc;
    lay.wleaf = (double *)h->ws;
    lay.winv = (double *)((char *)h->ws + leaf);
    lay.tri = upper ? (double *)((char *)h->ws + leaf + inv) : nullptr;
    return 0;
}

constexpr size_t kSyncHeader = mi::kSyncWords;

int reserve_sync(mi_llt_handle h, size_t flag_words)
{
    const size_t need = kSyncHeader + flag_words;
    if (need <= h->sync_words) return 0;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(h, e, "hipStreamSynchronize");
    if (h->sync) (void)hipFree(h->sync);
    h->sync = nullptr;
    h->sync_words = 0;
    const size_t words = need < 4096 ? 4096 : need;
    if (hipMalloc(&h->sync, words * sizeof(uint32_t)) != hipSuccess) {
        h->sync = nullptr;
        return MI_LLT_E_ALLOC;
    }
    // zeroed on the handle's own stream, so that the kernels queued after it on that stream see
    // the zeros (a memset on the null stream is not ordered with a non-blocking stream)
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(256), 0, h->stream, h->sync, (int)words);
    e = hipGetLastError();
    if (e != hipSuccess) return fail_hip(h, e, "zero sync words");
    h->sync_words = words;
    return 0;
}

template <typename T>
int factor_lower(mi_llt_handle h, T *A, int64_t lda, int64_t n, double *wleaf, int *d_info, int prec = MI_PREC_FP64,
                 __bf16 *wsplit = nullptr)
{
    int rc = reserve_sync(h, 0);
    if (rc) return rc;
    hipStream_t s = h->stream;
    mi::enqueue_factorization<T>(s, A, lda, n, wleaf, d_info, h->sync, (int *)h->sync + mi::kSyncStatus, prec, wsplit,
                                 [&](double fl, auto &&launch) { traced(h, s, MI_TRACE_POTRF, fl, launch); });
    return 0;
}

// Diagonal-block inverses of a factor (needs its leaf inverses), for solve_only.
template <typename T>
void block_inverses(mi_llt_handle h, const T *L, int64_t ldl, int64_t n, const double *wleaf, double *winv,
                    const int *d_info, bool couplings = true)
{
    traced(h, h->stream, MI_TRACE_TRTRI, (double)n * mi::kNB * mi::kNB / 3.0,
           [&] { mi::launch_block_inverses<T>(L, ldl, n, wleaf, winv, d_info, h->stream, couplings); });
}

// Forward + backward substitution in place on B with a factor whose block inverses are in winv.
template <typename T>
int solve_only(mi_llt_handle h, const T *L, int64_t ldl, int64_t n, double *B, int64_t ldb, int64_t nrhs,
               const double *winv, const int *d_info, int sweeps = 3, bool upper_copy = false)
{
    int rc = reserve_sync(h, 0);
    if (rc) return rc;
    const size_t pk = (size_t)mi::solve_pack_doubles(n, nrhs) * sizeof(double);  // both sweeps, in turn
    if (pk > h->ypk_bytes) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return fail_hip(h, e, "hipStreamSynchronize");
        if (h->ypk) (void)hipFree(h->ypk);
        h->ypk = nullptr;
        h->ypk_bytes = 0;
        e = hipMalloc(&h->ypk, pk);
        if (e != hipSuccess) {
            h->ypk = nullptr;
            snprintf(h->err, sizeof h->err, "solve scratch of %zu bytes: %s", pk, hipGetErrorString(e));
            return MI_LLT_E_ALLOC;
        }
        h->ypk_bytes = pk;
    }
    const double per_sweep = (double)n * (double)n * (double)nrhs;
    traced(h, h->stream, MI_TRACE_SOLVE, per_sweep * ((sweeps & 1) + ((sweeps >> 1) & 1)), [&] {
        mi::launch_solve_persistent<T>(L, ldl, n, B, ldb, nrhs, winv, h->ypk, (int *)h->sync + mi::kSyncSolveTickets, d_info,
                                       (int *)h->sync + mi::kSyncStatus,
                                       h->stream, sweeps, upper_copy);
    });
    return 0;
}

template <typename T>
int solve_lower(mi_llt_handle h, const T *L, int64_t ldl, int64_t n, double *B, int64_t ldb, int64_t nrhs,
                const double *wleaf, double *winv, const int *d_info)
{
    block_inverses<T>(h, L, ldl, n, wleaf, winv, d_info);
    return solve_only<T>(h, L, ldl, n, B, ldb, nrhs, winv, d_info);
}

// Workspace of mi_dposv_mixed: [ W (FP32, or FP64 for the plain path) | Wleaf | Winv | R | D |
// residual partials | norm partials | scalars: anorm, berr, rr[nrhs], xx[nrhs] | split-panel
// ring (FP32 path) ]
struct MixedLayout {
    float *w32 = nullptr;
    double *w64 = nullptr;
    int64_t ldw = 0;
    double *wleaf = nullptr, *winv = nullptr, *R = nullptr, *D = nullptr, *part = nullptr, *tpart = nullptr;
    double *anorm = nullptr, *berr = nullptr, *rr = nullptr, *xx = nullptr;
    __bf16 *split = nullptr;
};

int prepare_mixed(mi_llt_handle h, int64_t n, int64_t nrhs, bool fp64, MixedLayout &m)
{
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    const size_t wbytes = round_up((size_t)n * (size_t)n * (fp64 ? sizeof(double) : sizeof(float)));
    const size_t leaf = round_up((size_t)nblk * mi::kPanelDoubles * sizeof(double));
    const size_t inv = round_up((size_t)nblk * mi::kInvDoubles * sizeof(double));
    const size_t vec = round_up((size_t)n * (size_t)nrhs * sizeof(double));
    const size_t part = round_up((size_t)mi::residual_chunks(n, nrhs) * (size_t)n * (size_t)nrhs * sizeof(double));
    const size_t tpart = round_up((size_t)mi::sym_tiles64(n) * sizeof(double));
    const size_t scal = round_up((size_t)(2 + 2 * nrhs) * sizeof(double));
    const size_t split = fp64 ? 0 : round_up((size_t)(mi::kSplitRing * mi::split_slot_elems(n, 3)) * sizeof(__bf16));
    int rc = reserve(h, wbytes + leaf + inv + 2 * vec + part + tpart + scal + split);
    if (rc) return rc;
    char *p = (char *)h->ws;
    m = MixedLayout();
    if (fp64)
        m.w64 = (double *)p;
    else
        m.w32 = (float *)p;
    m.ldw = n;
    p += wbytes;
    m.wleaf = (double *)p;
    p += leaf;
    m.winv = (double *)p;
    p += inv;
    m.R = (double *)p;
    p += vec;
    m.D = (double *)p;
    p += vec;
    m.part = (double *)p;
    p += part;
    m.tpart = (double *)p;
    p += tpart;
    m.anorm = (double *)p;
    m.berr = m.anorm + 1;
    m.rr = m.anorm + 2;
    m.xx = m.rr + nrhs;
    p += scal;
    if (!fp64) m.split = (__bf16 *)p;
    return 0;
}

template <typename T>
int vbatched(mi_llt_handle h, char uplo, int64_t batch, const int *d_n, T *const *d_A, const int *d_lda,
             const int *d_nrhs, T *const *d_B, const int *d_ldb, int *d_info)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (batch < 0) return -3;
    if (batch == 0) return 0;
    if (!d_n) return -4;
    if (!d_A) return -5;
    if (!d_lda) return -6;
    if (!d_nrhs) return -7;
    if (!d_B) return -8;
    if (!d_ldb) return -9;
    if (!d_info) return -10;
    int64_t grid = (int64_t)h->cus * h->vb_wg_per_cu;
    const int64_t need_wg = (batch + 3) / 4;  // one system per worker wave at least
    if (grid > need_wg) grid = need_wg;
    if (grid < 1) grid = 1;
    const size_t order_b = round_up((size_t)batch * sizeof(int));
    const size_t ctr_b = 256;
    const size_t scr_b = round_up((size_t)mi::vb_scratch_elems((int)grid) * sizeof(T));
    const size_t need = order_b + ctr_b + scr_b;
    if (need > h->vb_bytes) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return fail_hip(h, e, "hipStreamSynchronize");
        if (h->vb_ws) (void)hipFree(h->vb_ws);
        h->vb_ws = nullptr;
        h->vb_bytes = 0;
        if (hipMalloc(&h->vb_ws, need) != hipSuccess) {
            h->vb_ws = nullptr;
            snprintf(h->err, sizeof h->err, "batched workspace of %zu bytes", need);
            return MI_LLT_E_ALLOC;
        }
        h->vb_bytes = need;
    }
    char *p = (char *)h->vb_ws;
    int *order = (int *)p;
    int *counter = (int *)(p + order_b);
    T *scratch = (T *)(p + order_b + ctr_b);
    const double fl_est = 0.0;
    traced(h, h->stream, MI_TRACE_BATCH, fl_est, [&] {
        mi::launch_vbatched<T>(batch, d_n, d_A, d_lda, d_nrhs, d_B, d_ldb, d_info, order, counter, scratch, (int)grid,
                               h->vb_group ? h->vb_group : (sizeof(T) == 8 ? 2 : 4), h->vb_lpt, h->stream);
    });
    return check_launch(h, "vbatched");
}

// Waits for the stream, then reads and clears the status word (0 when none is allocated yet).
int read_status(mi_llt_handle h, int *status)
{
    *status = 0;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(h, e, "hipStreamSynchronize");
    if (!h->sync) return 0;
    int v = 0;
    e = hipMemcpyAsync(&v, h->sync + mi::kSyncStatus, sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(h, e, "status readback");
    if (v != 0) {  // cleared on the stream: ordered before any later work of the handle
        hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, h->stream, (int *)h->sync + mi::kSyncStatus, 0);
        e = hipGetLastError();
        if (e != hipSuccess) return fail_hip(h, e, "status clear");
    }
    *status = v;
    return 0;
}

// MI_LLT_E_DEVICE with the status bits spelled out in the handle's error text.
int device_failure(mi_llt_handle h, int bits)
{
    snprintf(h->err, sizeof h->err, "device status %d:%s%s%s", bits,
             (bits & MI_LLT_STATUS_SOLVE_TIMEOUT) ? " solve sweep timed out;" : "",
             (bits & MI_LLT_STATUS_PANEL_TIMEOUT) ? " panel solve timed out waiting for its diagonal block;" : "",
             (bits & MI_LLT_STATUS_FACTOR_TICKET) ? " factorization ticket outside its launch;" : "");
    return MI_LLT_E_DEVICE;
}

// Waits for the stream and reports a device failure recorded since the last check.
int check_device(mi_llt_handle h)
{
    int st = 0;
    int rc = read_status(h, &st);
    if (rc) return rc;
    return st ? device_failure(h, st) : 0;
}

// Synchronous copy to the host of a result on the stream; fails on a device failure recorded
// in the status word (the result is then invalid).
int read_back(mi_llt_handle h, void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(h, e, "readback");
    return check_device(h);
}

}  // namespace

extern "C" {

int mi_llt_create(mi_llt_handle *h)
{
    if (!h) return -1;
    mi_llt_context *c = new (std::nothrow) mi_llt_context();
    if (!c) return MI_LLT_E_ALLOC;
    if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || c->cus <= 0)
        c->cus = 256;
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
    (void)hipStreamSynchronize(h->stream);  // destroying: nothing to report an error to
    if (h->sync) (void)hipFree(h->sync);
    if (h->ypk) (void)hipFree(h->ypk);
    if (h->vb_ws) (void)hipFree(h->vb_ws);
    if (h->ws) (void)hipFree(h->ws);
    if (h->d_info) (void)hipFree(h->d_info);
    h->magic = 0;
    delete h;
    return 0;
}

int mi_llt_set_stream(mi_llt_handle h, void *stream)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    hipStream_t next = (hipStream_t)stream;
    if (next != h->stream) {
        // the workspace and counters serve one stream's calls in order: let the old stream's
        // work finish before another stream's work can reuse them.  Not while either stream is
        // being captured into a graph: nothing captured runs yet, synchronising is not allowed
        // then, and the capturing caller orders the streams itself.
        hipStreamCaptureStatus cs_old = hipStreamCaptureStatusNone, cs_new = hipStreamCaptureStatusNone;
        const bool q_old = hipStreamIsCapturing(h->stream, &cs_old) == hipSuccess;
        const bool q_new = hipStreamIsCapturing(next, &cs_new) == hipSuccess;
        if (q_old && q_new && cs_old == hipStreamCaptureStatusNone && cs_new == hipStreamCaptureStatusNone) {
            hipError_t e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) return fail_hip(h, e, "hipStreamSynchronize (previous stream)");
        }
        h->stream = next;
    }
    return 0;
}

const char *mi_llt_last_error(mi_llt_handle h) { return valid(h) ? h->err : "invalid handle"; }

int64_t mi_llt_workspace_bytes(mi_llt_handle h)
{
    if (!valid(h)) return -1;
    return (int64_t)(h->ws_bytes + h->ypk_bytes + h->vb_bytes + h->sync_words * sizeof(uint32_t));
}

int mi_llt_status(mi_llt_handle h, int *status)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    if (!status) return -2;
    return read_status(h, status);
}

int mi_llt_inject_status(mi_llt_handle h, int bits)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    int rc = reserve_sync(h, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(or_int_kernel, dim3(1), dim3(1), 0, h->stream, (int *)h->sync + mi::kSyncStatus, bits);
    return check_launch(h, "inject status");
}

int mi_llt_trace(mi_llt_handle h, int enable)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    h->trace = enable != 0;
    return 0;
}

int mi_llt_trace_read(mi_llt_handle h, int kind, double *ms, double *flops, int64_t *launches)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    if (kind < 0 || kind >= MI_TRACE_KINDS) return -2;
    if (!h->recs.empty()) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return fail_hip(h, e, "trace sync");
        for (auto &r : h->recs) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, r.e0, r.e1) == hipSuccess) {
                h->t_ms[r.kind] += t;
                h->t_flops[r.kind] += r.flops;
                h->t_calls[r.kind] += 1;
            }
            (void)hipEventDestroy(r.e0);
            (void)hipEventDestroy(r.e1);
        }
        h->recs.clear();
    }
    if (ms) *ms = h->t_ms[kind];
    if (flops) *flops = h->t_flops[kind];
    if (launches) *launches = h->t_calls[kind];
    return 0;
}

int mi_llt_trace_reset(mi_llt_handle h)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    double ms;
    mi_llt_trace_read(h, 0, &ms, nullptr, nullptr);
    for (int k = 0; k < MI_TRACE_KINDS; ++k) h->t_ms[k] = h->t_flops[k] = 0.0, h->t_calls[k] = 0;
    return 0;
}

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
        rc = factor_lower(h, dA, lda, n, lay.wleaf, d_info);
        if (rc) return rc;
    } else {
        mi::launch_upper_to_lower(dA, lda, n, lay.tri, n, s);
        rc = factor_lower(h, lay.tri, n, n, lay.wleaf, d_info);
        if (rc) return rc;
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
    return read_back(h, info, h->d_info, sizeof(int));
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
    mi::launch_leaf_inverses<double>(L, ldl, n, lay.wleaf, s);
    rc = solve_lower(h, L, ldl, n, dB, ldb, nrhs, lay.wleaf, lay.winv, nullptr);
    if (rc) return rc;
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
    rc = factor_lower(h, L, ldl, n, lay.wleaf, d_info);
    if (rc) return rc;
    if (nrhs > 0) {
        rc = solve_lower(h, L, ldl, n, dB, ldb, nrhs, lay.wleaf, lay.winv, d_info);
        if (rc) return rc;
    }
    if (up == 1) mi::launch_lower_to_upper(lay.tri, n, n, dA, lda, s);
    return check_launch(h, "posv");
}

int mi_spotrf_async(mi_llt_handle h, char uplo, int64_t n, float *dA, int64_t lda, int *d_info)
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
    Layout lay;  // the 'U' workspace (n * n doubles) holds the n * n floats of the lower copy
    int rc = prepare(h, n, up == 1, lay);
    if (rc) return rc;
    float *L = dA;
    int64_t ldl = lda;
    if (up == 1) {
        L = (float *)lay.tri;
        ldl = n;
        mi::launch_upper_to_lower_t<float>(dA, lda, n, L, ldl, s);
    }
    rc = factor_lower<float>(h, L, ldl, n, lay.wleaf, d_info, MI_PREC_FP32);
    if (rc) return rc;
    if (up == 1) mi::launch_lower_to_upper<float>(L, ldl, n, dA, lda, s);
    return check_launch(h, "spotrf");
}

int mi_spotrf(mi_llt_handle h, char uplo, int64_t n, float *dA, int64_t lda, int *info)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    if (info == nullptr) return -6;
    int rc = mi_spotrf_async(h, uplo, n, dA, lda, h->d_info);
    if (rc) return rc;
    return read_back(h, info, h->d_info, sizeof(int));
}

int mi_spotrs(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, const float *dA, int64_t lda, float *dB,
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
    // [ Wleaf | Winv | lower copy (n * n floats, uplo 'U') | B in FP64 (n * nrhs) ]
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    const size_t leaf = round_up((size_t)nblk * mi::kPanelDoubles * sizeof(double));
    const size_t inv = round_up((size_t)nblk * mi::kInvDoubles * sizeof(double));
    const size_t tri = up == 1 ? round_up((size_t)n * n * sizeof(float)) : 0;
    const size_t bw = round_up((size_t)n * nrhs * sizeof(double));
    int rc = reserve(h, leaf + inv + tri + bw);
    if (rc) return rc;
    char *p = (char *)h->ws;
    double *wleaf = (double *)p, *winv = (double *)(p + leaf);
    float *Lc = (float *)(p + leaf + inv);
    double *Bd = (double *)(p + leaf + inv + tri);
    hipStream_t s = h->stream;
    const float *L = dA;
    int64_t ldl = lda;
    if (up == 1) {
        mi::launch_upper_to_lower_t<float>(dA, lda, n, Lc, n, s);
        L = Lc;
        ldl = n;
    }
    mi::launch_convert<float, double>(dB, ldb, n, nrhs, Bd, n, s);
    mi::launch_leaf_inverses<float>(L, ldl, n, wleaf, s);
    rc = solve_lower<float>(h, L, ldl, n, Bd, n, nrhs, wleaf, winv, nullptr);
    if (rc) return rc;
    mi::launch_convert<double, float>(Bd, n, n, nrhs, dB, ldb, s);
    return check_launch(h, "spotrs");
}

int mi_dtrsm(mi_llt_handle h, char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha,
             const double *dA, int64_t lda, double *dB, int64_t ldb)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const bool left = side == 'L' || side == 'l';
    if (!left && side != 'R' && side != 'r') return -2;
    const int up = uplo_code(uplo);
    if (up < 0) return -3;
    const bool tr = transa == 'T' || transa == 't' || transa == 'C' || transa == 'c';
    if (!tr && transa != 'N' && transa != 'n') return -4;
    const bool unit = diag == 'U' || diag == 'u';
    if (!unit && diag != 'N' && diag != 'n') return -5;
    if (m < 0) return -6;
    if (n < 0) return -7;
    const int64_t k = left ? m : n;  // order of A
    if (dA == nullptr && k > 0 && alpha != 0.0) return -9;
    if (lda < (k > 1 ? k : 1)) return -10;
    if (dB == nullptr && m > 0 && n > 0) return -11;
    if (ldb < (m > 1 ? m : 1)) return -12;
    if (m == 0 || n == 0) return 0;
    hipStream_t s = h->stream;
    if (alpha == 0.0) {  // BLAS: A is not referenced, B is set to zero
        mi::launch_scale(dB, ldb, m, n, 0.0, s);
        return check_launch(h, "trsm");
    }
    // Every case is one sweep of the LLT solve on a lower operand L: forward (L Y = C) when the
    // operator applied to C is lower triangular, backward (L^T Y = C) otherwise.  Left side:
    // C = B, operator op(A); right side: C = B^T, operator op(A)^T.
    const bool upper = up == 1;
    const bool forward = left ? (upper == tr) : (upper != tr);
    const bool stage = upper || unit;  // the solve reads a lower, non-unit triangle in place
    const int64_t nrhs = left ? n : m;
    const int64_t nblk = (k + mi::kNB - 1) / mi::kNB;
    const size_t leaf = round_up((size_t)nblk * mi::kPanelDoubles * sizeof(double));
    const size_t inv = round_up((size_t)nblk * mi::kInvDoubles * sizeof(double));
    const size_t tri = stage ? round_up((size_t)k * k * sizeof(double)) : 0;
    const size_t ct = left ? 0 : round_up((size_t)k * nrhs * sizeof(double));
    int rc = reserve(h, leaf + inv + tri + ct);
    if (rc) return rc;
    char *p = (char *)h->ws;
    double *wleaf = (double *)p, *winv = (double *)(p + leaf);
    double *Lw = (double *)(p + leaf + inv), *Ct = (double *)(p + leaf + inv + tri);
    const double *L = dA;
    int64_t ldl = lda;
    if (stage) {
        mi::launch_tri_stage(dA, lda, k, upper, unit, Lw, k, s);
        L = Lw;
        ldl = k;
    }
    double *C = dB;
    int64_t ldc = ldb;
    if (left) {
        if (alpha != 1.0) mi::launch_scale(dB, ldb, m, n, alpha, s);
    } else {
        mi::launch_transpose_scale(dB, ldb, m, n, alpha, Ct, k, s);  // C = alpha B^T (n x m)
        C = Ct;
        ldc = k;
    }
    mi::launch_leaf_inverses<double>(L, ldl, k, wleaf, s);
    block_inverses<double>(h, L, ldl, k, wleaf, winv, nullptr);
    if ((rc = solve_only<double>(h, L, ldl, k, C, ldc, nrhs, winv, nullptr, forward ? 1 : 2))) return rc;
    if (!left) mi::launch_transpose_scale(Ct, k, k, m, 1.0, dB, ldb, s);  // B = C^T
    return check_launch(h, "trsm");
}

int mi_dsyrk(mi_llt_handle h, char uplo, char trans, int64_t n, int64_t k, double alpha, const double *dA,
             int64_t lda, double beta, double *dC, int64_t ldc)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    const bool tr = trans == 'T' || trans == 't' || trans == 'C' || trans == 'c';
    if (!tr && trans != 'N' && trans != 'n') return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    const int64_t arows = tr ? k : n;
    if (dA == nullptr && n > 0 && k > 0 && alpha != 0.0) return -7;
    if (lda < (arows > 1 ? arows : 1)) return -8;
    if (dC == nullptr && n > 0) return -10;
    if (ldc < (n > 1 ? n : 1)) return -11;
    if (n == 0) return 0;
    double *W = nullptr;
    const int64_t kp = mi::syrk_depth_padded(k);
    if (alpha != 0.0 && k > 0 && (tr || kp != k)) {
        int rc = reserve(h, round_up((size_t)n * kp * sizeof(double)));
        if (rc) return rc;
        W = (double *)h->ws;
    }
    traced(h, h->stream, MI_TRACE_SYRK, (double)n * (double)(n + 1) * (double)k,
           [&] { mi::launch_syrk(up == 1, tr, n, k, alpha, dA, lda, beta, dC, ldc, W, h->stream); });
    return check_launch(h, "syrk");
}

int mi_dllt_rank_update(mi_llt_handle h, char uplo, int64_t n, double *dL, int64_t ldl, const double *dv,
                        int64_t incv, double sigma, int *d_info)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (n < 0) return -3;
    if (dL == nullptr && n > 0) return -4;
    if (ldl < (n > 1 ? n : 1)) return -5;
    if (dv == nullptr && n > 0) return -6;
    if (incv < 1) return -7;
    if (!(sigma == sigma)) return -8;
    if (d_info == nullptr) return -9;
    hipStream_t s = h->stream;
    if (n == 0) {
        hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, d_info, 0);
        return check_launch(h, "rank_update");
    }
    int rc = reserve_sync(h, 0);
    if (rc) return rc;
    // [ coefficients, flags | lower copy for uplo 'U' ]
    const size_t scr = round_up((size_t)mi::rank_update_scratch_bytes(n));
    const size_t tri = up == 1 ? round_up((size_t)n * n * sizeof(double)) : 0;
    if ((rc = reserve(h, scr + tri))) return rc;
    double *L = dL;
    int64_t ld = ldl;
    if (up == 1) {  // the upper factor is the lower factor of the transpose (LLT.h:377-399)
        L = (double *)((char *)h->ws + scr);
        ld = n;
        mi::launch_upper_to_lower(dL, ldl, n, L, ld, s);
    }
    mi::launch_rank_update(L, ld, n, dv, incv, sigma, h->ws, (int *)h->sync + mi::kSyncRankTicket, d_info,
                           (int *)h->sync + mi::kSyncStatus, s);
    if (up == 1) mi::launch_lower_to_upper<double>(L, ld, n, dL, ldl, s);
    return check_launch(h, "rank_update");
}

int mi_dpocon(mi_llt_handle h, char uplo, int64_t n, const double *dL, int64_t ldl, double anorm, double *rcond)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (n < 0) return -3;
    if (dL == nullptr && n > 0) return -4;
    if (ldl < (n > 1 ? n : 1)) return -5;
    if (!(anorm >= 0.0)) return -6;
    if (rcond == nullptr) return -7;
    // rcond_estimate_helper, ConditionEstimator.h:158-169
    if (n == 0) {
        *rcond = std::numeric_limits<double>::infinity();
        return 0;
    }
    if (anorm == 0.0) {
        *rcond = 0.0;
        return 0;
    }
    if (n == 1) {
        *rcond = 1.0;
        return 0;
    }
    hipStream_t s = h->stream;
    // [ Wleaf | Winv | lower copy for uplo 'U' | one vector ]
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    const size_t leaf = round_up((size_t)nblk * mi::kPanelDoubles * sizeof(double));
    const size_t inv = round_up((size_t)nblk * mi::kInvDoubles * sizeof(double));
    const size_t tri = up == 1 ? round_up((size_t)n * n * sizeof(double)) : 0;
    int rc = reserve(h, leaf + inv + tri + round_up((size_t)n * sizeof(double)));
    if (rc) return rc;
    char *p = (char *)h->ws;
    double *wleaf = (double *)p, *winv = (double *)(p + leaf), *dv = (double *)(p + leaf + inv + tri);
    const double *L = dL;
    int64_t ld = ldl;
    if (up == 1) {
        double *Lc = (double *)(p + leaf + inv);
        mi::launch_upper_to_lower(dL, ldl, n, Lc, n, s);
        L = Lc;
        ld = n;
    }
    mi::launch_leaf_inverses<double>(L, ld, n, wleaf, s);
    block_inverses<double>(h, L, ld, n, wleaf, winv, nullptr);
    std::vector<double> v((size_t)n), sgn((size_t)n), old_sgn((size_t)n);
    // v <- A^{-1} v with the factor (the same solve serves dec.solve and dec.adjoint().solve:
    // A is symmetric, LLT.h:204)
    auto solve = [&]() -> int {
        hipError_t e = hipMemcpyAsync(dv, v.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return fail_hip(h, e, "rcond vector");
        int r = solve_only<double>(h, L, ld, n, dv, n, 1, winv, nullptr);
        if (r) return r;
        return read_back(h, v.data(), dv, (size_t)n * sizeof(double));
    };
    auto norm1 = [&]() {
        double a = 0.0;
        for (int64_t i = 0; i < n; ++i) a += fabs(v[i]);
        return a;
    };
    // rcond_invmatrix_L1_norm_estimate, ConditionEstimator.h:55-141 (real case)
    for (int64_t i = 0; i < n; ++i) v[i] = 1.0 / (double)n;
    if ((rc = solve())) return rc;
    double lower_bound = norm1();
    double old_lower_bound = lower_bound;
    int64_t old_j = -1;
    for (int k = 0; k < 4; ++k) {
        for (int64_t i = 0; i < n; ++i) sgn[i] = v[i] < 0.0 ? -1.0 : 1.0;
        if (k > 0 && sgn == old_sgn) break;  // the solution stagnated
        v = sgn;
        if ((rc = solve())) return rc;
        int64_t j = 0;  // first index of max |v_i|
        for (int64_t i = 1; i < n; ++i)
            if (fabs(v[i]) > fabs(v[j])) j = i;
        if (j == old_j) break;
        std::fill(v.begin(), v.end(), 0.0);  // v = A^{-1} e_j
        v[j] = 1.0;
        if ((rc = solve())) return rc;
        lower_bound = norm1();
        if (lower_bound <= old_lower_bound) break;  // no increase
        old_sgn = sgn;
        old_j = j;
        old_lower_bound = lower_bound;
    }
    // Higham's alternating-sign vector v_i = (-1)^i (1 + i / (n - 1))
    double alt = 1.0;
    for (int64_t i = 0; i < n; ++i) {
        v[i] = alt * (1.0 + (double)i / (double)(n - 1));
        alt = -alt;
    }
    if ((rc = solve())) return rc;
    const double alternate = (2.0 * norm1()) / (3.0 * (double)n);
    const double inv_norm = lower_bound > alternate ? lower_bound : alternate;
    *rcond = inv_norm == 0.0 ? 0.0 : (1.0 / inv_norm) / anorm;
    return 0;
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
        hipLaunchKernelGGL(set_double_kernel, dim3(1), dim3(1), 0, h->stream, d_norm, 0.0);
        return check_launch(h, "l1norm");
    }
    int rc = reserve(h, round_up((size_t)n * sizeof(double)) > h->ws_bytes ? round_up((size_t)n * sizeof(double))
                                                                            : h->ws_bytes);
    if (rc) return rc;
    mi::launch_sy_l1norm(dA, lda, n, up == 1, (double *)h->ws, d_norm, h->stream);
    return check_launch(h, "l1norm");
}

int mi_dposv_mixed(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, const double *dA, int64_t lda,
                   const double *dB, int64_t ldb, double *dX, int64_t ldx, int prec, double tol, int maxit,
                   int *iters, double *berr, int *info)
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
    if (dX == nullptr && n > 0 && nrhs > 0) return -9;
    if (ldx < (n > 1 ? n : 1)) return -10;
    if (prec < MI_PREC_FP64 || prec > MI_PREC_BF16X6) return -11;
    if (!(tol >= 0.0)) return -12;
    if (maxit < 0) return -13;
    if (iters == nullptr) return -14;
    if (berr == nullptr) return -15;
    if (info == nullptr) return -16;
    *iters = 0;
    *berr = 0.0;
    *info = 0;
    if (n == 0 || nrhs == 0) return 0;
    hipStream_t s = h->stream;
    const bool upper = up == 1;
    MixedLayout m;
    int rc;

    auto copy_b_to_x = [&]() -> int {
        hipError_t e = hipMemcpy2DAsync(dX, (size_t)ldx * sizeof(double), dB, (size_t)ldb * sizeof(double),
                                        (size_t)n * sizeof(double), (size_t)nrhs, hipMemcpyDeviceToDevice, s);
        return e == hipSuccess ? 0 : fail_hip(h, e, "hipMemcpy2DAsync");
    };
    auto residual = [&]() {
        traced(h, s, MI_TRACE_GEMM, 2.0 * (double)n * (double)n * (double)nrhs, [&] {
            mi::launch_residual(dA, lda, n, upper, dX, ldx, dB, ldb, nrhs, m.R, n, m.part, m.rr, m.xx, m.anorm,
                                m.berr, s);
        });
    };
    // the plain FP64 factorization of a copy of A; `code` goes to *iters
    auto fp64_path = [&](int code) -> int {
        int r = prepare_mixed(h, n, nrhs, true, m);
        if (r) return r;
        hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, h->d_info, 0);
        mi::launch_sym_copy_lower<double>(dA, lda, n, upper, m.w64, m.ldw, m.tpart, m.anorm, s);
        r = factor_lower<double>(h, m.w64, m.ldw, n, m.wleaf, h->d_info);
        if (r) return r;
        int inf = 0;
        if ((r = read_back(h, &inf, h->d_info, sizeof(int)))) return r;
        *iters = code;
        if (inf != 0) {
            *info = inf;
            return 0;
        }
        if ((r = copy_b_to_x())) return r;
        if ((r = solve_lower<double>(h, m.w64, m.ldw, n, dX, ldx, nrhs, m.wleaf, m.winv, h->d_info))) return r;
        residual();
        if ((r = read_back(h, berr, m.berr, sizeof(double)))) return r;
        return check_launch(h, "posv_mixed fp64");
    };
    if (prec == MI_PREC_FP64) return fp64_path(0);

    if ((rc = prepare_mixed(h, n, nrhs, false, m))) return rc;
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, h->d_info, 0);
    mi::launch_sym_copy_lower<float>(dA, lda, n, upper, m.w32, m.ldw, m.tpart, m.anorm, s);
    if ((rc = factor_lower<float>(h, m.w32, m.ldw, n, m.wleaf, h->d_info, prec, m.split))) return rc;
    int inf = 0;
    if ((rc = read_back(h, &inf, h->d_info, sizeof(int)))) return rc;
    if (inf != 0) return fp64_path(-1);  // not positive definite in the low-precision factor
    block_inverses<float>(h, m.w32, m.ldw, n, m.wleaf, m.winv, h->d_info);
    // L^T into the working copy's unused upper triangle, for the backward sweeps of every solve
    mi::launch_lower_to_upper<float>(m.w32, m.ldw, n, m.w32, m.ldw, s);
    if ((rc = copy_b_to_x())) return rc;
    if ((rc = solve_only<float>(h, m.w32, m.ldw, n, dX, ldx, nrhs, m.winv, h->d_info, 3, true))) return rc;
    for (int it = 0;; ++it) {
        residual();
        double b = 0.0;
        if ((rc = read_back(h, &b, m.berr, sizeof(double)))) return rc;
        *berr = b;
        if (b <= tol) {
            *iters = it;
            return check_launch(h, "posv_mixed");
        }
        if (it == maxit) break;
        hipError_t e = hipMemcpyAsync(m.D, m.R, (size_t)n * (size_t)nrhs * sizeof(double), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return fail_hip(h, e, "hipMemcpyAsync");
        if ((rc = solve_only<float>(h, m.w32, m.ldw, n, m.D, n, nrhs, m.winv, h->d_info, 3, true))) return rc;
        mi::launch_add_in_place(dX, ldx, m.D, n, n, nrhs, s);
    }
    return fp64_path(-2);  // refinement did not reach tol
}

int mi_dposv_vbatched(mi_llt_handle h, char uplo, int64_t batch, const int *d_n, double *const *d_A,
                      const int *d_lda, const int *d_nrhs, double *const *d_B, const int *d_ldb, int *d_info)
{
    return vbatched<double>(h, uplo, batch, d_n, d_A, d_lda, d_nrhs, d_B, d_ldb, d_info);
}

int mi_sposv_vbatched(mi_llt_handle h, char uplo, int64_t batch, const int *d_n, float *const *d_A,
                      const int *d_lda, const int *d_nrhs, float *const *d_B, const int *d_ldb, int *d_info)
{
    return vbatched<float>(h, uplo, batch, d_n, d_A, d_lda, d_nrhs, d_B, d_ldb, d_info);
}

int mi_llt_set_vbatched_config(mi_llt_handle h, int wg_per_cu, int group, int longest_first)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    if (wg_per_cu < 1 || wg_per_cu > 64) return -2;
    if (group != 0 && group != 2 && group != 4) return -3;
    h->vb_wg_per_cu = wg_per_cu;
    h->vb_group = group;
    h->vb_lpt = longest_first ? 1 : 0;
    return 0;
}

int mi_llt_get_vbatched_config(mi_llt_handle h, int *wg_per_cu, int *group, int *longest_first)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    if (wg_per_cu) *wg_per_cu = h->vb_wg_per_cu;
    if (group) *group = h->vb_group;
    if (longest_first) *longest_first = h->vb_lpt;
    return 0;
}

int mi_dpotrf_adjoint(mi_llt_handle h, char uplo, int64_t n, const double *dL, int64_t ldl, double *dG, int64_t ldg)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (n < 0) return -3;
    if (dL == nullptr && n > 0) return -4;
    if (ldl < (n > 1 ? n : 1)) return -5;
    if (dG == nullptr && n > 0) return -6;
    if (ldg < (n > 1 ? n : 1)) return -7;
    if (n == 0) return 0;
    hipStream_t s = h->stream;
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    const size_t leaf = round_up((size_t)nblk * mi::kPanelDoubles * sizeof(double));
    const size_t inv = round_up((size_t)nblk * mi::kInvDoubles * sizeof(double));
    const size_t sq = round_up((size_t)n * (size_t)n * sizeof(double));
    int rc = reserve(h, leaf + inv + 3 * sq);
    if (rc) return rc;
    double *wleaf = (double *)h->ws;
    double *winv = (double *)((char *)h->ws + leaf);
    double *W2 = (double *)((char *)h->ws + leaf + inv);
    double *W3 = (double *)((char *)h->ws + leaf + inv + sq);
    double *Wt = (double *)((char *)h->ws + leaf + inv + 2 * sq);  // tril(L)^T
    const double nd = (double)n;
    mi::launch_sym_tile(4, dL, ldl, n, Wt, n, s);  // tril(L)^T
    mi::launch_sym_tile(4, dG, ldg, n, W3, n, s);  // tril(Lbar)^T, then overwritten by P
    traced(h, s, MI_TRACE_GEMM, nd * nd * nd / 3.0, [&] { mi::launch_trmm_tn_lower(Wt, n, W3, n, n, W3, n, W2, s); });
    mi::launch_sym_tile(0, W3, n, n, W2, n, s);              // M
    mi::launch_leaf_inverses<double>(dL, ldl, n, wleaf, s);  // recomputed from L
    block_inverses<double>(h, dL, ldl, n, wleaf, winv, nullptr, false);  // no solve here
    traced(h, s, MI_TRACE_SOLVE, nd * nd * nd,
           [&] { mi::launch_trsm_xl_multi(Wt, n, n, W2, n, n, winv, false, s); });  // Y = M L^{-1}
    mi::launch_sym_tile(1, W2, n, n, W3, n, s);                                      // Y^T = L^{-T} M
    traced(h, s, MI_TRACE_SOLVE, nd * nd * nd / 2.0,
           [&] { mi::launch_trsm_xl_multi(Wt, n, n, W3, n, n, winv, true, s); });   // S = Y^T L^{-1}, block-upper
    mi::launch_sym_tile(3, W3, n, n, dG, ldg, s);                                    // Abar = sym(S)
    return check_launch(h, "potrf_adjoint");
}

int mi_dpotrs_adjoint(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, const double *dL, int64_t ldl,
                      const double *dX, int64_t ldx, double *dXbar, int64_t ldxb, double *dLbar, int64_t ldlb)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (dL == nullptr && n > 0) return -5;
    if (ldl < (n > 1 ? n : 1)) return -6;
    if (dX == nullptr && n > 0 && nrhs > 0) return -7;
    if (ldx < (n > 1 ? n : 1)) return -8;
    if (dXbar == nullptr && n > 0 && nrhs > 0) return -9;
    if (ldxb < (n > 1 ? n : 1)) return -10;
    if (dLbar == nullptr && n > 0) return -11;
    if (ldlb < (n > 1 ? n : 1)) return -12;
    if (n == 0 || nrhs == 0) return 0;
    hipStream_t s = h->stream;
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    const size_t leaf = round_up((size_t)nblk * mi::kPanelDoubles * sizeof(double));
    const size_t inv = round_up((size_t)nblk * mi::kInvDoubles * sizeof(double));
    const size_t vec = round_up((size_t)n * (size_t)nrhs * sizeof(double));
    int rc = reserve(h, leaf + inv + 2 * vec);
    if (rc) return rc;
    double *wleaf = (double *)h->ws;
    double *winv = (double *)((char *)h->ws + leaf);
    double *U = (double *)((char *)h->ws + leaf + inv);
    double *V = (double *)((char *)h->ws + leaf + inv + vec);
    mi::launch_leaf_inverses<double>(dL, ldl, n, wleaf, s);
    block_inverses<double>(h, dL, ldl, n, wleaf, winv, nullptr);
    if ((rc = solve_only<double>(h, dL, ldl, n, dXbar, ldxb, nrhs, winv, nullptr))) return rc;  // Bbar
    mi::launch_solve_adjoint_update(dL, ldl, n, nrhs, dX, ldx, dXbar, ldxb, dLbar, ldlb, U, V, s);
    return check_launch(h, "potrs_adjoint");
}

}  // extern "C"

// ---- batched, uniform order -----------------------------------------------------------------
// `batch` systems of one order n, system z at dA + z * strideA (likewise every other array): the
// same kernels as the single-system entry points, each launch covering every system
// (blockIdx.z), so that a batch of small systems fills the GPU.  Per system, the workspace holds
// its own panel images, block inverses, packed solve copies and adjoint scratch, and the sync
// words its own ticket counters and panel flag (the status word is shared).
namespace {

// An n x cols array per system at p + z * stride, leading dimension ld.
int check_batch_args(int64_t n, int64_t cols, int64_t batch, int64_t ld, int64_t stride, const void *p)
{
    if (batch < 0 || batch > 65535) return 1;
    if (n > 0 && cols > 0 && batch > 0 && p == nullptr) return 2;
    if (ld < (n > 1 ? n : 1)) return 3;
    if (batch > 1 && stride < ld * cols) return 4;  // systems must not overlap
    return 0;
}

struct BatchLayout {
    double *leaf = nullptr, *inv = nullptr, *pack = nullptr, *aux = nullptr;
    mi::Batch bt;
};

// Workspace [ leaf x batch | inv x batch | pack x batch | aux x batch ] and the per-system sync
// blocks (after the handle's own); strides in elements.
int prepare_batch(mi_llt_handle h, int64_t n, int64_t nrhs, int64_t batch, bool inv, size_t aux_doubles,
                  BatchLayout &bl)
{
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    const size_t leaf_d = round_up((size_t)nblk * mi::kPanelDoubles * sizeof(double)) / sizeof(double);
    const size_t inv_d = inv ? round_up((size_t)nblk * mi::kInvDoubles * sizeof(double)) / sizeof(double) : 0;
    const size_t pack_d = nrhs > 0 ? round_up((size_t)mi::solve_pack_doubles(n, nrhs) * sizeof(double)) / sizeof(double) : 0;
    const size_t aux_d = round_up(aux_doubles * sizeof(double)) / sizeof(double);
    const size_t per = leaf_d + inv_d + pack_d + aux_d;
    int rc = reserve(h, per * (size_t)batch * sizeof(double));
    if (rc) return rc;
    rc = reserve_sync(h, (size_t)batch * mi::kSyncWords);
    if (rc) return rc;
    double *p = (double *)h->ws;
    bl.leaf = p;
    bl.inv = p + leaf_d * batch;
    bl.pack = bl.inv + inv_d * batch;
    bl.aux = bl.pack + pack_d * batch;
    bl.bt.count = (unsigned)batch;
    bl.bt.leaf = (int64_t)leaf_d;
    bl.bt.inv = (int64_t)inv_d;
    bl.bt.pack = (int64_t)pack_d;
    bl.bt.aux = (int64_t)aux_d;
    bl.bt.sync = mi::kSyncWords;
    return 0;
}

uint32_t *batch_sync(mi_llt_handle h) { return h->sync + mi::kSyncWords; }  // system 0's block
int *shared_status(mi_llt_handle h) { return (int *)h->sync + mi::kSyncStatus; }

__global__ void set_ints_kernel(int *p, int count, int v)
{
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < count; i += (int)(gridDim.x * blockDim.x)) p[i] = v;
}

}  // namespace

extern "C" {

int mi_dpotrf_batched_async(mi_llt_handle h, char uplo, int64_t n, int64_t batch, double *dA, int64_t lda,
                            int64_t strideA, int *d_info)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (n < 0) return -3;
    if (batch < 0 || batch > 65535) return -4;
    if (int e = check_batch_args(n, n, batch, lda, strideA, dA)) return e == 2 ? -5 : (e == 3 ? -6 : -7);
    if (d_info == nullptr && batch > 0) return -8;
    if (batch == 0) return 0;
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(set_ints_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s, d_info, (int)batch, 0);
    if (n == 0) return check_launch(h, "set_info");
    BatchLayout bl;
    int rc = prepare_batch(h, n, 0, batch, false, 0, bl);
    if (rc) return rc;
    bl.bt.a = strideA;
    mi::enqueue_factorization<double>(
        s, dA, lda, n, bl.leaf, d_info, batch_sync(h), shared_status(h), MI_PREC_FP64, nullptr,
        [&](double fl, auto &&launch) { traced(h, s, MI_TRACE_POTRF, fl * (double)batch, launch); }, bl.bt);
    return check_launch(h, "potrf_batched");
}

int mi_dpotrs_batched(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, int64_t batch, const double *dL,
                      int64_t ldl, int64_t strideL, double *dB, int64_t ldb, int64_t strideB)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (batch < 0 || batch > 65535) return -5;
    if (int e = check_batch_args(n, n, batch, ldl, strideL, dL)) return e == 2 ? -6 : (e == 3 ? -7 : -8);
    if (nrhs > 0) {
        if (int e = check_batch_args(n, nrhs, batch, ldb, strideB, dB)) return e == 2 ? -9 : (e == 3 ? -10 : -11);
    }
    if (n == 0 || nrhs == 0 || batch == 0) return 0;
    BatchLayout bl;
    int rc = prepare_batch(h, n, nrhs, batch, true, 0, bl);
    if (rc) return rc;
    hipStream_t s = h->stream;
    bl.bt.a = strideL;
    bl.bt.b = strideB;
    mi::launch_leaf_inverses<double>(dL, ldl, n, bl.leaf, s, bl.bt);
    traced(h, s, MI_TRACE_TRTRI, (double)n * mi::kNB * mi::kNB / 3.0 * (double)batch,
           [&] { mi::launch_block_inverses<double>(dL, ldl, n, bl.leaf, bl.inv, nullptr, s, true, bl.bt); });
    traced(h, s, MI_TRACE_SOLVE, 2.0 * (double)n * (double)n * (double)nrhs * (double)batch, [&] {
        mi::launch_solve_persistent<double>(dL, ldl, n, dB, ldb, nrhs, bl.inv, bl.pack,
                                            (int *)batch_sync(h) + mi::kSyncSolveTickets, nullptr, shared_status(h), s,
                                            3, false, bl.bt);
    });
    return check_launch(h, "potrs_batched");
}

int mi_dpotrf_adjoint_batched(mi_llt_handle h, char uplo, int64_t n, int64_t batch, const double *dL, int64_t ldl,
                              int64_t strideL, double *dG, int64_t ldg, int64_t strideG)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (n < 0) return -3;
    if (batch < 0 || batch > 65535) return -4;
    if (int e = check_batch_args(n, n, batch, ldl, strideL, dL)) return e == 2 ? -5 : (e == 3 ? -6 : -7);
    if (int e = check_batch_args(n, n, batch, ldg, strideG, dG)) return e == 2 ? -8 : (e == 3 ? -9 : -10);
    if (n == 0 || batch == 0) return 0;
    const size_t sq = (size_t)n * (size_t)n;
    BatchLayout bl;
    int rc = prepare_batch(h, n, 0, batch, true, 3 * sq, bl);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const unsigned nb = (unsigned)batch;
    const int64_t sa = bl.bt.aux;  // per-system block of 3 n^2: W2 | W3 | Wt
    double *W2 = bl.aux, *W3 = bl.aux + sq, *Wt = bl.aux + 2 * sq;
    const double nd = (double)n, bd = (double)batch;
    bl.bt.a = strideL;
    mi::launch_sym_tile(4, dL, ldl, n, Wt, n, s, nb, strideL, sa);  // tril(L)^T
    mi::launch_sym_tile(4, dG, ldg, n, W3, n, s, nb, strideG, sa);  // tril(Lbar)^T, then P
    traced(h, s, MI_TRACE_GEMM, nd * nd * nd / 3.0 * bd,
           [&] { mi::launch_trmm_tn_lower(Wt, n, W3, n, n, W3, n, W2, s, nb, sa, sa, sa, sa); });
    mi::launch_sym_tile(0, W3, n, n, W2, n, s, nb, sa, sa);  // M
    mi::launch_leaf_inverses<double>(dL, ldl, n, bl.leaf, s, bl.bt);
    traced(h, s, MI_TRACE_TRTRI, nd * mi::kNB * mi::kNB / 3.0 * bd,
           [&] { mi::launch_block_inverses<double>(dL, ldl, n, bl.leaf, bl.inv, nullptr, s, false, bl.bt); });
    traced(h, s, MI_TRACE_SOLVE, nd * nd * nd * bd,
           [&] { mi::launch_trsm_xl_multi(Wt, n, n, W2, n, n, bl.inv, false, s, nb, sa, sa, bl.bt.inv); });
    mi::launch_sym_tile(1, W2, n, n, W3, n, s, nb, sa, sa);  // Y^T
    traced(h, s, MI_TRACE_SOLVE, nd * nd * nd / 2.0 * bd,
           [&] { mi::launch_trsm_xl_multi(Wt, n, n, W3, n, n, bl.inv, true, s, nb, sa, sa, bl.bt.inv); });
    mi::launch_sym_tile(3, W3, n, n, dG, ldg, s, nb, sa, strideG);  // Abar = sym(S)
    return check_launch(h, "potrf_adjoint_batched");
}

int mi_dpotrs_adjoint_batched(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, int64_t batch, const double *dL,
                              int64_t ldl, int64_t strideL, const double *dX, int64_t ldx, int64_t strideX,
                              double *dXbar, int64_t ldxb, int64_t strideXbar, double *dLbar, int64_t ldlb,
                              int64_t strideLbar)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (batch < 0 || batch > 65535) return -5;
    if (int e = check_batch_args(n, n, batch, ldl, strideL, dL)) return e == 2 ? -6 : (e == 3 ? -7 : -8);
    if (nrhs > 0) {
        if (int e = check_batch_args(n, nrhs, batch, ldx, strideX, dX)) return e == 2 ? -9 : (e == 3 ? -10 : -11);
        if (int e = check_batch_args(n, nrhs, batch, ldxb, strideXbar, dXbar)) return e == 2 ? -12 : (e == 3 ? -13 : -14);
    }
    if (int e = check_batch_args(n, n, batch, ldlb, strideLbar, dLbar)) return e == 2 ? -15 : (e == 3 ? -16 : -17);
    if (n == 0 || nrhs == 0 || batch == 0) return 0;
    const size_t vec = (size_t)n * (size_t)nrhs;
    BatchLayout bl;
    int rc = prepare_batch(h, n, nrhs, batch, true, 2 * vec, bl);
    if (rc) return rc;
    hipStream_t s = h->stream;
    bl.bt.a = strideL;
    bl.bt.b = strideXbar;
    mi::launch_leaf_inverses<double>(dL, ldl, n, bl.leaf, s, bl.bt);
    mi::launch_block_inverses<double>(dL, ldl, n, bl.leaf, bl.inv, nullptr, s, true, bl.bt);
    traced(h, s, MI_TRACE_SOLVE, 2.0 * (double)n * (double)n * (double)nrhs * (double)batch, [&] {  // Bbar
        mi::launch_solve_persistent<double>(dL, ldl, n, dXbar, ldxb, nrhs, bl.inv, bl.pack,
                                            (int *)batch_sync(h) + mi::kSyncSolveTickets, nullptr, shared_status(h), s,
                                            3, false, bl.bt);
    });
    double *U = bl.aux, *V = bl.aux + vec;
    mi::launch_solve_adjoint_update(dL, ldl, n, nrhs, dX, ldx, dXbar, ldxb, dLbar, ldlb, U, V, s, (unsigned)batch,
                                    strideL, strideX, strideXbar, strideLbar, bl.bt.aux);
    return check_launch(h, "potrs_adjoint_batched");
}

}  // extern "C"

extern "C" {

int mi_dposv_batched_async(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, int64_t batch, double *dA,
                           int64_t lda, int64_t strideA, double *dB, int64_t ldb, int64_t strideB, int *d_info)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (batch < 0 || batch > 65535) return -5;
    if (int e = check_batch_args(n, n, batch, lda, strideA, dA)) return e == 2 ? -6 : (e == 3 ? -7 : -8);
    if (int e = check_batch_args(n, nrhs, batch, ldb, strideB, dB)) return e == 2 ? -9 : (e == 3 ? -10 : -11);
    if (d_info == nullptr && batch > 0) return -12;
    if (batch == 0) return 0;
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(set_ints_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s, d_info, (int)batch, 0);
    if (n == 0) return check_launch(h, "set_info");
    BatchLayout bl;
    int rc = prepare_batch(h, n, nrhs, batch, nrhs > 0, 0, bl);
    if (rc) return rc;
    bl.bt.a = strideA;
    bl.bt.b = strideB;
    mi::enqueue_factorization<double>(
        s, dA, lda, n, bl.leaf, d_info, batch_sync(h), shared_status(h), MI_PREC_FP64, nullptr,
        [&](double fl, auto &&launch) { traced(h, s, MI_TRACE_POTRF, fl * (double)batch, launch); }, bl.bt);
    if (nrhs > 0) {
        // the factor's panel images carry the leaf inverses the block inverses start from
        traced(h, s, MI_TRACE_TRTRI, (double)n * mi::kNB * mi::kNB / 3.0 * (double)batch,
               [&] { mi::launch_block_inverses<double>(dA, lda, n, bl.leaf, bl.inv, d_info, s, true, bl.bt); });
        traced(h, s, MI_TRACE_SOLVE, 2.0 * (double)n * (double)n * (double)nrhs * (double)batch, [&] {
            mi::launch_solve_persistent<double>(dA, lda, n, dB, ldb, nrhs, bl.inv, bl.pack,
                                                (int *)batch_sync(h) + mi::kSyncSolveTickets, d_info,
                                                shared_status(h), s, 3, false, bl.bt);
        });
    }
    return check_launch(h, "posv_batched");
}

int mi_dposv_adjoint_batched(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, int64_t batch, const double *dL,
                             int64_t ldl, int64_t strideL, const double *dX, int64_t ldx, int64_t strideX,
                             double *dXbar, int64_t ldxb, int64_t strideXbar, double *dLbar, int64_t ldlb,
                             int64_t strideLbar)
{
    if (!valid(h)) return MI_LLT_E_HANDLE;
    const int up = uplo_code(uplo);
    if (up < 0) return -2;
    if (up == 1) return MI_LLT_E_UNSUPPORTED;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (batch < 0 || batch > 65535) return -5;
    if (int e = check_batch_args(n, n, batch, ldl, strideL, dL)) return e == 2 ? -6 : (e == 3 ? -7 : -8);
    if (int e = check_batch_args(n, nrhs, batch, ldx, strideX, dX)) return e == 2 ? -9 : (e == 3 ? -10 : -11);
    if (int e = check_batch_args(n, nrhs, batch, ldxb, strideXbar, dXbar)) return e == 2 ? -12 : (e == 3 ? -13 : -14);
    if (int e = check_batch_args(n, n, batch, ldlb, strideLbar, dLbar)) return e == 2 ? -15 : (e == 3 ? -16 : -17);
    if (n == 0 || batch == 0) return 0;
    const size_t sq = (size_t)n * (size_t)n, vec = (size_t)n * (size_t)nrhs;
    BatchLayout bl;
    int rc = prepare_batch(h, n, nrhs, batch, true, (3 * sq > 2 * vec ? 3 * sq : 2 * vec), bl);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const unsigned nb = (unsigned)batch;
    const double nd = (double)n, bd = (double)batch;
    bl.bt.a = strideL;
    bl.bt.b = strideXbar;
    // one set of diagonal-block inverses (with the solve's couplings) serves both adjoints
    mi::launch_leaf_inverses<double>(dL, ldl, n, bl.leaf, s, bl.bt);
    traced(h, s, MI_TRACE_TRTRI, nd * mi::kNB * mi::kNB / 3.0 * bd,
           [&] { mi::launch_block_inverses<double>(dL, ldl, n, bl.leaf, bl.inv, nullptr, s, true, bl.bt); });
    if (nrhs > 0) {  // the solve's adjoint: Bbar = A^{-1} Xbar, Lbar -= tril(Bbar (L^T X)^T + X (L^T Bbar)^T)
        traced(h, s, MI_TRACE_SOLVE, 2.0 * nd * nd * (double)nrhs * bd, [&] {
            mi::launch_solve_persistent<double>(dL, ldl, n, dXbar, ldxb, nrhs, bl.inv, bl.pack,
                                                (int *)batch_sync(h) + mi::kSyncSolveTickets, nullptr,
                                                shared_status(h), s, 3, false, bl.bt);
        });
        mi::launch_solve_adjoint_update(dL, ldl, n, nrhs, dX, ldx, dXbar, ldxb, dLbar, ldlb, bl.aux, bl.aux + vec, s,
                                        nb, strideL, strideX, strideXbar, strideLbar, bl.bt.aux);
    }
    // the factor's adjoint of the accumulated Lbar, into dLbar (mi_dpotrf_adjoint_batched's steps)
    const int64_t sa = bl.bt.aux;
    double *W2 = bl.aux, *W3 = bl.aux + sq, *Wt = bl.aux + 2 * sq;
    mi::launch_sym_tile(4, dL, ldl, n, Wt, n, s, nb, strideL, sa);
    mi::launch_sym_tile(4, dLbar, ldlb, n, W3, n, s, nb, strideLbar, sa);
    traced(h, s, MI_TRACE_GEMM, nd * nd * nd / 3.0 * bd,
           [&] { mi::launch_trmm_tn_lower(Wt, n, W3, n, n, W3, n, W2, s, nb, sa, sa, sa, sa); });
    mi::launch_sym_tile(0, W3, n, n, W2, n, s, nb, sa, sa);
    traced(h, s, MI_TRACE_SOLVE, nd * nd * nd * bd,
           [&] { mi::launch_trsm_xl_multi(Wt, n, n, W2, n, n, bl.inv, false, s, nb, sa, sa, bl.bt.inv); });
    mi::launch_sym_tile(1, W2, n, n, W3, n, s, nb, sa, sa);
    traced(h, s, MI_TRACE_SOLVE, nd * nd * nd / 2.0 * bd,
           [&] { mi::launch_trsm_xl_multi(Wt, n, n, W3, n, n, bl.inv, true, s, nb, sa, sa, bl.bt.inv); });
    mi::launch_sym_tile(3, W3, n, n, dLbar, ldlb, s, nb, sa, strideLbar);
    return check_launch(h, "posv_adjoint_batched");
}

}  // extern "C"
