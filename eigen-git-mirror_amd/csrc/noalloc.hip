// noalloc.hip -- mi_dposv_noalloc: the factor + solve of mi_dposv_async (lower storage) with
// every scratch buffer in one file-scope device array, for callers that cannot allocate
// (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The kernels are the handle path's (factor_plan.hpp, launch_block_inverses,
// launch_solve_persistent); only where their scratch lives differs.  The array `scratch` is
// resolved once per device with hipGetSymbolAddress (the host-side name is a stand-in, not a
// device address), and its device address is what every kernel receives.  Layout of one call:
//   [ sync words (512 B): status, solve tickets, factor tickets, panel flag, each on its own line
//   | panel images (kPanelDoubles per 128-block)
//   | block inverses + couplings (kInvDoubles per 128-block; only when nrhs > 0)
//   | the solve's packed block copies (solve_pack_doubles(n, nrhs); only when nrhs > 0) ]
// Every call clears what must start at zero (status, tickets, flag; the solve resets its own
// tickets and packed copies) with small kernels on the caller's stream -- no memset, so the
// call can be captured and replayed.  All calls on one GPU share that GPU's array: calls that
// overlap on one GPU (two streams, two threads) race and give wrong results.
#include <mutex>

#include <hip/hip_runtime.h>

#include "../../include/mi_llt.h"
#include "factor_plan.hpp"
#include "kernels.hpp"

// Every GPU the library runs on holds its own copy of this array until the process ends.
static __device__ __align__(256) unsigned char scratch[MI_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kScratchBytes = (size_t)MI_NOALLOC_SCRATCH_BYTES;
constexpr size_t kSyncBytes = 4 * mi::kSyncWords;
constexpr int kMaxDevices = 64;

// Device address of `scratch` per device, resolved on the first call made with that device
// current.
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {nullptr};

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

__global__ void noalloc_begin_kernel(uint32_t *sync, int *info)
{
    // status, the solve tickets, the factor tickets and the panel flag start at zero
    if (threadIdx.x < mi::kSyncWords) sync[threadIdx.x] = 0u;
    if (threadIdx.x == 0) *info = 0;
}

// A device failure recorded during the call (a timeout or a ticket outside its launch)
// becomes the call's info, as no status query exists for this entry point.
__global__ void noalloc_end_kernel(const uint32_t *sync, int *info)
{
    if (threadIdx.x == 0 && __hip_atomic_load(sync + mi::kSyncStatus, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
        *info = MI_LLT_E_DEVICE;
}

}  // namespace

extern "C" int mi_dposv_noalloc(char uplo, int64_t n, int64_t nrhs, double *dA, int64_t lda, double *dB,
                                int64_t ldb, int *d_info, void *stream)
{
    if (uplo != 'L' && uplo != 'l') return (uplo == 'U' || uplo == 'u') ? MI_LLT_E_UNSUPPORTED : -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (dA == nullptr && n > 0) return -4;
    if (lda < (n > 1 ? n : 1)) return -5;
    if (dB == nullptr && n > 0 && nrhs > 0) return -6;
    if (ldb < (n > 1 ? n : 1)) return -7;
    if (d_info == nullptr) return -8;
    const int64_t nblk = (n + mi::kNB - 1) / mi::kNB;
    const size_t leaf = up256((size_t)nblk * mi::kPanelDoubles * sizeof(double));
    const size_t inv = nrhs > 0 ? up256((size_t)nblk * mi::kInvDoubles * sizeof(double)) : 0;
    const size_t pack = nrhs > 0 ? up256((size_t)mi::solve_pack_doubles(n, nrhs) * sizeof(double)) : 0;
    if (kSyncBytes + leaf + inv + pack > kScratchBytes) return MI_LLT_E_TOO_LARGE;

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return MI_LLT_E_HIP;
    char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_scratch[dev]) {
            void *p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return MI_LLT_E_HIP;
            g_scratch[dev] = p;
        }
        base = (char *)g_scratch[dev];
    }
    hipStream_t s = (hipStream_t)stream;
    uint32_t *sync = (uint32_t *)base;
    double *wleaf = (double *)(base + kSyncBytes);
    double *winv = (double *)(base + kSyncBytes + leaf);
    double *ypk = (double *)(base + kSyncBytes + leaf + inv);

    hipLaunchKernelGGL(noalloc_begin_kernel, dim3(1), dim3(mi::kSyncWords), 0, s, sync, d_info);
    if (n > 0) {
        mi::enqueue_factorization<double>(s, dA, lda, n, wleaf, d_info, sync, (int *)sync + mi::kSyncStatus,
                                          MI_PREC_FP64, nullptr,
                                          [](double, auto &&launch) { launch(); });
        if (nrhs > 0) {
            mi::launch_block_inverses<double>(dA, lda, n, wleaf, winv, d_info, s);
            mi::launch_solve_persistent<double>(dA, lda, n, dB, ldb, nrhs, winv, ypk, (int *)sync + mi::kSyncSolveTickets,
                                                d_info, (int *)sync + mi::kSyncStatus, s);
        }
    }
    hipLaunchKernelGGL(noalloc_end_kernel, dim3(1), dim3(64), 0, s, sync, d_info);
    return hipGetLastError() == hipSuccess ? 0 : MI_LLT_E_HIP;
}
