// rank_update.hip -- rank-one update / downdate of a Cholesky factor: L L^T + sigma v v^T.
//
// Reference: LLT::rankUpdate (Eigen/src/Cholesky/LLT.h:467-480) -> llt_rank_update_lower
// (LLT.h:239-307), both branches:
//   sigma > 0 : temp = sqrt(sigma) v; for each column j a Givens rotation G_j =
//               makeGivens(L_jj, -temp_j) (Jacobi.h:231-266) gives the new L_jj, and is applied
//               to (L(j+1:, j), temp(j+1:)) as x' = c x + s y, y' = -s x + c y
//               (apply_rotation_in_the_plane, Jacobi.h:331-332);
//   sigma <= 0: temp = v, beta = 1; for each column j: d = L_jj^2, w = temp_j, sw2 = sigma w^2,
//               gamma = d beta + sw2, x = d + sw2 / beta (x <= 0: fail at column j), new L_jj =
//               sqrt(x), beta += sw2 / d, then temp(j+1:) -= (w / L_jj) L(j+1:, j) and, when
//               gamma != 0, L(j+1:, j) = (L'_jj / L_jj) L(j+1:, j) + (L'_jj sigma w / gamma) temp(j+1:).
//
// Parallel form: column j's scalars depend only on row j's own history, and a row's entries
// and temp only on the scalars of the columns before it.  So rows go in blocks of 128, one
// wave each (two rows per lane), and a block takes a ticket in row order: it applies the
// published scalars of every earlier column block to its rows (a chain of two FMAs per column
// through its temp, the L loads running ahead), then runs the columns of its own diagonal block
// in order (the pivot row's values broadcast across the wave, no barriers) and publishes that
// block's scalars and flag.  Every element of L is read and written once.  The waits are for
// blocks that drew earlier tickets, which are resident: they always end (a bound records a
// timeout in the status word anyway).  After a failing column (sigma <= 0) the columns before
// it are updated in every row and the rest is untouched, as the reference leaves it.
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

constexpr int kRuRows = 128;                 // rows per work item (one wave, two per lane)
constexpr uint32_t kRuDone = 1u << 30;       // flag: block published (low bits: valid columns)
constexpr uint32_t kRuFail = 1u << 29;       // flag: the block's diagonal stage found x <= 0
constexpr uint32_t kRuSpinLimit = 1u << 22;

__device__ __forceinline__ double ru_bcast(double x, int lane)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The scalars of one column: GIVENS (c, s); otherwise (w / L_jj, L'_jj / L_jj, L'_jj sigma w /
// gamma, gamma != 0).
template <bool GIVENS>
__device__ __forceinline__ void ru_apply(double &x, double &t, const double *k)
{
    if (GIVENS) {
        const double c = k[0], s = k[1];
        const double xn = c * x + s * t;
        t = -s * x + c * t;
        x = xn;
    } else {
        t -= k[0] * x;
        if (k[3] != 0.0) x = k[1] * x + k[2] * t;
    }
}

// Waits until flags[b] is published (all lanes poll the same word); 0 on a timeout or when
// the status word reports a failure elsewhere.
__device__ __forceinline__ uint32_t ru_wait(const uint32_t *flag, int *status)
{
    uint32_t spins = 0;
    for (;;) {
        const uint32_t f = __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (f & kRuDone) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return f;
        }
        __builtin_amdgcn_s_sleep(1);
        ++spins;
        if ((spins & 255) == 0 && device_failed(status)) return 0;
        if (spins > kRuSpinLimit) {
            if (lane_id() == 0) atomicOr(status, MI_SOLVE_TIMEOUT);
            return 0;
        }
    }
}

// coef: 4 doubles per column; bbeta: beta after each block (sigma <= 0); flags: one per block,
// ticket: one int -- all zeroed before the launch.
template <bool GIVENS>
__global__ __launch_bounds__(64) void rank_update_kernel(double *__restrict__ L, int64_t ldl, int64_t n,
                                                         const double *__restrict__ v, int64_t incv, double sigma,
                                                         double *coef, double *bbeta, uint32_t *flags, int *ticket,
                                                         int *info, int *status)
{
    __shared__ double ks[kRuRows * 4];
    if (device_failed(status)) return;
    const int l = lane_id();
    int t = atomicAdd(ticket, l == 0 ? 1 : 0);  // one ticket per wave, read from lane 0
    t = __builtin_amdgcn_readlane(t, 0);
    const int nblk = (int)((n + kRuRows - 1) / kRuRows);
    if (t >= nblk) return;
    const int ib = t;
    const int64_t r0 = (int64_t)ib * kRuRows;
    const int bs = (int)((n - r0) < kRuRows ? (n - r0) : kRuRows);
    const int64_t ra = r0 + l, rb = r0 + 64 + l;  // this lane's two rows
    const bool ina = ra < n, inb = rb < n;
    const int64_t rca = ina ? ra : n - 1, rcb = inb ? rb : n - 1;
    double ta = ina ? v[ra * incv] : 0.0, tb = inb ? v[rb * incv] : 0.0;
    if (GIVENS) {
        const double sq = sqrt(sigma);
        ta *= sq;
        tb *= sq;
    }
    double beta = 1.0;

    // Columns go in chunks of 16 (two rows each), the next chunk's loads in flight while one
    // is computed: its values do not change before their own column is processed (a column is
    // written only at its own step), so they can be read that far ahead.
    constexpr int CH = 16;
    double x0a[CH], x0b[CH], x1a[CH], x1b[CH];
    auto load = [&](double (&pa)[CH], double (&pb)[CH], int64_t c0, int j0, int ncol) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const int64_t col = c0 + (j0 + u < ncol ? j0 + u : ncol - 1);
            pa[u] = L[rca + col * ldl];
            pb[u] = L[rcb + col * ldl];
        }
    };

    // ---- the column blocks before this one
    for (int b = 0; b < ib; ++b) {
        const uint32_t f = ru_wait(flags + b, status);
        if (f == 0) return;
        const int ncol = (int)(f & 0xffffu);
        for (int e = l; e < 4 * ncol; e += 64) ks[e] = __hip_atomic_load(coef + (int64_t)b * kRuRows * 4 + e,
                                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!GIVENS) beta = __hip_atomic_load(bbeta + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int64_t c0 = (int64_t)b * kRuRows;
        auto run = [&](double (&pa)[CH], double (&pb)[CH], int j0) {
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                if (j0 + u < ncol) {
                    ru_apply<GIVENS>(pa[u], ta, ks + 4 * (j0 + u));
                    ru_apply<GIVENS>(pb[u], tb, ks + 4 * (j0 + u));
                    const int64_t col = c0 + j0 + u;
                    if (ina) L[ra + col * ldl] = pa[u];
                    if (inb) L[rb + col * ldl] = pb[u];
                }
            }
        };
        if (ncol > 0) {
            load(x0a, x0b, c0, 0, ncol);
            for (int j0 = 0; j0 < ncol; j0 += 2 * CH) {
                if (j0 + CH < ncol) load(x1a, x1b, c0, j0 + CH, ncol);
                run(x0a, x0b, j0);
                if (j0 + CH >= ncol) break;
                if (j0 + 2 * CH < ncol) load(x0a, x0b, c0, j0 + 2 * CH, ncol);
                run(x1a, x1b, j0 + CH);
            }
        }
        __syncthreads();  // ks is rewritten for the next block
        if (f & kRuFail) return;  // the reference stops at the failing column
    }

    // ---- this block's diagonal: its columns in order
    int ncol = bs;
    bool failed = false;
    // column jj = j0 + u of the block, its values of this lane's rows in (pa[u], pb[u])
    auto diag = [&](double (&pa)[CH], double (&pb)[CH], int j0) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const int jj = j0 + u;
            if (failed || jj >= bs) continue;
            const int64_t j = r0 + jj;
            double xa = pa[u], xb = pb[u];
            const int o = jj & 63;
            const double ljj = jj < 64 ? ru_bcast(xa, o) : ru_bcast(xb, o);
            const double tj = jj < 64 ? ru_bcast(ta, o) : ru_bcast(tb, o);
            double k[4];
            double ndiag;
            if (GIVENS) {  // makeGivens(p = L_jj, q = -temp_j, &r), Jacobi.h:231-266
                const double p = ljj, q = -tj;
                double c, sn;
                if (q == 0.0) {
                    c = p < 0.0 ? -1.0 : 1.0;
                    sn = 0.0;
                    ndiag = fabs(p);
                } else if (p == 0.0) {
                    c = 0.0;
                    sn = q < 0.0 ? 1.0 : -1.0;
                    ndiag = fabs(q);
                } else if (fabs(p) > fabs(q)) {
                    const double tt = q / p;
                    double w = sqrt(1.0 + tt * tt);
                    if (p < 0.0) w = -w;
                    c = 1.0 / w;
                    sn = -tt * c;
                    ndiag = p * w;
                } else {
                    const double tt = p / q;
                    double w = sqrt(1.0 + tt * tt);
                    if (q < 0.0) w = -w;
                    sn = -1.0 / w;
                    c = -tt * sn;
                    ndiag = q * w;
                }
                k[0] = c;
                k[1] = sn;
                k[2] = k[3] = 0.0;
            } else {  // LLT.h:271-296
                const double dj = ljj * ljj, wj = tj, swj2 = sigma * (wj * wj);
                const double gamma = dj * beta + swj2;
                const double x = dj + swj2 / beta;
                if (!(x > 0.0)) {  // x <= 0: fail at column j (a NaN pivot fails too)
                    failed = true;
                    ncol = jj;
                    continue;
                }
                ndiag = sqrt(x);
                beta += swj2 / dj;
                k[0] = wj / ljj;
                k[1] = ndiag / ljj;
                k[2] = gamma != 0.0 ? (ndiag * sigma * wj) / gamma : 0.0;
                k[3] = gamma != 0.0 ? 1.0 : 0.0;
            }
            if (l == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) coef[j * 4 + q] = k[q];
            }
            // the pivot row takes the new diagonal, the rows below the column operations
            if (jj < 64) {
                if (l == o) L[j + j * ldl] = ndiag;
                if (l > o && ina) {
                    ru_apply<GIVENS>(xa, ta, k);
                    L[ra + j * ldl] = xa;
                }
                if (inb) {
                    ru_apply<GIVENS>(xb, tb, k);
                    L[rb + j * ldl] = xb;
                }
            } else {
                if (l == o) L[j + j * ldl] = ndiag;
                if (l > o && inb) {
                    ru_apply<GIVENS>(xb, tb, k);
                    L[rb + j * ldl] = xb;
                }
            }
        }
    };
    load(x0a, x0b, r0, 0, bs);
    for (int j0 = 0; j0 < bs && !failed; j0 += 2 * CH) {
        if (j0 + CH < bs) load(x1a, x1b, r0, j0 + CH, bs);
        diag(x0a, x0b, j0);
        if (j0 + CH >= bs || failed) break;
        if (j0 + 2 * CH < bs) load(x0a, x0b, r0, j0 + 2 * CH, bs);
        diag(x1a, x1b, j0 + CH);
    }
    if (failed && l == 0) atomicCAS(info, 0, (int)(r0 + ncol + 1));
    if (!GIVENS && l == 0) bbeta[ib] = beta;
    // publish: the wave's stores (coefficients, beta) complete before the flag
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (l == 0)
        __hip_atomic_store(flags + ib, kRuDone | (failed ? kRuFail : 0u) | (uint32_t)ncol, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void ru_reset_kernel(uint32_t *flags, int nblk, int *ticket, int *info)
{
    for (int i = threadIdx.x; i < nblk; i += 256) flags[i] = 0u;
    if (threadIdx.x == 0) {
        *ticket = 0;
        *info = 0;
    }
}

int64_t rank_update_scratch_bytes(int64_t n)
{
    const int64_t nblk = (n + kRuRows - 1) / kRuRows;
    return nblk * kRuRows * 4 * 8 + nblk * 8 + nblk * 4 + 256;
}

void launch_rank_update(double *L, int64_t ldl, int64_t n, const double *v, int64_t incv, double sigma, void *scratch,
                        int *ticket, int *info, int *status, hipStream_t s)
{
    const int nblk = (int)((n + kRuRows - 1) / kRuRows);
    char *p = (char *)scratch;
    double *coef = (double *)p;
    double *bbeta = (double *)(p + (int64_t)nblk * kRuRows * 4 * 8);
    uint32_t *flags = (uint32_t *)(p + (int64_t)nblk * kRuRows * 4 * 8 + (int64_t)nblk * 8);
    hipLaunchKernelGGL(ru_reset_kernel, dim3(1), dim3(256), 0, s, flags, nblk, ticket, info);
    if (sigma > 0.0)
        hipLaunchKernelGGL(rank_update_kernel<true>, dim3(nblk), dim3(64), 0, s, L, ldl, n, v, incv, sigma, coef,
                           bbeta, flags, ticket, info, status);
    else
        hipLaunchKernelGGL(rank_update_kernel<false>, dim3(nblk), dim3(64), 0, s, L, ldl, n, v, incv, sigma, coef,
                           bbeta, flags, ticket, info, status);
}

}  // namespace mi
