// vbatched.hip -- thousands of independent SPD factor + solve problems of uneven order
// (0..256) in one persistent launch.
//
// Reference counterpart: a loop of LLT<MatrixXd>(A_i).solve(B_i), i.e. llt_inplace::blocked /
// unblocked (Eigen/src/Cholesky/LLT.h:312-368; below 32 columns Eigen runs the unblocked
// column loop on the whole matrix, LLT.h:345) and LLT::_solve_impl (LLT.h:483-498) per system.
//
// Every wave is an independent worker: it draws the next system from a ticket counter (the
// systems sorted longest first by a one-workgroup plan kernel) and factors and solves it
// alone, so no wave ever waits for another and the launch always drains.  Per system
// (T = ceil(n/16) tiles of 16), left-looking by tile columns k = 0..T-1:
//   C_kk = A_kk - sum_{j<k} L_kj L_kj^T        (MFMA, operands read back from the system)
//   leaf: L_kk = chol(C_kk) and X_kk = L_kk^{-1} in one pass of the column loop of
//         LLT.h:320-334 -- lanes 0..15 hold the rows of C_kk, lanes 16..31 rows of the
//         identity, and the same column operations turn the identity rows into L_kk^{-T};
//   L_ik = (A_ik - sum_{j<k} L_ij L_kj^T) X_kk^T, tile rows i > k in groups of G
// then the forward and backward substitutions by tiles with the stored X_kk, all products
// on MFMA 16x16x4 (FP64 or FP32).  A partial last tile is padded with the identity.
// The MFMA accumulator layout (element r of lane l: row l%16, column kx(r, l/16)) is the
// operand layout of k-step r of the next product, so C_ik feeds C_ik X_kk^T directly.
#include "common.hpp"
#include "kernels.hpp"

namespace mi {

constexpr int kVbMaxN = 256;
constexpr int kVbMaxT = kVbMaxN / 16;
constexpr int kVbWaves = 4;  // waves (workers) per workgroup

__device__ __forceinline__ double vb_readlane(double x, int lane)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ float vb_readlane(float x, int lane)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
}

// Global-memory views of a system's pointers.  A pointer rebuilt from readfirstlane'd integers
// is a generic one, and generic (flat) loads count in both the vector-memory and the LDS
// counters, which the compiler can only wait on as a whole: with flat loads it waited for the
// prefetched k-step together with the current one.  Global loads are counted in order.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p)
{
    return (gptr<T>)p;
}
// Load at a byte offset from a wave-uniform base (scalar base + 32-bit vector offset form).
template <typename T>
__device__ __forceinline__ T load_at(gptr<const T> base, uint32_t off)
{
    return *(gptr<const T>)((__attribute__((address_space(1))) const char *)base + off);
}

template <typename T>
struct VbArith;
// FP64: v_mfma_f64_16x16x4f64; k-step s, lane group g supplies k = 4s + g
template <>
struct VbArith<double> {
    using acc_t = dvec4;
    __device__ static acc_t mma(double q, double p, acc_t c) { return mfma_f64(q, p, c); }
    __device__ static int kx(int s, int g) { return 4 * s + g; }
    __device__ static constexpr int step(int s) { return 4 * s; }  // kx(s, g) - kx(0, g)
    __device__ static double rsqrt(double x)
    {
        const double r = __builtin_amdgcn_rsq(x);
        return r * (1.5 - 0.5 * x * r * r);
    }
};
// FP32: v_mfma_f32_16x16x4f32, whose result element r of lane l is column 4(l/16) + r; the
// k index of the operands is permuted the same way (any common order gives the same sums)
template <>
struct VbArith<float> {
    using acc_t = fvec4;
    __device__ static acc_t mma(float q, float p, acc_t c) { return mfma_f32(q, p, c); }
    __device__ static int kx(int s, int g) { return 4 * g + s; }
    __device__ static constexpr int step(int s) { return s; }  // kx(s, g) - kx(0, g)
    __device__ static float rsqrt(float x)
    {
        const float r = __builtin_amdgcn_rsqf(x);
        return r * (1.5f - 0.5f * x * r * r);
    }
};

// Memory ordering between the lanes of one wave for data it stored and reads back: the
// wave's global stores complete (s_waitcnt 0) before any later load is issued.
__device__ __forceinline__ void vb_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Factor + solve one system with one wave.  Ls, Xs: 256 elements of LDS each (this wave's);
// Xg: kVbMaxT * 256 elements of global scratch (this wave's).  Returns LAPACK info.
//
// Loads of a k-step s of column block j use a wave-uniform base (A + (16j + step(s)) * lda)
// and a per-lane 32-bit byte offset of (row, lane group's column), the scalar-base form of
// the global loads.  Every left-looking loop over j is software-pipelined: the loads of step
// j+1 go out before the MFMAs of step j (two register buffers used alternately by name, so
// no buffer is indexed at run time).
template <typename T, int G>
__device__ int vb_system(int n, T *A_, int64_t lda, int nrhs, T *B_, int64_t ldb, T *Ls, T *Xs, T *Xg_)
{
    using M = VbArith<T>;
    using acc_t = typename M::acc_t;
    const gptr<T> A = to_global(A_), B = to_global(B_), Xg = to_global(Xg_);
    const int l = lane_id(), li = l & 15, g = l >> 4;
    const int nt = (n + 15) >> 4;
    const T zero = (T)0, one = (T)1;
    const int kx0 = M::kx(0, g);  // kx(s, g) = kx0 + M::step(s)
    auto at = [](gptr<const T> base, uint32_t off) -> T { return load_at<T>(base, off); };
    auto lane_off = [&](int row, int64_t ld) -> uint32_t { return (uint32_t)((row + (int64_t)kx0 * ld) * (int64_t)sizeof(T)); };

    for (int k = 0; k < nt; ++k) {
        const int k0 = 16 * k;
        const int rk = k0 + li < n ? k0 + li : n - 1;
        const uint32_t offq = lane_off(rk, lda);
        // ---- C_kk = A_kk - sum_{j<k} L_kj L_kj^T -> Ls (identity beyond n)
        {
            acc_t acc = {zero, zero, zero, zero};
            const bool kin = k0 + li < n;
            T qa[4], qb[4];
#define VB_LOADD(q_, j_)                                                                           \
    do {                                                                                           \
        _Pragma("unroll") for (int s = 0; s < 4; ++s) q_[s] = at(A + (int64_t)(16 * (j_) + M::step(s)) * lda, offq); \
    } while (0)
#define VB_MMAD(q_)                                                                                \
    do {                                                                                           \
        _Pragma("unroll") for (int s = 0; s < 4; ++s) acc = M::mma(kin ? q_[s] : zero, kin ? q_[s] : zero, acc); \
    } while (0)
            if (k > 0) {
                VB_LOADD(qa, 0);
                for (int j = 0; j < k; j += 2) {
                    if (j + 1 < k) VB_LOADD(qb, j + 1);
                    VB_MMAD(qa);
                    if (j + 1 >= k) break;
                    if (j + 2 < k) VB_LOADD(qa, j + 2);
                    VB_MMAD(qb);
                }
            }
#undef VB_LOADD
#undef VB_MMAD
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = li, col = M::kx(r, g);
                const int gi = k0 + row, gj = k0 + col;
                const bool in = gi < n && gj < n;
                const T a = A[(in ? gi : 0) + (int64_t)(in ? gj : 0) * lda];
                Ls[row + 16 * col] = in ? (T)(a - acc[r]) : (row == col ? one : zero);
            }
        }
        // ---- leaf: lanes 0..15 rows of C_kk, lanes 16..31 identity rows; the column operations
        // run on every lane without lane predicates (entries above a leaf lane's diagonal change
        // but are never broadcast nor stored), column j+1 -- the next pivot -- first
        {
            T v[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) v[c] = l < 16 ? Ls[li + 16 * c] : ((l < 32 && c == l - 16) ? one : zero);
            int fail = -1;
            T x = vb_readlane(v[0], 0);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (fail < 0 && x <= zero) fail = j;  // LLT.h:327-330; a NaN pivot passes
                v[j] *= M::rsqrt(x);
                if (j + 1 < 16) {
                    v[j + 1] -= v[j] * vb_readlane(v[j], j + 1);
                    x = vb_readlane(v[j + 1], j + 1);
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int c = j + 2; c < 16; ++c) v[c] -= v[j] * vb_readlane(v[j], c);
            }
            if (fail >= 0) return k0 + fail + 1;
            // L_kk to the system (lower part), X_kk = L_kk^{-1} to Xs / Xg (column-major)
            if (l < 16 && k0 + l < n) {
#pragma unroll
                for (int c = 0; c < 16; ++c)
                    if (c <= l) A[(k0 + l) + (int64_t)(k0 + c) * lda] = v[c];
            }
            if (l >= 16 && l < 32) {
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    Xs[c + 16 * (l - 16)] = v[c];  // lane 16+i holds row i of L^{-T} = column i of X
                    Xg[k * 256 + c + 16 * (l - 16)] = v[c];
                }
            }
        }
        // ---- panel: L_ik = (A_ik - sum_{j<k} L_ij L_kj^T) X_kk^T, i > k, G tile rows at a time
        T xq[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) xq[s] = Xs[li + 16 * M::kx(s, g)];
        for (int i0 = k + 1; i0 < nt; i0 += G) {
            uint32_t offp[G];
            bool pin[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const int row = 16 * (i0 + u) + li;
                pin[u] = row < n;
                offp[u] = lane_off(row < n ? row : n - 1, lda);
            }
            acc_t pa[G];
#pragma unroll
            for (int u = 0; u < G; ++u) pa[u] = acc_t{zero, zero, zero, zero};
            T qa[4], qb[4], xa[G][4], xb[G][4];
#define VB_LOAD(q_, x_, j_)                                                                        \
    do {                                                                                           \
        _Pragma("unroll") for (int s = 0; s < 4; ++s)                                              \
        {                                                                                          \
            const gptr<const T> base_ = A + (int64_t)(16 * (j_) + M::step(s)) * lda;               \
            q_[s] = at(base_, offq);                                                               \
            _Pragma("unroll") for (int u = 0; u < G; ++u) x_[u][s] = at(base_, offp[u]);            \
        }                                                                                          \
    } while (0)
#define VB_MMA(q_, x_)                                                                             \
    do {                                                                                           \
        _Pragma("unroll") for (int u = 0; u < G; ++u)                                              \
        {                                                                                          \
            _Pragma("unroll") for (int s = 0; s < 4; ++s) pa[u] = M::mma(q_[s], pin[u] ? x_[u][s] : zero, pa[u]); \
        }                                                                                          \
    } while (0)
            if (k > 0) {
                VB_LOAD(qa, xa, 0);
                for (int j = 0; j < k; j += 2) {
                    if (j + 1 < k) VB_LOAD(qb, xb, j + 1);
                    VB_MMA(qa, xa);
                    if (j + 1 >= k) break;
                    if (j + 2 < k) VB_LOAD(qa, xa, j + 2);
                    VB_MMA(qb, xb);
                }
            }
#undef VB_LOAD
#undef VB_MMA
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (i0 + u < nt) {
                    const int row = 16 * (i0 + u) + li;
                    const int rc = row < n ? row : n - 1;
                    T c[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const T a = A[rc + (int64_t)(k0 + M::kx(r, g)) * lda];
                        c[r] = row < n ? (T)(a - pa[u][r]) : zero;
                    }
                    acc_t o = {zero, zero, zero, zero};
#pragma unroll
                    for (int s = 0; s < 4; ++s) o = M::mma(xq[s], c[s], o);
                    if (row < n) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) A[row + (int64_t)(k0 + M::kx(r, g)) * lda] = o[r];
                    }
                }
            }
        }
        vb_fence();
    }

    // ---- solve, 16 right-hand sides at a time; products on the transposed blocks
    // (C[c][i] = sum_m P[c][m] Q[i][m] with P from B, Q from L or X)
    for (int c0 = 0; c0 < nrhs; c0 += 16) {
        const int col = c0 + li;
        const bool cin = col < nrhs;
        const int64_t bcol = (int64_t)(cin ? col : c0) * ldb;
        const gptr<const T> Bc = B + bcol;
        // forward: Y_k^T = (B_k^T - sum_{j<k} Y_j^T L_kj^T) X_kk^T
        for (int k = 0; k < nt; ++k) {
            const int k0 = 16 * k;
            const int rk = k0 + li < n ? k0 + li : n - 1;
            const uint32_t offq = lane_off(rk, lda);
            const uint32_t offb = (uint32_t)(kx0 * (int64_t)sizeof(T));
            acc_t acc = {zero, zero, zero, zero};
            T pa_[4], qa_[4], pb_[4], qb_[4];
#define VB_LOADF(p_, q_, j_)                                                                       \
    do {                                                                                           \
        _Pragma("unroll") for (int s = 0; s < 4; ++s)                                              \
        {                                                                                          \
            p_[s] = at(Bc + 16 * (j_) + M::step(s), offb);                                         \
            q_[s] = at(A + (int64_t)(16 * (j_) + M::step(s)) * lda, offq);                         \
        }                                                                                          \
    } while (0)
#define VB_MMAF(p_, q_)                                                                            \
    do {                                                                                           \
        _Pragma("unroll") for (int s = 0; s < 4; ++s) acc = M::mma(k0 + li < n ? q_[s] : zero, cin ? p_[s] : zero, acc); \
    } while (0)
            if (k > 0) {
                VB_LOADF(pa_, qa_, 0);
                for (int j = 0; j < k; j += 2) {
                    if (j + 1 < k) VB_LOADF(pb_, qb_, j + 1);
                    VB_MMAF(pa_, qa_);
                    if (j + 1 >= k) break;
                    if (j + 2 < k) VB_LOADF(pa_, qa_, j + 2);
                    VB_MMAF(pb_, qb_);
                }
            }
#undef VB_LOADF
#undef VB_MMAF
            T c[4], xq2[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = k0 + M::kx(r, g);
                const T bv = B[(row < n ? row : n - 1) + bcol];
                c[r] = (row < n && cin) ? (T)(bv - acc[r]) : zero;
                xq2[r] = Xg[k * 256 + li + 16 * M::kx(r, g)];
            }
            acc_t y = {zero, zero, zero, zero};
#pragma unroll
            for (int s = 0; s < 4; ++s) y = M::mma(xq2[s], c[s], y);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = k0 + M::kx(r, g);
                if (row < n && cin) B[row + bcol] = y[r];
            }
            vb_fence();
        }
        // backward: X_k^T = (Y_k^T - sum_{j>k} X_j^T L_jk) X_kk
        for (int k = nt - 1; k >= 0; --k) {
            const int k0 = 16 * k;
            // column k0 + li of L: below n whenever the loop runs (k < nt - 1); clamped anyway
            const int64_t colkc = (int64_t)(k0 + li < n ? k0 + li : n - 1) * lda;
            acc_t acc = {zero, zero, zero, zero};
            T pa_[4], qa_[4], pb_[4], qb_[4];
            bool oka[4], okb[4];
#define VB_LOADB(p_, q_, ok_, j_)                                                                  \
    do {                                                                                           \
        _Pragma("unroll") for (int s = 0; s < 4; ++s)                                              \
        {                                                                                          \
            const int m = 16 * (j_) + M::kx(s, g);                                                 \
            ok_[s] = m < n;                                                                        \
            const int mc = ok_[s] ? m : n - 1;                                                     \
            p_[s] = B[mc + bcol];                                                                  \
            q_[s] = A[mc + colkc];                                                                 \
        }                                                                                          \
    } while (0)
#define VB_MMAB(p_, q_, ok_)                                                                       \
    do {                                                                                           \
        _Pragma("unroll") for (int s = 0; s < 4; ++s) acc = M::mma(ok_[s] ? q_[s] : zero, (ok_[s] && cin) ? p_[s] : zero, acc); \
    } while (0)
            if (k + 1 < nt) {
                VB_LOADB(pa_, qa_, oka, k + 1);
                for (int j = k + 1; j < nt; j += 2) {
                    if (j + 1 < nt) VB_LOADB(pb_, qb_, okb, j + 1);
                    VB_MMAB(pa_, qa_, oka);
                    if (j + 1 >= nt) break;
                    if (j + 2 < nt) VB_LOADB(pa_, qa_, oka, j + 2);
                    VB_MMAB(pb_, qb_, okb);
                }
            }
#undef VB_LOADB
#undef VB_MMAB
            T c[4], xq2[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = k0 + M::kx(r, g);
                const T bv = B[(row < n ? row : n - 1) + bcol];
                c[r] = (row < n && cin) ? (T)(bv - acc[r]) : zero;
                xq2[r] = Xg[k * 256 + M::kx(r, g) + 16 * li];
            }
            acc_t y = {zero, zero, zero, zero};
#pragma unroll
            for (int s = 0; s < 4; ++s) y = M::mma(xq2[s], c[s], y);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = k0 + M::kx(r, g);
                if (row < n && cin) B[row + bcol] = y[r];
            }
            vb_fence();
        }
    }
    return 0;
}

// Plan: check every system's arguments, sort the systems longest first (counting sort by
// tile count; LPT order keeps the tail of the launch short) or keep the given order, and
// reset the work counter.  One workgroup.
__global__ __launch_bounds__(1024) void vb_plan_kernel(int64_t batch, const int *__restrict__ d_n,
                                                       const int *__restrict__ d_lda, const int *__restrict__ d_nrhs,
                                                       const int *__restrict__ d_ldb, int *__restrict__ d_info,
                                                       int *__restrict__ order, int lpt, int *counter)
{
    __shared__ int hist[kVbMaxT + 2], start[kVbMaxT + 2];
    const int tid = threadIdx.x;
    if (tid < kVbMaxT + 2) hist[tid] = 0;
    __syncthreads();
    for (int64_t i = tid; i < batch; i += blockDim.x) {
        const int n = d_n[i], lda = d_lda[i], nr = d_nrhs[i], ldb = d_ldb[i];
        int bad = 0;
        if (n < 0 || n > kVbMaxN) bad = -1;
        else if (lda < (n > 1 ? n : 1)) bad = -2;
        else if (nr < 0) bad = -3;
        else if (nr > 0 && ldb < (n > 1 ? n : 1)) bad = -4;
        d_info[i] = bad;
        const int b = bad ? kVbMaxT + 1 : (n + 15) / 16;  // bucket; invalid systems last
        atomicAdd(&hist[b], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int s = 0;
        for (int b = kVbMaxT; b >= 0; --b) {  // longest first
            start[b] = s;
            s += hist[b];
        }
        start[kVbMaxT + 1] = s;
        *counter = 0;
    }
    __syncthreads();
    for (int64_t i = tid; i < batch; i += blockDim.x) {
        if (!lpt) {
            order[i] = (int)i;
            continue;
        }
        const int n = d_n[i];
        const int b = d_info[i] ? kVbMaxT + 1 : (n + 15) / 16;
        order[atomicAdd(&start[b], 1)] = (int)i;
    }
}

template <typename T, int G>
__global__ __launch_bounds__(64 * kVbWaves, 2) void vb_persistent_kernel(int64_t batch, const int *__restrict__ d_n,
                                                                      T *const *__restrict__ d_A,
                                                                      const int *__restrict__ d_lda,
                                                                      const int *__restrict__ d_nrhs,
                                                                      T *const *__restrict__ d_B,
                                                                      const int *__restrict__ d_ldb, int *d_info,
                                                                      const int *__restrict__ order, int *counter,
                                                                      T *__restrict__ scratch)
{
    __shared__ T lds[kVbWaves][2][256];
    const int w = threadIdx.x >> 6, l = lane_id();
    const int worker = blockIdx.x * kVbWaves + w;
    T *Xg = scratch + (size_t)worker * kVbMaxT * 256;
    // The loop has no lane-dependent branch: every value that steers it is wave-uniform
    // (scalar).  The ticket is drawn by all lanes adding (lane == 0) -- one atomic add of 1
    // after the wave reduction -- and read from lane 0; the per-system parameters go through
    // readfirstlane; the info store is done by every lane (same address, same value).  (A
    // branch on `lane == 0` both at the top and at the bottom of the body lets the compiler
    // merge the two across the back edge and split the loop so that lanes leave it one at a
    // time, with the others re-reading a ticket they never drew: that never terminates.)
    // Each pass draws one ticket and tickets only grow, so no worker needs more than batch + 1
    // passes; the bound keeps the loop finite whatever the counter holds.
    for (int64_t pass = 0; pass <= batch; ++pass) {
        int t = atomicAdd(counter, l == 0 ? 1 : 0);
        t = __builtin_amdgcn_readlane(t, 0);
        if (t >= batch) break;
        const int i = __builtin_amdgcn_readfirstlane(order[t]);
        const int pre = __builtin_amdgcn_readfirstlane(d_info[i]);  // argument error set by the plan
        if (pre == 0) {
            const int n = __builtin_amdgcn_readfirstlane(d_n[i]);
            const int lda = __builtin_amdgcn_readfirstlane(d_lda[i]);
            const int nr = __builtin_amdgcn_readfirstlane(d_nrhs[i]);
            const int ldb = __builtin_amdgcn_readfirstlane(d_ldb[i]);
            const uint64_t pa = (uint64_t)(uintptr_t)d_A[i], pb = (uint64_t)(uintptr_t)d_B[i];
            T *A = (T *)(uintptr_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(pa >> 32)) << 32) |
                                    (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pa));
            T *B = (T *)(uintptr_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(pb >> 32)) << 32) |
                                    (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pb));
            const int info = vb_system<T, G>(n, A, lda, nr, B, ldb, lds[w][0], lds[w][1], Xg);
            d_info[i] = info;
        }
    }
}

int64_t vb_scratch_elems(int grid) { return (int64_t)grid * kVbWaves * kVbMaxT * 256; }

template <typename T>
void launch_vbatched(int64_t batch, const int *d_n, T *const *d_A, const int *d_lda, const int *d_nrhs,
                     T *const *d_B, const int *d_ldb, int *d_info, int *order, int *counter, T *scratch, int grid,
                     int group, int lpt, hipStream_t s)
{
    hipLaunchKernelGGL(vb_plan_kernel, dim3(1), dim3(1024), 0, s, batch, d_n, d_lda, d_nrhs, d_ldb, d_info, order,
                       lpt, counter);
    if (group == 2)
        hipLaunchKernelGGL((vb_persistent_kernel<T, 2>), dim3(grid), dim3(64 * kVbWaves), 0, s, batch, d_n, d_A,
                           d_lda, d_nrhs, d_B, d_ldb, d_info, order, counter, scratch);
    else
        hipLaunchKernelGGL((vb_persistent_kernel<T, 4>), dim3(grid), dim3(64 * kVbWaves), 0, s, batch, d_n, d_A,
                           d_lda, d_nrhs, d_B, d_ldb, d_info, order, counter, scratch);
}

template void launch_vbatched<double>(int64_t, const int *, double *const *, const int *, const int *,
                                      double *const *, const int *, int *, int *, int *, double *, int, int, int,
                                      hipStream_t);
template void launch_vbatched<float>(int64_t, const int *, float *const *, const int *, const int *, float *const *,
                                     const int *, int *, int *, int *, float *, int, int, int, hipStream_t);

}  // namespace mi
