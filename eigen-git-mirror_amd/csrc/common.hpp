// common.hpp -- device helpers shared by the LLT kernels (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // 16 raw bytes

// Arithmetic of the trailing updates (the values of mi_prec in mi_llt.h).
enum : int { kPrecF64 = 0, kPrecF32 = 1, kPrecBF16 = 2, kPrecBF16x3 = 3, kPrecBF16x6 = 4 };

// FP64 MFMA 16x16x4.  Lane layout of v_mfma_f64_16x16x4_f64, measured on gfx950:
//   operand a: lane l supplies a[i = l%16][k = l/16]
//   operand b: lane l supplies b[k = l/16][j = l%16]
//   result   : lane l, element r holds d[i = 4r + l/16][j = l%16]
// The kernels compute C[i][j] += sum_k P[i][k] Q[j][k] by passing Q as `a` and P as `b`:
// both operands are then loaded as x[row0 + l%16][k0 + l/16], and lane l, element r of
// the result is C[i0 + l%16][j0 + 4r + l/16] -- 16 consecutive rows of one column per
// group of 16 lanes, which coalesces on column-major stores.
__device__ __forceinline__ dvec4 mfma_f64(double q, double p, dvec4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(q, p, c, 0, 0, 0);
}

// FP32 MFMA 16x16x4 and BF16 MFMA 16x16x32 (FP32 accumulation).  Lane layout on gfx950:
//   operand a: lane l supplies a[i = l%16][k = E*(l/16) + e], e < E (E = 1 for FP32, 8 for BF16)
//   operand b: lane l supplies b[k = E*(l/16) + e][j = l%16]
//   result   : lane l, element r holds d[i = 4*(l/16) + r][j = l%16]
// With Q passed as `a` and P as `b` (C[i][j] += sum_k P[i][k] Q[j][k]), lane l, element r of
// the result is C[i0 + l%16][j0 + 4*(l/16) + r].
__device__ __forceinline__ fvec4 mfma_f32(float q, float p, fvec4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(q, p, c, 0, 0, 0);
}
__device__ __forceinline__ fvec4 mfma_bf16(bf16x8 q, bf16x8 p, fvec4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(q, p, c, 0, 0, 0);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Workgroup barrier after which the global stores of every wave are complete (in L2): each
// wave drains its vector-memory counter (s_waitcnt 0) before the barrier.  Needed (1) before
// waves read global data that other waves of the workgroup stored, and (2) before one thread
// publishes the workgroup's stores to other workgroups with an agent-scope release, whose L2
// writeback then covers all of them.  A plain __syncthreads() orders LDS only here: it does
// not wait for outstanding global stores.
__device__ __forceinline__ void global_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Load of data that another workgroup published (a flag release after its stores): an
// agent-scope atomic load is served from the coherence point, so it cannot hit a stale line
// in this CU's L1 or this XCD's L2.  (One thread's acquire after the flag invalidates the
// caches, but the other waves of the workgroup, which only pass a barrier, are not ordered
// after that invalidation.)
template <typename T>
__device__ __forceinline__ T coherent_load(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup barrier ordering LDS accesses only (outstanding global loads stay in flight).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Early-exit test used by every kernel of a factorization: a previous kernel on the
// stream recorded a failing pivot.  Uniform across the workgroup.
__device__ __forceinline__ bool factor_failed(const int *info)
{
    return info != nullptr && __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// The handle's status word (timeouts, inconsistent tickets) is set: the kernels of the
// factorization and the solve then exit without work until the host reads and clears it.
__device__ __forceinline__ bool device_failed(const int *status)
{
    return status != nullptr && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// Lower-triangular tile enumeration: t -> (ti, tj), ti >= tj, row-major over the lower
// triangle (t = ti*(ti+1)/2 + tj).
__device__ __forceinline__ void lower_tile(int t, int &ti, int &tj)
{
    int i = (int)((__fsqrt_rn(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);  // estimate, then exact fix-up
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i > 0 && i * (i + 1) / 2 > t) --i;
    ti = i;
    tj = t - i * (i + 1) / 2;
}
