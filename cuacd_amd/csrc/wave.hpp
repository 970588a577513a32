// wave.hpp - wave64 primitives for gfx950 (CDNA4).
//
// Every kernel in this library is written for 64-lane wavefronts: ballots are
// 64-bit masks, reductions are 6 butterfly steps, and a single-wave workgroup
// uses the LDS as wave-private memory with __syncthreads() as its fence.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cuacd {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane (exclusive prefix of a ballot).
__device__ __forceinline__ int mask_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

template <class T> __device__ __forceinline__ T uniform(T v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ int read_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float read_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint64_t w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint64_t w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Inclusive prefix sum across the wave (Hillis-Steele over 64 lanes).
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int w = __shfl_up(v, o);
        if (l >= o) v += w;
    }
    return v;
}

// Order-preserving float <-> uint mapping (total order incl. negatives).
__device__ __forceinline__ uint32_t fkey(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Device-scope atomics used by the heap and the work queues.
template <class T> __device__ __forceinline__ T ld_acquire(T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ bool cas_acq_rel(T* p, T& expected, T desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQ_REL,
                                                __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T fetch_add_relaxed(T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// XCD-aware workgroup order. The dispatcher hands consecutive workgroups of a
// launch to the 8 XCDs in turn; this renumbers workgroup b of nb so that each
// XCD runs groups of G consecutive task slots (groups still go round the XCDs,
// keeping the load spread). Tasks that read the same data (the candidate
// planes of one part) then share one XCD's L2 instead of fetching the data
// into several. A bijection of [0, nb); the last nb mod 8G workgroups keep
// their own index.
__device__ __forceinline__ unsigned xcd_group_slot(unsigned b, unsigned nb, unsigned G) {
    const unsigned span = 8u * G;
    const unsigned full = nb / span * span;
    if (b >= full) return b;
    const unsigned x = b % 8u, j = b / 8u;
    return ((j / G) * 8u + x) * G + (j % G);
}

// err == nullptr: the caller handles the failure itself (nothing recorded)
__device__ __forceinline__ void err_or(int* err, int bits) {
    if (bits && err) __hip_atomic_fetch_or(err, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace cuacd
