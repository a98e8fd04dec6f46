// Wavefront-level helpers for one-point-per-wavefront kernels (CDNA, wave64).
#pragma once

#include <hip/hip_runtime.h>

namespace emx {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// All 64 lanes receive the same (bitwise identical) sum: a xor butterfly adds
// the same pair of operands on both partners at every level.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

// value of `v` held by lane `src` (src uniform across the wave)
__device__ __forceinline__ double bcast(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ int bcast_i(int v, int src) { return __shfl(v, src, 64); }

// Order LDS traffic between lanes of one wavefront: wait for this wave's LDS
// operations and keep the compiler from moving LDS accesses across the point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace emx
