// Wavefront-level helpers for one-point-per-wavefront kernels (CDNA, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace emx {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
// index of this wavefront within its workgroup, as a wave-uniform (scalar) value
__device__ __forceinline__ int wave_in_block() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
}

// ---- DPP data movement on 64-bit values (two 32-bit halves moved with one pattern).
// `identity` is written where the pattern has no source lane.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v, double identity) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t o = __builtin_bit_cast(uint64_t, identity);
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)o, (int)(uint32_t)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(o >> 32), (int)(uint32_t)(b >> 32), CTRL,
                                               0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The same moves for reductions whose identity is +0.0: lanes without a source read 0
// (bound_ctrl), and the destination's previous contents are not needed, so no copy of an
// identity value precedes each move.
template <int CTRL>
__device__ __forceinline__ double dpp_move0(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// A DPP move that writes only the rows in ROWMASK; lanes without a source, and the rows
// left out, read +0.0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_move_rows0(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWMASK, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Inclusive prefix sum over the 64 lanes in lane order: shifts by 1, 2, 4, 8 within each
// row of 16 lanes, then row 0's total into row 1 and row 2's into row 3 (row_bcast:15),
// then rows 0-1's total into rows 2 and 3 (row_bcast:31).  Fixed combination order.
__device__ __forceinline__ double wave_inclusive_scan(double v) {
    v += dpp_move_rows0<0x111, 0xF>(v);  // row_shr:1
    v += dpp_move_rows0<0x112, 0xF>(v);  // row_shr:2
    v += dpp_move_rows0<0x114, 0xF>(v);  // row_shr:4
    v += dpp_move_rows0<0x118, 0xF>(v);  // row_shr:8
    v += dpp_move_rows0<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_move_rows0<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// value of `v` held by lane `src` (src uniform across the wave): v_readlane
__device__ __forceinline__ double bcast(double v, int src) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int bcast_i(int v, int src) { return __builtin_amdgcn_readlane(v, src); }

// Reductions over all 64 lanes: quad permutes, row shifts and row broadcasts leave the
// total in lane 63, which is then read into a scalar, so every lane receives the same
// bits.  The combination order is fixed (independent of the data).
template <class Op>
__device__ __forceinline__ double wave_reduce(double v, double identity, Op op) {
    v = op(v, dpp_move<0xB1>(v, identity));   // quad_perm [1,0,3,2]
    v = op(v, dpp_move<0x4E>(v, identity));   // quad_perm [2,3,0,1]
    v = op(v, dpp_move<0x114>(v, identity));  // row_shr:4
    v = op(v, dpp_move<0x118>(v, identity));  // row_shr:8
    v = op(v, dpp_move<0x142>(v, identity));  // row_bcast:15
    v = op(v, dpp_move<0x143>(v, identity));  // row_bcast:31
    return bcast(v, 63);
}

// wave_reduce with identity +0.0 (sums; maxima of values whose result is only used
// through fmax(result, positive floor))
template <class Op>
__device__ __forceinline__ double wave_reduce0(double v, Op op) {
    v = op(v, dpp_move0<0xB1>(v));
    v = op(v, dpp_move0<0x4E>(v));
    v = op(v, dpp_move0<0x114>(v));
    v = op(v, dpp_move0<0x118>(v));
    v = op(v, dpp_move0<0x142>(v));
    v = op(v, dpp_move0<0x143>(v));
    return bcast(v, 63);
}

struct AddOp {
    __device__ __forceinline__ double operator()(double a, double b) const { return a + b; }
};
struct MaxOp {
    __device__ __forceinline__ double operator()(double a, double b) const { return fmax(a, b); }
};
struct MinOp {
    __device__ __forceinline__ double operator()(double a, double b) const { return fmin(a, b); }
};

__device__ __forceinline__ double wave_sum(double v) { return wave_reduce0(v, AddOp()); }
// max(0, max over lanes): equal to fmax(floor, wave_max(v)) for any floor > 0
__device__ __forceinline__ double wave_max0(double v) { return wave_reduce0(v, MaxOp()); }
__device__ __forceinline__ double wave_max(double v) { return wave_reduce(v, -__builtin_huge_val(), MaxOp()); }
__device__ __forceinline__ double wave_min(double v) { return wave_reduce(v, __builtin_huge_val(), MinOp()); }

__device__ __forceinline__ int wave_or(int v) { return __ballot(v != 0) != 0ull; }

// Order LDS traffic between lanes of one wavefront: wait for this wave's LDS
// operations and keep the compiler from moving LDS accesses across the point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace emx
