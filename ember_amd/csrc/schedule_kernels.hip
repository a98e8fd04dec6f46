// Cost-ordered work queue for the chemistry stage kernels.
//
// The stage kernels record a cost per point: the integration time of its last stage (the
// QSS, BDF and adjoint forward stages, in 640 ns units of the real-time counter).  Before the next stage over the same points,
// one workgroup bins the points by cost (1/16-octave bins, costs up to 2^16 resolved) and
// writes them in descending cost order, so the
// persistent waves take the most expensive points first (longest-processing-time
// first).  The order only decides which wave integrates which point and when; every
// point's result is independent of it.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.h"

namespace emx {

namespace {

constexpr int kBins = 256;

__device__ __forceinline__ int cost_bin(int c) {
    if (c <= 1) return 0;
    const int b = (int)(16.0f * __log2f((float)c));
    return b < kBins - 1 ? b : kBins - 1;
}

__global__ void __launch_bounds__(1024) cost_order_kernel(const int* __restrict__ cost, int n,
                                                          int* __restrict__ order) {
    __shared__ int hist[kBins];
    __shared__ int start[kBins];
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&hist[cost_bin(cost[i])], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int b = kBins - 1; b >= 0; b--) {  // descending cost
            start[b] = acc;
            acc += hist[b];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int pos = atomicAdd(&start[cost_bin(cost[i])], 1);
        order[pos] = i;
    }
}

}  // namespace

int launch_cost_order(const int* cost, int n, int* order, hipStream_t stream) {
    hipLaunchKernelGGL(cost_order_kernel, dim3(1), dim3(1024), 0, stream, cost, n, order);
    return (int)hipGetLastError();
}

}  // namespace emx
