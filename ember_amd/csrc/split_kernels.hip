// Device-resident split step: the elementwise bookkeeping of SplitSolver
// (src/splitSolver.cpp:27-121) and FlameSolver::correctMassFractions
// (src/flameSolver.cpp:942-949), so that a whole Strang step runs on the GPU with the state
// resident between the diffusion and production stages.
//
// All arrays are Ember's nVars x nPoints (point-contiguous) matrices, processed as flat
// vectors of n = nVars * nPoints doubles.  The arithmetic is written operation by
// operation (products and sums rounded separately, no contraction) so that it matches the
// host-side SplitSolver bit for bit.
#include <hip/hip_runtime.h>

#include "chem_device.h"
#include "emx_common.h"
#include "kernels.h"
#include "wave.h"

namespace emx {

namespace {

constexpr int kSplitThreads = 256;

__device__ __forceinline__ size_t gid() {
    return (size_t)blockIdx.x * blockDim.x + threadIdx.x;
}

// SplitSolver::step (src/splitSolver.cpp:38-44): balanced split constants from the
// previous step's time derivatives; the convection stage is empty here (its delta stays
// zero), so this step's ddtConv = 0 / dt - splitConstConv = -splitConstConv is written
// directly.  Non-balanced: splitConstDiff = ddtCross, the other constants stay zero.
__global__ void split_constants_kernel(size_t n, int balanced, double* ddt_conv,
                                       const double* ddt_diff, const double* ddt_prod,
                                       const double* ddt_cross, double* sc_diff, double* sc_prod) {
    const size_t i = gid();
    if (i >= n) return;
    const double cross = ddt_cross ? ddt_cross[i] : 0.0;
    if (balanced) {
        const double conv = ddt_conv[i], diff = ddt_diff[i], prod = ddt_prod[i];
        // 0.25 * (prod + conv + cross - 3 * diff), each operation rounded
        sc_diff[i] = __dmul_rn(0.25, __dsub_rn(__dadd_rn(__dadd_rn(prod, conv), cross), __dmul_rn(3.0, diff)));
        const double scc = __dmul_rn(0.25, __dsub_rn(__dadd_rn(__dadd_rn(prod, diff), cross), __dmul_rn(3.0, conv)));
        sc_prod[i] = __dmul_rn(0.5, __dsub_rn(__dadd_rn(__dadd_rn(conv, diff), cross), prod));
        ddt_conv[i] = 0.0 - scc;
    } else {
        sc_diff[i] = cross;
        sc_prod[i] = 0.0;
        ddt_conv[i] = 0.0;
    }
}

// The same constants for a step with its convection stages: splitConstConv is written out
// (the convection stages read it) and ddtConv stays the previous step's until the step's
// own is formed from deltaConv (calculateTimeDerivatives); Strang: splitConstConv = 0.
__global__ void split_constants_conv_kernel(size_t n, int balanced, const double* ddt_conv,
                                            const double* ddt_diff, const double* ddt_prod,
                                            const double* ddt_cross, double* sc_diff, double* sc_prod,
                                            double* sc_conv) {
    const size_t i = gid();
    if (i >= n) return;
    const double cross = ddt_cross ? ddt_cross[i] : 0.0;
    if (balanced) {
        const double conv = ddt_conv[i], diff = ddt_diff[i], prod = ddt_prod[i];
        sc_diff[i] = __dmul_rn(0.25, __dsub_rn(__dadd_rn(__dadd_rn(prod, conv), cross), __dmul_rn(3.0, diff)));
        sc_conv[i] = __dmul_rn(0.25, __dsub_rn(__dadd_rn(__dadd_rn(prod, diff), cross), __dmul_rn(3.0, conv)));
        sc_prod[i] = __dmul_rn(0.5, __dsub_rn(__dadd_rn(__dadd_rn(conv, diff), cross), prod));
    } else {
        sc_diff[i] = cross;
        sc_prod[i] = 0.0;
        sc_conv[i] = 0.0;
    }
}

// FlameSolver::prepareIntegrators (src/flameSolver.cpp:203-213): the density derivative of
// the convection stages from the previous step's split time derivatives,
// ddt = ddtConv + ddtDiff + ddtProd (+ ddtCross with balanced splitting), and the step's
// starting state: drhodt = -rho (ddt_T / T + Wmx sum_k ddt_Yk / W_k) (the time derivative of
// the ideal-gas density rho = P Wmx / (R T) at constant pressure).  One thread per point, the
// species sum in species order.
__global__ void density_derivative_kernel(int N, int nv, int balanced, const double* state,
                                          const double* ddt_conv, const double* ddt_diff,
                                          const double* ddt_prod, const double* ddt_cross,
                                          const double* rho, const double* Wmx, const double* rW,
                                          double* drhodt) {
    const size_t j = gid();
    if (j >= (size_t)N) return;
    const size_t o = j * nv;
    auto ddt = [&](int v) {
        double d = __dadd_rn(__dadd_rn(ddt_conv[o + v], ddt_diff[o + v]), ddt_prod[o + v]);
        if (balanced && ddt_cross) d = __dadd_rn(d, ddt_cross[o + v]);
        return d;
    };
    double tmp = 0.0;
    for (int k = 0; k < nv - 2; k++) tmp = __dadd_rn(tmp, __dmul_rn(rW[k], ddt(2 + k)));
    const double T = state[o + 1];
    drhodt[j] = __dmul_rn(-rho[j], __dadd_rn(__ddiv_rn(ddt(1), T), __dmul_rn(tmp, Wmx[j])));
}

// SplitSolver::calculateTimeDerivatives (src/splitSolver.cpp:29): ddtConv of a step with
// its convection stages
__global__ void conv_time_derivative_kernel(size_t n, double dt, const double* delta_conv,
                                            const double* sc_conv, double* ddt_conv) {
    const size_t i = gid();
    if (i >= n) return;
    ddt_conv[i] = __dsub_rn(__ddiv_rn(delta_conv[i], dt), sc_conv[i]);
}

// FlameSolver::prepareIntegrators (src/flameSolver.cpp:160-172): B of the momentum and
// species systems is rho.inverse(), of the energy system (rho * cp).inverse(); B holds rho
// in row 0 and cp in row 1 on entry
__global__ void diffusion_B_kernel(int n, double* B) {
    const size_t i = gid();
    if (i >= (size_t)n) return;
    const double rho = B[i], cp = B[n + i];
    B[i] = 1.0 / rho;
    B[n + i] = 1.0 / (rho * cp);
}

// ConvectionSystemSplit::setSplitConstants (src/convectionSystem.cpp:573-586)
__global__ void conv_split_constants_kernel(int N, int nv, const double* __restrict__ sc,
                                            const double* __restrict__ W, double* scU, double* scT,
                                            double* scW) {
    const size_t j = gid();
    if (j >= (size_t)N) return;
    if (!sc) {
        scU[j] = scT[j] = scW[j] = 0.0;
        return;
    }
    const double* r = sc + j * nv;
    scU[j] = r[0];
    scT[j] = r[1];
    double w = 0.0;
    for (int k = 0; k < nv - 2; k++) w += r[2 + k] / W[k];
    scW[j] = w;
}

__global__ void fill_kernel(size_t n, double v, double* a, double* b) {
    const size_t i = gid();
    if (i >= n) return;
    a[i] = v;
    if (b) b[i] = v;
}

__global__ void copy_kernel(size_t n, const double* src, double* dst) {
    const size_t i = gid();
    if (i < n) dst[i] = src[i];
}

// SplitSolver stage bookkeeping (src/splitSolver.cpp:77-121): delta += state - startState
__global__ void accumulate_kernel(size_t n, const double* state, const double* start, double* delta) {
    const size_t i = gid();
    if (i < n) delta[i] = __dadd_rn(delta[i], __dsub_rn(state[i], start[i]));
}

// SplitSolver::calculateTimeDerivatives (src/splitSolver.cpp:27-32) for the diffusion and
// production terms
__global__ void time_derivatives_kernel(size_t n, double dt, const double* delta_diff,
                                        const double* delta_prod, const double* sc_diff,
                                        const double* sc_prod, double* ddt_diff, double* ddt_prod) {
    const size_t i = gid();
    if (i >= n) return;
    ddt_diff[i] = __dsub_rn(__ddiv_rn(delta_diff[i], dt), sc_diff[i]);
    ddt_prod[i] = __dsub_rn(__ddiv_rn(delta_prod[i], dt), sc_prod[i]);
}

// FlameSolver::correctMassFractions: setStateMass then getMassFractions, i.e. Cantera's
// Phase::setMassFractions (negative mass fractions clipped to zero, then scaled by the
// reciprocal of their sum).  One wavefront per point, lane = state variable.
__global__ void __launch_bounds__(256) correct_mass_fractions_kernel(int nVars, int nPoints,
                                                                     double* state) {
    const int lane = lane_id();
    const int p = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (p >= nPoints) return;
    const bool sp = lane >= 2 && lane < nVars;
    const size_t o = (size_t)p * nVars + lane;
    const double y = sp ? fmax(state[o], 0.0) : 0.0;
    const double norm = wave_sum(y);
    if (sp) state[o] = y * (1.0 / norm);
}

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + kSplitThreads - 1) / kSplitThreads)); }

}  // namespace

int launch_split_constants(size_t n, int balanced, double* ddt_conv, const double* ddt_diff,
                           const double* ddt_prod, const double* ddt_cross, double* sc_diff,
                           double* sc_prod, hipStream_t st) {
    hipLaunchKernelGGL(split_constants_kernel, grid_for(n), dim3(kSplitThreads), 0, st, n, balanced,
                       ddt_conv, ddt_diff, ddt_prod, ddt_cross, sc_diff, sc_prod);
    return (int)hipGetLastError();
}

int launch_split_constants_conv(size_t n, int balanced, const double* ddt_conv, const double* ddt_diff,
                                const double* ddt_prod, const double* ddt_cross, double* sc_diff,
                                double* sc_prod, double* sc_conv, hipStream_t st) {
    hipLaunchKernelGGL(split_constants_conv_kernel, grid_for(n), dim3(kSplitThreads), 0, st, n, balanced,
                       ddt_conv, ddt_diff, ddt_prod, ddt_cross, sc_diff, sc_prod, sc_conv);
    return (int)hipGetLastError();
}

int launch_density_derivative(const DevMech& m, int N, int balanced, const double* state,
                              const double* ddt_conv, const double* ddt_diff, const double* ddt_prod,
                              const double* ddt_cross, const double* rho, const double* Wmx,
                              double* drhodt, hipStream_t st) {
    hipLaunchKernelGGL(density_derivative_kernel, grid_for((size_t)N), dim3(kSplitThreads), 0, st, N,
                       m.nVars, balanced, state, ddt_conv, ddt_diff, ddt_prod, ddt_cross, rho, Wmx, m.rW,
                       drhodt);
    return (int)hipGetLastError();
}

int launch_conv_time_derivative(size_t n, double dt, const double* delta_conv, const double* sc_conv,
                                double* ddt_conv, hipStream_t st) {
    hipLaunchKernelGGL(conv_time_derivative_kernel, grid_for(n), dim3(kSplitThreads), 0, st, n, dt,
                       delta_conv, sc_conv, ddt_conv);
    return (int)hipGetLastError();
}

int launch_diffusion_B(int n, double* B, hipStream_t stream) {
    const int blocks = (n + kSplitThreads - 1) / kSplitThreads;
    hipLaunchKernelGGL(diffusion_B_kernel, dim3(blocks), dim3(kSplitThreads), 0, stream, n, B);
    return (int)hipGetLastError();
}

int launch_conv_split_constants(const DevMech& m, int N, const double* sc, double* scU, double* scT,
                                double* scW, hipStream_t stream) {
    const int blocks = (N + kSplitThreads - 1) / kSplitThreads;
    hipLaunchKernelGGL(conv_split_constants_kernel, dim3(blocks), dim3(kSplitThreads), 0, stream, N, m.nVars,
                       sc, m.W, scU, scT, scW);
    return (int)hipGetLastError();
}

int launch_fill(size_t n, double v, double* a, double* b, hipStream_t st) {
    hipLaunchKernelGGL(fill_kernel, grid_for(n), dim3(kSplitThreads), 0, st, n, v, a, b);
    return (int)hipGetLastError();
}

int launch_copy(size_t n, const double* src, double* dst, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, grid_for(n), dim3(kSplitThreads), 0, st, n, src, dst);
    return (int)hipGetLastError();
}

int launch_accumulate(size_t n, const double* state, const double* start, double* delta,
                      hipStream_t st) {
    hipLaunchKernelGGL(accumulate_kernel, grid_for(n), dim3(kSplitThreads), 0, st, n, state, start,
                       delta);
    return (int)hipGetLastError();
}

int launch_time_derivatives(size_t n, double dt, const double* delta_diff, const double* delta_prod,
                            const double* sc_diff, const double* sc_prod, double* ddt_diff,
                            double* ddt_prod, hipStream_t st) {
    hipLaunchKernelGGL(time_derivatives_kernel, grid_for(n), dim3(kSplitThreads), 0, st, n, dt,
                       delta_diff, delta_prod, sc_diff, sc_prod, ddt_diff, ddt_prod);
    return (int)hipGetLastError();
}

int launch_correct_mass_fractions(int nVars, int nPoints, double* state, hipStream_t st) {
    hipLaunchKernelGGL(correct_mass_fractions_kernel, dim3((nPoints + 3) / 4), dim3(256), 0, st,
                       nVars, nPoints, state);
    return (int)hipGetLastError();
}

}  // namespace emx
