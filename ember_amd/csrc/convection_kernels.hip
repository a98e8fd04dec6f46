// Convection stage of the split step: ConvectionSystemSplit::integrateToTime
// (src/convectionSystem.cpp:588-627) on the GPU.
//
// Two kernels, as the reference orders the work:
//   utw_kernel      one workgroup integrates the coupled (U, T, Wmx) system of all grid
//                   points (ConvectionSystemUTW::f, :14-140, continuity condition
//                   fixedLeft) and records v = V / rho at its start and after every
//                   accepted step (vInterp);
//   species_kernel  one workgroup per species integrates Y_k' = -v(t) dY_k/dx + splitConst
//                   (ConvectionSystemY::f, :318-366) with v(t) interpolated from the
//                   record as update_v does (:378-403: the interval starts at the first
//                   recorded time >= t, one interval back at the end of the record).
// CVODE (Adams, bandwidth-0 Newton) is replaced by the explicit Dormand-Prince 5(4) pair
// with scipy RK45's step control (error norm rms(err / (atol + rtol max(|y|, |y_new|))),
// factors 0.9 err^(-1/5) in [0.2, 10], no growth after a rejection, select_initial_step),
// with Ember's tolerances; the CPU oracle runs the same method.
//
// Layout: a workgroup of kThreads threads; thread t owns the contiguous grid points
// [t P, (t + 1) P), P = ceil(N / kThreads).  The integrator's vectors live in a global
// scratch as structure of arrays (component c = var * N + j), each thread touching only
// its own components; a right-hand side copies its input into LDS so that every thread
// can read its neighbours, and the continuity equation's running sum is a workgroup scan
// over the threads' chunk sums.
#include <hip/hip_runtime.h>

#include "../../include/ember_amd.h"
#include "emx_common.h"
#include "kernels.h"
#include "wave.h"

namespace emx {

namespace {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;

// Dormand-Prince 5(4): nodes, stage matrix (the last row are the 5th-order weights),
// error weights
__constant__ double kRkC[7] = {0.0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
__constant__ double kRkA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
__constant__ double kRkE[7] = {-71.0 / 57600, 0, 71.0 / 16695, -71.0 / 1920, 17253.0 / 339200, -22.0 / 525,
                               1.0 / 40};

// Sum over the workgroup (every thread gets the same bits; fixed order)
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < kWaves; i++) s += red[i];
    return s;
}

// Exclusive scan of the threads' values in thread order (LDS, Hillis-Steele)
__device__ double block_exclusive_scan(double v, double* buf) {
    const int t = threadIdx.x;
    __syncthreads();
    buf[t] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const double add = t >= off ? buf[t - off] : 0.0;
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    return buf[t] - v;
}

__device__ __forceinline__ double strain_a(const ConvParams& p, double t) {
    return (t <= p.T0) ? p.aInitial : (t >= p.T0 + p.Dt) ? p.aFinal
                                                        : p.aInitial + (p.aFinal - p.aInitial) * (t - p.T0) / p.Dt;
}
__device__ __forceinline__ double strain_dadt(const ConvParams& p, double t) {
    return (t >= p.T0 && t < p.T0 + p.Dt) ? (p.aFinal - p.aInitial) / p.Dt : 0.0;
}

// This thread's components of an nvar x N structure-of-arrays vector: points [j0, j1)
// of every variable.
struct Comps {
    int N, nvar, j0, j1;
    template <class Fn>
    __device__ __forceinline__ void each(Fn&& fn) const {
        for (int v = 0; v < nvar; v++)
            for (int j = j0; j < j1; j++) fn(v * N + j);
    }
};

// y' = f(t, y) over [t0, t1] by one workgroup (scipy RK45 semantics).  y, tmp: nc
// components; K: 7 stage vectors of nc.  f(t, in, out, last) evaluates the right-hand
// side for the whole workgroup (`last`: the stage at the end of the step, whose input is
// the new solution); after(t) runs after each accepted step.  Returns the accepted steps,
// -1 beyond max_steps, -3 on a step-size underflow, -4 on a non-finite error estimate.
template <class F, class After>
__device__ int rk45_block(F& f, After& after, const Comps& cm, double* y, double* tmp, double* K,
                          const double* atol, double rtol, double t0, double t1, int max_steps, double* red) {
    const size_t nc = (size_t)cm.N * cm.nvar;
    const double ncd = (double)nc;
    double t = t0;
    f(t, y, K, false);
    double s0 = 0.0, s1 = 0.0;
    cm.each([&](int c) {
        const double sc = atol[c] + fabs(y[c]) * rtol;
        s0 += (y[c] / sc) * (y[c] / sc);
        s1 += (K[c] / sc) * (K[c] / sc);
    });
    const double d0 = sqrt(block_sum(s0, red) / ncd), d1 = sqrt(block_sum(s1, red) / ncd);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    h0 = fmin(h0, t1 - t0);
    cm.each([&](int c) { tmp[c] = y[c] + h0 * K[c]; });
    f(t + h0, tmp, K + nc, false);
    double s2 = 0.0;
    cm.each([&](int c) {
        const double sc = atol[c] + fabs(y[c]) * rtol;
        const double q = (K[nc + c] - K[c]) / sc;
        s2 += q * q;
    });
    const double d2 = sqrt(block_sum(s2, red) / ncd) / h0;
    const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / fmax(d1, d2), 1.0 / 5.0);
    double h_abs = fmin(100 * h0, fmin(h1, t1 - t0));
    int nsteps = 0;
    while (t < t1) {
        const double min_step = 10 * fabs(nextafter(t, __builtin_huge_val()) - t);
        if (h_abs < min_step) h_abs = min_step;
        bool rejected = false;
        double t_new;
        while (true) {
            if (h_abs < min_step) return -3;
            t_new = t + h_abs;
            if (t_new - t1 > 0) t_new = t1;
            const double h = t_new - t;
            h_abs = fabs(h);
            for (int s = 1; s < 7; s++) {
                cm.each([&](int c) {
                    double dy = 0.0;
                    for (int q = 0; q < s; q++) dy += kRkA[s][q] * K[q * nc + c];
                    tmp[c] = y[c] + h * dy;
                });
                f(t + kRkC[s] * h, tmp, K + s * nc, s == 6);
            }
            double se = 0.0;
            cm.each([&](int c) {
                double e = 0.0;
                for (int q = 0; q < 7; q++) e += kRkE[q] * K[q * nc + c];
                const double sc = atol[c] + fmax(fabs(y[c]), fabs(tmp[c])) * rtol;
                const double r = e * h / sc;
                se += r * r;
            });
            const double err = sqrt(block_sum(se, red) / ncd);
            if (isfinite(err) && err < 1.0) {
                double factor = (err == 0.0) ? 10.0 : fmin(10.0, 0.9 * pow(err, -1.0 / 5.0));
                if (rejected) factor = fmin(1.0, factor);
                h_abs *= factor;
                cm.each([&](int c) {
                    y[c] = tmp[c];
                    K[c] = K[6 * nc + c];  // first same as last
                });
                break;
            }
            if (!isfinite(err)) return -4;
            h_abs *= fmax(0.2, 0.9 * pow(err, -1.0 / 5.0));
            rejected = true;
        }
        t = t_new;
        if (++nsteps > max_steps) return -1;
        after(t);
    }
    return nsteps;
}

__device__ __forceinline__ void owned_points(int N, int& j0, int& j1) {
    const int P = (N + kThreads - 1) / kThreads;
    j0 = min(N, (int)threadIdx.x * P);
    j1 = min(N, j0 + P);
}

// ---------------------------------------------------------------------------
// UTW system
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) utw_kernel(ConvParams p, ConvArrays a) {
    __shared__ double sU[kConvMaxPoints], sT[kConvMaxPoints], sW[kConvMaxPoints];
    __shared__ double scan[kThreads];
    __shared__ double red[kWaves];
    const int N = p.N, jj = N - 1;
    int j0, j1;
    owned_points(N, j0, j1);
    const Comps cm{N, 3, j0, j1};
    const size_t nc = 3 * (size_t)N;
    double* y = a.work;
    double* tmp = y + nc;
    double* atol = tmp + nc;
    double* K = atol + nc;  // 7 x nc
    for (int j = j0; j < j1; j++) {
        y[j] = a.state[(size_t)j * p.nv + 0];
        y[N + j] = a.state[(size_t)j * p.nv + 1];
        y[2 * N + j] = a.Wmx[j];
        atol[j] = p.atolU;
        atol[N + j] = p.atolT;
        atol[2 * N + j] = p.atolW;
    }
    int nrec = 0;
    // ConvectionSystemUTW::f (continuity fixedLeft); with `rec`, v = V / rho of this
    // evaluation goes to record slot nrec (kept only if the step is accepted)
    auto f = [&](double t, const double* in, double* out, bool rec) {
        __syncthreads();  // the previous evaluation's LDS reads are done
        double csum = 0.0;
        for (int j = j0; j < j1; j++) {
            const double U = in[j], T = in[N + j], W = in[2 * N + j];
            sU[j] = U;
            sT[j] = T;
            sW[j] = W;
            if (j < N - 1) {
                const double rho = p.P * W / (kGasConstant * T);
                const double hh = a.x[j + 1] - a.x[j];
                const double rph = p.alpha == 0 ? 1.0 : 0.5 * (a.x[j] + a.x[j + 1]);
                csum += hh * rph * (a.drhodt[j] + rho * p.beta * U);
            }
        }
        double rV = p.rVzero - block_exclusive_scan(csum, scan);  // rV at j0 (the scan syncs)
        const double at = strain_a(p, t), dadt = strain_dadt(p, t);
        const double strain = dadt / p.beta + at * at / (p.beta * p.beta);
        for (int j = j0; j < j1; j++) {
            const double U = sU[j], T = sT[j], W = sW[j];
            const double rho = p.P * W / (kGasConstant * T);
            const double V = (j == 0 || p.alpha == 0) ? rV : rV / a.x[j];
            double dU, dT, dW;
            if (j == 0) {
                dU = a.scU[0] - U * U + p.rhou / rho * strain;
                if (p.leftBC == EMX_BC_CONTROL_VOLUME || p.leftBC == EMX_BC_WALL_FLUX) {
                    const double x1 = a.x[1];
                    const double centerVol = (p.alpha == 0 ? x1 : x1 * x1) / (p.alpha + 1);
                    const double rVz = fmax(rV, 0.0);
                    dT = -rVz * (T - p.Tleft) / (rho * centerVol) + a.scT[0];
                    dW = -rVz * (W - p.Wleft) / (rho * centerVol) - W * W * a.scW[0];
                } else {
                    dT = a.scT[0];
                    dW = -W * W * a.scW[0];
                }
            } else if (j < jj) {
                double dUdx, dTdx, dWdx;
                if (rV < 0) {
                    const double hh = a.x[j + 1] - a.x[j];
                    dTdx = (sT[j + 1] - T) / hh;
                    dUdx = (sU[j + 1] - U) / hh;
                    dWdx = (sW[j + 1] - W) / hh;
                } else {
                    const double hh = a.x[j] - a.x[j - 1];
                    dTdx = (T - sT[j - 1]) / hh;
                    dUdx = (U - sU[j - 1]) / hh;
                    dWdx = (W - sW[j - 1]) / hh;
                }
                dU = -V * dUdx / rho - U * U + p.rhou / rho * strain + a.scU[j];
                dT = -V * dTdx / rho + a.scT[j];
                dW = -V * dWdx / rho - W * W * a.scW[j];
            } else if (rV < 0 || p.rightBC == EMX_BC_FIXED_VALUE) {
                dU = a.scU[jj] - U * U + p.rhou / rho * strain;
                dT = a.scT[jj];
                dW = -W * W * a.scW[jj];
            } else {
                const double hh = a.x[jj] - a.x[jj - 1];
                dU = a.scU[jj] - V * (U - sU[jj - 1]) / hh / rho - U * U + p.rhou / rho * strain;
                dT = a.scT[jj] - V * (T - sT[jj - 1]) / hh / rho;
                dW = -W * W * a.scW[jj] - V * (W - sW[jj - 1]) / hh / rho;
            }
            out[j] = dU;
            out[N + j] = dT;
            out[2 * N + j] = dW;
            if (rec) a.rec_v[(size_t)nrec * N + j] = V / rho;
            if (j < N - 1) {  // continuity sum to the next point
                const double hh = a.x[j + 1] - a.x[j];
                const double rph = p.alpha == 0 ? 1.0 : 0.5 * (a.x[j] + a.x[j + 1]);
                rV = rV - hh * rph * (a.drhodt[j] + rho * p.beta * U);
            }
        }
    };
    // slot 0: v at the stage start (integrateToTime's first insert)
    f(p.t0, y, K, true);
    if (threadIdx.x == 0) a.rec_t[0] = p.t0;
    nrec = 1;
    auto after = [&](double t) {
        // the accepted step's last stage evaluated f at the new solution into slot nrec
        if (threadIdx.x == 0) a.rec_t[nrec] = t;
        nrec++;
    };
    // a step's last stage writes slot (accepted steps + 1) before the step count is checked
    const int r = rk45_block(f, after, cm, y, tmp, K, atol, p.rtol, p.t0, p.t1, min(p.maxSteps, p.maxRec - 2),
                             red);
    if (threadIdx.x == 0) {
        a.steps[0] = r;
        a.nrec[0] = nrec;
    }
    if (r >= 0)
        for (int j = j0; j < j1; j++) {
            a.state_out[(size_t)j * p.nv + 0] = y[j];
            a.state_out[(size_t)j * p.nv + 1] = y[N + j];
        }
}

// ---------------------------------------------------------------------------
// species systems: workgroup k integrates Y_k
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) species_kernel(ConvParams p, ConvArrays a) {
    __shared__ double sY[kConvMaxPoints];
    __shared__ double red[kWaves];
    __shared__ int sInterval[2];
    const int k = blockIdx.x;
    const int N = p.N, jj = N - 1;
    int j0, j1;
    owned_points(N, j0, j1);
    const Comps cm{N, 1, j0, j1};
    const size_t nc = (size_t)N;
    double* y = a.work + (size_t)k * 10 * nc;
    double* tmp = y + nc;
    double* atol = tmp + nc;
    double* K = atol + nc;  // 7 x N
    const int nrec = a.nrec[0];
    const bool utw_ok = a.steps[0] >= 0;
    for (int j = j0; j < j1; j++) {
        y[j] = a.state[(size_t)j * p.nv + 2 + k];
        atol[j] = p.atolY;
    }
    const double Yl = a.Yleft ? a.Yleft[k] : 0.0;
    auto f = [&](double t, const double* in, double* out, bool) {
        __syncthreads();
        // update_v: iL = first recorded time >= t (the end of the record: the last one),
        // iR = iL + 1, one interval back when iR is past the end
        if (threadIdx.x == 0) {
            int iL = nrec - 1;
            for (int i = 0; i < nrec; i++)
                if (a.rec_t[i] >= t) {
                    iL = i;
                    break;
                }
            int iR = iL + 1;
            if (nrec > 1 && iR >= nrec) {
                iL--;
                iR--;
            }
            sInterval[0] = iL;
            sInterval[1] = nrec > 1 ? iR : iL;
        }
        for (int j = j0; j < j1; j++) sY[j] = in[j];
        __syncthreads();
        const int iL = sInterval[0], iR = sInterval[1];
        const double s = (iL == iR) ? 0.0 : (t - a.rec_t[iL]) / (a.rec_t[iR] - a.rec_t[iL]);
        const double* vL = a.rec_v + (size_t)iL * N;
        const double* vR = a.rec_v + (size_t)iR * N;
        for (int j = j0; j < j1; j++) {
            const double v = (iL == iR) ? vL[j] : vL[j] * (1 - s) + vR[j] * s;
            const double sc = a.sc ? a.sc[(size_t)j * p.nv + 2 + k] : 0.0;
            double yd;
            if (j == 0) {
                if (p.leftBC == EMX_BC_CONTROL_VOLUME || p.leftBC == EMX_BC_WALL_FLUX) {
                    const double x1 = a.x[1];
                    const double centerVol = (p.alpha == 0 ? x1 : x1 * x1) / (p.alpha + 1);
                    yd = -fmax(v, 0.0) * (sY[0] - Yl) / centerVol + sc;
                } else {
                    yd = sc;
                }
            } else if (j < jj) {
                const double dYdx = v < 0 ? (sY[j + 1] - sY[j]) / (a.x[j + 1] - a.x[j])
                                          : (sY[j] - sY[j - 1]) / (a.x[j] - a.x[j - 1]);
                yd = -v * dYdx + sc;
            } else if (v < 0 || p.rightBC == EMX_BC_FIXED_VALUE) {
                yd = sc;
            } else {
                yd = sc - v * (sY[jj] - sY[jj - 1]) / (a.x[jj] - a.x[jj - 1]);
            }
            out[j] = yd;
        }
    };
    auto after = [](double) {};
    int r = -6;  // the UTW system failed: no velocity record
    if (utw_ok) r = rk45_block(f, after, cm, y, tmp, K, atol, p.rtol, p.t0, p.t1, p.maxSteps, red);
    if (threadIdx.x == 0) a.steps[1 + k] = r;
    if (r >= 0)
        for (int j = j0; j < j1; j++) a.state_out[(size_t)j * p.nv + 2 + k] = y[j];
}

// ---------------------------------------------------------------------------
// Register path (N <= kConvRegMax): a workgroup of ceil(N / 64) waves, thread j owns grid
// point j; the integrator's vectors (solution, 7 stages) live in registers, the right-hand
// side exchanges neighbours through LDS, and the continuity sum is a wave scan by
// shuffles with the wave totals through LDS (two workgroup barriers per scan).
// ---------------------------------------------------------------------------
constexpr int kConvRecLds = 1024;  // record times the species kernel keeps in LDS
constexpr double kA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
constexpr double kC[7] = {0.0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
constexpr double kE[7] = {-71.0 / 57600, 0, 71.0 / 16695, -71.0 / 1920, 17253.0 / 339200, -22.0 / 525,
                          1.0 / 40};

// Workgroup-wide operations with one barrier each: the per-wave partials go to one of two
// LDS buffers chosen by the caller's alternating parity, so a buffer is rewritten only
// after the barrier of the operation in between (no barrier before the write).
// Sum over the workgroup, the same bits in every thread (wave sums, then the wave totals
// in a fixed order).
__device__ double wg_sum(double v, double* red, int& par) {
    v = wave_sum(v);
    double* r = red + 16 * (par ^= 1);
    if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = v;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += r[i];
    return s;
}

// Exclusive prefix sum in thread order (a DPP scan within each wave plus the totals of
// the preceding waves); `sync`: the barrier, which also publishes the caller's other LDS
// writes made before the call.
__device__ double wg_exclusive_scan(double v, double* wtot, int& par) {
    const double incl = wave_inclusive_scan(v);
    double* t = wtot + 16 * (par ^= 1);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) t[w] = incl;
    __syncthreads();
    double pre = 0.0;
    for (int i = 0; i < w; i++) pre += t[i];
    return pre + (incl - v);
}

// y' = f(t, y) over [t0, t1] with NV components per thread in registers (scipy RK45
// semantics, as rk45_block); `own`: the thread holds a grid point (the others carry zeros
// and take part in the workgroup operations); ncd: number of components in the system.
template <int NV, class F, class After>
__device__ int rk45_reg(F& f, After& after, bool own, double (&y)[NV], const double (&atol)[NV], double rtol,
                        double t0, double t1, int max_steps, double ncd, double* red) {
    int par = 0;  // LDS buffer parity of wg_sum (red holds two buffers of 16)
    double K[7][NV], tmp[NV];
    double t = t0;
    f(t, y, K[0], false);
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const double sc = atol[v] + fabs(y[v]) * rtol;
        if (own) {
            s0 += (y[v] / sc) * (y[v] / sc);
            s1 += (K[0][v] / sc) * (K[0][v] / sc);
        }
    }
    const double d0 = sqrt(wg_sum(s0, red, par) / ncd), d1 = sqrt(wg_sum(s1, red, par) / ncd);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    h0 = fmin(h0, t1 - t0);
#pragma unroll
    for (int v = 0; v < NV; v++) tmp[v] = y[v] + h0 * K[0][v];
    f(t + h0, tmp, K[1], false);
    double s2 = 0.0;
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const double sc = atol[v] + fabs(y[v]) * rtol;
        const double q = (K[1][v] - K[0][v]) / sc;
        if (own) s2 += q * q;
    }
    const double d2 = sqrt(wg_sum(s2, red, par) / ncd) / h0;
    const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / fmax(d1, d2), 1.0 / 5.0);
    double h_abs = fmin(100 * h0, fmin(h1, t1 - t0));
    int nsteps = 0;
    while (t < t1) {
        const double min_step = 10 * fabs(nextafter(t, __builtin_huge_val()) - t);
        if (h_abs < min_step) h_abs = min_step;
        bool rejected = false;
        double t_new;
        while (true) {
            if (h_abs < min_step) return -3;
            t_new = t + h_abs;
            if (t_new - t1 > 0) t_new = t1;
            const double h = t_new - t;
            h_abs = fabs(h);
#pragma unroll
            for (int s = 1; s < 7; s++) {
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    double dy = 0.0;
#pragma unroll
                    for (int q = 0; q < s; q++) dy += kA[s][q] * K[q][v];
                    tmp[v] = y[v] + h * dy;
                }
                f(t + kC[s] * h, tmp, K[s], s == 6);
            }
            double se = 0.0;
#pragma unroll
            for (int v = 0; v < NV; v++) {
                double e = 0.0;
#pragma unroll
                for (int q = 0; q < 7; q++) e += kE[q] * K[q][v];
                const double sc = atol[v] + fmax(fabs(y[v]), fabs(tmp[v])) * rtol;
                const double r = e * h / sc;
                if (own) se += r * r;
            }
            const double err = sqrt(wg_sum(se, red, par) / ncd);
            if (isfinite(err) && err < 1.0) {
                double factor = (err == 0.0) ? 10.0 : fmin(10.0, 0.9 * pow(err, -1.0 / 5.0));
                if (rejected) factor = fmin(1.0, factor);
                h_abs *= factor;
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    y[v] = tmp[v];
                    K[0][v] = K[6][v];  // first same as last
                }
                break;
            }
            if (!isfinite(err)) return -4;
            h_abs *= fmax(0.2, 0.9 * pow(err, -1.0 / 5.0));
            rejected = true;
        }
        t = t_new;
        if (++nsteps > max_steps) return -1;
        after(t);
    }
    return nsteps;
}

// MAXT: the workgroup size bound (512: a register budget of 256 per thread, no spills;
// 1024 for grids of 513 to 1024 points).  Wmx and the split constants come from
// conv_prep_kernel.  The neighbour exchange alternates between two LDS buffers, so the
// continuity scan's barrier is the only one of a right-hand side.
template <int MAXT>
__global__ void __launch_bounds__(MAXT) utw_reg_kernel(ConvParams p, ConvArrays a) {
    __shared__ double sN[2][3][MAXT];
    __shared__ double wtot[2 * 16], red[2 * 16];
    const int N = p.N, jj = N - 1, j = threadIdx.x, nv = p.nv;
    const bool own = j < N;
    const int jc = own ? j : 0;  // clamped index for the constant loads
    const double xj = a.x[jc];
    const double hhm = jc > 0 ? xj - a.x[jc - 1] : 1.0;
    const double hhp = jc < N - 1 ? a.x[jc + 1] - xj : 1.0;
    const double rph = p.alpha == 0 ? 1.0 : (jc < N - 1 ? 0.5 * (xj + a.x[jc + 1]) : 1.0);
    const double drho = a.drhodt ? a.drhodt[jc] : 0.0;
    const double scU = a.scU[jc], scT = a.scT[jc], scW = a.scW[jc];
    const double x1 = a.x[1];
    // reciprocals of the grid factors, so that a right-hand side divides only by rho
    const double rhhm = 1.0 / hhm, rhhp = 1.0 / hhp;
    const double rcv = (p.alpha + 1) / (p.alpha == 0 ? x1 : x1 * x1);  // 1 / centerVol
    const double rxj = 1.0 / xj;
    double y[3] = {own ? a.state[(size_t)j * nv + 0] : 0.0, own ? a.state[(size_t)j * nv + 1] : 1.0,
                   own ? a.Wmx[j] : 1.0};
    const double atol[3] = {p.atolU, p.atolT, p.atolW};
    int nrec = 0, par = 0;
    // ConvectionSystemUTW::f (continuity fixedLeft); `rec`: v = V / rho of this evaluation
    // goes to record slot nrec (kept only if the step is accepted)
    auto f = [&](double t, const double (&in)[3], double (&out)[3], bool rec) {
        const double U = in[0], T = in[1], W = in[2];
        double* sU = sN[par ^ 1][0];
        double* sT = sN[par ^ 1][1];
        double* sW = sN[par ^ 1][2];
        if (own) {
            sU[j] = U;
            sT[j] = T;
            sW[j] = W;
        }
        const double rho = p.P * W / (kGasConstant * T);
        const double rrho = 1.0 / rho;
        const double cs = (own && j < N - 1) ? hhp * rph * (drho + rho * p.beta * U) : 0.0;
        // rV at this point: rVzero minus the continuity sum over the points before it (the
        // scan's barrier also publishes the neighbour values written above)
        const double rV = p.rVzero - wg_exclusive_scan(cs, wtot, par);
        const double at = strain_a(p, t), dadt = strain_dadt(p, t);
        const double strain = dadt / p.beta + at * at / (p.beta * p.beta);
        const double V = (j == 0 || p.alpha == 0) ? rV : rV * rxj;
        const double sr = p.rhou * rrho * strain;
        double dU = 0.0, dT = 0.0, dW = 0.0;
        if (own) {
            if (j == 0) {
                dU = scU - U * U + sr;
                if (p.leftBC == EMX_BC_CONTROL_VOLUME || p.leftBC == EMX_BC_WALL_FLUX) {
                    const double rVz = fmax(rV, 0.0) * rrho * rcv;
                    dT = -rVz * (T - p.Tleft) + scT;
                    dW = -rVz * (W - p.Wleft) - W * W * scW;
                } else {
                    dT = scT;
                    dW = -W * W * scW;
                }
            } else if (j < jj) {
                double dUdx, dTdx, dWdx;
                if (rV < 0) {
                    dTdx = (sT[j + 1] - T) * rhhp;
                    dUdx = (sU[j + 1] - U) * rhhp;
                    dWdx = (sW[j + 1] - W) * rhhp;
                } else {
                    dTdx = (T - sT[j - 1]) * rhhm;
                    dUdx = (U - sU[j - 1]) * rhhm;
                    dWdx = (W - sW[j - 1]) * rhhm;
                }
                const double vr = V * rrho;
                dU = -vr * dUdx - U * U + sr + scU;
                dT = -vr * dTdx + scT;
                dW = -vr * dWdx - W * W * scW;
            } else if (rV < 0 || p.rightBC == EMX_BC_FIXED_VALUE) {
                dU = scU - U * U + sr;
                dT = scT;
                dW = -W * W * scW;
            } else {
                const double vr = V * rhhm * rrho;
                dU = scU - vr * (U - sU[jj - 1]) - U * U + sr;
                dT = scT - vr * (T - sT[jj - 1]);
                dW = -W * W * scW - vr * (W - sW[jj - 1]);
            }
            if (rec) a.rec_v[(size_t)nrec * N + j] = V * rrho;
        }
        out[0] = dU;
        out[1] = dT;
        out[2] = dW;
    };
    // slot 0: v at the stage start (integrateToTime's first insert)
    {
        double k0[3];
        f(p.t0, y, k0, true);
    }
    if (threadIdx.x == 0) a.rec_t[0] = p.t0;
    nrec = 1;
    auto after = [&](double t) {
        if (threadIdx.x == 0) a.rec_t[nrec] = t;
        nrec++;
    };
    const int r = rk45_reg<3>(f, after, own, y, atol, p.rtol, p.t0, p.t1, min(p.maxSteps, p.maxRec - 2),
                              3.0 * N, red);
    if (threadIdx.x == 0) {
        a.steps[0] = r;
        a.nrec[0] = nrec;
    }
    if (r >= 0 && own) {
        a.state_out[(size_t)j * nv + 0] = y[0];
        a.state_out[(size_t)j * nv + 1] = y[1];
    }
}

// Inputs of the register path, one wavefront per grid point (lane = species): Wmx of the
// state (Phase::setMassFractions: clip, normalise; then the mean molecular weight) and the
// UTW split constants U, T and sum_k splitConst_Yk / W_k (setSplitConstants).
__global__ void __launch_bounds__(256) conv_prep_kernel(ConvParams p, const double* __restrict__ state,
                                                        const double* __restrict__ sc,
                                                        const double* __restrict__ rW, double* Wmx,
                                                        double* scU, double* scT, double* scW) {
    const int lane = lane_id();
    const int j = blockIdx.x * (blockDim.x / 64) + wave_in_block();
    if (j >= p.N) return;
    const bool spl = lane >= 2 && lane < p.nv;
    const double yv = lane < p.nv ? state[(size_t)j * p.nv + lane] : 0.0;
    const double sv = (sc && lane < p.nv) ? sc[(size_t)j * p.nv + lane] : 0.0;
    const double rw = spl ? rW[lane - 2] : 0.0;
    const double yc = spl ? fmax(yv, 0.0) : 0.0;
    const double norm = wave_sum(yc);
    const double ym = yc * (1.0 / norm) * rw;
    const double mmw = 1.0 / wave_sum(ym);
    const double w = wave_sum(spl ? sv * rw : 0.0);
    if (lane == 0) {
        Wmx[j] = mmw;
        scU[j] = sv;
        scW[j] = w;
    }
    if (lane == 1) scT[j] = sv;
}

__global__ void __launch_bounds__(kConvRegMax) species_reg_kernel(ConvParams p, ConvArrays a) {
    __shared__ double sY[2][kConvRegMax];
    __shared__ double red[2 * 16];
    __shared__ int sIv[2][2];
    __shared__ double sRt[kConvRecLds];  // the record's times (the first kConvRecLds)
    const int k = blockIdx.x, N = p.N, jj = N - 1, j = threadIdx.x, nv = p.nv;
    const bool own = j < N;
    const int jc = own ? j : 0;
    const double xj = a.x[jc];
    const double hhm = jc > 0 ? xj - a.x[jc - 1] : 1.0;
    const double hhp = jc < N - 1 ? a.x[jc + 1] - xj : 1.0;
    const double sc = (own && a.sc) ? a.sc[(size_t)j * nv + 2 + k] : 0.0;
    const double Yl = a.Yleft ? a.Yleft[k] : 0.0;
    const int nrec = a.nrec[0];
    const bool utw_ok = a.steps[0] >= 0;
    const double x1 = a.x[1];
    const double rhhm = 1.0 / hhm, rhhp = 1.0 / hhp;
    const double rcv = (p.alpha + 1) / (p.alpha == 0 ? x1 : x1 * x1);  // 1 / centerVol
    for (int i = j; i < min(nrec, kConvRecLds); i += blockDim.x) sRt[i] = a.rec_t[i];
    __syncthreads();
    auto rec_t = [&](int i) { return i < kConvRecLds ? sRt[i] : a.rec_t[i]; };
    double y[1] = {own ? a.state[(size_t)j * nv + 2 + k] : 0.0};
    const double atol[1] = {p.atolY};
    int cL = -1, cR = -1, par = 0;
    double vL = 0.0, vR = 0.0;  // this point's velocities at the interval ends
    auto f = [&](double t, const double (&in)[1], double (&out)[1], bool) {
        par ^= 1;
        // update_v: iL = first recorded time >= t (the end of the record: the last one),
        // iR = iL + 1, one interval back when iR is past the end
        if (threadIdx.x == 0) {
            int iL = nrec - 1;
            for (int i = 0; i < nrec; i++)
                if (rec_t(i) >= t) {
                    iL = i;
                    break;
                }
            int iR = iL + 1;
            if (nrec > 1 && iR >= nrec) {
                iL--;
                iR--;
            }
            sIv[par][0] = iL;
            sIv[par][1] = nrec > 1 ? iR : iL;
        }
        double* Ys = sY[par];
        if (own) Ys[j] = in[0];
        __syncthreads();
        const int iL = sIv[par][0], iR = sIv[par][1];
        if (iL != cL || iR != cR) {  // uniform: the record rows of a new interval
            vL = own ? a.rec_v[(size_t)iL * N + j] : 0.0;
            vR = own ? a.rec_v[(size_t)iR * N + j] : 0.0;
            cL = iL;
            cR = iR;
        }
        const double s = (iL == iR) ? 0.0 : (t - rec_t(iL)) / (rec_t(iR) - rec_t(iL));
        const double v = (iL == iR) ? vL : vL * (1 - s) + vR * s;
        const double Y = in[0];
        double yd = 0.0;
        if (own) {
            if (j == 0) {
                if (p.leftBC == EMX_BC_CONTROL_VOLUME || p.leftBC == EMX_BC_WALL_FLUX) {
                    yd = -fmax(v, 0.0) * rcv * (Y - Yl) + sc;
                } else {
                    yd = sc;
                }
            } else if (j < jj) {
                const double dYdx = v < 0 ? (Ys[j + 1] - Y) * rhhp : (Y - Ys[j - 1]) * rhhm;
                yd = -v * dYdx + sc;
            } else if (v < 0 || p.rightBC == EMX_BC_FIXED_VALUE) {
                yd = sc;
            } else {
                yd = sc - v * rhhm * (Y - Ys[jj - 1]);
            }
        }
        out[0] = yd;
    };
    auto after = [](double) {};
    int r = -6;  // the UTW system failed: no velocity record
    if (utw_ok) r = rk45_reg<1>(f, after, own, y, atol, p.rtol, p.t0, p.t1, p.maxSteps, (double)N, red);
    if (threadIdx.x == 0) a.steps[1 + k] = r;
    if (r >= 0 && own) a.state_out[(size_t)j * nv + 2 + k] = y[0];
}

}  // namespace

size_t convection_work_doubles(int N, int K) {
    // UTW: y, tmp, atol, 7 stages of 3N; species: 10 N each
    return (size_t)10 * 3 * N + (size_t)10 * N * K;
}

int launch_convection(const ConvParams& p, const ConvArrays& a, hipStream_t stream) {
    if (p.N <= kConvRegMax) {
        const int threads = (p.N + 63) / 64 * 64;
        hipLaunchKernelGGL(conv_prep_kernel, dim3((p.N + 3) / 4), dim3(256), 0, stream, p, a.state, a.sc, a.rW,
                           const_cast<double*>(a.Wmx), const_cast<double*>(a.scU), const_cast<double*>(a.scT),
                           const_cast<double*>(a.scW));
        if (threads <= 512)
            hipLaunchKernelGGL(utw_reg_kernel<512>, dim3(1), dim3(threads), 0, stream, p, a);
        else
            hipLaunchKernelGGL(utw_reg_kernel<kConvRegMax>, dim3(1), dim3(threads), 0, stream, p, a);
        hipLaunchKernelGGL(species_reg_kernel, dim3(p.K), dim3(threads), 0, stream, p, a);
        return (int)hipGetLastError();
    }
    ConvArrays au = a;
    hipLaunchKernelGGL(utw_kernel, dim3(1), dim3(kThreads), 0, stream, p, au);
    ConvArrays as = a;
    as.work = a.work + (size_t)10 * 3 * p.N;
    hipLaunchKernelGGL(species_kernel, dim3(p.K), dim3(kThreads), 0, stream, p, as);
    return (int)hipGetLastError();
}

}  // namespace emx
