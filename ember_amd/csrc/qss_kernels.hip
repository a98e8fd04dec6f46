// Batched chemistry production stage, QSS integrator: one wavefront per grid point,
// persistent waves pulling points from a device work queue.
//
// Per point this is SourceSystemQSS::setState / integrateToTime
// (src/sourceSystem.cpp:434-443, src/sourceSystem.h:252) driving
// QssIntegrator::integrateToTime / integrateOneStep (src/qssintegrator.cpp:98-265)
// with SourceSystemQSS::odefun (src/sourceSystem.cpp:445-499) as the ODE.  The caller
// policy of FlameSolver::integrateProductionTerms (src/flameSolver.cpp:794-834) is
// applied in the kernel: a failed point keeps its input state.
#include <hip/hip_runtime.h>

#include "chem_device.h"
#include "kernels.h"

namespace emx {

namespace {

// Per-lane view of SourceSystemQSS: the variable this lane holds plus the
// thermo quantities refreshed on non-corrector calls.
struct QssSystem {
    double W, rW;      // species lanes
    double sc;         // splitConst for this lane's variable
    double qloss;
    double T;          // temperature used by the rates (frozen in correctors)
    double rho, cp;    // from the last non-corrector call (SourceSystem::updateThermo)
    double hk;         // J/kmol, species lanes
    double qdot;
};

template <int NR, bool INTERP>
__device__ inline bool qss_odefun(const DevMech& m, int lane, double y, bool corrector,
                                  QssSystem& S, RateRegs<NR>& R, const double* perturb,
                                  const WaveLds& L, double& q, double& d) {
    if (!corrector) S.T = bcast(y, 1);  // unroll_y: T not updated in correctors
    GasState g = set_state(m, lane, y, S.rW, S.T, false, L);
    bool ok = true;
    if (!corrector) {
        double hRT, cpR;
        ok = update_T<NR, INTERP>(m, lane, g, R, L, hRT, cpR);
        // SourceSystem::updateThermo: h_k = RT h/RT, rho, cp_mass
        S.hk = is_species_lane(lane, m) ? g.T * kGasConstant * hRT : 0.0;
        S.rho = g.dens;
        double s = wave_sum(g.ym * cpR);
        double cp_mole = kGasConstant * (g.mmw * s);
        S.cp = cp_mole / g.mmw;
    }
    rates_of_progress<NR>(m, lane, g, R, perturb, L);
    double wQ, wD;
    species_rates(m, lane, L, wQ, wD);
    S.qdot = -wave_sum((wQ - wD) * S.hk);
    const double rho = S.rho, cp = S.cp;
    if (lane == 0) {
        q = S.sc;
        d = 0.0;
    } else if (lane == 1) {
        q = S.qdot / (rho * cp) + S.sc;
        d = S.qloss / (rho * cp);
    } else if (is_species_lane(lane, m)) {
        q = wQ * S.W / rho + S.sc;
        d = wD * S.W / rho;
    } else {
        q = 0.0;
        d = 0.0;
    }
    return ok;
}

__device__ __forceinline__ double isign(double x) { return (x > 0) ? 1.0 : (x < 0) ? -1.0 : 0.0; }

// QssIntegrator state for this lane's component
struct QssLane {
    double y, q, d, rtau, rtaus, y1, ys, qs, scratch, ym1, ym2;
    double ymin;
    bool enforce;
};

constexpr double kNegInf = -__builtin_huge_val();

// Returns 0 on success, -1 for a too-small step (QssIntegrator), -2 for T out of the
// interpolation range, -3 when the internal step budget is exhausted, -4 on NaN.
//
// QssIntegrator::integrateToTime / integrateOneStep written as a state machine with
// a single call site of the ODE function: the predictor evaluation starts a step,
// every further evaluation is a corrector.  The control flow and arithmetic follow
// src/qssintegrator.cpp:98-265 line by line.
template <int NR, bool INTERP>
__device__ int qss_integrate(const DevMech& m, const DevOpts& o, int lane, bool active,
                             QssLane& Z, QssSystem& S, RateRegs<NR>& R, const double* perturb,
                             const WaveLds& L, double tf, int& gcount, int& rcount,
                             double& qint) {
    const double tfd = 1.000008;
    const bool stab_on = o.qss_stability != 0;
    double tn = 0.0, ts = 0.0, dt = 0.0;
    bool firstStep = true;
    gcount = 0;
    rcount = 0;
    qint = 0.0;
    if (!(tfd * tn < tf)) return 0;
    long budget = 20000000L;
    bool predict = true;
    int iter = 0;
    double qdot_step = 0.0;
    while (true) {
        if (--budget < 0) return -3;
        if (!qss_odefun<NR, INTERP>(m, lane, Z.y, !predict, S, R, perturb, L, Z.q, Z.d)) return -2;
        gcount += 1;
        bool start_attempt = false;
        if (predict) {
            // ---- integrateOneStep: start of a step
            qdot_step = S.qdot;
            if (firstStep) {
                firstStep = false;
                double c = kNegInf;
                if (active && fabs(Z.y) > o.qss_abstol) {
                    double absq = fabs(Z.q);
                    double scr2 = fabs(1 / Z.y) * isign(0.1 * o.qss_epsmin * absq - Z.d);
                    double scr1 = scr2 * Z.d;
                    c = fmax(scr1, -fabs(absq - Z.d) * scr2);
                }
                double scrtch = fmax(wave_max(c), 1.0e-25);
                const double sqreps = 0.5;
                dt = fmin(sqreps / scrtch, tf);
                dt = fmin(dt, o.qss_dtmax);
            }
            ts = tn;
            Z.rtau = Z.enforce ? dt * Z.d / Z.y : 0.0;
            Z.qs = Z.q;
            Z.ys = Z.y;
            Z.rtaus = Z.rtau;
            predict = false;
            start_attempt = true;
        } else {
            // ---- after a corrector evaluation
            double rtaub = Z.enforce ? 0.5 * (Z.rtaus + dt * Z.d / Z.y) : 0.0;
            double alpha = (180. + rtaub * (60. + rtaub * (11. + rtaub))) /
                           (360. + rtaub * (60. + rtaub * (12. + rtaub)));
            Z.scratch = (Z.qs * (1.0 - alpha) + Z.q * alpha - Z.ys * rtaub / dt) /
                        (1.0 + alpha * rtaub);
            iter++;
            if (iter < o.qss_itermax) {
                if (stab_on) {
                    Z.ym2 = Z.ym1;
                    Z.ym1 = Z.y;
                }
                Z.y = fmax(Z.ymin, Z.ys + dt * Z.scratch);
                continue;
            }
            // convergence test and step size control
            double eps = 1.0e-10;
            double c = kNegInf;
            {
                double scr2 = Z.ys + dt * Z.scratch;
                if (Z.enforce) scr2 = fmax(scr2, Z.ymin);
                double scr1 = fabs(scr2 - Z.y1);
                Z.y = fmax(scr2, Z.ymin);
                if (active && fabs(Z.y) > o.qss_abstol && 0.25 * (Z.ys + Z.y) > Z.ymin) {
                    scr1 = scr1 / Z.y;
                    c = .5 * (scr1 + fmin(fabs(Z.q - Z.d) / (Z.q + Z.d + 1e-30), scr1));
                }
            }
            eps = fmax(eps, wave_max(c));
            if (stab_on) {
                Z.ym2 = Z.ym1;
                Z.ym1 = Z.y;
            }
            eps /= o.qss_epsmin;
            if (dt <= o.qss_dtmin + 1e-16 * tn) return -1;
            if (!(eps == eps) || !(dt == dt)) return -4;
            double stab = 0;
            if (stab_on && o.qss_itermax >= 3) {
                double sc = kNegInf;
                if (active && fabs(Z.y) > o.qss_abstol)
                    sc = fabs(Z.y - Z.ym1) / (fabs(Z.ym1 - Z.ym2) + 1e-20 * Z.y);
                stab = fmax(0.01, wave_max(sc));
            }
            const bool ok = eps <= o.qss_epsmax && stab <= 1.0;
            if (ok) {
                qint += dt * qdot_step;
                if (tf <= tn * tfd) return 0;
            } else {
                tn = ts;
            }
            double rteps = 0.5 * (eps + 1.0);
            rteps = 0.5 * (rteps + eps / rteps);
            rteps = 0.5 * (rteps + eps / rteps);
            double dto = dt;
            dt = fmin(dt * (1.0 / rteps + 0.005), tfd * (tf - tn));
            dt = fmin(dt, o.qss_dtmax);
            if (stab_on) dt = fmin(dt, dto / (stab + .001));
            if (!ok) {
                rcount += 1;
                Z.rtaus *= dt / dto;
                start_attempt = true;
            } else {
                // accepted: integrateToTime's loop condition
                if (!(tfd * tn < tf)) return 0;
                predict = true;
                continue;
            }
        }
        if (start_attempt) {
            // predictor terms, then the first corrector iterate
            double r = Z.rtau;
            Z.scratch = (Z.q - Z.d) /
                        (1.0 + r * (180 + r * (60 + r * (11 + r))) / (360 + r * (60 + r * (12 + r))));
            iter = 0;
            if (stab_on) {
                Z.ym2 = Z.ym1;
                Z.ym1 = Z.y;
            }
            Z.y = fmax(Z.ymin, Z.ys + dt * Z.scratch);
            tn = ts + dt;
            Z.y1 = Z.y;
        }
    }
}

// OCC = waves per SIMD the register allocation targets: 2 (up to 8 waves per workgroup,
// no spills) or 3 (9..12 waves per workgroup, register budget 168).
template <int NR, bool INTERP, int OCC>
__global__ void __launch_bounds__(OCC == 2 ? 512 : 768, OCC) qss_stage_kernel(
    DevMech mg, DevOpts o, const double* __restrict__ state_in,
    const double* __restrict__ split_const, const double* __restrict__ qloss,
    const double* __restrict__ perturb_g, double dt, int nPoints, const int* __restrict__ order,
    unsigned int* __restrict__ queue, double* __restrict__ state_out, int* __restrict__ status,
    int* __restrict__ stat1, int* __restrict__ stat2, int* __restrict__ cost,
    double* __restrict__ qint_out) {
    extern __shared__ double lds[];
    const double* perturb;
    const DevMech m = stage_mech(mg, perturb_g, lds, perturb);
    __syncthreads();
    const int lane = lane_id();
    const int wave = threadIdx.x / 64;
    const int per_wave = wave_lds_doubles(m.nRounds, m.nChunkRounds);
    WaveLds L = carve_lds(lds + mech_lds_doubles(m) + wave * per_wave, m.nRounds, m.nChunkRounds);
    init_wave_lds(L, lane, m.nRounds);
    const int nv = m.nVars;
    const bool active = lane < nv;
    QssSystem S;
    S.W = is_species_lane(lane, m) ? m.W[lane - 2] : 0.0;
    S.rW = is_species_lane(lane, m) ? m.rW[lane - 2] : 0.0;
    RateRegs<NR> R;
    while (true) {
        unsigned int t = 0;
        if (lane == 0) t = atomicAdd(queue, 1u);
        t = (unsigned int)bcast_i((int)t, 0);
        if (t >= (unsigned int)nPoints) break;
        const int p = order ? order[t] : (int)t;
        const double* yin = state_in + (size_t)p * nv;
        double y0 = active ? yin[lane] : 0.0;
        S.sc = (active && split_const) ? split_const[(size_t)p * nv + lane] : 0.0;
        S.qloss = qloss ? qloss[p] : 0.0;
        R.T = -1.0;
        QssLane Z;
        Z.enforce = active && lane != 0;
        Z.ymin = (lane == 0) ? -1e4 : o.qss_minval;
        // QssIntegrator::setState
        Z.y = Z.enforce ? fmax(y0, Z.ymin) : y0;
        Z.ym1 = Z.ym2 = 0.0;
        int g = 0, r = 0;
        double qi = 0.0;
        int err = qss_integrate<NR, INTERP>(m, o, lane, active, Z, S, R, perturb, L, dt, g, r, qi);
        double* yout = state_out + (size_t)p * nv;
        if (active) yout[lane] = (err >= 0) ? Z.y : y0;
        if (lane == 0) {
            if (status) status[p] = err;
            if (stat1) stat1[p] = g;
            if (stat2) stat2[p] = r;
            if (cost) cost[p] = g;
            if (qint_out) qint_out[p] = qi;
        }
    }
}

__global__ void reset_queue_kernel(unsigned int* q) {
    if (threadIdx.x == 0) *q = 0u;
}

// Rates probe: creation / destruction / net production rates at given states
template <int NR, bool INTERP>
__global__ void __launch_bounds__(256) rates_kernel(DevMech mg, const double* __restrict__ T,
                                                    const double* __restrict__ Y, int nonorm,
                                                    const double* __restrict__ perturb_g, int n,
                                                    double* wQ_out, double* wD_out,
                                                    double* wnet_out, int* status) {
    extern __shared__ double lds[];
    const double* perturb;
    const DevMech m = stage_mech(mg, perturb_g, lds, perturb);
    __syncthreads();
    const int lane = lane_id();
    const int wave = threadIdx.x / 64;
    const int per_wave = wave_lds_doubles(m.nRounds, m.nChunkRounds);
    WaveLds L = carve_lds(lds + mech_lds_doubles(m) + wave * per_wave, m.nRounds, m.nChunkRounds);
    init_wave_lds(L, lane, m.nRounds);
    const int p = blockIdx.x * (blockDim.x / 64) + wave;
    if (p >= n) return;
    const int K = m.K;
    double y = is_species_lane(lane, m) ? Y[(size_t)p * K + lane - 2] : 0.0;
    double rW = is_species_lane(lane, m) ? m.rW[lane - 2] : 0.0;
    RateRegs<NR> R;
    R.T = -1.0;
    GasState g = set_state(m, lane, y, rW, T[p], nonorm != 0, L);
    double hRT, cpR;
    bool ok = update_T<NR, INTERP>(m, lane, g, R, L, hRT, cpR);
    rates_of_progress<NR>(m, lane, g, R, perturb, L);
    double wQ, wD;
    species_rates(m, lane, L, wQ, wD);
    if (is_species_lane(lane, m)) {
        size_t o = (size_t)p * K + lane - 2;
        if (wQ_out) wQ_out[o] = wQ;
        if (wD_out) wD_out[o] = wD;
        if (wnet_out) wnet_out[o] = wQ - wD;
    }
    if (lane == 0 && status) status[p] = ok ? 0 : -2;
}

}  // namespace

// compile-time reaction-round bound for a mechanism with nRounds rounds
#define EMX_DISPATCH_NR(nr, BODY)                  \
    switch ((nr) <= 1 ? 1 : (nr) <= 2 ? 2 : (nr) <= 4 ? 4 : (nr) <= 6 ? 6 : 8) { \
        case 1: { constexpr int NR = 1; BODY; } break; \
        case 2: { constexpr int NR = 2; BODY; } break; \
        case 4: { constexpr int NR = 4; BODY; } break; \
        case 6: { constexpr int NR = 6; BODY; } break; \
        default: { constexpr int NR = 8; BODY; } break; \
    }

size_t qss_lds_bytes(const DevMech& m, int waves_per_block) {
    return ((size_t)mech_lds_doubles(m) + (size_t)waves_per_block * wave_lds_doubles(m.nRounds, m.nChunkRounds)) * sizeof(double);
}

int launch_qss_stage(const DevMech& m, const DevOpts& o, const double* state_in,
                     const double* split_const, const double* qloss, const double* perturb,
                     double dt, int nPoints, const int* order, unsigned int* queue,
                     double* state_out, int* status, int* stat1, int* stat2, int* cost, double* qint,
                     int grid_blocks, int waves_per_block, hipStream_t stream) {
    size_t lds = ((size_t)mech_lds_doubles(m) + (size_t)waves_per_block * wave_lds_doubles(m.nRounds, m.nChunkRounds)) * sizeof(double);
    if (waves_per_block < 1 || waves_per_block > 12) return (int)hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, stream, queue);
#define EMX_QSS_LAUNCH(INTERP, OCC)                                                                    \
    hipLaunchKernelGGL((qss_stage_kernel<NR, INTERP, OCC>), dim3(grid_blocks), dim3(64 * waves_per_block), \
                       lds, stream, m, o, state_in, split_const, qloss, perturb, dt, nPoints, order,  \
                       queue, state_out, status, stat1, stat2, cost, qint)
    const bool occ3 = waves_per_block > 8;
    EMX_DISPATCH_NR(m.nRounds, {
        if (m.interp) {
            if (occ3) EMX_QSS_LAUNCH(true, 3); else EMX_QSS_LAUNCH(true, 2);
        } else {
            if (occ3) EMX_QSS_LAUNCH(false, 3); else EMX_QSS_LAUNCH(false, 2);
        }
    });
#undef EMX_QSS_LAUNCH
    return (int)hipGetLastError();
}

int launch_rates(const DevMech& m, const double* T, const double* Y, int nonorm,
                 const double* perturb, int n, double* wQ, double* wD, double* wnet, int* status,
                 hipStream_t stream) {
    const int wpb = 4;
    size_t lds = ((size_t)mech_lds_doubles(m) + (size_t)wpb * wave_lds_doubles(m.nRounds, m.nChunkRounds)) * sizeof(double);
    int blocks = (n + wpb - 1) / wpb;
    EMX_DISPATCH_NR(m.nRounds, {
        if (m.interp)
            hipLaunchKernelGGL((rates_kernel<NR, true>), dim3(blocks), dim3(64 * wpb), lds, stream, m, T,
                               Y, nonorm, perturb, n, wQ, wD, wnet, status);
        else
            hipLaunchKernelGGL((rates_kernel<NR, false>), dim3(blocks), dim3(64 * wpb), lds, stream, m, T,
                               Y, nonorm, perturb, n, wQ, wD, wnet, status);
    });
    return (int)hipGetLastError();
}

}  // namespace emx
