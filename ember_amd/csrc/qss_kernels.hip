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

#include "kernels.h"
#include "qss_device.h"

namespace emx {

namespace {

// OCC = waves per SIMD the register allocation targets: 2 (up to 8 waves per workgroup,
// no spills) or 3 (9..12 waves per workgroup, register budget 168).  EXT: Ember's rate
// multiplier function (o.ratefn) or the igniter (o.qign) is on, or the mechanism has
// reactions with explicit forward orders.
template <int NR, bool INTERP, int OCC, bool EXT>
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
    const int wave = wave_in_block();
    const int per_wave = wave_lds_doubles(m.nRounds, m.nChunkRounds);
    WaveLds L = carve_lds(lds + mech_lds_doubles(m) + wave * per_wave, m.nRounds, m.nChunkRounds);
    init_wave_lds(L, lane, m.nRounds, m.nChunkRounds);
    const int nv = m.nVars;
    const bool active = lane < nv;
    QssSystem S;
    S.prio = 0;
    S.W = is_species_lane(lane, m) ? m.W[lane - 2] : 0.0;
    S.rW = is_species_lane(lane, m) ? m.rW[lane - 2] : 0.0;
    RateRegs<NR> R;
    // queue == NULL: one point per wave, p = global wave index (non-persistent launch)
    bool taken = false;
    while (true) {
        unsigned int t = 0;
        if (queue) {
            if (lane == 0) t = atomicAdd(queue, 1u);
            t = (unsigned int)bcast_i((int)t, 0);
        } else {
            if (taken) break;
            taken = true;
            t = blockIdx.x * (blockDim.x / 64) + wave;
        }
        if (t >= (unsigned int)nPoints) break;
        const int p = order ? order[t] : (int)t;
        const double* yin = state_in + (size_t)p * nv;
        double y0 = active ? yin[lane] : 0.0;
        S.sc = (active && split_const) ? split_const[(size_t)p * nv + lane] : 0.0;
        S.qloss = qloss ? qloss[p] : 0.0;
        S.qign = (EXT && o.qign) ? o.qign[p] : 0.0;
        R.T = -1.0;
        QssLane Z;
        Z.enforce = active && lane != 0;
        Z.ymin = (lane == 0) ? -1e4 : o.qss_minval;
        // QssIntegrator::setState
        Z.y = Z.enforce ? fmax(y0, Z.ymin) : y0;
        Z.ym1 = Z.ym2 = 0.0;
        int g = 0, r = 0;
        double qi = 0.0;
        const unsigned long long t_point = point_clock();
        NoHooks H;
        int err = kInvalidState;
        if (valid_input_state(lane, active, y0))
            err = qss_integrate<NR, INTERP, NoHooks, EXT>(m, o, lane, active, Z, S, R, perturb, L, dt,
                                                          QssResume{0.0, 0.0, true}, H, g, r, qi);
        double* yout = state_out + (size_t)p * nv;
        if (active) yout[lane] = (err >= 0) ? Z.y : y0;
        if (lane == 0) {
            if (status) status[p] = err;
            if (stat1) stat1[p] = g;
            if (stat2) stat2[p] = r;
            if (cost) cost[p] = point_cost(t_point);
            if (qint_out) qint_out[p] = qi;
        }
    }
}

__global__ void reset_queue_kernel(unsigned int* q) {
    if (threadIdx.x == 0) *q = 0u;
}

// Rates probe: creation / destruction / net production rates at given states
template <int NR, bool INTERP, bool ORD>
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
    const int wave = wave_in_block();
    const int per_wave = wave_lds_doubles(m.nRounds, m.nChunkRounds);
    WaveLds L = carve_lds(lds + mech_lds_doubles(m) + wave * per_wave, m.nRounds, m.nChunkRounds);
    init_wave_lds(L, lane, m.nRounds, m.nChunkRounds);
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
    rates_of_progress<NR, ORD>(m, lane, g, R, perturb, 1.0, L);
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

// FlameSolver::calculateQdot (src/flameSolver.cpp:926-939) per point: the rate multiplier
// function's a(t_now) when one is set, setStateMass, getEnthalpies, getReactionRates,
// qdot = -sum_k wdot_k h_k [W/m^3].  One wavefront per point over Ember's
// point-contiguous state (U, T, Y).
template <int NR, bool INTERP, bool ORD>
__global__ void __launch_bounds__(256) heat_release_kernel(DevMech mg, const double* __restrict__ state,
                                                           const double* __restrict__ perturb_g, int n,
                                                           DevOpts o, double* qdot, int* status) {
    extern __shared__ double lds[];
    const double* perturb;
    const DevMech m = stage_mech(mg, perturb_g, lds, perturb);
    __syncthreads();
    const int lane = lane_id();
    const int wave = wave_in_block();
    const int per_wave = wave_lds_doubles(m.nRounds, m.nChunkRounds);
    WaveLds L = carve_lds(lds + mech_lds_doubles(m) + wave * per_wave, m.nRounds, m.nChunkRounds);
    init_wave_lds(L, lane, m.nRounds, m.nChunkRounds);
    const int p = blockIdx.x * (blockDim.x / 64) + wave;
    if (p >= n) return;
    const int nv = m.nVars;
    const bool sp = is_species_lane(lane, m);
    const double y = lane < nv ? state[(size_t)p * nv + lane] : 0.0;
    const double T = bcast(y, 1);
    const double rW = sp ? m.rW[lane - 2] : 0.0;
    RateRegs<NR> R;
    R.T = -1.0;
    GasState g = set_state(m, lane, y, rW, T, false, L);
    double hRT, cpR;
    const bool ok = update_T<NR, INTERP>(m, lane, g, R, L, hRT, cpR);
    rates_of_progress<NR, ORD>(m, lane, g, R, perturb, fn_mult(o, 0.0), L);
    double wQ, wD;
    species_rates(m, lane, L, wQ, wD);
    const double hk = sp ? hRT * kGasConstant * T : 0.0;  // J/kmol
    const double q = -wave_sum(sp ? (wQ - wD) * hk : 0.0);
    if (lane == 0) {
        qdot[p] = q;
        if (status) status[p] = ok ? 0 : -2;
    }
}

// FlameSolver::updateChemicalProperties, thermodynamic block (src/flameSolver.cpp:667-673):
// CanteraGas::setStateMass (Phase::setMassFractions clips negatives and normalises), then
// density, meanMolecularWeight, cp_mass, getPartialMolarCp and getPartialMolarEnthalpies,
// with the arithmetic of the stage kernels' set_state, nasa() and odefun thermo.  Lane =
// state variable; each wavefront takes kThermoPointsPerWave consecutive points, issuing
// all their state loads first and reading its species' NASA7 coefficients (both ranges)
// once.  No LDS.
constexpr int kThermoPointsPerWave = 4;

__global__ void __launch_bounds__(256) thermo_props_kernel(DevMech m, const double* __restrict__ state,
                                                           int n, double* __restrict__ rho,
                                                           double* __restrict__ Wmx,
                                                           double* __restrict__ cp,
                                                           double* __restrict__ cpSpec,
                                                           double* __restrict__ hk) {
    constexpr int PPW = kThermoPointsPerWave;
    const int lane = lane_id();
    const long p0 = ((long)blockIdx.x * (blockDim.x / 64) + wave_in_block()) * PPW;
    if (p0 >= n) return;
    const int nv = m.nVars, K = nv - 2;
    const bool sp = is_species_lane(lane, m);
    double y[PPW];
#pragma unroll
    for (int i = 0; i < PPW; i++)
        y[i] = (lane < nv && p0 + i < n) ? state[(size_t)(p0 + i) * nv + lane] : 0.0;
    const int k = sp ? lane - 2 : 0;
    double lo[6], hi[6];
#pragma unroll
    for (int j = 0; j < 6; j++) {
        lo[j] = sp ? m.nasa[14 * k + j] : 0.0;
        hi[j] = sp ? m.nasa[14 * k + 7 + j] : 0.0;
    }
    const double Tmid = m.Tmid[k];
    const double rW = sp ? m.rW[k] : 0.0;
#pragma unroll
    for (int i = 0; i < PPW; i++) {
        const long p = p0 + i;
        if (p >= n) break;  // uniform: p is the wave's
        const double T = bcast(y[i], 1);
        const double yc = sp ? fmax(y[i], 0.0) : 0.0;
        const double norm = wave_sum(yc);
        const double ym = sp ? yc * (1.0 / norm) * rW : 0.0;
        const double mmw = 1.0 / wave_sum(ym);
        // NasaPoly1::updateProperties (cp/R and h/RT), as nasa()
        const bool low = T <= Tmid;
        double c[6];
#pragma unroll
        for (int j = 0; j < 6; j++) c[j] = low ? lo[j] : hi[j];
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
        const double ct0 = c[0];
        const double ct1 = c[1] * T;
        const double ct2 = c[2] * T2;
        const double ct3 = c[3] * T3;
        const double ct4 = c[4] * T4;
        const double cpR = ct0 + ct1 + ct2 + ct3 + ct4;
        const double hRT = ct0 + 0.5 * ct1 + 1.0 / 3.0 * ct2 + 0.25 * ct3 + 0.2 * ct4 + c[5] * (1.0 / T);
        const double s = wave_sum(ym * cpR);
        if (lane == 0) {
            rho[p] = m.P * mmw / (T * kGasConstant);
            if (Wmx) Wmx[p] = mmw;
            cp[p] = kGasConstant * (mmw * s) / mmw;
        }
        if (sp) {
            if (cpSpec) cpSpec[(size_t)p * K + k] = kGasConstant * cpR;  // J/kmol/K
            if (hk) hk[(size_t)p * K + k] = hRT * kGasConstant * T;      // J/kmol
        }
    }
}

}  // namespace

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
    if (queue) hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, stream, queue);
#define EMX_QSS_LAUNCH(INTERP, OCC, EXT)                                                              \
    hipLaunchKernelGGL((qss_stage_kernel<NR, INTERP, OCC, EXT>), dim3(grid_blocks),                    \
                       dim3(64 * waves_per_block), lds, stream, m, o, state_in, split_const, qloss,   \
                       perturb, dt, nPoints, order, queue, state_out, status, stat1, stat2, cost, qint)
    const bool occ3 = waves_per_block > 8;
    const bool fn = o.ratefn != nullptr || o.qign != nullptr || m.has_orders;
    if (fn && occ3) return (int)hipErrorInvalidConfiguration;  // external terms: up to 8 waves
    EMX_DISPATCH_NR(m.nRounds, {
        if (m.interp) {
            if (fn) EMX_QSS_LAUNCH(true, 2, true);
            else if (occ3) EMX_QSS_LAUNCH(true, 3, false);
            else EMX_QSS_LAUNCH(true, 2, false);
        } else {
            if (fn) EMX_QSS_LAUNCH(false, 2, true);
            else if (occ3) EMX_QSS_LAUNCH(false, 3, false);
            else EMX_QSS_LAUNCH(false, 2, false);
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
#define EMX_RATES_LAUNCH(INTERP, ORD)                                                              \
    hipLaunchKernelGGL((rates_kernel<NR, INTERP, ORD>), dim3(blocks), dim3(64 * wpb), lds, stream, m, T, \
                       Y, nonorm, perturb, n, wQ, wD, wnet, status)
    EMX_DISPATCH_NR(m.nRounds, {
        if (m.has_orders) {
            if (m.interp) EMX_RATES_LAUNCH(true, true);
            else EMX_RATES_LAUNCH(false, true);
        } else {
            if (m.interp) EMX_RATES_LAUNCH(true, false);
            else EMX_RATES_LAUNCH(false, false);
        }
    });
#undef EMX_RATES_LAUNCH
    return (int)hipGetLastError();
}

int launch_heat_release(const DevMech& m, const DevOpts& o, const double* state, const double* perturb,
                        int n, double* qdot, int* status, hipStream_t stream) {
    const int wpb = 4;
    const size_t lds = qss_lds_bytes(m, wpb);
    const int blocks = (int)(((long)n + wpb - 1) / wpb);
#define EMX_QDOT_LAUNCH(INTERP, ORD)                                                               \
    hipLaunchKernelGGL((heat_release_kernel<NR, INTERP, ORD>), dim3(blocks), dim3(64 * wpb), lds, stream, \
                       m, state, perturb, n, o, qdot, status)
    EMX_DISPATCH_NR(m.nRounds, {
        if (m.has_orders) {
            if (m.interp) EMX_QDOT_LAUNCH(true, true);
            else EMX_QDOT_LAUNCH(false, true);
        } else {
            if (m.interp) EMX_QDOT_LAUNCH(true, false);
            else EMX_QDOT_LAUNCH(false, false);
        }
    });
#undef EMX_QDOT_LAUNCH
    return (int)hipGetLastError();
}

int launch_thermo_props(const DevMech& m, const double* state, int n, double* rho, double* Wmx,
                        double* cp, double* cpSpec, double* hk, hipStream_t stream) {
    const int wpb = 4;
    const int per_block = wpb * kThermoPointsPerWave;
    const int blocks = (int)(((long)n + per_block - 1) / per_block);
    hipLaunchKernelGGL(thermo_props_kernel, dim3(blocks), dim3(64 * wpb), 0, stream, m, state, n, rho,
                       Wmx, cp, cpSpec, hk);
    return (int)hipGetLastError();
}

}  // namespace emx
