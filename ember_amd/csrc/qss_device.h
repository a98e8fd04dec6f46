// QSS integration of one point per wavefront: SourceSystemQSS's right-hand side and
// QssIntegrator's step control, shared by the stage kernel and the adjoint kernels.
//
// Per point this is SourceSystemQSS::setState / integrateToTime
// (src/sourceSystem.cpp:434-443, src/sourceSystem.h:252) driving
// QssIntegrator::integrateToTime / integrateOneStep (src/qssintegrator.cpp:98-265)
// with SourceSystemQSS::odefun (src/sourceSystem.cpp:445-499) as the ODE.
#pragma once

#include "chem_device.h"

namespace emx {

// compile-time reaction-round bound for a mechanism with nRounds rounds
#define EMX_DISPATCH_NR(nr, BODY)                  \
    switch ((nr) <= 1 ? 1 : (nr) <= 2 ? 2 : (nr) <= 4 ? 4 : (nr) <= 6 ? 6 : 8) { \
        case 1: { constexpr int NR = 1; BODY; } break; \
        case 2: { constexpr int NR = 2; BODY; } break; \
        case 4: { constexpr int NR = 4; BODY; } break; \
        case 6: { constexpr int NR = 6; BODY; } break; \
        default: { constexpr int NR = 8; BODY; } break; \
    }


struct QssResume {
    double tn, dt;
    bool firstStep;
};

struct NoHooks {
    template <class Z>
    __device__ __forceinline__ bool step_start(int, const Z&, double, double, bool) { return true; }
    template <class Z, class Sys>
    __device__ __forceinline__ void predicted(int, const Z&, const Sys&) {}
    template <class Z, class Sys>
    __device__ __forceinline__ bool corrected(int, const Z&, const Sys&, double) { return true; }
};

// Per-lane view of SourceSystemQSS: the variable this lane holds plus the
// thermo quantities refreshed on non-corrector calls.
struct QssSystem {
    double W, rW;      // species lanes
    double sc;         // splitConst for this lane's variable
    double qloss;
    double qign;       // igniter amplitude at this point's position (0: none)
    double T;          // temperature used by the rates (frozen in correctors)
    double rho, cp;    // from the last non-corrector call (SourceSystem::updateThermo)
    double rrho, rrc;  // 1 / rho and 1 / (rho cp) of the same call
    double hk;         // J/kmol, species lanes
    double qdot;
    double wQ, wD;     // creation / destruction rate of this lane's species (last evaluation)
    int prio;          // issue-priority toggle (alternate_priority)
};

// Alternates this wave's issue priority between 0 and 1 at each call (once per right-hand
// side).  The SIMD otherwise favours the older of its two waves, so whichever point a
// later-dispatched wave holds would progress more slowly; with the toggle the two waves
// take turns, and points progress at about the same rate (which the cost-ordered queue
// assumes).
__device__ __forceinline__ void alternate_priority(int& prio) {
    prio ^= 1;
    if (prio) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// EXT: the call carries the time-dependent external terms (rate multiplier factor fa,
// igniter heat release qi) and the explicit-order concentration products; without them
// fa = 1 and qi = 0 compile away and every product is the plain mass-action one.
template <int NR, bool INTERP, bool EXT = false>
__device__ inline bool qss_odefun(const DevMech& m, int lane, double y, bool corrector,
                                  QssSystem& S, RateRegs<NR>& R, const double* perturb,
                                  double fa, double qi, const WaveLds& L, double& q, double& d) {
    alternate_priority(S.prio);
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
        S.cp = emx_div(cp_mole, g.mmw);
        S.rrho = emx_div(1.0, S.rho);
        S.rrc = emx_div(1.0, S.rho * S.cp);
    }
    rates_of_progress<NR, EXT>(m, lane, g, R, perturb, fa, L);
    double wQ, wD;
    species_rates(m, lane, L, wQ, wD);
    S.wQ = wQ;
    S.wD = wD;
    S.qdot = -wave_sum((wQ - wD) * S.hk);
    if (EXT) S.qdot += qi;  // + getQdotIgniter(t) (src/sourceSystem.cpp:465)
    // q and d of every lane without divergent branches: the energy lane scales by
    // 1 / (rho cp), the species lanes by W_k / rho
    const bool sp = is_species_lane(lane, m);
    const bool en = lane == 1;
    const double f = en ? S.rrc : S.rrho;
    const double nq = en ? S.qdot : wQ * S.W;
    const double nd = en ? S.qloss : wD * S.W;
    q = (lane == 0) ? S.sc : (en || sp) ? nq * f + S.sc : 0.0;
    d = (en || sp) ? nd * f : 0.0;
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
// interpolation range, -3 when the internal step budget is exhausted, -4 on NaN, -5 when
// the hooks refuse an attempt (record capacity), and 1 when the hooks stop the integration
// at the start of a step.
//
// `start` resumes from a recorded step start (tn, dt, firstStep); a fresh integration
// starts from {0, 0, true}.  The hooks see every step start (before the predictor
// evaluation), the predictor result, and every corrector evaluation; NoHooks compiles away.
//
// QssIntegrator::integrateToTime / integrateOneStep written as a state machine with
// a single call site of the ODE function: the predictor evaluation starts a step,
// every further evaluation is a corrector.  The control flow and arithmetic follow
// src/qssintegrator.cpp:98-265 line by line.
// EXT: a rate multiplier function (o.ratefn) or the igniter (o.qign) is on; without them
// the multiplier factor 1 and the igniter term compile away.
// Always inlined: in a called function the LDS pointers lose their address space (every
// staged-table access becomes a flat load) and the rate-constant registers go to the stack.
template <int NR, bool INTERP, class Hooks, bool EXT = false>
__device__ __forceinline__ int qss_integrate(const DevMech& m, const DevOpts& o, int lane, bool active,
                             QssLane& Z, QssSystem& S, RateRegs<NR>& R, const double* perturb,
                             const WaveLds& L, double tf, QssResume start, Hooks& H, int& gcount,
                             int& rcount, double& qint) {
    const double tfd = 1.000008;
    const bool stab_on = o.qss_stability != 0;
    double tn = start.tn, ts = 0.0, dt = start.dt;
    bool firstStep = start.firstStep;
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
        if (predict && !H.step_start(lane, Z, tn, dt, firstStep)) return 1;
        // SourceSystemQSS::odefun(tn + tstart) (src/qssintegrator.cpp:113, :159)
        const double fa = EXT ? fn_mult(o, tn) : 1.0;
        const double qi = EXT ? igniter(o, S.qign, tn) : 0.0;
        if (!qss_odefun<NR, INTERP, EXT>(m, lane, Z.y, !predict, S, R, perturb, fa, qi, L, Z.q, Z.d))
            return -2;
        gcount += 1;
        if (predict) H.predicted(lane, Z, S);
        else if (!H.corrected(lane, Z, S, dt)) return -5;
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
                double scrtch = fmax(wave_max0(c), 1.0e-25);
                const double sqreps = 0.5;
                dt = fmin(sqreps / scrtch, tf);
                dt = fmin(dt, o.qss_dtmax);
            }
            ts = tn;
            Z.rtau = Z.enforce ? emx_div(dt * Z.d, Z.y) : 0.0;
            Z.qs = Z.q;
            Z.ys = Z.y;
            Z.rtaus = Z.rtau;
            predict = false;
            start_attempt = true;
        } else {
            // ---- after a corrector evaluation
            double rtaub = Z.enforce ? 0.5 * (Z.rtaus + emx_div(dt * Z.d, Z.y)) : 0.0;
            double alpha = emx_div(180. + rtaub * (60. + rtaub * (11. + rtaub)),
                                   360. + rtaub * (60. + rtaub * (12. + rtaub)));
            Z.scratch = emx_div(Z.qs * (1.0 - alpha) + Z.q * alpha - emx_div(Z.ys * rtaub, dt),
                                1.0 + alpha * rtaub);
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
                    scr1 = emx_div(scr1, Z.y);
                    c = .5 * (scr1 + fmin(emx_div(fabs(Z.q - Z.d), Z.q + Z.d + 1e-30), scr1));
                }
            }
            eps = fmax(eps, wave_max0(c));
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
                stab = fmax(0.01, wave_max0(sc));
            }
            const bool ok = eps <= o.qss_epsmax && stab <= 1.0;
            if (ok) {
                qint += dt * qdot_step;
                if (tf <= tn * tfd) return 0;
            } else {
                tn = ts;
            }
            double rteps = 0.5 * (eps + 1.0);
            rteps = 0.5 * (rteps + emx_div(eps, rteps));
            rteps = 0.5 * (rteps + emx_div(eps, rteps));
            double dto = dt;
            dt = fmin(dt * (emx_div(1.0, rteps) + 0.005), tfd * (tf - tn));
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
            Z.scratch = emx_div(Z.q - Z.d,
                                1.0 + emx_div(r * (180 + r * (60 + r * (11 + r))), 360 + r * (60 + r * (12 + r))));
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

}  // namespace emx
