// Discrete adjoint of the batched QSS production stage (one wavefront per point).
//
// Objective per point (no counterpart in the reference; the heat release rate is
// SourceSystemQSS::odefun's qDot, src/sourceSystem.cpp:465):
//     J = Qbar * Q + ybar . y_final,   Q = sum over accepted internal steps of dt * qDot,
// qDot from the step's predictor evaluation.  The derivative is that of the integrator's
// arithmetic with the step sequence frozen (attempt sizes, accept/reject decisions and the
// active sides of the max(ymin, .) clips are the forward run's), for qss_itermax = 1 without
// the stability check (src/qssintegrator.cpp:110-265).  Parameters: the initial state and
// theta_i = ln m_i, m_i the rate multiplier of reaction i (Cantera m_perturb; for falloff
// reactions it scales both pre-exponential factors).
//
// Memory: the forward kernel stores the integrator state (y, tn, dt) at the start of every
// kCkptInterval-th accepted step.  The backward kernel walks the checkpoints from last to
// first; for each it re-integrates up to kCkptInterval steps from the checkpoint, recording
// the step starts, predictor outputs and every attempt in a per-wave record in global
// memory, then runs the reverse sweep over those steps.  The re-integration is the forward
// code from a recorded start, so it reproduces the forward decisions bit for bit.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "qss_device.h"

namespace emx {

namespace {

constexpr int kCkptInterval = 16;  // accepted steps per checkpoint
constexpr int kTapeAttempts = 64;  // attempts recorded per checkpoint interval
constexpr int kRowsPerEval = 5;     // y, q, d, creation rate, destruction rate
constexpr int kTapeRows = kRowsPerEval * (kCkptInterval + kTapeAttempts);
constexpr int kRecordDoubles = kTapeRows * 64;

// Adjoint scratch per wave (doubles)
struct AdjLds {
    double* wQb;    // [64] adjoint of creation rates (index 63: zero, stoichiometry padding)
    double* wDb;    // [64] adjoint of destruction rates
    double* hrt;    // [64] h/RT of each species (index 63: zero)
    double* Mgb;    // [64] adjoint of third-body concentrations per group
    double* MB;     // [nTbRounds*64] per-slot third-body adjoint contributions
    double* att_dt; // [kTapeAttempts] attempt step sizes of the recorded interval
    double* att_qd; // [kTapeAttempts] heat release rate of each attempt's corrector evaluation
    double* stp_qd; // [kCkptInterval] heat release rate of each step's predictor evaluation
    int* att0;      // [kCkptInterval + 1] first attempt of each recorded step
};

// MB covers the rounds holding third-body slots (they come first in device order)
__host__ __device__ inline int adj_lds_doubles(const DevMech& m) {
    return 4 * 64 + m.nTbRounds * 64 + 2 * kTapeAttempts + kCkptInterval + (kCkptInterval + 2) / 2 + 1;
}

__device__ inline AdjLds carve_adj(double* base, const DevMech& m) {
    AdjLds a;
    a.wQb = base;
    a.wDb = base + 64;
    a.hrt = base + 128;
    a.Mgb = base + 192;
    a.MB = base + 256;
    a.att_dt = a.MB + m.nTbRounds * 64;
    a.att_qd = a.att_dt + kTapeAttempts;
    a.stp_qd = a.att_qd + kTapeAttempts;
    a.att0 = reinterpret_cast<int*>(a.stp_qd + kCkptInterval);
    return a;
}

__device__ inline void init_adj_lds(const AdjLds& A, int lane) {
    if (lane == 0) {
        A.wQb[kPadSpecies] = 0.0;
        A.wDb[kPadSpecies] = 0.0;
        A.hrt[kPadSpecies] = 0.0;
    }
    wave_lds_sync();
}

// Adjoint tables (group -> slots, species -> groups, chunk -> species) staged in LDS by the
// backward kernel after the per-wave regions: every VJP walks them.
__host__ __device__ inline int adj_tables_doubles(const DevMech& m) {
    const int ints = (m.nGroups + 1) + m.nGrpSlots + (m.K + 1) + m.nSpGrp + m.nChunkRounds * 64;
    return m.nSpGrp + (ints + 1) / 2;
}

// Block-cooperative copy; the caller synchronises the block.
__device__ inline void stage_adj_tables(DevMech& v, const DevMech& m, double* base) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < m.nSpGrp; i += nt) base[i] = m.sp_grp_eff[i];
    v.sp_grp_eff = base;
    int32_t* q = reinterpret_cast<int32_t*>(base + m.nSpGrp);
    auto cpi = [&](const int32_t* src, int n) {
        for (int i = tid; i < n; i += nt) q[i] = src[i];
        const int32_t* r = q;
        q += n;
        return r;
    };
    v.grp_slot_off = cpi(m.grp_slot_off, m.nGroups + 1);
    v.grp_slot = cpi(m.grp_slot, m.nGrpSlots);
    v.sp_grp_off = cpi(m.sp_grp_off, m.K + 1);
    v.sp_grp = cpi(m.sp_grp, m.nSpGrp);
    v.chunk_sp = cpi(m.chunk_sp, m.nChunkRounds * 64);
}

__global__ void reset_queue_kernel_adj(unsigned int* q) {
    if (threadIdx.x == 0) *q = 0u;
}

__device__ __forceinline__ double dcpR_dT(const DevMech& m, int k, double T) {
    const double* c = m.nasa + 14 * k + ((T <= m.Tmid[k]) ? 0 : 7);
    return c[1] + T * (2.0 * c[2] + T * (3.0 * c[3] + T * 4.0 * c[4]));
}

// Thermo of the predictor evaluation (SourceSystem::updateThermo at the step start),
// bit for bit as qss_odefun computes it.
__device__ inline void predictor_thermo(const DevMech& m, int lane, double y, double T,
                                        const WaveLds& L, double& rho, double& cp, double& hk) {
    const bool sp = is_species_lane(lane, m);
    const double rW = sp ? m.rW[lane - 2] : 0.0;
    GasState g = set_state(m, lane, y, rW, T, false, L);
    double cpR = 0.0, hRT = 0.0, sR = 0.0;
    if (sp) nasa(m, lane - 2, T, emx_log(T), cpR, hRT, sR);
    hk = sp ? g.T * kGasConstant * hRT : 0.0;
    rho = g.dens;
    const double s = wave_sum(g.ym * cpR);
    const double cp_mole = kGasConstant * (g.mmw * s);
    cp = cp_mole / g.mmw;
}

// Thermo adjoints carried from a step's corrector evaluations to its predictor evaluation
struct ThermoBar {
    double rho, cp;  // uniform
    double hk;       // this lane's species
};

// VJP of SourceSystemQSS::odefun at this lane's state value yy.
//  corrector: rates at T = Tn with the given predictor thermo (rho, cp, hk); the adjoints
//             of Tn and of the thermo accumulate into Trb and thb.
//  predictor: T = yy[T lane] and the thermo is computed from yy; thb holds the step's
//             corrector contributions on entry and is consumed; Trb receives the T adjoint.
//  wQ, wD, qdot: the evaluation's creation and destruction rates (this lane's species) and
//  heat release rate, as recorded by the forward evaluation.
//  qb, db: adjoints of the outputs q, d (this lane); qdotb: adjoint of qDot (uniform).
//  yb accumulates the composition adjoint (species lanes); thetab[j] the ln-multiplier
//  adjoint of slot j*64 + lane.
//  Always inlined, as qss_integrate: outlined, its LDS reads become flat loads.
template <int NR, bool INTERP>
__device__ __forceinline__ bool rhs_vjp(const DevMech& m, int lane, double yy, bool corrector, double Tn,
                        double rho, double cp, double hk, double wQ, double wD, double qdot,
                        double qb, double db, double qdotb, double qloss, ThermoBar& thb,
                        double& Trb, double& yb, double (&thetab)[NR], RateRegs<NR>& R,
                        const double* perturb, const WaveLds& L, const AdjLds& A) {
    const bool sp = is_species_lane(lane, m);
    const int S = m.nRounds * 64;
    const double T = corrector ? Tn : bcast(yy, 1);
    const double rW = sp ? m.rW[lane - 2] : 0.0;
    const double W = sp ? m.W[lane - 2] : 0.0;
    // ---- forward recomputation (as qss_odefun)
    GasState g = set_state(m, lane, yy, rW, T, false, L);
    double hRT, cpR;
    if (!update_T<NR, INTERP>(m, lane, g, R, L, hRT, cpR)) return false;
    if (sp) A.hrt[lane - 2] = hRT;
    if (!corrector) {
        hk = sp ? g.T * kGasConstant * hRT : 0.0;
        rho = g.dens;
        const double s = wave_sum(g.ym * cpR);
        const double cp_mole = kGasConstant * (g.mmw * s);
        cp = cp_mole / g.mmw;
    }
    // rates of progress are not needed again: the reverse sweep below works from the
    // concentrations, the third-body sums and the cached rate constants, and the
    // evaluation's creation/destruction rates and heat release rate come from the record
    third_body_conc(m, lane, g, L);
    // ---- outputs -> creation/destruction rates and thermo
    const double qbT = bcast(qb, 1), dbT = bcast(db, 1);
    const double rc = rho * cp;
    const double a = qbT / rc + qdotb;
    const double s1 = wave_sum(sp ? (qb * wQ + db * wD) * W : 0.0);
    thb.rho += -(qbT * qdot + dbT * qloss) / (rho * rc) - s1 / (rho * rho);
    thb.cp += -(qbT * qdot + dbT * qloss) / (rc * cp);
    if (sp) {
        A.wQb[lane - 2] = qb * W / rho - a * hk;
        A.wDb[lane - 2] = db * W / rho + a * hk;
        thb.hk += -a * (wQ - wD);
    }
    wave_lds_sync();
    // ---- reaction slots: rates of progress -> base rate, rate constants, concentrations
    const double logT = emx_log(T), recipT = 1.0 / T;
    size_t nI = 0;
    if (INTERP) nI = (size_t)floor((T - m.interpTmin) / (m.interpTmax - m.interpTmin) * (m.interpN - 1));
    double Trl = 0.0;  // this lane's T adjoint contributions
    const double* C = L.C;
#pragma unroll
    for (int j = 0; j < NR; j++) {
        if (j < m.nRounds) {
            const int s = j * 64 + lane;
            const int w = m.info[s];
            const int kind = info_kind(w);
            // the slot's Arrhenius parameters (global memory), issued before the gathers
            // below so that their latency overlaps them
            const double sb = m.b[s], sE = m.EaR[s], sdn = m.dn[s];
            double pf = 0.0, pr = 0.0, mb = 0.0;
            if (kind != kNone) {
                const bool rev = info_rev(w);
                const uint32_t rs = m.reac[s], ps = m.prod[s];
                const int r0 = st_k(rs, 0), r1 = st_k(rs, 1), r2 = st_k(rs, 2);
                const int p0 = st_k(ps, 0), p1 = st_k(ps, 1), p2 = st_k(ps, 2);
                const double ropfb = A.wQb[p0] + A.wQb[p1] + A.wQb[p2] + A.wDb[r0] + A.wDb[r1] + A.wDb[r2];
                double roprb = A.wQb[r0] + A.wQb[r1] + A.wQb[r2];
                if (rev) roprb += A.wDb[p0] + A.wDb[p1] + A.wDb[p2];
                const double PCr = C[r0] * C[r1] * C[r2];
                const double PCp = C[p0] * C[p1] * C[p2];
                const double Mgv = (kind >= kThreeBody && kind <= kTroe) ? L.Mg[info_group(w)] : 1.0;
                // k_f as the forward computed it
                double kf = R.kf[j];
                double klow = 0.0, khigh = 0.0, work = 0.0, pr0 = 0.0, F = 1.0, f1 = 0.0, u = 0.0,
                       nn = 0.0;
                const bool fall = j < RateRegs<NR>::NF && (kind == kLindemann || kind == kTroe);  // SRI: refused
                if (fall) {
                    klow = R.klow[j];
                    khigh = R.khigh[j];
                    work = R.work[j];
                    pr0 = Mgv * klow / (khigh + kSmallNumber);
                    if (kind == kTroe) {
                        const double lpr = emx_log(fmax(pr0, kSmallNumber)) * kInvLn10;
                        const double cc = -0.4 - 0.67 * work;
                        nn = 0.75 - 1.27 * work;
                        u = lpr + cc;
                        f1 = u / (nn - 0.14 * u);
                        const double lf = work / (1.0 + f1 * f1);
                        F = emx_exp(kLn10 * lf, m.exp2t);
                    }
                    kf = pr0 * (F / (1.0 + pr0)) * khigh;
                }
                const double pert = perturb[s];
                const double rfn = kf * pert;
                const double base = (kind == kThreeBody) ? rfn * Mgv : rfn;
                const double rk = R.rkc[j];
                const double baseb = ropfb * PCr + roprb * rk * (rev ? PCp : 1.0);
                pf = ropfb * base;
                pr = rev ? roprb * base * rk : 0.0;
                double rfnb = baseb;
                if (kind == kThreeBody) {
                    rfnb = baseb * Mgv;
                    mb = baseb * rfn;
                }
                thetab[j] += rfnb * rfn;
                const double kfb = rfnb * pert;
                if (!fall) {
                    if (INTERP)
                        Trl += kfb * m.kf_s[nI * S + s];
                    else
                        Trl += kfb * kf * (sb * recipT + sE * recipT * recipT);
                } else {
                    const double prf = pr0 * (F / (1.0 + pr0));
                    double khighb = kfb * prf;
                    const double prfb = kfb * khigh;
                    const double Fb = prfb * pr0 / (1.0 + pr0);
                    double pr0b = prfb * F / ((1.0 + pr0) * (1.0 + pr0));
                    if (kind == kTroe) {
                        const double ln10 = kLn10;
                        const double lfb = Fb * F * ln10;
                        const double gg = 1.0 + f1 * f1;
                        double workb = lfb / gg;
                        const double f1b = -lfb * work * 2.0 * f1 / (gg * gg);
                        const double den = nn - 0.14 * u;
                        const double ub = f1b * nn / (den * den);
                        const double nnb = -f1b * u / (den * den);
                        workb += -0.67 * ub - 1.27 * nnb;
                        if (pr0 > kSmallNumber) pr0b += ub / (pr0 * ln10);
                        const double a3 = m.troeA[s], rT3 = m.rT3[s], rT1 = m.rT1[s], T2 = m.T2[s];
                        const double e3 = emx_exp(-T * rT3, m.exp2t), e1 = emx_exp(-T * rT1, m.exp2t);
                        double Fcent = (1.0 - a3) * e3 + a3 * e1;
                        double dF = 0.0;
                        if (rT3 < kBigNumber) dF += -(1.0 - a3) * rT3 * e3;
                        if (rT1 < kBigNumber) dF += -a3 * rT1 * e1;
                        if (T2 != 0.0) {
                            const double e2 = emx_exp(-T2 / T, m.exp2t);
                            Fcent += e2;
                            dF += T2 * recipT * recipT * e2;
                        }
                        if (Fcent > kSmallNumber) Trl += workb / (Fcent * ln10) * dF;
                    }
                    const double kd = khigh + kSmallNumber;
                    mb += pr0b * klow / kd;
                    const double klowb = pr0b * Mgv / kd;
                    khighb += -pr0b * Mgv * klow / (kd * kd);
                    Trl += klowb * klow * (m.b0[s] * recipT + m.EaR0[s] * recipT * recipT);
                    Trl += khighb * khigh * (sb * recipT + sE * recipT * recipT);
                }
                if (rev) {
                    const double rkcb = roprb * base * PCp;
                    if (INTERP) {
                        Trl += rkcb * m.kc_s[nI * S + s];
                    } else if (rk < kBigNumber) {
                        const double sh = sdn + A.hrt[r0] + A.hrt[r1] + A.hrt[r2] - A.hrt[p0] -
                                          A.hrt[p1] - A.hrt[p2];
                        Trl += rkcb * rk * sh * recipT;
                    }
                }
            }
            L.ropf[s] = pf;  // reactant-side product adjoint
            L.ropr[s] = pr;  // product-side product adjoint
            if (j < m.nTbRounds) A.MB[s] = mb;  // zero in the later rounds
        }
    }
    wave_lds_sync();
    // ---- third-body groups
    double ctotb_part = 0.0;
    if (lane < m.nGroups) {
        double mg = 0.0;
        for (int e = m.grp_slot_off[lane]; e < m.grp_slot_off[lane + 1]; e++) mg += A.MB[m.grp_slot[e]];
        A.Mgb[lane] = mg;
        ctotb_part = mg * m.grp_def[lane];
    }
    const double ctotb = wave_sum(ctotb_part);
    wave_lds_sync();
    // ---- concentration adjoints: fixed-length chunks over the species-sum entries
#pragma unroll
    for (int cr = 0; cr < kMaxChunkRounds; cr++) {
        if (cr < m.nChunkRounds) {
            const int c = cr * 64 + lane;
            const int kc = m.chunk_sp[c];
            const double Ck = C[kc];
            double part = 0.0;
#pragma unroll
            for (int e = 0; e < kChunkLen; e++) {
                const uint32_t en = m.entry[c * kChunkLen + e];
                const int s = en_slot(en);
                const int nu = en_nu(en);
                const bool side = en_prod(en);
                const double P = side ? L.ropr[s] : L.ropf[s];
                const uint32_t st = side ? m.prod[s < S ? s : 0] : m.reac[s < S ? s : 0];
                const int k0 = st_k(st, 0), k1 = st_k(st, 1), k2 = st_k(st, 2);
                const double o = (k0 == kc ? 1.0 : C[k0]) * (k1 == kc ? 1.0 : C[k1]) * (k2 == kc ? 1.0 : C[k2]);
                const double cpw = nu == 1 ? 1.0 : nu == 2 ? 2.0 * Ck : nu == 3 ? 3.0 * Ck * Ck : 0.0;
                part += P * cpw * o;
            }
            L.partq[c] = part;
        }
    }
    wave_lds_sync();
    double Cb = 0.0;
    if (sp) {
        const int k = lane - 2;
        const int c0 = m.sp_chunk0[k], nc = m.sp_nchunk[k];
        for (int i = 0; i < nc; i++) Cb += L.partq[c0 + i];
        for (int e = m.sp_grp_off[k]; e < m.sp_grp_off[k + 1]; e++) Cb += A.Mgb[m.sp_grp[e]] * m.sp_grp_eff[e];
    }
    // ---- state setter: C_k = ym_k dens, ctot = dens / mmw, dens = P mmw / (R T)
    const double dens = g.dens, mmw = g.mmw;
    double densb = wave_sum(sp ? Cb * g.ym : 0.0) + ctotb / mmw;
    double mmwb = -ctotb * dens / (mmw * mmw);
    double ymb = sp ? Cb * dens : 0.0;
    double Tb = 0.0;
    if (!corrector) {
        densb += thb.rho;
        const double sb = thb.cp * kGasConstant;
        ymb += sp ? sb * cpR : 0.0;
        const double dcp = sp ? dcpR_dT(m, lane - 2, T) : 0.0;
        Tb += wave_sum(sp ? sb * g.ym * dcp + thb.hk * kGasConstant * cpR : 0.0);
    }
    mmwb += densb * m.P / (T * kGasConstant);
    Tb += -densb * dens / T;
    const double sumb = -mmwb * mmw * mmw;
    ymb += sp ? sumb : 0.0;
    // Phase::setMassFractions: y' = max(y, 0) / sum max(y, 0), ym = y' / W
    const double yc = sp ? fmax(yy, 0.0) : 0.0;
    const double norm = wave_sum(yc);
    const double yk = yc * (1.0 / norm);
    const double ypb = ymb * rW;
    const double dot = wave_sum(ypb * yk);
    if (sp && yy >= 0.0) yb += (ypb - dot) / norm;
    Trb += Tb + wave_sum(Trl);
    wave_lds_sync();
    return true;
}

// Hooks: checkpoint the step starts every kCkptInterval steps (forward kernel)
struct CkptHooks {
    double* ck_y;  // [cap][nv]
    double* ck_s;  // [cap][2]: tn, dt
    int cap, nv, nsteps, nckpt;
    template <class Z>
    __device__ bool step_start(int lane, const Z& z, double tn, double dt, bool) {
        if (nsteps % kCkptInterval == 0) {
            if (nckpt == cap) return false;
            if (lane < nv) ck_y[(size_t)nckpt * nv + lane] = z.y;
            if (lane == 0) {
                ck_s[2 * nckpt] = tn;
                ck_s[2 * nckpt + 1] = dt;
            }
            nckpt++;
        }
        nsteps++;
        return true;
    }
    template <class Z, class Sys>
    __device__ void predicted(int, const Z&, const Sys&) {}
    template <class Z, class Sys>
    __device__ bool corrected(int, const Z&, const Sys&, double) { return true; }
};

// Hooks: record up to kCkptInterval steps from a checkpoint (backward kernel)
struct RecordHooks {
    double* rows;  // this wave's record, kTapeRows x 64
    AdjLds A;
    int nsteps, natt;
    double stop_tn, stop_dt, stop_y;  // integrator state where the interval ends
    template <class Z>
    __device__ bool step_start(int, const Z& z, double tn, double dt, bool) {
        if (nsteps < kCkptInterval) return true;
        stop_tn = tn;
        stop_dt = dt;
        stop_y = z.y;
        return false;
    }
    template <class Z, class Sys>
    __device__ void predicted(int lane, const Z& z, const Sys& S) {
        double* r = rows + (size_t)(kRowsPerEval * nsteps) * 64;
        r[lane] = z.y;
        r[64 + lane] = z.q;
        r[128 + lane] = z.d;
        r[192 + lane] = S.wQ;
        r[256 + lane] = S.wD;
        if (lane == 0) {
            A.att0[nsteps] = natt;
            A.stp_qd[nsteps] = S.qdot;
        }
        nsteps++;
    }
    template <class Z, class Sys>
    __device__ bool corrected(int lane, const Z& z, const Sys& S, double dt) {
        if (natt == kTapeAttempts) return false;
        double* r = rows + (size_t)(kRowsPerEval * (kCkptInterval + natt)) * 64;
        r[lane] = z.y;
        r[64 + lane] = z.q;
        r[128 + lane] = z.d;
        r[192 + lane] = S.wQ;
        r[256 + lane] = S.wD;
        if (lane == 0) {
            A.att_dt[natt] = dt;
            A.att_qd[natt] = S.qdot;
        }
        natt++;
        return true;
    }
};

__device__ __forceinline__ void qss_alpha(double r, double& a, double& da) {
    const double p = 180. + r * (60. + r * (11. + r));
    const double q = 360. + r * (60. + r * (12. + r));
    const double dp = 60. + r * (22. + 3. * r);
    const double dq = 60. + r * (24. + 3. * r);
    a = p / q;
    da = (dp * q - p * dq) / (q * q);
}

template <int NR, bool INTERP>
__global__ void __launch_bounds__(512, 2) qss_adjoint_forward_kernel(
    DevMech mg, DevOpts o, const double* __restrict__ state_in,
    const double* __restrict__ split_const, const double* __restrict__ qloss,
    const double* __restrict__ perturb_g, double dt, int nPoints, const int* __restrict__ order,
    unsigned int* __restrict__ queue, double* __restrict__ state_out, int* __restrict__ status,
    double* __restrict__ Q_out, int* __restrict__ cost, double* __restrict__ ck_y,
    double* __restrict__ ck_s, int* __restrict__ nckpt, int cap) {
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
        Z.y = Z.enforce ? fmax(y0, Z.ymin) : y0;
        Z.ym1 = Z.ym2 = 0.0;
        CkptHooks H{ck_y + (size_t)p * cap * nv, ck_s + (size_t)p * cap * 2, cap, nv, 0, 0};
        int g = 0, r = 0;
        double qi = 0.0;
        const unsigned long long t_point = point_clock();
        int err = qss_integrate<NR, INTERP>(m, o, lane, active, Z, S, R, perturb, L, dt,
                                            QssResume{0.0, 0.0, true}, H, g, r, qi);
        if (err == 1) err = -5;  // checkpoint capacity
        if (active) state_out[(size_t)p * nv + lane] = (err >= 0) ? Z.y : y0;
        if (lane == 0) {
            if (status) status[p] = err;
            if (Q_out) Q_out[p] = qi;
            if (cost) cost[p] = point_cost(t_point);
            nckpt[p] = err >= 0 ? H.nckpt : 0;
        }
    }
}

template <int NR, bool INTERP>
__global__ void __launch_bounds__(512, 2) qss_adjoint_backward_kernel(
    DevMech mg, DevOpts o, const double* __restrict__ state_in,
    const double* __restrict__ split_const, const double* __restrict__ qloss,
    const double* __restrict__ perturb_g, double dt, int nPoints, const int* __restrict__ order,
    unsigned int* __restrict__ queue, const double* __restrict__ ck_y,
    const double* __restrict__ ck_s, const int* __restrict__ nckpt, int cap,
    const double* __restrict__ Qbar, const double* __restrict__ ybar, double* __restrict__ record,
    double* __restrict__ grad_state, double* __restrict__ grad_lnk, int* __restrict__ status) {
    extern __shared__ double lds[];
    const double* perturb;
    DevMech ms = stage_mech(mg, perturb_g, lds, perturb);
    {
        const int pw = wave_lds_doubles(mg.nRounds, mg.nChunkRounds) + adj_lds_doubles(mg);
        stage_adj_tables(ms, mg, lds + mech_lds_doubles(mg) + (int)(blockDim.x / 64) * pw);
    }
    const DevMech m = ms;
    __syncthreads();
    const int lane = lane_id();
    const int wave = wave_in_block();
    const int per_wave = wave_lds_doubles(m.nRounds, m.nChunkRounds) + adj_lds_doubles(m);
    double* wbase = lds + mech_lds_doubles(m) + wave * per_wave;
    WaveLds L = carve_lds(wbase, m.nRounds, m.nChunkRounds);
    init_wave_lds(L, lane, m.nRounds, m.nChunkRounds);
    AdjLds A = carve_adj(wbase + wave_lds_doubles(m.nRounds, m.nChunkRounds), m);
    init_adj_lds(A, lane);
    double* rows = record + (size_t)(blockIdx.x * (blockDim.x / 64) + wave) * kRecordDoubles;
    const int nv = m.nVars;
    const bool active = lane < nv;
    const bool sp = is_species_lane(lane, m);
    QssSystem S;
    S.prio = 0;
    S.W = sp ? m.W[lane - 2] : 0.0;
    S.rW = sp ? m.rW[lane - 2] : 0.0;
    RateRegs<NR> R;
    while (true) {
        unsigned int t = 0;
        if (lane == 0) t = atomicAdd(queue, 1u);
        t = (unsigned int)bcast_i((int)t, 0);
        if (t >= (unsigned int)nPoints) break;
        const int p = order ? order[t] : (int)t;
        const double y0 = active ? state_in[(size_t)p * nv + lane] : 0.0;
        S.sc = (active && split_const) ? split_const[(size_t)p * nv + lane] : 0.0;
        S.qloss = qloss ? qloss[p] : 0.0;
        const double qloss_p = S.qloss;
        const double Qb = Qbar ? Qbar[p] : 1.0;
        double yb = (active && ybar) ? ybar[(size_t)p * nv + lane] : 0.0;
        double thetab[NR];
#pragma unroll
        for (int j = 0; j < NR; j++) thetab[j] = 0.0;
        const bool enforce = active && lane != 0;
        const double ymin = (lane == 0) ? -1e4 : o.qss_minval;
        const int nck = nckpt[p];
        int err = nck > 0 ? 0 : -5;
        R.T = -1.0;
        for (int c = nck - 1; c >= 0 && err == 0; c--) {
            // ---- re-integrate the interval from checkpoint c, recording it
            QssLane Z;
            Z.enforce = enforce;
            Z.ymin = ymin;
            Z.y = active ? ck_y[((size_t)p * cap + c) * nv + lane] : 0.0;
            Z.ym1 = Z.ym2 = 0.0;
            const double tn0 = ck_s[((size_t)p * cap + c) * 2];
            const double dtn0 = ck_s[((size_t)p * cap + c) * 2 + 1];
            RecordHooks H{rows, A, 0, 0, 0.0, 0.0, 0.0};
            int g = 0, r = 0;
            double qi = 0.0;
            int ret = qss_integrate<NR, INTERP>(m, o, lane, active, Z, S, R, perturb, L, dt,
                                                QssResume{tn0, dtn0, c == 0}, H, g, r, qi);
            if (ret < 0) {
                err = ret;
                break;
            }
            // the re-integration must end where the forward run stored the next checkpoint
            if (c + 1 < nck) {
                const size_t nx = (size_t)p * cap + c + 1;
                bool same = ret == 1 && H.stop_tn == ck_s[nx * 2] && H.stop_dt == ck_s[nx * 2 + 1];
                same = same && !wave_or(active && H.stop_y != ck_y[nx * nv + lane]);
                if (!same) {
                    err = -7;
                    break;
                }
            }
            if (lane == 0) A.att0[H.nsteps] = H.natt;
            wave_lds_sync();
            // ---- reverse sweep over the recorded steps
            for (int s = H.nsteps - 1; s >= 0; s--) {
                const double* rs = rows + (size_t)(kRowsPerEval * s) * 64;
                const double yn = rs[lane], q0 = rs[64 + lane], d0 = rs[128 + lane];
                const int t0 = A.att0[s];
                const int na = A.att0[s + 1] - t0;
                const double Tn = bcast(yn, 1);
                double rho, cp, hk;
                predictor_thermo(m, lane, yn, Tn, L, rho, cp, hk);
                const double dt0 = A.att_dt[t0];
                const double rtau = enforce ? dt0 * d0 / yn : 0.0;
                double al0, dal0;
                qss_alpha(rtau, al0, dal0);
                const double G = 1.0 + rtau * al0;
                const double dG = al0 + rtau * dal0;
                // rtaus of the accepted attempt as the forward computed it, and its factor
                double rts = rtau, fac = 1.0;
                for (int a = 1; a < na; a++) {
                    const double ratio = A.att_dt[t0 + a] / A.att_dt[t0 + a - 1];
                    fac *= ratio;
                    rts *= ratio;
                }
                double ysb = 0.0, qsb = 0.0, rtauacc = 0.0, qinb = 0.0, dinb = 0.0;
                ThermoBar thb{0.0, 0.0, 0.0};
                double TrbStep = 0.0;
                for (int a = na - 1; a >= 0; a--) {
                    const double* ra = rows + (size_t)(kRowsPerEval * (kCkptInterval + t0 + a)) * 64;
                    const double yp = ra[lane], qc = ra[64 + lane], dc = ra[128 + lane];
                    const double dta = A.att_dt[t0 + a];
                    double qcb = qinb, dcb = dinb, ypb = 0.0, rtausb = 0.0;
                    if (a == na - 1) {
                        const double rb = enforce ? 0.5 * (rts + dta * dc / yp) : 0.0;
                        double al, dal;
                        qss_alpha(rb, al, dal);
                        const double Den = 1.0 + al * rb;
                        const double Num = q0 * (1.0 - al) + qc * al - yn * rb / dta;
                        const double s2 = Num / Den;
                        const double v = yn + dta * s2;
                        const double gsel = (active && !(v < ymin)) ? yb : 0.0;
                        ysb += gsel;
                        const double s2b = gsel * dta;
                        qsb += s2b * (1.0 - al) / Den;
                        qcb += s2b * al / Den;
                        ysb += -s2b * rb / (dta * Den);
                        const double alb = s2b * ((qc - q0) / Den - Num * rb / (Den * Den));
                        const double rbb = s2b * (-yn / (dta * Den) - Num * al / (Den * Den)) + alb * dal;
                        if (enforce) {
                            rtausb = 0.5 * rbb;
                            dcb += 0.5 * rbb * dta / yp;
                            ypb += -0.5 * rbb * dta * dc / (yp * yp);
                        }
                    }
                    if (!rhs_vjp<NR, INTERP>(m, lane, yp, true, Tn, rho, cp, hk, ra[192 + lane],
                                             ra[256 + lane], A.att_qd[t0 + a], active ? qcb : 0.0,
                                             active ? dcb : 0.0, 0.0, qloss_p, thb, TrbStep, ypb,
                                             thetab, R, perturb, L, A)) {
                        err = -2;
                        break;
                    }
                    // yp = max(ymin, y_n + dt s), s = (q_a - d_a) / G(rtau)
                    double qa = q0, da = d0;
                    if (a > 0) {
                        const double* rp = rows + (size_t)(kRowsPerEval * (kCkptInterval + t0 + a - 1)) * 64;
                        qa = rp[64 + lane];
                        da = rp[128 + lane];
                    }
                    const double sv = (qa - da) / G;
                    const double v = yn + dta * sv;
                    const double gsel = (active && ymin < v) ? ypb : 0.0;
                    ysb += gsel;
                    const double sb = gsel * dta;
                    qinb = sb / G;
                    dinb = -sb / G;
                    rtauacc += -sb * (qa - da) / (G * G) * dG + rtausb * fac;
                }
                if (err) break;
                // attempt 0 read the predictor outputs; qs = q0, ys = y_n, rtau from d0
                double q0b = qinb + qsb, d0b = dinb, ynb = ysb;
                if (enforce) {
                    d0b += rtauacc * dt0 / yn;
                    ynb += -rtauacc * dt0 * d0 / (yn * yn);
                }
                double TrbPred = 0.0;
                const double dtA = A.att_dt[t0 + na - 1];
                if (!rhs_vjp<NR, INTERP>(m, lane, yn, false, Tn, 0.0, 0.0, 0.0, rs[192 + lane],
                                         rs[256 + lane], A.stp_qd[s], active ? q0b : 0.0,
                                         active ? d0b : 0.0, Qb * dtA, qloss_p, thb, TrbPred, ynb,
                                         thetab, R, perturb, L, A)) {
                    err = -2;
                    break;
                }
                if (lane == 1) ynb += TrbPred + TrbStep;
                yb = active ? ynb : 0.0;
            }
            wave_lds_sync();
        }
        // QssIntegrator::setState clip
        if (enforce && y0 < ymin) yb = 0.0;
        if (err) yb = 0.0;
        if (active) grad_state[(size_t)p * nv + lane] = yb;
        if (grad_lnk) {
#pragma unroll
            for (int j = 0; j < NR; j++) {
                if (j < m.nRounds) {
                    const int rx = m.orig[j * 64 + lane];
                    if (rx >= 0) grad_lnk[(size_t)p * m.I + rx] = err ? 0.0 : thetab[j];
                }
            }
        }
        if (lane == 0 && status) status[p] = err;
    }
}

}  // namespace

size_t adjoint_record_doubles_per_wave() { return (size_t)kRecordDoubles; }
int adjoint_checkpoint_interval() { return kCkptInterval; }

size_t adjoint_lds_bytes(const DevMech& m, int waves_per_block, bool backward) {
    size_t per = wave_lds_doubles(m.nRounds, m.nChunkRounds) + (backward ? adj_lds_doubles(m) : 0);
    return ((size_t)mech_lds_doubles(m) + (size_t)waves_per_block * per + (backward ? adj_tables_doubles(m) : 0)) *
           sizeof(double);
}

int launch_qss_adjoint_forward(const DevMech& m, const DevOpts& o, const double* state_in,
                               const double* split_const, const double* qloss,
                               const double* perturb, double dt, int nPoints, const int* order,
                               unsigned int* queue, double* state_out, int* status, double* Q,
                               int* cost, double* ck_y, double* ck_s, int* nckpt, int cap,
                               int grid_blocks, int waves_per_block, hipStream_t stream) {
    if (waves_per_block < 1 || waves_per_block > 8) return (int)hipErrorInvalidConfiguration;
    const size_t lds = adjoint_lds_bytes(m, waves_per_block, false);
    hipLaunchKernelGGL(reset_queue_kernel_adj, dim3(1), dim3(64), 0, stream, queue);
    EMX_DISPATCH_NR(m.nRounds, {
        if (m.interp)
            hipLaunchKernelGGL((qss_adjoint_forward_kernel<NR, true>), dim3(grid_blocks),
                               dim3(64 * waves_per_block), lds, stream, m, o, state_in, split_const,
                               qloss, perturb, dt, nPoints, order, queue, state_out, status, Q,
                               cost, ck_y, ck_s, nckpt, cap);
        else
            hipLaunchKernelGGL((qss_adjoint_forward_kernel<NR, false>), dim3(grid_blocks),
                               dim3(64 * waves_per_block), lds, stream, m, o, state_in, split_const,
                               qloss, perturb, dt, nPoints, order, queue, state_out, status, Q,
                               cost, ck_y, ck_s, nckpt, cap);
    });
    return (int)hipGetLastError();
}

int launch_qss_adjoint_backward(const DevMech& m, const DevOpts& o, const double* state_in,
                                const double* split_const, const double* qloss,
                                const double* perturb, double dt, int nPoints, const int* order,
                                unsigned int* queue, const double* ck_y, const double* ck_s,
                                const int* nckpt, int cap, const double* Qbar, const double* ybar,
                                double* record, double* grad_state, double* grad_lnk, int* status,
                                int grid_blocks, int waves_per_block, hipStream_t stream) {
    if (waves_per_block < 1 || waves_per_block > 8) return (int)hipErrorInvalidConfiguration;
    const size_t lds = adjoint_lds_bytes(m, waves_per_block, true);
    hipLaunchKernelGGL(reset_queue_kernel_adj, dim3(1), dim3(64), 0, stream, queue);
    EMX_DISPATCH_NR(m.nRounds, {
        if (m.interp)
            hipLaunchKernelGGL((qss_adjoint_backward_kernel<NR, true>), dim3(grid_blocks),
                               dim3(64 * waves_per_block), lds, stream, m, o, state_in, split_const,
                               qloss, perturb, dt, nPoints, order, queue, ck_y, ck_s, nckpt, cap,
                               Qbar, ybar, record, grad_state, grad_lnk, status);
        else
            hipLaunchKernelGGL((qss_adjoint_backward_kernel<NR, false>), dim3(grid_blocks),
                               dim3(64 * waves_per_block), lds, stream, m, o, state_in, split_const,
                               qloss, perturb, dt, nPoints, order, queue, ck_y, ck_s, nckpt, cap,
                               Qbar, ybar, record, grad_state, grad_lnk, status);
    });
    return (int)hipGetLastError();
}

}  // namespace emx
