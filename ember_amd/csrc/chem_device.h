// Wavefront-cooperative gas kinetics and thermo for one grid point.
//
// One wavefront holds one point.  Lane v holds state variable v (0 = U,
// 1 = T, 2 + k = Y_k, Ember's kMomentum/kEnergy/kSpecies rows,
// src/readConfig.h:7-9).  For the reaction work, lane l owns reaction slots
// l, l + 64, ... ("rounds"); the device slot order groups falloff, then
// three-body, then elementary reactions so that each round runs one code path.
//
// The arithmetic restates Cantera 3.2 as called by Ember's CanteraGas
// (src/chemistry0d.cpp:487-672): Phase::setMassFractions /
// setMassFractions_NoNorm, IdealGasPhase density and standard chemical
// potentials, NasaPoly2, BulkKinetics::updateROP (third bodies, falloff
// Lindemann/Troe, reverse rates through 1/Kc capped at BigNumber), and
// Kinetics::getCreationRates / getDestructionRates / getNetProductionRates.
// Ember's InterpKinetics (src/chemistry0d.cpp:271-436) is the INTERP=true
// variant: kf of Arrhenius/three-body reactions and 1/Kc come from linear
// tables over [273, 3500] K.
#pragma once

#include "emx_common.h"
#include "emx_math.h"
#include "wave.h"

namespace emx {

// Status of a point whose input state Cantera's state setters would refuse
// (Phase::setTemperature needs T > 0; a NaN or infinite value anywhere in the state).
constexpr int kInvalidState = -9;

// A point's cost for the next stage's queue order, from the real-time counter (100 MHz)
// read at its start: the integration time in 640 ns units (1/16-octave cost bins resolve
// times up to 42 ms).  Time rather than a count of right-hand sides: a predictor
// evaluation (rate constants, thermo) costs about twice a corrector's, a Jacobian column
// less than a Newton iteration with its solves, and the mix differs between points.
__device__ __forceinline__ unsigned long long point_clock() { return __builtin_amdgcn_s_memrealtime(); }
__device__ __forceinline__ int point_cost(unsigned long long t0) {
    const unsigned long long d = (point_clock() - t0) >> 6;
    return (int)(d >= (1ull << 30) ? (1ull << 30) : (d | 1ull));
}

// The whole wave agrees whether this point's input state can be integrated: every state
// variable finite and T (lane 1) positive.
__device__ __forceinline__ bool valid_input_state(int lane, bool active, double y0) {
    const bool bad = active && (!isfinite(y0) || (lane == 1 && !(y0 > 0.0)));
    return !wave_or(bad ? 1 : 0);
}

// Third-body cell of the slots without a third body (info_group 255): Mg[kNoGroup] = 1.
// Mechanisms have at most kNoGroup efficiency groups, so no group writes it.
constexpr int kNoGroup = 63;

// per-wave LDS scratch (doubles)
struct WaveLds {
    double* C;      // [64] concentrations
    double* ge;     // [64] emx_exp(mu_k / RT - ln c0): standard-state factors of 1/Kc
    double* gi;     // [64] 1 / ge
    double* Mg;     // [64] third-body concentrations per efficiency group
    double* ropf;   // [nRounds*64 + 1]: the last entry is the zero slot
    double* ropr;   // [nRounds*64 + 1]
    double* partq;  // [nChunkRounds*64 + 1]: the last entry is a zero cell
    double* partd;  // [nChunkRounds*64 + 1]
};

__host__ __device__ inline int wave_lds_doubles(int nRounds, int nChunkRounds) {
    return 4 * 64 + 2 * (nRounds * 64 + 1) + 2 * (nChunkRounds * 64 + 1);
}

__device__ inline WaveLds carve_lds(double* base, int nRounds, int nChunkRounds) {
    WaveLds w;
    w.C = base;
    w.ge = base + 64;
    w.gi = base + 128;
    w.Mg = base + 192;
    w.ropf = base + 256;
    w.ropr = w.ropf + nRounds * 64 + 1;
    w.partq = w.ropr + nRounds * 64 + 1;
    w.partd = w.partq + nChunkRounds * 64 + 1;
    return w;
}

// Constant cells of the per-wave scratch: C[kPadSpecies] = ge = gi = 1 (stoichiometry
// padding), ropf/ropr zero slot (species-sum padding), the chunk partials' zero cells.
// Species lanes never write index kPadSpecies since K <= 62.
__device__ inline void init_wave_lds(const WaveLds& L, int lane, int nRounds, int nChunkRounds) {
    if (lane == 0) {
        L.C[kPadSpecies] = 1.0;
        L.ge[kPadSpecies] = 1.0;
        L.gi[kPadSpecies] = 1.0;
        L.ropf[nRounds * 64] = 0.0;
        L.ropr[nRounds * 64] = 0.0;
        L.partq[nChunkRounds * 64] = 0.0;
        L.partd[nChunkRounds * 64] = 0.0;
        L.Mg[kNoGroup] = 1.0;
    }
    wave_lds_sync();
}

// ---------------------------------------------------------------------------
// Mechanism tables staged in LDS once per workgroup (read on every right-hand side):
// per-slot info/stoichiometry/multipliers, species data, third-body groups and the
// species-sum entries.  Arrhenius and falloff parameters (used once per step) stay in
// global memory.
// ---------------------------------------------------------------------------
__host__ __device__ inline int mech_lds_doubles(const DevMech& m) {
    const int S = m.nRounds * 64;
    int n = S + 3 * m.K + 14 * m.K + 16 + m.nGrpEntries + (m.nGroups > 0 ? m.nGroups : 1);
    int ints = 3 * S + m.nEntries + 2 * m.K + (m.nGroups + 1) + m.nGrpEntries;
    return n + (ints + 1) / 2;
}

// Copies the tables into `base` (block-cooperative; the caller synchronises the block)
// and returns a view whose staged pointers refer to LDS.
__device__ inline DevMech stage_mech(const DevMech& m, const double* perturb, double* base,
                                     const double*& perturb_lds) {
    DevMech v = m;
    const int S = m.nRounds * 64;
    const int tid = threadIdx.x, nt = blockDim.x;
    double* p = base;
    auto cpd = [&](const double* src, int n) {
        for (int i = tid; i < n; i += nt) p[i] = src[i];
        const double* r = p;
        p += n;
        return r;
    };
    perturb_lds = cpd(perturb, S);
    v.W = cpd(m.W, m.K);
    v.rW = cpd(m.rW, m.K);
    v.Tmid = cpd(m.Tmid, m.K);
    v.nasa = cpd(m.nasa, 14 * m.K);
    v.exp2t = cpd(m.exp2t, 16);
    v.grp_eff = cpd(m.grp_eff, m.nGrpEntries);
    v.grp_def = cpd(m.grp_def, m.nGroups > 0 ? m.nGroups : 1);
    int32_t* q = reinterpret_cast<int32_t*>(p);
    auto cpi = [&](const int32_t* src, int n) {
        for (int i = tid; i < n; i += nt) q[i] = src[i];
        const int32_t* r = q;
        q += n;
        return r;
    };
    v.info = cpi(m.info, S);
    v.reac = reinterpret_cast<const uint32_t*>(cpi(reinterpret_cast<const int32_t*>(m.reac), S));
    v.prod = reinterpret_cast<const uint32_t*>(cpi(reinterpret_cast<const int32_t*>(m.prod), S));
    v.entry = reinterpret_cast<const uint32_t*>(cpi(reinterpret_cast<const int32_t*>(m.entry), m.nEntries));
    v.sp_chunk0 = cpi(m.sp_chunk0, m.K);
    v.sp_nchunk = cpi(m.sp_nchunk, m.K);
    v.grp_off = cpi(m.grp_off, m.nGroups + 1);
    v.grp_sp = cpi(m.grp_sp, m.nGrpEntries);
    return v;
}

// ChebyshevFunction::a (src/scalarFunction.cpp:53-72): t mapped onto [-1, 1] over
// f = [T0, T1, c_0 .. c_{nc-1}] (nc >= 2), then sum c_i T_i(x) by the three-term
// recurrence.  Wave-uniform.
__device__ inline double rate_fn_eval(const double* f, int nc, double t) {
    const double x = 2 * (t - f[0]) / (f[1] - f[0]) - 1;
    double Cnm1 = 1, Cn = x;
    double y = f[2] + x * f[3];
    for (int i = 2; i < nc; i++) {
        const double Cnp1 = 2 * x * Cn - Cnm1;
        y += Cnp1 * f[2 + i];
        Cnm1 = Cn;
        Cn = Cnp1;
    }
    return y;
}

// The rate multipliers of a right-hand side are perturb[s] * fa: the per-slot multipliers
// (Cantera's m_perturb) with fa = 1, or, while Ember's rate multiplier function is set,
// unit per-slot multipliers with fa = a(t) for every reaction (CanteraGas::setRateMultiplier,
// src/chemistry0d.cpp:634-642, called from src/sourceSystem.cpp:133-135 and :453-455).
// Multiplying by an exact 1 keeps either product bit for bit.
__device__ __forceinline__ double fn_mult(const DevOpts& o, double t) {
    return o.ratefn ? rate_fn_eval(o.ratefn, o.ratefn_n, o.t_start + t) : 1.0;
}

// SourceSystem::getQdotIgniter (src/sourceSystem.cpp:33-45) at time t from the stage start,
// for a point whose pulse amplitude energy / (stddev sqrt(2 pi) duration) *
// exp(-(x - center)^2 / (2 stddev^2)) is `amp` (precomputed on the host per position)
__device__ __forceinline__ double igniter(const DevOpts& o, double amp, double t) {
    const double ta = o.t_start + t;
    return (ta >= o.ign_t0 && ta < o.ign_t1) ? amp : 0.0;
}

// k_f of a PLOG or Chebyshev reaction slot at the context pressure (tables from
// build_pressure_tables): Cantera's PlogRate (log-linear interpolation in ln P between the
// bracketing levels, each level's rates summed) and ChebyshevRate (log10 k as a Chebyshev
// series in the reduced inverse temperature, the pressure sums precomputed).
__device__ inline double pdep_kf(const DevMech& m, int s, int kind, double logT, double recipT) {
    const double* d = m.auxd + m.aux[s];
    if (kind == kPlog) {
        const int n1 = (int)d[1], n2 = (int)d[2];
        auto level_log = [&](const double* q, int n) {  // q: (A, ln A, b, Ea/R) per rate
            if (n == 1) return q[1] + q[2] * logT - q[3] * recipT;
            double k = 1e-300;
            for (int i = 0; i < n; i++) k += q[4 * i] * emx_exp(q[4 * i + 2] * logT - q[4 * i + 3] * recipT, m.exp2t);
            return emx_log(k);
        };
        const double l1 = level_log(d + 3, n1);
        const double l2 = level_log(d + 3 + 4 * n1, n2);
        return emx_exp(l1 + (l2 - l1) * d[0], m.exp2t);
    }
    const int nT = (int)d[0];
    const double Tr = (2 * recipT + d[1]) * d[2];
    double Cnm1 = Tr, Cn = 1.0, logk = 0.0;
    for (int i = 0; i < nT; i++) {
        const double Cnp1 = 2 * Tr * Cn - Cnm1;
        logk += Cn * d[3 + i];
        Cnm1 = Cn;
        Cn = Cnp1;
    }
    return emx_exp(logk * kLn10, m.exp2t);  // 10^logk
}

// T-dependent reaction data held in registers (one set per owned slot).
// NR = compile-time bound on reaction rounds (>= DevMech::nRounds).
template <int NR>
struct RateRegs {
    static constexpr int NF = NR < kMaxFallRounds ? NR : kMaxFallRounds;
    double kf[NR];
    double rkc[NR];
    double klow[NF];
    double khigh[NF];
    double work[NF];   // log10(Fcent) for Troe; log10 of a exp(-b/T) + exp(-T/c) for SRI
    double work2[NF];  // d T^e for SRI
    double T;         // temperature these were computed at (-1: none)
};

// Composition state produced by set_state (uniform values are identical in all lanes)
struct GasState {
    double T, mmw, dens, ctot;
    double ym;  // this lane's Y_k / W_k (species lanes), 0 otherwise
};

__device__ __forceinline__ bool is_species_lane(int lane, const DevMech& m) {
    return lane >= 2 && lane < m.nVars;
}

// Phase::setMassFractions (+ setState_TP): clip negatives, normalise.
// nonorm: Phase::setMassFractions_NoNorm.  y = this lane's state variable.
__device__ inline GasState set_state(const DevMech& m, int lane, double y, double rW, double T,
                                     bool nonorm, const WaveLds& L) {
    GasState g;
    g.T = T;
    const bool sp = is_species_lane(lane, m);
    double yk = sp ? y : 0.0;
    if (!nonorm) {
        double yc = fmax(yk, 0.0);
        double norm = wave_sum(yc);
        yk = yc * emx_div(1.0, norm);
    }
    g.ym = sp ? yk * rW : 0.0;
    double sum = wave_sum(g.ym);
    g.mmw = emx_div(1.0, sum);
    g.dens = emx_div(m.P * g.mmw, T * kGasConstant);
    g.ctot = emx_div(g.dens, g.mmw);
    if (sp) L.C[lane - 2] = g.ym * g.dens;
    return g;
}

// NasaPoly1::updateProperties for this lane's species
__device__ inline void nasa(const DevMech& m, int k, double T, double logT, double& cpR,
                            double& hRT, double& sR) {
    const double* c = m.nasa + 14 * k + ((T <= m.Tmid[k]) ? 0 : 7);
    double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
    double ct0 = c[0];
    double ct1 = c[1] * T;
    double ct2 = c[2] * T2;
    double ct3 = c[3] * T3;
    double ct4 = c[4] * T4;
    cpR = ct0 + ct1 + ct2 + ct3 + ct4;
    hRT = ct0 + 0.5 * ct1 + 1.0 / 3.0 * ct2 + 0.25 * ct3 + 0.2 * ct4 + c[5] * (1.0 / T);
    sR = ct0 * logT + ct1 + 0.5 * ct2 + 1.0 / 3.0 * ct3 + 0.25 * ct4 + c[6];
}

__device__ __forceinline__ double arrhenius(double A, double b, double EaR, double logT,
                                            double recipT, const double* e2t) {
    return A * emx_exp(b * logT - EaR * recipT, e2t);
}

// Products over the (padded) three stoichiometry indices: the padding cells (C = 1)
// leave the one- and two-species results unchanged bit for bit.
__device__ __forceinline__ double conc_product(uint32_t st, const double* C) {
    return C[st_k(st, 0)] * C[st_k(st, 1)] * C[st_k(st, 2)];
}

// 1/Kc = exp(sum_k nu_k mu_k / RT - dn ln c0) (reactants negative, dn = sum of the product
// minus the reactant coefficients), as the product of the per-species factors
// ge_k = exp(mu_k / RT - ln c0) of the products and 1/ge_k of the reactants: one exp per
// species instead of one per reaction.  Pairing product and reactant factors keeps the
// partial products in range; the padding cells are 1.
__device__ __forceinline__ double inv_kc(uint32_t reac, uint32_t prod, const WaveLds& L) {
    return (L.ge[st_k(prod, 0)] * L.gi[st_k(reac, 0)]) * (L.ge[st_k(prod, 1)] * L.gi[st_k(reac, 1)]) *
           (L.ge[st_k(prod, 2)] * L.gi[st_k(reac, 2)]);
}

// Thermo for this lane's species plus the T-dependent rate data, cached on T.
// Returns false when T leaves the interpolation table range (INTERP only).
// Outputs hRT/cpR for the lane's species.
//
// Every round's parameters are loaded before any is used (one memory latency instead of
// one per round: the tables stay in global memory), the rate constants are formed without
// branches on the slot kind (padding slots have A = 0 and padding stoichiometry, so they
// give exact zeros), and a falloff slot's k_inf is the Arrhenius value of its round.
template <int NR, bool INTERP>
__device__ inline bool update_T(const DevMech& m, int lane, const GasState& g, RateRegs<NR>& R,
                                const WaveLds& L, double& hRT, double& cpR) {
    constexpr int NF = RateRegs<NR>::NF;
    const double T = g.T;
    const double logT = emx_log(T);
    const double recipT = 1.0 / T;
    const bool sp = is_species_lane(lane, m);
    double sR = 0.0;
    hRT = 0.0;
    cpR = 0.0;
    if (sp) nasa(m, lane - 2, T, logT, cpR, hRT, sR);
    if (R.T == T) return true;
    // PLOG and Chebyshev k_f into the forward-rate LDS row (free until rates_of_progress),
    // in one rolled loop so that their code exists once instead of once per round
    if (m.has_pdep) {
        for (int j = 0; j < m.nRounds; j++) {
            const int s = j * 64 + lane;
            const int kind = info_kind(m.info[s]);
            if (kind == kPlog || kind == kCheb) L.ropf[s] = pdep_kf(m, s, kind, logT, recipT);
        }
    }
    // rounds past the mechanism's are not loaded (uniform branches: the loads keep one
    // address per table and the round in the instruction's offset)
    double khigh_of[NF];
    if (INTERP) {
        // InterpKinetics::updateROP's index and offset (src/chemistry0d.cpp:392-397); a T
        // outside [Tmin, Tmax] would make Ember rebuild its tables (:279-291), here the
        // point fails with -2 and the host grows the tables
        const double Tlo = m.interpTmin, Thi = m.interpTmax;
        if (!(T >= Tlo && T <= Thi)) return false;
        size_t n = (size_t)floor((T - Tlo) / (Thi - Tlo) * (m.interpN - 1));
        double dT = T - n * (Thi - Tlo) / (m.interpN - 1) - Tlo;
        const int stride = m.nRounds * 64;
        double kc0[NR], kc1[NR], kf0[NR], kf1[NR];
        const size_t o0 = n * stride + lane;
#pragma unroll
        for (int j = 0; j < NR; j++) {
            kc0[j] = kc1[j] = kf0[j] = kf1[j] = 0.0;
            if (j < m.nRounds) {
                kc0[j] = m.kc_c[o0 + j * 64];
                kc1[j] = m.kc_s[o0 + j * 64];
                kf0[j] = m.kf_c[o0 + j * 64];
                kf1[j] = m.kf_s[o0 + j * 64];
            }
        }
#pragma unroll
        for (int j = 0; j < NR; j++) {
            if (j < m.nRounds) {
                const int s = j * 64 + lane;
                R.rkc[j] = kc0[j] + kc1[j] * dT;
                R.kf[j] = kf0[j] + kf1[j] * dT;
                // PLOG and Chebyshev rates keep their own evaluation (only Arrhenius rates
                // are interpolated, src/chemistry0d.cpp:264-269)
                if (m.has_pdep) {
                    const int kind = info_kind(m.info[s]);
                    if (kind == kPlog || kind == kCheb) R.kf[j] = L.ropf[s];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NF; j++) {
            const int s = j * 64 + lane;
            khigh_of[j] = (j < m.nFallRounds) ? arrhenius(m.A[s], m.b[s], m.EaR[s], logT, recipT, m.exp2t) : 0.0;
        }
    } else {
        double pA[NR], pb[NR], pE[NR];
#pragma unroll
        for (int j = 0; j < NR; j++) {
            pA[j] = pb[j] = pE[j] = 0.0;
            if (j < m.nRounds) {
                pA[j] = m.A[j * 64 + lane];
                pb[j] = m.b[j * 64 + lane];
                pE[j] = m.EaR[j * 64 + lane];
            }
        }
        // IdealGasPhase::getStandardChemPotentials: mu_k = RT (h_k/RT - s_k/R) + RT ln(P/Pref);
        // with the standard concentration c0 = P / RT the pressure cancels from
        // mu_k / RT - ln c0 = h_k/RT - s_k/R + ln T + ln(R / Pref)
        if (sp) {
            const double x = (hRT - sR) + (logT + m.lnRPref);
            const double e = emx_exp(x, m.exp2t);
            L.ge[lane - 2] = e;
            L.gi[lane - 2] = emx_div(1.0, e);
        }
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < NR; j++) {
            if (j < m.nRounds) {
                const int s = j * 64 + lane;
                const int w = m.info[s];
                const int kind = info_kind(w);
                const uint32_t rs = m.reac[s], ps = m.prod[s];
                const double g0 = L.ge[st_k(ps, 0)], g1 = L.gi[st_k(rs, 0)];
                const double g2 = L.ge[st_k(ps, 1)], g3 = L.gi[st_k(rs, 1)];
                const double g4 = L.ge[st_k(ps, 2)], g5 = L.gi[st_k(rs, 2)];
                const double ka = arrhenius(pA[j], pb[j], pE[j], logT, recipT, m.exp2t);
                const double ikc = (g0 * g1) * (g2 * g3) * (g4 * g5);  // inv_kc's pairing
                R.rkc[j] = (kind != kNone && info_rev(w)) ? fmin(ikc, kBigNumber) : 0.0;
                double kf = 0.0;
                if (kind <= kThreeBody) kf = ka;
                else if (kind == kPlog || kind == kCheb) kf = L.ropf[s];
                R.kf[j] = kf;
                if (j < NF) khigh_of[j] = ka;
            }
        }
    }
    // falloff: k_low, k_high and the T-dependent parts of F (FalloffRate, TroeRate and
    // SriRate::updateTemp)
#pragma unroll
    for (int j = 0; j < NF; j++) {
        if (j < m.nFallRounds) {
            const int s = j * 64 + lane;
            const int kind = info_kind(m.info[s]);
            double klow = 0.0, khigh = 0.0, work = 0.0, work2 = 0.0;
            // the round's falloff parameters, loaded together
            const double A0 = m.A0[s], b0 = m.b0[s], E0 = m.EaR0[s];
            const double Ta = m.troeA[s], T3 = m.rT3[s], T1 = m.rT1[s], T2 = m.T2[s];
            if (kind == kLindemann || kind == kTroe || kind == kSri) {
                klow = arrhenius(A0, b0, E0, logT, recipT, m.exp2t);
                khigh = khigh_of[j];
                if (kind == kTroe) {
                    double Fcent = (1.0 - Ta) * emx_exp(-T * T3, m.exp2t) + Ta * emx_exp(-T * T1, m.exp2t);
                    if (T2 != 0.0) Fcent += emx_exp(-T2 / T, m.exp2t);
                    work = emx_log(fmax(Fcent, kSmallNumber)) * kInvLn10;
                } else if (kind == kSri) {
                    const double* q = m.auxd + m.aux[s];  // a, b, c, d, e
                    const double base = q[0] * emx_exp(-q[1] * recipT, m.exp2t) + emx_exp(-T / q[2], m.exp2t);
                    work = emx_log(fmax(base, kSmallNumber)) * kInvLn10;
                    work2 = q[4] != 0.0 ? q[3] * emx_exp(q[4] * logT, m.exp2t) : q[3];
                }
            }
            R.klow[j] = klow;
            R.khigh[j] = khigh;
            R.work[j] = work;
            R.work2[j] = work2;
        }
    }
    R.T = T;
    return true;
}

// ThirdBodyCalc::update per efficiency group: Mg = default * ctot + sum (eff - default) C.
// Lane g < nGroups sums group g's entries in order, kGrpBatch at a time with the species
// indices, efficiencies and concentrations of a batch loaded together; entries past the
// group's last add exact zeros (efficiency 0 on the unit padding concentration).
__device__ inline void third_body_conc(const DevMech& m, int lane, const GasState& g,
                                       const WaveLds& L) {
    constexpr int kGrpBatch = 8;
    wave_lds_sync();  // concentrations written by set_state
    const bool act = lane < m.nGroups;
    const int e0 = act ? m.grp_off[lane] : 0;
    const int ne = act ? m.grp_off[lane + 1] - e0 : 0;
    const int elast = m.nGrpEntries > 0 ? m.nGrpEntries - 1 : 0;
    double sum = 0.0;
    for (int i = 0; i < m.grpMaxLen; i += kGrpBatch) {
        int sp[kGrpBatch];
        double ef[kGrpBatch];
#pragma unroll
        for (int u = 0; u < kGrpBatch; u++) {
            const int e = min(e0 + i + u, elast);
            const bool in = i + u < ne;
            const int k = m.grp_sp[e];
            const double f = m.grp_eff[e];
            sp[u] = in ? k : kPadSpecies;
            ef[u] = in ? f : 0.0;
        }
        double c[kGrpBatch];
#pragma unroll
        for (int u = 0; u < kGrpBatch; u++) c[u] = L.C[sp[u]];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kGrpBatch; u++) sum += ef[u] * c[u];
    }
    if (act) L.Mg[lane] = m.grp_def[lane] * g.ctot + sum;
    wave_lds_sync();
}

// Rates of progress into LDS (composition-dependent part of updateROP).
// Rate multipliers perturb[s] * fa (Cantera m_perturb; fa: the rate multiplier function).
// Each round's table words and concentration gathers are issued before its arithmetic,
// and the slot kinds are applied with selects rather than branches (a padding slot has
// k_f = 0 and padding stoichiometry, and is written as exact zeros; a slot without third
// body reads the unit cell Mg[kNoGroup]), so the loads of a round are in flight together.
// Only the falloff rounds keep a branch around their pressure-dependent factor.
// Cantera's C_AnyN::multiply for a reaction with explicit forward orders: the product of
// c_j^o_j over the reactant positions with a nonzero order, zero when any of those
// concentrations is <= 0.  c_j^o_j is formed as exp(sum_j o_j ln c_j).
__device__ __forceinline__ double ordered_conc_product(double c0, double c1, double c2, double o0,
                                                       double o1, double o2, const double* exp2t) {
    const bool zero = (o0 != 0.0 && !(c0 > 0.0)) || (o1 != 0.0 && !(c1 > 0.0)) || (o2 != 0.0 && !(c2 > 0.0));
    const double l = (o0 != 0.0 ? o0 * emx_log(fmax(c0, kSmallNumber)) : 0.0) +
                     (o1 != 0.0 ? o1 * emx_log(fmax(c1, kSmallNumber)) : 0.0) +
                     (o2 != 0.0 ? o2 * emx_log(fmax(c2, kSmallNumber)) : 0.0);
    return zero ? 0.0 : emx_exp(l, exp2t);
}

// ORD: the mechanism has reactions with explicit forward orders (m.has_orders); the
// instantiation without it is the plain mass-action kernel.
template <int NR, bool ORD = false>
__device__ inline void rates_of_progress(const DevMech& m, int lane, const GasState& g,
                                         const RateRegs<NR>& R, const double* perturb, double fa,
                                         const WaveLds& L) {
    third_body_conc(m, lane, g, L);
    // the table words of every round first (one LDS latency for all of them)
    int wd[NR];
    uint32_t rsw[NR], psw[NR];
    double ptw[NR];
    double ow[ORD ? 3 * NR : 1];
#pragma unroll
    for (int j = 0; j < NR; j++) {
        wd[j] = 0;
        rsw[j] = psw[j] = 0;
        ptw[j] = 0.0;
        if (j < m.nRounds) {
            wd[j] = m.info[j * 64 + lane];
            rsw[j] = m.reac[j * 64 + lane];
            psw[j] = m.prod[j * 64 + lane];
            ptw[j] = perturb[j * 64 + lane];
        }
        if (ORD) {
            const int S = m.nRounds * 64;
#pragma unroll
            for (int q = 0; q < 3; q++) ow[ORD ? 3 * j + q : 0] = (j < m.nRounds) ? m.ord[q * S + j * 64 + lane] : 0.0;
        }
    }
#pragma unroll
    for (int j = 0; j < NR; j++) {
        if (j < m.nRounds) {
            const int s = j * 64 + lane;
            const int w = wd[j];
            const int kind = info_kind(w);
            const uint32_t rs = rsw[j], ps = psw[j];
            const double pt = ptw[j];
            // the round's gathers, all issued before its arithmetic
            const double mg = L.Mg[min(info_group(w), kNoGroup)];
            const double c0 = L.C[st_k(rs, 0)], c1 = L.C[st_k(rs, 1)], c2 = L.C[st_k(rs, 2)];
            const double d0 = L.C[st_k(ps, 0)], d1 = L.C[st_k(ps, 1)], d2 = L.C[st_k(ps, 2)];
            __builtin_amdgcn_sched_barrier(0);
            double cr = c0 * c1 * c2;
            if (ORD && info_ord(w))
                cr = ordered_conc_product(c0, c1, c2, ow[ORD ? 3 * j : 0], ow[ORD ? 3 * j + 1 : 0],
                                          ow[ORD ? 3 * j + 2 : 0], m.exp2t);
            const double cp = d0 * d1 * d2;
            double kf = R.kf[j];
            if (j < RateRegs<NR>::NF && (kind == kLindemann || kind == kTroe || kind == kSri)) {
                double khigh = R.khigh[j];
                double pr = emx_div(mg * R.klow[j], khigh + kSmallNumber);
                double F = 1.0;
                if (kind == kTroe) {
                    double work = R.work[j];
                    double lpr = emx_log(fmax(pr, kSmallNumber)) * kInvLn10;
                    double cc = -0.4 - 0.67 * work;
                    double nn = 0.75 - 1.27 * work;
                    double f1 = emx_div(lpr + cc, nn - 0.14 * (lpr + cc));
                    double lf = emx_div(work, 1.0 + f1 * f1);
                    F = emx_exp(kLn10 * lf, m.exp2t);
                } else if (kind == kSri) {
                    // SriRate: F = d T^e [a exp(-b/T) + exp(-T/c)]^X, X = 1/(1 + log10^2 Pr)
                    const double lpr = emx_log(fmax(pr, kSmallNumber)) * kInvLn10;
                    const double X = emx_div(1.0, 1.0 + lpr * lpr);
                    F = R.work2[j] * emx_exp(kLn10 * (X * R.work[j]), m.exp2t);
                }
                pr *= emx_div(F, 1.0 + pr);
                kf = pr * khigh;
            }
            const double rfn = kf * pt * fa;
            double ropf = (kind == kThreeBody) ? rfn * mg : rfn;
            double ropr = ropf * R.rkc[j];
            ropf *= cr;
            if (info_rev(w)) ropr *= cp;
            const bool none = kind == kNone;
            L.ropf[s] = none ? 0.0 : ropf;
            L.ropr[s] = none ? 0.0 : ropr;
        }
    }
    wave_lds_sync();
}

// Species creation / destruction rates: fixed-length chunk sums (chunk c = cr*64 + lane,
// kChunkLen entries, padding entries add exact zeros), then per-species totals over the
// species' consecutive chunks.  Returns (wQ, wD) for this lane's species (0 elsewhere).
// The gathers of a chunk are all issued before its sums (the scheduling barrier keeps the
// compiler from interleaving each load with its use, which serialised the LDS latencies);
// the sums keep their sequential order.
__device__ inline void species_rates(const DevMech& m, int lane, const WaveLds& L, double& wQ,
                                     double& wD) {
#pragma unroll
    for (int cr = 0; cr < kMaxChunkRounds; cr++) {
        if (cr < m.nChunkRounds) {
            const int c = cr * 64 + lane;
            const uint32_t* ep = m.entry + c * kChunkLen;
            uint32_t en[kChunkLen];
#pragma unroll
            for (int e = 0; e < kChunkLen; e++) en[e] = ep[e];
            double vq[kChunkLen], vd[kChunkLen];
#pragma unroll
            for (int e = 0; e < kChunkLen; e++) {
                vq[e] = L.ropf[en_q(en[e])];  // ropf, zero slot and ropr are one array
                vd[e] = L.ropf[en_d(en[e])];
            }
            __builtin_amdgcn_sched_barrier(0);
            double q = 0.0, d = 0.0;
#pragma unroll
            for (int e = 0; e < kChunkLen; e++) {
                const double nu = (double)en_nu(en[e]);
                q += nu * vq[e];
                d += nu * vd[e];
            }
            L.partq[c] = q;
            L.partd[c] = d;
        }
    }
    wave_lds_sync();
    wQ = 0.0;
    wD = 0.0;
    // per-species totals, kSumBatch chunks per batch with their loads issued together; a
    // batch's chunks past the species' last read the zero cell (x + 0 = x exactly, so the
    // totals equal the one-by-one sums)
    constexpr int kSumBatch = 4;
    const bool sp = is_species_lane(lane, m);
    const int k = sp ? lane - 2 : 0;
    const int c0 = sp ? m.sp_chunk0[k] : 0, nc = sp ? m.sp_nchunk[k] : 0;
    const int zero = m.nChunkRounds * 64;
    const int nmax = (int)wave_max0((double)nc);
    for (int i = 0; i < nmax; i += kSumBatch) {
        double aq[kSumBatch], ad[kSumBatch];
#pragma unroll
        for (int u = 0; u < kSumBatch; u++) {
            const int c = (i + u < nc) ? c0 + i + u : zero;
            aq[u] = L.partq[c];
            ad[u] = L.partd[c];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kSumBatch; u++) {
            wQ += aq[u];
            wD += ad[u];
        }
    }
    if (!sp) {
        wQ = 0.0;
        wD = 0.0;
    }
    wave_lds_sync();
}

}  // namespace emx
