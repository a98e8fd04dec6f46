// CPU FP64 oracle for the Ember split-step hot path -- TEST INFRASTRUCTURE ONLY.
//
// Restates, in plain C++17, the algorithms of:
//   src/qssintegrator.cpp            (QssIntegrator, CHEMEQ2-style QSS)
//   src/sourceSystem.cpp:24-31,398-531 (SourceSystem::updateThermo, SourceSystemQSS)
//   src/sourceSystem.cpp:104-164,268-338 (SourceSystemCVODE::f, fdJacobian)
//   src/chemistry0d.cpp:264-436,487-672 (InterpKinetics, MultiArrheniusInterp, CanteraGas)
//   src/diffusionSystem.cpp:10-76     (DiffusionSystem::get_A, get_k)
//   src/integrator.cpp:37-47,133-204  (Integrator::integrateToTime, TridiagonalIntegrator::step)
//   src/grid.cpp:43-69                (OneDimGrid::updateValues metrics)
//   src/flameSolver.cpp:702-718,754-854 (production / diffusion stage drivers)
// and the parts of Cantera 3.2 those call (ideal-gas Phase state setters, NasaPoly2,
// BulkKinetics::updateROP, ThirdBodyCalc, FalloffRate/TroeRate, StoichManagerN),
// restated from Cantera's published algorithms.  SUNDIALS CVODE is replaced by a
// variable-order BDF (Shampine & Reichelt, "The MATLAB ODE Suite", SIAM J. Sci.
// Comput. 18 (1997), NDF family with kappa = 0) restated from scipy's
// scipy.integrate._ivp.bdf (SciPy 1.15, BSD-3-Clause), with CVODE's per-step failure
// bounds, applied with Ember's tolerances and Ember's finite-difference Jacobian.
//
// Only the tests, the smoke check and the benchmark's CPU-baseline leg load this.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kGasConstant = 6.02214076e26 * 1.380649e-23;  // J/kmol/K (Cantera ct_defs.h)
constexpr double kSmallNumber = 1e-300;                        // Cantera SmallNumber
constexpr double kBigNumber = 1e300;                           // Cantera BigNumber
// failed attempts allowed per BDF step before the point fails (CVODE's defaults maxncf and
// maxnef, CVodeSetMaxConvFails / CVodeSetMaxErrTestFails)
constexpr int kBdfMaxConvFail = 10;
constexpr int kBdfMaxErrFail = 7;

// ---------------------------------------------------------------------------
// Mechanism tables (written by ember_amd/mechanism.py)
// ---------------------------------------------------------------------------
struct Rxn {
    int type = 0, rev = 1, nr = 0, np = 0;
    int r[3] = {-1, -1, -1}, p[3] = {-1, -1, -1};
    double A = 0, b = 0, EaR = 0, A0 = 0, b0 = 0, EaR0 = 0;
    double troeA = 0, T3 = 0, T1 = 0, T2 = 0;
    int hasT2 = 0;
    double defeff = 1.0;
    std::vector<std::pair<int, double>> eff;
    // type 4, pressure-dependent Arrhenius (Cantera PlogRate): (P, A, b, Ea/R) per rate
    std::vector<double> plog;
    // type 5, Cantera ChebyshevRate: nT x nP coefficients of log10 k (row-major)
    int chebT = 0, chebP = 0;
    double chebTmin = 0, chebTmax = 0, chebPmin = 0, chebPmax = 0;
    std::vector<double> cheb;
    // type 6, Cantera SriRate falloff: a, b, c, d, e
    double sri[5] = {0, 0, 0, 1, 0};
    // explicit forward orders of reactant positions 0..2 (irreversible reactions; the
    // first position of a species carries its order, repeated positions 0)
    int hasOrd = 0;
    double ord[3] = {0, 0, 0};
};

struct Mech {
    int K = 0, I = 0, E = 0;
    double pref = 101325.0;
    std::vector<std::string> names;
    std::vector<double> W, Tmid, low, high;  // low/high: K x 7
    std::vector<Rxn> rxn;
    std::vector<double> perturb;
    // Ember's rate multiplier function (a ChebyshevFunction): [T0, T1, c_0, c_1, ...] as
    // ChebyshevFunction::setCoefficients stores them (src/scalarFunction.cpp:33-51); empty
    // when none is set
    std::vector<double> ratefn;
    std::vector<double> dn;
    // Cantera StoichManagerN lists: grouped by molecule count (1, 2, 3)
    struct Stoich { int rxn; int n; int k[3]; };
    std::vector<Stoich> reac, prod, revprod;
    std::vector<int> thirdbody;  // reactions with a third body (three-body + falloff)
    std::vector<int> tb_slot;    // reaction -> slot in concm, or -1
    // interp tables (InterpKinetics / MultiArrheniusInterp) over [TminInt, TmaxInt]
    int nInterp = 200;
    double TminInt = 273.0, TmaxInt = 3500.0;
    std::vector<double> kf_const, kf_slope, kc_const, kc_slope;  // [n][I]
    bool interp_ready = false;
    double interp_P = -1;
    // transport fits (TRANSPORT records): coefficients of polynomials in ln T for
    // sqrt(mu_k)/T^(1/4) and lambda_k/sqrt(T) (K x 5) and D_jk(T, 1 Pa)/T^(3/2) (K x K x 5)
    double trTmin = 0, trTmax = 0;
    std::vector<double> visc, cond, diff;
};

// ChebyshevFunction::a (src/scalarFunction.cpp:53-72): t mapped onto [-1, 1] over
// [T0, T1], then the sum of c_i T_i(x) with the three-term recurrence
double rate_fn_eval(const Mech& m, double t) {
    const double* f = m.ratefn.data();
    const int nc = (int)m.ratefn.size() - 2;
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

void build_stoich(Mech& m) {
    m.reac.clear(); m.prod.clear(); m.revprod.clear();
    for (int n = 1; n <= 3; n++) {
        for (int i = 0; i < m.I; i++) {
            const Rxn& r = m.rxn[i];
            if (r.nr == n) m.reac.push_back({i, n, {r.r[0], r.r[1], r.r[2]}});
        }
    }
    for (int n = 1; n <= 3; n++) {
        for (int i = 0; i < m.I; i++) {
            const Rxn& r = m.rxn[i];
            if (r.np == n) {
                m.prod.push_back({i, n, {r.p[0], r.p[1], r.p[2]}});
                if (r.rev) m.revprod.push_back({i, n, {r.p[0], r.p[1], r.p[2]}});
            }
        }
    }
    m.dn.assign(m.I, 0.0);
    m.thirdbody.clear();
    m.tb_slot.assign(m.I, -1);
    for (int i = 0; i < m.I; i++) {
        m.dn[i] = double(m.rxn[i].np) - double(m.rxn[i].nr);
        if ((m.rxn[i].type >= 1 && m.rxn[i].type <= 3) || m.rxn[i].type == 6) {  // three-body and falloff
            m.tb_slot[i] = (int)m.thirdbody.size();
            m.thirdbody.push_back(i);
        }
    }
}

Mech* load_mech(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) return nullptr;
    auto* m = new Mech();
    char tag[64];
    char buf[256];
    bool ok = true;
    if (fscanf(f, "%63s %255s", tag, buf) != 2 || strcmp(tag, "EMXM") != 0) ok = false;
    while (ok && fscanf(f, "%63s", tag) == 1) {
        if (!strcmp(tag, "END")) break;
        if (!strcmp(tag, "name")) { ok = fscanf(f, "%255s", buf) == 1; }
        else if (!strcmp(tag, "counts")) {
            ok = fscanf(f, "%d %d %d", &m->K, &m->I, &m->E) == 3;
            m->names.reserve(m->K);
        } else if (!strcmp(tag, "pref")) { ok = fscanf(f, "%lf", &m->pref) == 1; }
        else if (!strcmp(tag, "elements")) {
            for (int e = 0; e < m->E && ok; e++) ok = fscanf(f, "%255s", buf) == 1;
        } else if (!strcmp(tag, "S")) {
            double v[1 + 3 + 14];
            ok = fscanf(f, "%255s", buf) == 1;
            for (int j = 0; j < 18 && ok; j++) ok = fscanf(f, "%lf", &v[j]) == 1;
            for (int e = 0; e < m->E && ok; e++) { double c; ok = fscanf(f, "%lf", &c) == 1; }
            m->names.push_back(buf);
            m->W.push_back(v[0]);
            m->Tmid.push_back(v[2]);
            for (int j = 0; j < 7; j++) m->low.push_back(v[4 + j]);
            for (int j = 0; j < 7; j++) m->high.push_back(v[11 + j]);
        } else if (!strcmp(tag, "R")) {
            Rxn r;
            int neff = 0;
            ok = fscanf(f, "%d %d %d %d %d %d %d %d %d %d", &r.type, &r.rev, &r.nr, &r.r[0], &r.r[1],
                        &r.r[2], &r.np, &r.p[0], &r.p[1], &r.p[2]) == 10;
            ok = ok && fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %d %lf %d", &r.A, &r.b,
                              &r.EaR, &r.A0, &r.b0, &r.EaR0, &r.troeA, &r.T3, &r.T1, &r.T2,
                              &r.hasT2, &r.defeff, &neff) == 13;
            for (int j = 0; j < neff && ok; j++) {
                int k; double e;
                ok = fscanf(f, "%d %lf", &k, &e) == 2;
                r.eff.emplace_back(k, e);
            }
            m->rxn.push_back(r);
        } else if (!strcmp(tag, "PLOG")) {
            int n = 0;
            ok = !m->rxn.empty() && m->rxn.back().type == 4 && fscanf(f, "%d", &n) == 1 && n >= 1;
            if (ok) m->rxn.back().plog.resize(4 * (size_t)n);
            for (int j = 0; j < 4 * n && ok; j++) ok = fscanf(f, "%lf", &m->rxn.back().plog[j]) == 1;
        } else if (!strcmp(tag, "ORD")) {
            Rxn* r = m->rxn.empty() ? nullptr : &m->rxn.back();
            ok = r && !r->rev && fscanf(f, "%lf %lf %lf", &r->ord[0], &r->ord[1], &r->ord[2]) == 3;
            if (ok) r->hasOrd = 1;
        } else if (!strcmp(tag, "SRI")) {
            Rxn* r = m->rxn.empty() ? nullptr : &m->rxn.back();
            ok = r && r->type == 6 &&
                 fscanf(f, "%lf %lf %lf %lf %lf", &r->sri[0], &r->sri[1], &r->sri[2], &r->sri[3], &r->sri[4]) == 5;
        } else if (!strcmp(tag, "CHEB")) {
            Rxn* r = m->rxn.empty() ? nullptr : &m->rxn.back();
            ok = r && r->type == 5 &&
                 fscanf(f, "%d %d %lf %lf %lf %lf", &r->chebT, &r->chebP, &r->chebTmin, &r->chebTmax,
                        &r->chebPmin, &r->chebPmax) == 6 && r->chebT >= 1 && r->chebP >= 1;
            if (ok) r->cheb.resize((size_t)r->chebT * r->chebP);
            for (size_t j = 0; ok && j < r->cheb.size(); j++) ok = fscanf(f, "%lf", &r->cheb[j]) == 1;
        } else if (!strcmp(tag, "TRANSPORT")) {
            ok = fscanf(f, "%lf %lf", &m->trTmin, &m->trTmax) == 2 && m->K > 0;
            m->visc.assign((size_t)m->K * 5, 0.0);
            m->cond.assign((size_t)m->K * 5, 0.0);
            m->diff.assign((size_t)m->K * m->K * 5, 0.0);  // full symmetric K x K
        } else if (!strcmp(tag, "TV") || !strcmp(tag, "TC") || !strcmp(tag, "TD")) {
            int j = 0, k = 0;
            const bool pair = tag[1] == 'D';
            ok = !m->visc.empty() && fscanf(f, "%d", &j) == 1 && (!pair || fscanf(f, "%d", &k) == 1) &&
                 j >= 0 && j < m->K && k >= 0 && k < m->K;
            double c[5];
            for (int i = 0; i < 5 && ok; i++) ok = fscanf(f, "%lf", &c[i]) == 1;
            if (ok) {
                if (tag[1] == 'V') std::copy(c, c + 5, &m->visc[(size_t)j * 5]);
                else if (tag[1] == 'C') std::copy(c, c + 5, &m->cond[(size_t)j * 5]);
                else {
                    std::copy(c, c + 5, &m->diff[((size_t)j * m->K + k) * 5]);
                    std::copy(c, c + 5, &m->diff[((size_t)k * m->K + j) * 5]);
                }
            }
        } else {
            ok = false;
        }
    }
    fclose(f);
    if (!ok || (int)m->names.size() != m->K || (int)m->rxn.size() != m->I) {
        delete m;
        return nullptr;
    }
    m->perturb.assign(m->I, 1.0);
    build_stoich(*m);
    return m;
}

// ---------------------------------------------------------------------------
// Cantera NasaPoly2 / NasaPoly1::updateProperties
// ---------------------------------------------------------------------------
void nasa_props(const Mech& m, int k, const double* tt, double& cp_R, double& h_RT, double& s_R) {
    const double* c = (tt[0] <= m.Tmid[k]) ? &m.low[7 * k] : &m.high[7 * k];
    double ct0 = c[0];
    double ct1 = c[1] * tt[0];
    double ct2 = c[2] * tt[1];
    double ct3 = c[3] * tt[2];
    double ct4 = c[4] * tt[3];
    cp_R = ct0 + ct1 + ct2 + ct3 + ct4;
    h_RT = ct0 + 0.5 * ct1 + 1.0 / 3.0 * ct2 + 0.25 * ct3 + 0.2 * ct4 + c[5] * tt[4];
    s_R = ct0 * tt[5] + ct1 + 0.5 * ct2 + 1.0 / 3.0 * ct3 + 0.25 * ct4 + c[6];
}

// ---------------------------------------------------------------------------
// Ideal gas phase + bulk kinetics (Cantera restatement; one instance per thread,
// like Ember's per-thread CanteraGas, src/flameSolver.h:167)
// ---------------------------------------------------------------------------
struct Gas {
    Mech& m;
    int kinetics;
    double P;
    double T = 300.0, mmw = 0, dens = 0;
    std::vector<double> y, ym, conc;
    std::vector<double> cpR, hRT, sR, grt, mu0;
    std::vector<double> kf0, rkcn, concm, rfn, ropf, ropr, ropnet;
    double thermoT = -1.0, rateT = -1.0;
    const double* perturb = nullptr;  // per-thread rate multipliers (NULL: the mechanism's)
    // CanteraGas::setRateMultiplier(a) (src/chemistry0d.cpp:634-642): every reaction's
    // multiplier set to one value, in place of the per-reaction ones
    bool uniform_mult = false;
    double mult_all = 1.0;
    // the multipliers at time t when the mechanism has a rate multiplier function
    // (SourceSystemQSS::odefun / SourceSystemCVODE::f, src/sourceSystem.cpp:453-455, :133-135)
    void rate_function_at(double t) {
        uniform_mult = !m.ratefn.empty();
        if (uniform_mult) mult_all = rate_fn_eval(m, t);
    }

    Gas(Mech& mech, int kin, double pressure)
        : m(mech), kinetics(kin), P(pressure), y(m.K), ym(m.K), conc(m.K), cpR(m.K), hRT(m.K),
          sR(m.K), grt(m.K), mu0(m.K), kf0(m.I), rkcn(m.I), concm(m.thirdbody.size()),
          rfn(m.I), ropf(m.I), ropr(m.I), ropnet(m.I) {}

    // Phase::setMassFractions: clip negatives, normalise
    void setMassFractions(const double* Y) {
        for (int k = 0; k < m.K; k++) y[k] = std::max(Y[k], 0.0);
        double norm = 0.0;
        for (int k = 0; k < m.K; k++) norm += y[k];
        double s = 1.0 / norm;
        for (int k = 0; k < m.K; k++) y[k] = y[k] * s;
        double sum = 0.0;
        for (int k = 0; k < m.K; k++) { ym[k] = y[k] * (1.0 / m.W[k]); sum += ym[k]; }
        mmw = 1.0 / sum;
    }
    // Phase::setMassFractions_NoNorm
    void setMassFractions_NoNorm(const double* Y) {
        double sum = 0.0;
        for (int k = 0; k < m.K; k++) { y[k] = Y[k]; ym[k] = y[k] * (1.0 / m.W[k]); sum += ym[k]; }
        mmw = 1.0 / sum;
    }
    // ThermoPhase::setState_TP for IdealGasPhase: density from the pressure
    void setState_TP(double Tin) {
        T = Tin;
        dens = P * mmw / (T * kGasConstant);
    }
    double pressure() const { return kGasConstant * (dens / mmw) * T; }
    double RT() const { return T * kGasConstant; }

    void updateThermo() {
        if (thermoT == T) return;
        double tt[6] = {T, T * T, T * T * T, T * T * T * T, 1.0 / T, std::log(T)};
        for (int k = 0; k < m.K; k++) {
            nasa_props(m, k, tt, cpR[k], hRT[k], sR[k]);
            grt[k] = hRT[k] - sR[k];
        }
        thermoT = T;
    }
    double cp_mass() {
        updateThermo();
        double s = 0.0;
        for (int k = 0; k < m.K; k++) s += ym[k] * cpR[k];
        double cp_mole = kGasConstant * (mmw * s);
        return cp_mole / mmw;
    }
    void enthalpies(double* hk) {
        updateThermo();
        double rt = RT();
        for (int k = 0; k < m.K; k++) hk[k] = rt * hRT[k];
    }

    // reverse-rate factor 1/Kc at the current T (BulkKinetics::updateROP, first block)
    void rkc_at(double* out) {
        updateThermo();
        double rt = RT();
        double tmp = std::log(pressure() / m.pref) * rt;
        for (int k = 0; k < m.K; k++) mu0[k] = rt * grt[k] + tmp;
        double logStandConc = std::log(pressure() / rt);
        double rrt = 1.0 / rt;
        std::vector<double>& dg = rfn;  // scratch
        std::fill(dg.begin(), dg.end(), 0.0);
        for (const auto& s : m.revprod) {
            if (s.n == 1) dg[s.rxn] += mu0[s.k[0]];
            else if (s.n == 2) dg[s.rxn] += mu0[s.k[0]] + mu0[s.k[1]];
            else dg[s.rxn] += mu0[s.k[0]] + mu0[s.k[1]] + mu0[s.k[2]];
        }
        for (const auto& s : m.reac) {
            if (s.n == 1) dg[s.rxn] -= mu0[s.k[0]];
            else if (s.n == 2) dg[s.rxn] -= (mu0[s.k[0]] + mu0[s.k[1]]);
            else dg[s.rxn] -= (mu0[s.k[0]] + mu0[s.k[1]] + mu0[s.k[2]]);
        }
        for (int i = 0; i < m.I; i++) {
            out[i] = m.rxn[i].rev ? std::min(std::exp(dg[i] * rrt - m.dn[i] * logStandConc), kBigNumber)
                                  : 0.0;
        }
    }

    static double arrhenius(double A, double b, double EaR, double logT, double recipT) {
        return A * std::exp(b * logT - EaR * recipT);
    }

    // Cantera PlogRate at pressure P: ln k interpolated linearly in ln P between the rates
    // of the two listed pressures that bracket P (P1 <= P < P2); below the first or from
    // the last pressure on, that level alone.  A level's rates are summed (one rate: its
    // logarithm directly).
    static double plog_kf(const Rxn& r, double P, double logT, double recipT) {
        const int n = (int)r.plog.size() / 4;
        std::vector<int> lev;  // first rate of each distinct pressure
        for (int j = 0; j < n; j++)
            if (j == 0 || r.plog[4 * j] != r.plog[4 * (j - 1)]) lev.push_back(j);
        lev.push_back(n);
        const int nl = (int)lev.size() - 1;
        auto level_log = [&](int l) {
            const int a = lev[l], e = lev[l + 1];
            if (e - a == 1) return std::log(r.plog[4 * a + 1]) + r.plog[4 * a + 2] * logT - r.plog[4 * a + 3] * recipT;
            double k = 1e-300;
            for (int j = a; j < e; j++) k += arrhenius(r.plog[4 * j + 1], r.plog[4 * j + 2], r.plog[4 * j + 3], logT, recipT);
            return std::log(k);
        };
        const double logP = std::log(P);
        if (logP < std::log(r.plog[0])) return std::exp(level_log(0));
        if (logP >= std::log(r.plog[4 * lev[nl - 1]])) return std::exp(level_log(nl - 1));
        int hi = 0;
        while (std::log(r.plog[4 * lev[hi]]) <= logP) hi++;
        const double logP1 = std::log(r.plog[4 * lev[hi - 1]]), logP2 = std::log(r.plog[4 * lev[hi]]);
        const double k1 = level_log(hi - 1), k2 = level_log(hi);
        return std::exp(k1 + (k2 - k1) * ((logP - logP1) * (1.0 / (logP2 - logP1))));
    }

    // Cantera ChebyshevRate: log10 k = sum_t sum_p a_tp T_t(Tr) T_p(Pr) with
    // Tr = (2/T - 1/Tmin - 1/Tmax) / (1/Tmax - 1/Tmin) and
    // Pr = (2 log10 P - log10 Pmin - log10 Pmax) / (log10 Pmax - log10 Pmin)
    static double cheb_kf(const Rxn& r, double P, double recipT) {
        const double Pr = (2 * std::log10(P) - std::log10(r.chebPmin) - std::log10(r.chebPmax)) /
                          (std::log10(r.chebPmax) - std::log10(r.chebPmin));
        const double Tr = (2 * recipT - 1.0 / r.chebTmin - 1.0 / r.chebTmax) / (1.0 / r.chebTmax - 1.0 / r.chebTmin);
        std::vector<double> Tt(r.chebT), Tp(r.chebP);
        for (int i = 0; i < r.chebT; i++) Tt[i] = i == 0 ? 1.0 : i == 1 ? Tr : 2 * Tr * Tt[i - 1] - Tt[i - 2];
        for (int i = 0; i < r.chebP; i++) Tp[i] = i == 0 ? 1.0 : i == 1 ? Pr : 2 * Pr * Tp[i - 1] - Tp[i - 2];
        double logk = 0.0;
        for (int t = 0; t < r.chebT; t++)
            for (int p = 0; p < r.chebP; p++) logk += r.cheb[(size_t)t * r.chebP + p] * Tt[t] * Tp[p];
        return std::pow(10.0, logk);
    }

    double pdep_kf(const Rxn& r, double logT, double recipT) const {
        return r.type == 4 ? plog_kf(r, P, logT, recipT) : cheb_kf(r, P, recipT);
    }

    // InterpKinetics / MultiArrheniusInterp tables over [TminInt, TmaxInt]
    // (src/chemistry0d.cpp:291-331, 415-436); the last node keeps slope 0.
    void build_interp() {
        if (m.interp_ready && m.interp_P == P) return;
        int n = m.nInterp;
        m.kf_const.assign((size_t)n * m.I, 0.0);
        m.kf_slope.assign((size_t)n * m.I, 0.0);
        m.kc_const.assign((size_t)n * m.I, 0.0);
        m.kc_slope.assign((size_t)n * m.I, 0.0);
        double Tsave = T, dsave = dens;
        for (int j = 0; j < n; j++) {
            double Tn = m.TminInt + ((double)j) / (n - 1) * (m.TmaxInt - m.TminInt);
            T = Tn;
            dens = P * mmw / (T * kGasConstant);
            thermoT = -1;
            rkc_at(&m.kc_const[(size_t)j * m.I]);
            double logT = std::log(Tn), recipT = 1.0 / Tn;
            for (int i = 0; i < m.I; i++) {
                const Rxn& r = m.rxn[i];
                if (r.type <= 1) m.kf_const[(size_t)j * m.I + i] = arrhenius(r.A, r.b, r.EaR, logT, recipT);
            }
        }
        double dT = (m.TmaxInt - m.TminInt) / (n - 1);
        for (int j = 0; j < n - 1; j++) {
            for (int i = 0; i < m.I; i++) {
                size_t a = (size_t)j * m.I + i, b = (size_t)(j + 1) * m.I + i;
                m.kf_slope[a] = (m.kf_const[b] - m.kf_const[a]) / dT;
                m.kc_slope[a] = (m.kc_const[b] - m.kc_const[a]) / dT;
            }
        }
        T = Tsave; dens = dsave; thermoT = -1;
        m.interp_ready = true;
        m.interp_P = P;
    }

    // T-dependent rate data: kf for Arrhenius-type reactions and rkcn
    bool updateRateConstants() {
        if (rateT == T) return true;
        if (kinetics == 1) {
            if (T < m.TminInt || T > m.TmaxInt) return false;
            size_t n = (size_t)std::floor((T - m.TminInt) / (m.TmaxInt - m.TminInt) * (m.nInterp - 1));
            double dT = T - n * (m.TmaxInt - m.TminInt) / (m.nInterp - 1) - m.TminInt;
            const double logT = std::log(T), recipT = 1.0 / T;
            for (int i = 0; i < m.I; i++) {
                rkcn[i] = m.kc_const[n * m.I + i] + m.kc_slope[n * m.I + i] * dT;
                if (m.rxn[i].type <= 1)
                    kf0[i] = m.kf_const[n * m.I + i] + m.kf_slope[n * m.I + i] * dT;
                else if (m.rxn[i].type == 4 || m.rxn[i].type == 5)  // only Arrhenius rates are interpolated
                    kf0[i] = pdep_kf(m.rxn[i], logT, recipT);
            }
        } else {
            rkc_at(rkcn.data());
            double logT = std::log(T), recipT = 1.0 / T;
            for (int i = 0; i < m.I; i++) {
                const Rxn& r = m.rxn[i];
                if (r.type <= 1) kf0[i] = arrhenius(r.A, r.b, r.EaR, logT, recipT);
                else if (r.type == 4 || r.type == 5) kf0[i] = pdep_kf(r, logT, recipT);
            }
        }
        rateT = T;
        return true;
    }

    // BulkKinetics::updateROP (composition-dependent part every call)
    bool updateROP() {
        if (!updateRateConstants()) return false;
        for (int k = 0; k < m.K; k++) conc[k] = ym[k] * dens;
        double ctot = dens / mmw;
        // ThirdBodyCalc::update
        for (size_t s = 0; s < m.thirdbody.size(); s++) {
            const Rxn& r = m.rxn[m.thirdbody[s]];
            double sum = 0.0;
            for (const auto& e : r.eff) sum += (e.second - r.defeff) * conc[e.first];
            concm[s] = r.defeff * ctot + sum;
        }
        // falloff rates (FalloffRate::evalFromStruct)
        double logT = std::log(T), recipT = 1.0 / T;
        for (int i = 0; i < m.I; i++) {
            const Rxn& r = m.rxn[i];
            if (r.type != 2 && r.type != 3 && r.type != 6) continue;
            double work = 0.0;
            if (r.type == 3) {
                double rt3 = 1.0 / r.T3, rt1 = 1.0 / r.T1;
                double Fcent = (1.0 - r.troeA) * std::exp(-T * rt3) + r.troeA * std::exp(-T * rt1);
                if (r.T2 != 0.0) Fcent += std::exp(-r.T2 / T);
                work = std::log10(std::max(Fcent, kSmallNumber));
            }
            double klow = arrhenius(r.A0, r.b0, r.EaR0, logT, recipT);
            double khigh = arrhenius(r.A, r.b, r.EaR, logT, recipT);
            double pr = concm[m.tb_slot[i]] * klow / (khigh + kSmallNumber);
            double F = 1.0;
            if (r.type == 3) {
                double lpr = std::log10(std::max(pr, kSmallNumber));
                double cc = -0.4 - 0.67 * work;
                double nn = 0.75 - 1.27 * work;
                double f1 = (lpr + cc) / (nn - 0.14 * (lpr + cc));
                double lf = work / (1.0 + f1 * f1);
                F = std::pow(10.0, lf);
            } else if (r.type == 6) {
                // SriRate: F = d T^e [a exp(-b/T) + exp(-T/c)]^X, X = 1 / (1 + log10^2 Pr)
                const double base = r.sri[0] * std::exp(-r.sri[1] / T) + std::exp(-T / r.sri[2]);
                const double lpr = std::log10(std::max(pr, kSmallNumber));
                const double X = 1.0 / (1.0 + lpr * lpr);
                F = r.sri[3] * std::pow(T, r.sri[4]) * std::pow(10.0, X * std::log10(std::max(base, kSmallNumber)));
            }
            pr *= F / (1.0 + pr);
            kf0[i] = pr * khigh;
        }
        const double* pt = perturb ? perturb : m.perturb.data();
        if (uniform_mult)
            for (int i = 0; i < m.I; i++) rfn[i] = kf0[i] * mult_all;
        else
            for (int i = 0; i < m.I; i++) rfn[i] = kf0[i] * pt[i];
        for (int i = 0; i < m.I; i++) ropf[i] = rfn[i];
        for (int i = 0; i < m.I; i++)
            if (m.rxn[i].type == 1) ropf[i] *= concm[m.tb_slot[i]];
        for (int i = 0; i < m.I; i++) ropr[i] = ropf[i];
        for (const auto& s : m.reac) {
            const Rxn& rx = m.rxn[s.rxn];
            if (rx.hasOrd) {
                // Cantera C_AnyN::multiply (StoichManager.h): each reactant with a nonzero
                // order multiplies by c^order; a concentration <= 0 zeroes the rate
                for (int j = 0; j < rx.nr; j++) {
                    if (rx.ord[j] == 0.0) continue;
                    const double c = conc[rx.r[j]];
                    if (c > 0.0) ropf[s.rxn] *= std::pow(c, rx.ord[j]);
                    else ropf[s.rxn] = 0.0;
                }
            } else if (s.n == 1) ropf[s.rxn] *= conc[s.k[0]];
            else if (s.n == 2) ropf[s.rxn] *= conc[s.k[0]] * conc[s.k[1]];
            else ropf[s.rxn] *= conc[s.k[0]] * conc[s.k[1]] * conc[s.k[2]];
        }
        for (int i = 0; i < m.I; i++) ropr[i] *= rkcn[i];
        for (const auto& s : m.revprod) {
            if (s.n == 1) ropr[s.rxn] *= conc[s.k[0]];
            else if (s.n == 2) ropr[s.rxn] *= conc[s.k[0]] * conc[s.k[1]];
            else ropr[s.rxn] *= conc[s.k[0]] * conc[s.k[1]] * conc[s.k[2]];
        }
        for (int i = 0; i < m.I; i++) ropnet[i] = ropf[i] - ropr[i];
        return true;
    }

    static void increment(const std::vector<Mech::Stoich>& L, const double* in, double* out) {
        for (const auto& s : L)
            for (int j = 0; j < s.n; j++) out[s.k[j]] += in[s.rxn];
    }
    static void decrement(const std::vector<Mech::Stoich>& L, const double* in, double* out) {
        for (const auto& s : L)
            for (int j = 0; j < s.n; j++) out[s.k[j]] -= in[s.rxn];
    }
    bool creationRates(double* cdot) {
        if (!updateROP()) return false;
        std::fill(cdot, cdot + m.K, 0.0);
        increment(m.prod, ropf.data(), cdot);
        increment(m.reac, ropr.data(), cdot);
        return true;
    }
    void destructionRates(double* ddot) {  // after creationRates / updateROP
        std::fill(ddot, ddot + m.K, 0.0);
        increment(m.revprod, ropr.data(), ddot);
        increment(m.reac, ropf.data(), ddot);
    }
    bool netRates(double* wdot) {
        if (!updateROP()) return false;
        std::fill(wdot, wdot + m.K, 0.0);
        increment(m.prod, ropnet.data(), wdot);
        decrement(m.reac, ropnet.data(), wdot);
        return true;
    }
};

// ---------------------------------------------------------------------------
// QssIntegrator + SourceSystemQSS (src/qssintegrator.cpp, src/sourceSystem.cpp:398-531)
// ---------------------------------------------------------------------------
int isign(double x) { return (x > 0) ? 1 : (x < 0) ? -1 : 0; }

// QssOde (src/qssintegrator.h:6-21)
struct QssOde {
    virtual ~QssOde() {}
    virtual void odefun(double t, const double* y, double* q, double* d, bool corrector) = 0;
    virtual bool failed() const { return false; }
};

// Record of one accepted internal step, kept for the discrete adjoint: the state and the
// predictor outputs at the step start, then every attempt (predictor value yp and the
// corrector outputs at it); the last attempt is the accepted one.
struct QssAttempt {
    double dt;
    std::vector<double> yp, qc, dc;
};
struct QssStepRecord {
    std::vector<double> y, q, d;
    double qdot;
    std::vector<QssAttempt> att;
};

// QssIntegrator (src/qssintegrator.h:24-110, src/qssintegrator.cpp)
struct QssIntegrator {
    QssOde* ode = nullptr;
    int N = 0;
    std::vector<double> y, q, d, rtau, rtaus, y1, ys, qs, scratch, ym1, ym2, ymin, rtaub, alpha;
    std::vector<char> enforce;
    double tn = 0, ts = 0, dt = 0, tstart = 0, tfd = 1.000008;
    bool firstStep = true;
    int gcount = 0, rcount = 0;
    // options (constructor defaults of src/qssintegrator.cpp:9-30)
    double epsmax = 20, epsmin = 1e-2, dtmin = 1e-15, dtmax = 1e-6, abstol = 1e-8;
    int itermax = 2;
    bool stabilityCheck = true;
    // heat-release bookkeeping for accepted steps (set by the ODE on predictor calls)
    double qint = 0, qdot_step = 0;
    double* qdot_src = nullptr;
    std::vector<QssStepRecord>* tape = nullptr;  // adjoint record (itermax 1), or none

    void initialize(int n) {
        N = n;
        y.assign(N, 0); q.assign(N, 0); d.assign(N, 0); rtau.assign(N, 0); rtaus.assign(N, 0);
        y1.assign(N, 0); ys.assign(N, 0); qs.assign(N, 0); scratch.assign(N, 0);
        ym1.assign(N, 0); ym2.assign(N, 0); rtaub.assign(N, 0); alpha.assign(N, 0);
        ymin.assign(N, 1e-20);
        enforce.assign(N, 1);
    }
    void setState(const double* yIn, double tstart_) {
        for (int i = 0; i < N; i++) y[i] = enforce[i] ? std::max(yIn[i], ymin[i]) : yIn[i];
        gcount = 0; rcount = 0; tstart = tstart_; tn = 0; firstStep = true;
        qint = 0;
    }
    void getInitialStepSize(double tf) {
        firstStep = false;
        double scrtch = 1.0e-25;
        for (int i = 0; i < N; i++) {
            if (std::abs(y[i]) > abstol) {
                double absq = std::abs(q[i]);
                double scr2 = std::abs(1 / y[i]) * isign(0.1 * epsmin * absq - d[i]);
                double scr1 = scr2 * d[i];
                scrtch = std::max(std::max(scr1, -std::abs(absq - d[i]) * scr2), scrtch);
            }
        }
        double sqreps = 0.5;
        dt = std::min(sqreps / scrtch, tf);
        dt = std::min(dt, dtmax);
    }
    int integrateToTime(double tf) {
        while (tfd * tn < tf) {
            int ret = integrateOneStep(tf);
            if (ode->failed()) return -2;
            if (ret) return ret;
        }
        return 0;
    }
    int integrateOneStep(double tf) {
        ode->odefun(tn + tstart, y.data(), q.data(), d.data(), false);
        gcount += 1;
        qdot_step = qdot_src ? *qdot_src : 0.0;
        if (tape) tape->push_back({y, q, d, qdot_step, {}});
        if (firstStep) getInitialStepSize(tf);
        ts = tn;
        for (int i = 0; i < N; i++) rtau[i] = enforce[i] ? dt * d[i] / y[i] : 0.0;
        qs = q; ys = y; rtaus = rtau;
        while (true) {
            for (int i = 0; i < N; i++) {
                double r = rtau[i];
                scratch[i] = (q[i] - d[i]) /
                             (1.0 + r * (180 + r * (60 + r * (11 + r))) / (360 + r * (60 + r * (12 + r))));
            }
            double eps = 1e-10;
            for (int iter = 0; iter < itermax; iter++) {
                if (stabilityCheck) { ym2 = ym1; ym1 = y; }
                for (int i = 0; i < N; i++) y[i] = std::max(ymin[i], ys[i] + dt * scratch[i]);
                if (iter == 0) { tn = ts + dt; y1 = y; }
                ode->odefun(tn + tstart, y.data(), q.data(), d.data(), true);
                gcount += 1;
                if (tape && iter == 0) tape->back().att.push_back({dt, y, q, d});
                eps = 1.0e-10;
                for (int i = 0; i < N; i++) rtaub[i] = enforce[i] ? 0.5 * (rtaus[i] + dt * d[i] / y[i]) : 0.0;
                for (int i = 0; i < N; i++) {
                    double r = rtaub[i];
                    alpha[i] = (180. + r * (60. + r * (11. + r))) / (360. + r * (60. + r * (12. + r)));
                }
                for (int i = 0; i < N; i++)
                    scratch[i] = (qs[i] * (1.0 - alpha[i]) + q[i] * alpha[i] - ys[i] * rtaub[i] / dt) /
                                 (1.0 + alpha[i] * rtaub[i]);
            }
            for (int i = 0; i < N; i++) {
                double scr2 = ys[i] + dt * scratch[i];
                if (enforce[i]) scr2 = std::max(scr2, ymin[i]);
                double scr1 = std::abs(scr2 - y1[i]);
                y[i] = std::max(scr2, ymin[i]);
                if (std::abs(y[i]) > abstol && 0.25 * (ys[i] + y[i]) > ymin[i]) {
                    scr1 = scr1 / y[i];
                    eps = std::max(.5 * (scr1 + std::min(std::abs(q[i] - d[i]) / (q[i] + d[i] + 1e-30), scr1)), eps);
                }
            }
            if (stabilityCheck) { ym2 = ym1; ym1 = y; }
            eps /= epsmin;
            if (dt <= dtmin + 1e-16 * tn) return -1;
            double stab = 0;
            if (stabilityCheck && itermax >= 3) {
                stab = 0.01;
                for (int i = 0; i < N; i++)
                    if (std::abs(y[i]) > abstol)
                        stab = std::max(stab, std::abs(y[i] - ym1[i]) / (std::abs(ym1[i] - ym2[i]) + 1e-20 * y[i]));
            }
            if (eps <= epsmax && stab <= 1.0) {
                qint += dt * qdot_step;
                if (tf <= tn * tfd) return 0;
            } else {
                tn = ts;
            }
            double rteps = 0.5 * (eps + 1.0);
            rteps = 0.5 * (rteps + eps / rteps);
            rteps = 0.5 * (rteps + eps / rteps);
            double dto = dt;
            dt = std::min(dt * (1.0 / rteps + 0.005), tfd * (tf - tn));
            dt = std::min(dt, dtmax);
            if (stabilityCheck) dt = std::min(dt, dto / (stab + .001));
            if (eps > epsmax || stab > 1) {
                rcount += 1;
                for (int i = 0; i < N; i++) rtaus[i] *= dt / dto;
            } else {
                return 0;
            }
        }
    }
};

// SourceSystem::getQdotIgniter (src/sourceSystem.cpp:33-45) at position x, time t
double qdot_igniter(const orc_options& opt, double x, double t) {
    if (t >= opt.ignition_t_start && t < opt.ignition_t_start + opt.ignition_duration) {
        return opt.ignition_energy /
               (opt.ignition_stddev * std::sqrt(2 * M_PI) * opt.ignition_duration) *
               std::exp(-std::pow(x - opt.ignition_center, 2) / (2 * std::pow(opt.ignition_stddev, 2)));
    }
    return 0.0;
}

// SourceSystemQSS (src/sourceSystem.cpp:398-531)
struct SourceQSS : QssOde {
    Gas& gas;
    int K, N;
    QssIntegrator integrator;
    std::vector<double> splitConst, Yv, hk, wQ, wD;
    double U = 0, T = 0, rho = 0, cp = 0, qLoss = 0, qDot = 0;
    double x = 0;  // SourceSystem::setPosition
    bool failed_rates = false;
    const orc_options& opt;

    SourceQSS(Gas& g, const orc_options& o)
        : gas(g), K(g.m.K), N(g.m.K + 2), splitConst(N), Yv(K), hk(K), wQ(K), wD(K), opt(o) {
        integrator.ode = this;
        integrator.initialize(N);
        integrator.enforce[0] = 0;                 // initialize: enforce_ymin[kMomentum] = false
        integrator.epsmin = o.qss_epsmin;          // setOptions
        integrator.epsmax = o.qss_epsmax;
        integrator.dtmin = o.qss_dtmin;
        integrator.dtmax = o.qss_dtmax;
        integrator.itermax = o.qss_itermax;
        integrator.abstol = o.qss_abstol;
        integrator.stabilityCheck = o.qss_stability_check != 0;
        std::fill(integrator.ymin.begin(), integrator.ymin.end(), o.qss_minval);
        integrator.ymin[0] = -1e4;
        integrator.qdot_src = &qDot;
    }
    bool failed() const override { return failed_rates; }

    void unroll_y(const double* yy, bool corrector) {
        if (!corrector) T = yy[1];
        U = yy[0];
        for (int k = 0; k < K; k++) Yv[k] = yy[2 + k];
    }

    void odefun(double t, const double* yy, double* qq, double* dd, bool corrector) override {
        unroll_y(yy, corrector);
        gas.rate_function_at(t);
        gas.setMassFractions(Yv.data());
        gas.setState_TP(T);
        if (!gas.creationRates(wQ.data())) failed_rates = true;
        gas.destructionRates(wD.data());
        if (!corrector) {
            gas.enthalpies(hk.data());
            rho = gas.dens;
            cp = gas.cp_mass();
        }
        double s = 0.0;
        for (int k = 0; k < K; k++) s += (wQ[k] - wD[k]) * hk[k];
        qDot = -s + qdot_igniter(opt, x, t);
        double dUdtQ = splitConst[0], dUdtD = 0.0;
        double dTdtQ = qDot / (rho * cp) + splitConst[1];
        double dTdtD = qLoss / (rho * cp);
        qq[0] = dUdtQ; dd[0] = dUdtD;
        qq[1] = dTdtQ; dd[1] = dTdtD;
        for (int k = 0; k < K; k++) {
            qq[2 + k] = wQ[k] * gas.m.W[k] / rho + splitConst[2 + k];
            dd[2 + k] = wD[k] * gas.m.W[k] / rho;
        }
    }
    double tStart = 0.0;
    void setState(const double* yIn) { integrator.setState(yIn, tStart); }
    int integrateToTime(double tf) { return integrator.integrateToTime(tf); }
};

// CHEMEQ2 cesium/ion relaxation problem (test/old/test_qssintegrator.cpp:11-83)
struct CesiumOde : QssOde {
    void odefun(double, const double* y, double* q, double* d, bool) override {
        double o2m = y[0], csp = y[1], cs = y[2], cso2 = y[3], o2 = y[4], n2 = y[5];
        double ne = std::max(csp - o2m, 0.0);
        double cr1 = 5.00e-08 * o2m * csp;
        double cr2 = 1.00e-12 * csp * ne;
        double cr3 = 3.24e-03 * cs;
        double cr4 = 4.00e-01 * o2m;
        double cr5 = 1.00e-31 * o2 * cs * (cs + cso2 + n2 + o2);
        double cr6 = 1.24e-30 * o2 * o2 * ne;
        double cr7 = 1.00e-31 * o2 * n2 * ne;
        q[0] = cr6 + cr7; d[0] = cr1 + cr4;
        q[1] = cr3;       d[1] = cr1 + cr2;
        q[2] = cr1 + cr2; d[2] = cr3 + cr5;
        q[3] = cr5;       d[3] = 0.0;
        q[4] = cr1 + cr4; d[4] = cr5 + cr6 + cr7;
        q[5] = 0.0;       d[5] = 0.0;
    }
};

// y_i' = q_i - k_i y_i with constant q, k
struct LinearOde : QssOde {
    std::vector<double> qc, kc;
    void odefun(double, const double* y, double* q, double* d, bool) override {
        for (size_t i = 0; i < qc.size(); i++) { q[i] = qc[i]; d[i] = kc[i] * y[i]; }
    }
};

// ---------------------------------------------------------------------------
// SourceSystemCVODE semantics with a variable-order BDF
// ---------------------------------------------------------------------------
struct SourceBDF {
    Gas& gas;
    const orc_options& o;
    int K, N;
    std::vector<double> splitConst, Yv, hk, wdot, atol;
    double qLoss = 0, qDot = 0;
    double x = 0;  // SourceSystem::setPosition
    int nfe = 0, nsteps = 0, njac = 0, nlu = 0, nsolve = 0;
    bool failed_rates = false;

    SourceBDF(Gas& g, const orc_options& opt)
        : gas(g), o(opt), K(g.m.K), N(g.m.K + 2), splitConst(N), Yv(K), hk(K), wdot(K), atol(N) {
        atol[0] = o.atol_momentum;
        atol[1] = o.atol_energy;
        for (int k = 0; k < K; k++) atol[2 + k] = o.atol_species;
    }

    double tStart = 0.0;  // absolute time of the integration's t = 0 (CVODE's t0 = tInitial)

    // SourceSystemCVODE::f (src/sourceSystem.cpp:128-164) at t (from the stage start)
    void f(double t, const double* y, double* ydot) {
        nfe++;
        gas.rate_function_at(tStart + t);
        double T = y[1];
        for (int k = 0; k < K; k++) Yv[k] = y[2 + k];
        gas.setMassFractions_NoNorm(Yv.data());
        gas.setState_TP(T);
        if (!gas.netRates(wdot.data())) failed_rates = true;
        gas.enthalpies(hk.data());
        double rho = gas.dens;
        double cp = gas.cp_mass();
        double s = 0.0;
        for (int k = 0; k < K; k++) s += wdot[k] * hk[k];
        qDot = -s + qdot_igniter(o, x, tStart + t);
        qDot -= qLoss;
        ydot[0] = splitConst[0];
        ydot[1] = qDot / (rho * cp) + splitConst[1];
        for (int k = 0; k < K; k++) ydot[2 + k] = wdot[k] * gas.m.W[k] / rho + splitConst[2 + k];
    }

    // SourceSystemCVODE::fdJacobian (src/sourceSystem.cpp:268-293); J column-major
    void fdJacobian(double t, const double* y, const double* ydot, double* J) {
        njac++;
        std::vector<double> yp(y, y + N), yd2(N);
        const double eps = std::sqrt(std::numeric_limits<double>::epsilon());
        const double atl = std::numeric_limits<double>::epsilon();
        for (int i = 0; i < N; i++) {
            for (int k = 0; k < N; k++) yp[k] = y[k];
            double dy = (std::abs(y[i]) > atl) ? std::abs(y[i]) * eps : std::abs(y[i]) * eps + atl;
            yp[i] += dy;
            f(t, yp.data(), yd2.data());
            for (int k = 0; k < N; k++) J[(size_t)i * N + k] = (yd2[k] - ydot[k]) / dy;
        }
    }

    double rms(const double* v, const double* scale) const {
        double s = 0.0;
        for (int i = 0; i < N; i++) { double r = v[i] / scale[i]; s += r * r; }
        return std::sqrt(s / N);
    }

    // LU with partial pivoting of M = I - c J (row-major copy), returns false if singular
    static bool lu_factor(int n, std::vector<double>& A, std::vector<int>& piv) {
        for (int k = 0; k < n; k++) {
            int p = k;
            double best = std::abs(A[(size_t)k * n + k]);
            for (int i = k + 1; i < n; i++) {
                double v = std::abs(A[(size_t)i * n + k]);
                if (v > best) { best = v; p = i; }
            }
            piv[k] = p;
            if (best == 0.0) return false;
            if (p != k)
                for (int j = 0; j < n; j++) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
            double inv = 1.0 / A[(size_t)k * n + k];
            for (int i = k + 1; i < n; i++) {
                double l = A[(size_t)i * n + k] * inv;
                A[(size_t)i * n + k] = l;
                if (l != 0.0)
                    for (int j = k + 1; j < n; j++) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
            }
        }
        return true;
    }
    static void lu_solve(int n, const std::vector<double>& A, const std::vector<int>& piv, double* b) {
        for (int k = 0; k < n; k++) {
            if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        }
        for (int i = 0; i < n; i++) {
            double s = b[i];
            for (int j = 0; j < i; j++) s -= A[(size_t)i * n + j] * b[j];
            b[i] = s;
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = b[i];
            for (int j = i + 1; j < n; j++) s -= A[(size_t)i * n + j] * b[j];
            b[i] = s / A[(size_t)i * n + i];
        }
    }

    // Shampine-Reichelt difference-array step change: D[0..q] <- (R U)^T D[0..q]
    void change_D(std::vector<double>& D, int order, double factor) {
        int m1 = order + 1;
        auto R = [&](double fac, std::vector<double>& out) {
            std::vector<double> M((size_t)m1 * m1, 0.0);
            for (int j = 0; j < m1; j++) M[j] = 1.0;
            for (int i = 1; i < m1; i++)
                for (int j = 1; j < m1; j++) M[(size_t)i * m1 + j] = (i - 1 - fac * j) / (double)i;
            out.assign((size_t)m1 * m1, 0.0);
            for (int j = 0; j < m1; j++) {
                double c = 1.0;
                for (int i = 0; i < m1; i++) { c *= M[(size_t)i * m1 + j]; out[(size_t)i * m1 + j] = c; }
            }
        };
        std::vector<double> Rm, Um, RU((size_t)m1 * m1, 0.0);
        R(factor, Rm);
        R(1.0, Um);
        for (int i = 0; i < m1; i++)
            for (int j = 0; j < m1; j++) {
                double s = 0.0;
                for (int l = 0; l < m1; l++) s += Rm[(size_t)i * m1 + l] * Um[(size_t)l * m1 + j];
                RU[(size_t)i * m1 + j] = s;
            }
        std::vector<double> Dn((size_t)m1 * N, 0.0);
        for (int j = 0; j < m1; j++)
            for (int v = 0; v < N; v++) {
                double s = 0.0;
                for (int i = 0; i < m1; i++) s += RU[(size_t)i * m1 + j] * D[(size_t)i * N + v];
                Dn[(size_t)j * N + v] = s;
            }
        std::copy(Dn.begin(), Dn.end(), D.begin());
    }

    // integrate y from 0 to tf; returns 0 or <0
    int integrate(double* y, double tf, double* qint_out) {
        const int MAXO = 5, NEWTON_MAXITER = 4;
        const double MIN_FACTOR = 0.2, MAX_FACTOR = 10.0;
        const double rtol = o.rtol;
        double gamma[MAXO + 2], alpha_[MAXO + 2], errc[MAXO + 2];
        gamma[0] = 0.0;
        for (int k = 1; k <= MAXO + 1; k++) gamma[k] = gamma[k - 1] + 1.0 / k;
        for (int k = 0; k <= MAXO + 1; k++) { alpha_[k] = gamma[k]; errc[k] = 1.0 / (k + 1); }
        const double EPS = std::numeric_limits<double>::epsilon();
        const double newton_tol = std::max(10 * EPS / rtol, std::min(0.03, std::sqrt(rtol)));

        std::vector<double> f0(N), scale(N), f1(N), y1(N);
        double t = 0.0;
        f(t, y, f0.data());
        double qprev = qDot;
        double qint = 0.0;
        // initial step (Hairer/Shampine select_initial_step, order 1)
        for (int i = 0; i < N; i++) scale[i] = atol[i] + std::abs(y[i]) * rtol;
        double d0 = rms(y, scale.data()), d1 = rms(f0.data(), scale.data());
        double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
        h0 = std::min(h0, tf);
        for (int i = 0; i < N; i++) y1[i] = y[i] + h0 * f0[i];
        f(t + h0, y1.data(), f1.data());  // select_initial_step: f(t0 + h0, y0 + h0 f0)
        for (int i = 0; i < N; i++) y1[i] = f1[i] - f0[i];
        double d2 = rms(y1.data(), scale.data()) / h0;
        double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? std::max(1e-6, h0 * 1e-3)
                                                   : std::pow(0.01 / std::max(d1, d2), 1.0 / 2.0);
        double h_abs = std::min(100 * h0, h1);
        h_abs = std::min(h_abs, tf);

        int order = 1, n_equal_steps = 0;
        std::vector<double> D((size_t)(MAXO + 3) * N, 0.0);
        for (int i = 0; i < N; i++) { D[i] = y[i]; D[N + i] = f0[i] * h_abs; }
        std::vector<double> J((size_t)N * N), LU((size_t)N * N);
        std::vector<int> piv(N);
        bool lu_valid = false;
        fdJacobian(t, y, f0.data(), J.data());
        std::vector<double> ypred(N), psi(N), ynew(N), dvec(N), dy(N), fy(N), err(N);

        while (t < tf) {
            double min_step = 10 * std::abs(std::nextafter(t, std::numeric_limits<double>::infinity()) - t);
            min_step = std::max(min_step, o.min_step);
            if (h_abs < min_step) { h_abs = min_step; }
            bool accepted = false;
            bool current_jac = false;
            double error_norm = 0, safety = 0;
            int n_iter = 0;
            double t_new = 0;
            // failed attempts of this step, bounded as CVODE bounds them (maxncf = 10
            // convergence failures, maxnef = 7 error-test failures per step)
            int n_conv_fail = 0, n_err_fail = 0;
            while (!accepted) {
                if (h_abs < min_step) return -3;
                if (nsteps >= o.max_steps) return -4;
                double h = h_abs;
                t_new = t + h;
                if (t_new - tf > 0) {
                    t_new = tf;
                    change_D(D, order, std::abs(t_new - t) / h_abs);
                    n_equal_steps = 0;
                    lu_valid = false;
                }
                h = t_new - t;
                h_abs = std::abs(h);
                for (int v = 0; v < N; v++) {
                    double s = 0.0;
                    for (int i = 0; i <= order; i++) s += D[(size_t)i * N + v];
                    ypred[v] = s;
                    scale[v] = atol[v] + rtol * std::abs(ypred[v]);
                }
                for (int v = 0; v < N; v++) {
                    double s = 0.0;
                    for (int i = 1; i <= order; i++) s += D[(size_t)i * N + v] * gamma[i];
                    psi[v] = s / alpha_[order];
                }
                double c = h / alpha_[order];
                bool converged = false;
                while (!converged) {
                    if (!lu_valid) {
                        for (int i = 0; i < N; i++)
                            for (int j = 0; j < N; j++)
                                LU[(size_t)i * N + j] = (i == j ? 1.0 : 0.0) - c * J[(size_t)j * N + i];
                        nlu++;
                        if (!lu_factor(N, LU, piv)) return -5;
                        lu_valid = true;
                    }
                    // solve_bdf_system
                    std::fill(dvec.begin(), dvec.end(), 0.0);
                    ynew = ypred;
                    double dy_norm_old = -1.0;
                    converged = false;
                    int k;
                    for (k = 0; k < NEWTON_MAXITER; k++) {
                        f(t_new, ynew.data(), fy.data());
                        bool finite = true;
                        for (int v = 0; v < N; v++) finite = finite && std::isfinite(fy[v]);
                        if (!finite) break;
                        for (int v = 0; v < N; v++) dy[v] = c * fy[v] - psi[v] - dvec[v];
                        nsolve++;
                        lu_solve(N, LU, piv, dy.data());
                        double dy_norm = rms(dy.data(), scale.data());
                        double rate = (dy_norm_old < 0) ? -1.0 : dy_norm / dy_norm_old;
                        if (rate >= 0 && (rate >= 1 || std::pow(rate, NEWTON_MAXITER - k) / (1 - rate) * dy_norm > newton_tol))
                            break;
                        for (int v = 0; v < N; v++) { ynew[v] += dy[v]; dvec[v] += dy[v]; }
                        if (dy_norm == 0 || (rate >= 0 && rate / (1 - rate) * dy_norm < newton_tol)) {
                            converged = true;
                            break;
                        }
                        dy_norm_old = dy_norm;
                    }
                    n_iter = k + 1;
                    if (!converged) {
                        if (current_jac) break;
                        std::vector<double> fp(N);
                        f(t_new, ypred.data(), fp.data());
                        fdJacobian(t_new, ypred.data(), fp.data(), J.data());
                        lu_valid = false;
                        current_jac = true;
                    }
                }
                if (failed_rates) return -2;
                if (!converged) {
                    if (++n_conv_fail >= kBdfMaxConvFail) return -8;
                    double factor = 0.5;
                    h_abs *= factor;
                    change_D(D, order, factor);
                    n_equal_steps = 0;
                    lu_valid = false;
                    continue;
                }
                safety = 0.9 * (2 * NEWTON_MAXITER + 1) / (2 * NEWTON_MAXITER + n_iter);
                for (int v = 0; v < N; v++) scale[v] = atol[v] + rtol * std::abs(ynew[v]);
                for (int v = 0; v < N; v++) err[v] = errc[order] * dvec[v];
                error_norm = rms(err.data(), scale.data());
                // a non-finite error estimate (NaN or Inf in the Newton update) fails the
                // point instead of being taken as a small error
                if (!std::isfinite(error_norm)) return -6;
                if (error_norm > 1) {
                    if (++n_err_fail >= kBdfMaxErrFail) return -7;
                    double factor = std::max(MIN_FACTOR, safety * std::pow(error_norm, -1.0 / (order + 1)));
                    h_abs *= factor;
                    change_D(D, order, factor);
                    n_equal_steps = 0;
                } else {
                    accepted = true;
                }
            }
            nsteps++;
            n_equal_steps++;
            double hstep = t_new - t;
            t = t_new;
            for (int v = 0; v < N; v++) y[v] = ynew[v];
            // heat release on the accepted step (trapezoid); not counted in nfe
            {
                std::vector<double> tmp(N);
                f(t, y, tmp.data());
                nfe--;
                qint += 0.5 * hstep * (qprev + qDot);
                qprev = qDot;
            }
            for (int v = 0; v < N; v++) {
                D[(size_t)(order + 2) * N + v] = dvec[v] - D[(size_t)(order + 1) * N + v];
                D[(size_t)(order + 1) * N + v] = dvec[v];
            }
            for (int i = order; i >= 0; i--)
                for (int v = 0; v < N; v++) D[(size_t)i * N + v] += D[(size_t)(i + 1) * N + v];
            if (n_equal_steps < order + 1) continue;
            double em = std::numeric_limits<double>::infinity(), ep = std::numeric_limits<double>::infinity();
            if (order > 1) {
                for (int v = 0; v < N; v++) err[v] = errc[order - 1] * D[(size_t)order * N + v];
                em = rms(err.data(), scale.data());
            }
            if (order < MAXO) {
                for (int v = 0; v < N; v++) err[v] = errc[order + 1] * D[(size_t)(order + 2) * N + v];
                ep = rms(err.data(), scale.data());
            }
            double norms[3] = {em, error_norm, ep};
            double factors[3];
            int best = 0;
            for (int j = 0; j < 3; j++) {
                factors[j] = std::pow(norms[j], -1.0 / (order + j));
                if (std::isnan(factors[j])) factors[j] = 0;
            }
            for (int j = 1; j < 3; j++) if (factors[j] > factors[best]) best = j;
            order = std::max(1, order + best - 1);
            double factor = std::min(MAX_FACTOR, safety * factors[best]);
            h_abs *= factor;
            change_D(D, order, factor);
            n_equal_steps = 0;
            lu_valid = false;
        }
        if (qint_out) *qint_out = qint;
        return 0;
    }
};

// ---------------------------------------------------------------------------
// TridiagonalIntegrator (src/integrator.cpp)
// ---------------------------------------------------------------------------
struct Tridiag {
    int N;
    std::vector<double> a, b, c, k, lu_b, lu_c, lu_d, invDenom, y, yprev;
    double t = 0, h = 0;
    unsigned stepCount = 0;
    explicit Tridiag(int n)
        : N(n), a(n, 0.0), b(n, 0.0), c(n, 0.0), k(n, 0.0), lu_b(n), lu_c(n), lu_d(n), invDenom(n),
          y(n), yprev(n) {}
    void initialize(double t0, double h_) { t = t0; h = h_; stepCount = 0; }
    void factor(double shift) {
        for (int i = 0; i < N; i++) lu_b[i] = b[i] - shift;
        lu_c[0] = c[0] / lu_b[0];
        for (int i = 1; i < N; i++) {
            invDenom[i] = 1.0 / (lu_b[i] - lu_c[i - 1] * a[i]);
            lu_c[i] = c[i] * invDenom[i];
        }
    }
    void solve() {
        lu_d[0] = y[0] / lu_b[0];
        for (int i = 1; i < N; i++) lu_d[i] = (y[i] - lu_d[i - 1] * a[i]) * invDenom[i];
        y[N - 1] = lu_d[N - 1];
        for (int i = N - 2; i >= 0; i--) y[i] = lu_d[i] - lu_c[i] * y[i + 1];
    }
    void step() {
        if (stepCount == 0) {
            yprev = y;
            factor(1.0 / (h / 8.0));
            for (int j = 0; j < 8; j++) {
                for (int i = 0; i < N; i++) y[i] = -y[i] / (h / 8.0) - k[i];
                solve();
            }
        } else {
            if (stepCount == 1) factor(3.0 / (2.0 * h));
            std::vector<double> ynm1 = y;
            for (int i = 0; i < N; i++) y[i] = -2 * ynm1[i] / h + yprev[i] / (2 * h) - k[i];
            yprev = ynm1;
            solve();
        }
        stepCount++;
        t += h;
    }
    void integrateToTime(double tEnd) {
        int nSteps = static_cast<int>((tEnd - t) / h + 1e-5);
        nSteps = std::max(nSteps, 1);
        h = (tEnd - t) / nSteps;
        for (int i = 0; i < nSteps; i++) step();
    }
};

// DiffusionSystem::get_A / get_k with OneDimGrid metrics
void diffusion_coefficients(int N, const double* B, const double* D, const double* sc,
                            const double* x, int alpha, int leftBC, int rightBC, double yInf,
                            double wallConst, Tridiag& s) {
    std::vector<double> hh(N - 1), rphalf(N - 1), r(N), dlj(N, 0.0), c1(N, 0.0), c2(N, 0.0);
    for (int j = 0; j < N - 1; j++) {
        hh[j] = x[j + 1] - x[j];
        rphalf[j] = std::pow(0.5 * (x[j] + x[j + 1]), alpha);
    }
    for (int j = 0; j < N; j++) r[j] = std::pow(x[j], alpha);
    for (int j = 1; j < N - 1; j++) dlj[j] = 0.5 * (x[j + 1] - x[j - 1]);
    for (int j = 1; j <= N - 2; j++) {
        c1[j] = 0.5 * B[j] / (dlj[j] * r[j]);
        c2[j] = rphalf[j] * (D[j] + D[j + 1]) / hh[j];
    }
    int jStart;
    if (leftBC == 0) {
        jStart = 1;
    } else if (leftBC == 2) {
        jStart = 1;
        double c0 = B[0] * (alpha + 1) * (D[0] + D[1]) / (2 * hh[0] * hh[0]);
        s.b[0] = -c0;
        s.c[0] = c0;
    } else if (leftBC == 3) {
        jStart = 1;
        double c0 = B[0] * (alpha + 1) / hh[0];
        double d = 0.5 * (D[0] + D[1]);
        s.b[0] = -c0 * (d / hh[0] + wallConst);
        s.c[0] = d * c0 / hh[0];
    } else {
        jStart = 2;
        s.b[1] = -c1[1] * c2[1];
        s.c[1] = c1[1] * c2[1];
    }
    int jStop;
    if (rightBC == 0) {
        jStop = N - 1;
    } else {
        jStop = N - 2;
        s.a[N - 2] = c1[N - 2] * c2[N - 3];
        s.b[N - 2] = -c1[N - 2] * c2[N - 3];
    }
    for (int j = jStart; j < jStop; j++) {
        s.a[j] = c1[j] * c2[j - 1];
        s.b[j] = -c1[j] * (c2[j - 1] + c2[j]);
        s.c[j] = c1[j] * c2[j];
    }
    for (int j = 0; j < N; j++) s.k[j] = sc[j];
    if (leftBC == 3) s.k[0] += B[0] * (alpha + 1) / hh[0] * wallConst * yInf;
}

#include "qss_adjoint.inc"

}  // namespace

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

void orc_default_options(orc_options* o) {
    o->integrator = 0;
    o->kinetics = 0;
    o->pressure = 101325.0;
    o->qss_epsmin = 2e-2;
    o->qss_epsmax = 10.0;
    o->qss_dtmin = 1e-16;
    o->qss_dtmax = 1e-6;
    o->qss_itermax = 1;
    o->qss_abstol = 1e-11;
    o->qss_minval = 1e-60;
    o->qss_stability_check = 0;
    o->rtol = 1e-6;
    o->atol_momentum = 1e-7;
    o->atol_energy = 1e-8;
    o->atol_species = 1e-13;
    o->min_step = 1e-18;
    o->max_steps = 1000000;
    o->t_start = 0.0;
    o->ignition_t_start = 0.0;  // python/ember/input.py class Ignition defaults
    o->ignition_duration = 1e-3;
    o->ignition_energy = 0.0;
    o->ignition_center = 0.0;
    o->ignition_stddev = 1e-4;
    o->positions = nullptr;
}

void* orc_mech_load(const char* path) { return load_mech(path); }
void orc_mech_free(void* mech) { delete static_cast<Mech*>(mech); }
int orc_mech_nspecies(void* mech) { return static_cast<Mech*>(mech)->K; }
int orc_mech_nreactions(void* mech) { return static_cast<Mech*>(mech)->I; }
void orc_mech_set_multipliers(void* mech, const double* m) {
    Mech* me = static_cast<Mech*>(mech);
    for (int i = 0; i < me->I; i++) me->perturb[i] = m ? m[i] : 1.0;
}

int orc_mech_set_interp_range(void* mech, int n, double Tmin, double Tmax) {
    Mech& m = *static_cast<Mech*>(mech);
    if (n < 2 || !(Tmin > 0.0) || !(Tmax > Tmin)) return -1;
    m.nInterp = n;
    m.TminInt = Tmin;
    m.TmaxInt = Tmax;
    m.interp_ready = false;
    return 0;
}

int orc_mech_set_rate_function(void* mech, int n, const double* data) {
    Mech* me = static_cast<Mech*>(mech);
    me->ratefn.clear();
    if (n == 0 || !data) return 0;
    if (n < 2) return -1;
    // ChebyshevFunction::setCoefficients: at least two coefficients, zero-filled
    me->ratefn = {data[0], data[1]};
    if (n >= 4)
        me->ratefn.insert(me->ratefn.end(), data + 2, data + n);
    else if (n == 3)
        me->ratefn.insert(me->ratefn.end(), {data[2], 0.0});
    else
        me->ratefn.insert(me->ratefn.end(), {0.0, 0.0});
    return 0;
}

double orc_rate_function_eval(void* mech, double t) {
    const Mech* me = static_cast<Mech*>(mech);
    return me->ratefn.empty() ? 1.0 : rate_fn_eval(*me, t);
}

int orc_rates(void* mech, const orc_options* o, int n, const double* T, const double* Y, int nonorm,
              double* wdotQ, double* wdotD, double* wdot_net, double* hk, double* rho, double* cp,
              double* ropf, double* ropr) {
    Mech& m = *static_cast<Mech*>(mech);
    Gas g(m, o->kinetics, o->pressure);
    if (o->kinetics == 1) {
        std::vector<double> Y0(m.K, 1.0 / m.K);
        g.setMassFractions(Y0.data());
        g.setState_TP(300.0);
        g.build_interp();
    }
    std::vector<double> tmp(m.K);
    for (int j = 0; j < n; j++) {
        const double* Yj = Y + (size_t)j * m.K;
        if (nonorm) g.setMassFractions_NoNorm(Yj);
        else g.setMassFractions(Yj);
        g.setState_TP(T[j]);
        if (!g.creationRates(wdotQ ? wdotQ + (size_t)j * m.K : tmp.data())) return -2;
        if (wdotD) g.destructionRates(wdotD + (size_t)j * m.K);
        if (wdot_net) g.netRates(wdot_net + (size_t)j * m.K);
        if (hk) g.enthalpies(hk + (size_t)j * m.K);
        if (rho) rho[j] = g.dens;
        if (cp) cp[j] = g.cp_mass();
        if (ropf) std::copy(g.ropf.begin(), g.ropf.end(), ropf + (size_t)j * m.I);
        if (ropr) std::copy(g.ropr.begin(), g.ropr.end(), ropr + (size_t)j * m.I);
    }
    return 0;
}

// FlameSolver::updateChemicalProperties, thermodynamic block (src/flameSolver.cpp:667-673):
// setStateMass, then getDensity, getMixtureMolecularWeight, getSpecificHeatCapacity,
// getSpecificHeatCapacities (R cp_k/R) and getEnthalpies (RT h_k/RT) per point; the
// tbb::parallel_for over points (grain 1, :653-654) as OpenMP dynamic,1.
int orc_thermo_properties(void* mech, double pressure, int nPoints, const double* state,
                          double* rho, double* Wmx, double* cp, double* cpSpec, double* hk,
                          int nthreads) {
    Mech& m = *static_cast<Mech*>(mech);
    const int nv = m.K + 2;
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        Gas g(m, 0, pressure);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < nPoints; j++) {
            const double* s = state + (size_t)j * nv;
            g.setMassFractions(s + 2);
            g.setState_TP(s[1]);
            rho[j] = g.dens;
            if (Wmx) Wmx[j] = g.mmw;
            cp[j] = g.cp_mass();
            if (cpSpec)
                for (int k = 0; k < m.K; k++) cpSpec[(size_t)j * m.K + k] = kGasConstant * g.cpR[k];
            if (hk) g.enthalpies(hk + (size_t)j * m.K);
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Density derivative of the convection stage, FlameSolver::prepareIntegrators
// (src/flameSolver.cpp:203-213): ddt = ddtConv + ddtDiff + ddtProd (+ ddtCross with
// balanced splitting); tmp = W.inverse()^T ddt(species rows), summed here in species order;
// drhodt = -rho * (ddt(kEnergy) / T + tmp * Wmx).
int orc_density_derivative(void* mech, int nPoints, const double* state, const double* ddt_conv,
                           const double* ddt_diff, const double* ddt_prod, const double* ddt_cross,
                           int balanced, const double* rho, const double* Wmx, double* drhodt) {
    Mech& m = *static_cast<Mech*>(mech);
    const int nv = m.K + 2;
    for (int j = 0; j < nPoints; j++) {
        const size_t o = (size_t)j * nv;
        auto ddt = [&](int v) {
            double d = ddt_conv[o + v] + ddt_diff[o + v];
            d = d + ddt_prod[o + v];
            if (balanced && ddt_cross) d = d + ddt_cross[o + v];
            return d;
        };
        double tmp = 0.0;
        for (int k = 0; k < m.K; k++) tmp = tmp + (1.0 / m.W[k]) * ddt(2 + k);
        drhodt[j] = -rho[j] * (ddt(1) / state[o + 1] + tmp * Wmx[j]);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Transport properties: the transport block of FlameSolver::updateChemicalProperties
// (src/flameSolver.cpp:678-692) with Ember's default model, ApproxMixTransport
// (src/chemistry0d.cpp:12-261) over Cantera's MixTransport: mu (ApproxMixTransport::
// viscosity, Wilke's rule over the species with X_k >= threshold), lambda
// (MixTransport::thermalConductivity: 0.5 (sum X_k lambda_k + 1 / sum X_k / lambda_k)) and
// rho D_k (CanteraGas::getWeightedDiffusionCoefficientsMass: rho times
// ApproxMixTransport::getMixDiffCoeffsMass).  Pure-species and binary properties come
// from the table's temperature fits as MixTransport evaluates them (update_T: sqrt(T),
// T^(1/4), [1, ln T, ..., ln^4 T]; dot5).  Mole fractions carry MixTransport::update_C's
// floor of 1e-20 (Cantera's Tiny).
// ---------------------------------------------------------------------------
namespace {
double dot5(const double* x, const double* c) {
    return x[0] * c[0] + x[1] * c[1] + x[2] * c[2] + x[3] * c[3] + x[4] * c[4];
}
}  // namespace

int orc_transport_properties(void* mech, double pressure, double threshold, int nPoints,
                             const double* state, double* mu, double* lambda, double* rhoD,
                             int nthreads) {
    Mech& m = *static_cast<Mech*>(mech);
    if (m.visc.empty()) return -1;
    const int K = m.K, nv = K + 2;
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        Gas g(m, 0, pressure);
        std::vector<double> X(K), visc(K), sqvisc(K), cond(K), phi((size_t)K * K), bdiff((size_t)K * K);
        std::vector<int> major;
#pragma omp for schedule(dynamic, 1)
        for (int p = 0; p < nPoints; p++) {
            const double* s = state + (size_t)p * nv;
            g.setMassFractions(s + 2);
            g.setState_TP(s[1]);
            const double T = g.T;
            // MixTransport::update_T
            const double logt = std::log(T), sqrt_t = std::sqrt(T), t14 = std::sqrt(sqrt_t);
            const double ptv[5] = {1.0, logt, logt * logt, logt * logt * logt, logt * logt * logt * logt};
            // MixTransport::update_C and ApproxMixTransport::update_C
            major.clear();
            for (int k = 0; k < K; k++) {
                X[k] = std::max(1e-20, g.ym[k] * g.mmw);
                if (X[k] >= threshold) major.push_back(k);
            }
            // GasTransport::updateSpeciesViscosities, MixTransport::updateCond_T
            for (int k = 0; k < K; k++) {
                sqvisc[k] = t14 * dot5(ptv, &m.visc[(size_t)k * 5]);
                visc[k] = sqvisc[k] * sqvisc[k];
                cond[k] = sqrt_t * dot5(ptv, &m.cond[(size_t)k * 5]);
            }
            // ApproxMixTransport::updateViscosity_T (Wilke, Reid-Prausnitz-Poling 9-5.15)
            for (size_t ij = 0; ij < major.size(); ij++) {
                const int j = major[ij];
                for (size_t ik = ij; ik < major.size(); ik++) {
                    const int k = major[ik];
                    const double vratiokj = visc[k] / visc[j];
                    const double wratiojk = m.W[j] / m.W[k];
                    const double wratjk_kj = std::sqrt(std::sqrt(m.W[j] / m.W[k]));
                    const double wratkj1_jk = std::sqrt(1.0 + m.W[k] / m.W[j]);
                    const double factor1 = 1.0 + (sqvisc[k] / sqvisc[j]) * wratjk_kj;
                    phi[(size_t)k * K + j] = factor1 * factor1 / (std::sqrt(8.0) * wratkj1_jk);
                    phi[(size_t)j * K + k] = phi[(size_t)k * K + j] / (vratiokj * wratiojk);
                }
            }
            double vismix = 0.0;
            for (int k : major) {
                double sum = 0.0;
                for (int j : major) sum += X[j] * phi[(size_t)k * K + j];
                vismix += X[k] * visc[k] / sum;
            }
            mu[p] = vismix;
            // MixTransport::thermalConductivity
            double sum1 = 0.0, sum2 = 0.0;
            for (int k = 0; k < K; k++) {
                sum1 += X[k] * cond[k];
                sum2 += X[k] / cond[k];
            }
            lambda[p] = 0.5 * (sum1 + 1.0 / sum2);
            // ApproxMixTransport::updateDiff_T (rows and columns of the major species)
            for (int j : major)
                for (int k = 0; k < K; k++) {
                    const double b = T * sqrt_t * dot5(ptv, &m.diff[((size_t)j * K + k) * 5]);
                    bdiff[(size_t)j * K + k] = b;
                    bdiff[(size_t)k * K + j] = b;
                }
            // ApproxMixTransport::getMixDiffCoeffsMass, times rho
            const double mmw = g.mmw, P = pressure;
            for (int k = 0; k < K; k++) {
                double s1 = 0.0, s2 = 0.0;
                if (X[k] >= threshold) {
                    for (int i = 0; i < K; i++) {
                        if (i == k) continue;
                        s1 += X[i] / bdiff[(size_t)k * K + i];
                        s2 += X[i] * m.W[i] / bdiff[(size_t)k * K + i];
                    }
                } else {
                    for (int i : major) {
                        if (i == k) continue;
                        s1 += X[i] / bdiff[(size_t)k * K + i];
                        s2 += X[i] * m.W[i] / bdiff[(size_t)k * K + i];
                    }
                }
                s1 *= P;
                s2 *= P * X[k] / (mmw - m.W[k] * X[k]);
                rhoD[(size_t)p * K + k] = g.dens * (1.0 / (s1 + s2));
            }
        }
    }
    return 0;
}

int orc_source_integrate(void* mech, const orc_options* o, int nPoints, const double* state_in,
                         const double* split_const, const double* qloss, double dt,
                         double* state_out, int* status, int* stat1, int* stat2, double* qint,
                         int nthreads) {
    Mech& m = *static_cast<Mech*>(mech);
    const int N = m.K + 2;
    if (o->kinetics == 1) {
        Gas g(m, 1, o->pressure);
        std::vector<double> Y0(m.K, 1.0 / m.K);
        g.setMassFractions(Y0.data());
        g.setState_TP(300.0);
        g.build_interp();
    }
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        Gas gas(m, o->kinetics, o->pressure);  // per-thread gas (gases.local())
        SourceQSS qss(gas, *o);
        SourceBDF bdf(gas, *o);
        qss.tStart = o->t_start;
        bdf.tStart = o->t_start;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int j = 0; j < nPoints; j++) {
            const double* yin = state_in + (size_t)j * N;
            double* yout = state_out + (size_t)j * N;
            int err;
            double qi = 0.0;
            // Cantera's state setters refuse a non-finite or non-positive temperature and
            // non-finite mass fractions; the point fails before integrating
            bool valid = yin[1] > 0.0;
            for (int i = 0; i < N; i++) valid = valid && std::isfinite(yin[i]);
            if (!valid) {
                err = -9;
                if (stat1) stat1[j] = 0;
                if (stat2) stat2[j] = 0;
            } else if (o->integrator == 0) {
                for (int i = 0; i < N; i++) qss.splitConst[i] = split_const ? split_const[(size_t)j * N + i] : 0.0;
                qss.qLoss = qloss ? qloss[j] : 0.0;
                qss.x = o->positions ? o->positions[j] : 0.0;
                qss.failed_rates = false;
                qss.setState(yin);
                err = qss.integrateToTime(dt);
                if (err >= 0) std::copy(qss.integrator.y.begin(), qss.integrator.y.end(), yout);
                if (stat1) stat1[j] = qss.integrator.gcount;
                if (stat2) stat2[j] = qss.integrator.rcount;
                qi = qss.integrator.qint;
            } else {
                for (int i = 0; i < N; i++) bdf.splitConst[i] = split_const ? split_const[(size_t)j * N + i] : 0.0;
                bdf.qLoss = qloss ? qloss[j] : 0.0;
                bdf.x = o->positions ? o->positions[j] : 0.0;
                bdf.failed_rates = false;
                bdf.nfe = bdf.nsteps = bdf.njac = 0;
                std::vector<double> y(yin, yin + N);
                err = bdf.integrate(y.data(), dt, &qi);
                if (err >= 0) std::copy(y.begin(), y.end(), yout);
                if (stat1) stat1[j] = bdf.nsteps;
                if (stat2) stat2[j] = bdf.nfe;
            }
            if (err < 0) std::copy(yin, yin + N, yout);  // keep the old state
            if (status) status[j] = err;
            if (qint) qint[j] = qi;
        }
    }
    return 0;
}

int orc_bdf_work(void* mech, const orc_options* o, int nPoints, const double* state_in, double dt,
                 int* work, int nthreads) {
    Mech& m = *static_cast<Mech*>(mech);
    const int N = m.K + 2;
    if (o->kinetics == 1) return -1;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        Gas gas(m, o->kinetics, o->pressure);
        SourceBDF bdf(gas, *o);
        bdf.tStart = o->t_start;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int j = 0; j < nPoints; j++) {
            std::fill(bdf.splitConst.begin(), bdf.splitConst.end(), 0.0);
            bdf.qLoss = 0.0;
            bdf.x = o->positions ? o->positions[j] : 0.0;
            bdf.failed_rates = false;
            bdf.nfe = bdf.nsteps = bdf.njac = bdf.nlu = bdf.nsolve = 0;
            std::vector<double> y(state_in + (size_t)j * N, state_in + (size_t)(j + 1) * N);
            double qi = 0.0;
            int err = bdf.integrate(y.data(), dt, &qi);
            int* w = work + (size_t)j * 4;
            w[0] = err < 0 ? -1 : bdf.nfe;
            w[1] = bdf.njac;
            w[2] = bdf.nlu;
            w[3] = bdf.nsolve;
        }
    }
    return 0;
}

int orc_qss_test_problem(int problem, int n, const double* qc, const double* kc, double epsmin,
                         double epsmax, double dtmin, double dtmax, int itermax, double abstol,
                         double ymin, int stability, double tf, double* y, int* gcount, int* rcount) {
    CesiumOde ces;
    LinearOde lin;
    QssIntegrator it;
    if (problem == 0) {
        n = 6;
        it.ode = &ces;
    } else {
        lin.qc.assign(qc, qc + n);
        lin.kc.assign(kc, kc + n);
        it.ode = &lin;
    }
    it.initialize(n);
    std::fill(it.ymin.begin(), it.ymin.end(), ymin);
    it.epsmin = epsmin; it.epsmax = epsmax; it.dtmin = dtmin; it.dtmax = dtmax;
    it.itermax = itermax; it.abstol = abstol; it.stabilityCheck = stability != 0;
    it.setState(y, 0.0);
    int err = it.integrateToTime(tf);
    for (int i = 0; i < n; i++) y[i] = it.y[i];
    if (gcount) *gcount = it.gcount;
    if (rcount) *rcount = it.rcount;
    return err;
}

int orc_source_rhs(void* mech, const orc_options* o, const double* y, const double* split_const,
                   double qloss, double* ydot) {
    Mech& m = *static_cast<Mech*>(mech);
    Gas gas(m, o->kinetics, o->pressure);
    if (o->kinetics == 1) {
        std::vector<double> Y0(m.K, 1.0 / m.K);
        gas.setMassFractions(Y0.data());
        gas.setState_TP(300.0);
        gas.build_interp();
    }
    SourceBDF bdf(gas, *o);
    for (int i = 0; i < m.K + 2; i++) bdf.splitConst[i] = split_const ? split_const[i] : 0.0;
    bdf.qLoss = qloss;
    bdf.tStart = o->t_start;
    bdf.f(0.0, y, ydot);
    return bdf.failed_rates ? -2 : 0;
}

int orc_tridiag_integrate(int N, const double* a, const double* b, const double* c, const double* k,
                          double t0, double h, double t1, int mode, int nsteps, double* y,
                          double* t_out, double* h_out) {
    Tridiag s(N);
    for (int i = 0; i < N; i++) { s.a[i] = a[i]; s.b[i] = b[i]; s.c[i] = c[i]; s.k[i] = k[i]; s.y[i] = y[i]; }
    s.initialize(t0, h);
    if (mode == 0) {
        for (int i = 0; i < nsteps; i++) s.step();
    } else {
        s.integrateToTime(t1);
    }
    for (int i = 0; i < N; i++) y[i] = s.y[i];
    if (t_out) *t_out = s.t;
    if (h_out) *h_out = s.h;
    return 0;
}

int orc_diffusion_integrate(int nVars, int N, const double* B, int nB, const int* B_row,
                            const double* D, const double* split_const, const double* x, int alpha,
                            int leftBC, int rightBC, double yInf, double wallConst, int wallVar,
                            double multiplier, double dt_global, const double* h_per_var, double t0,
                            double t1, double* state, int nthreads) {
    (void)nB;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (int v = 0; v < nVars; v++) {
        const double* Bv = B + (size_t)(B_row ? B_row[v] : v) * N;
        std::vector<double> Dv(N), sc(N);
        for (int j = 0; j < N; j++) {
            Dv[j] = D[(size_t)j * nVars + v];
            sc[j] = split_const ? split_const[(size_t)j * nVars + v] : 0.0;
        }
        Tridiag s(N);
        // FlameSolver: only diffusionTerms[kEnergy] carries yInf / wallConst (src/flameSolver.cpp:79-82)
        const double wc = (v == wallVar) ? wallConst : 0.0, yi = (v == wallVar) ? yInf : 0.0;
        diffusion_coefficients(N, Bv, Dv.data(), sc.data(), x, alpha, leftBC, rightBC, yi, wc, s);
        double h;
        if (h_per_var) {
            h = h_per_var[v];
        } else {
            // FlameSolver::setDiffusionSolverState: min over j = 1..N-2 of dlj^2/(B D)
            double mn = std::numeric_limits<double>::infinity();
            for (int j = 1; j <= N - 2; j++) {
                double dl = 0.5 * (x[j + 1] - x[j - 1]);
                mn = std::min(mn, dl * dl / (Bv[j] * Dv[j]));
            }
            h = std::min(multiplier * mn, dt_global);
        }
        for (int j = 0; j < N; j++) s.y[j] = state[(size_t)j * nVars + v];
        s.initialize(t0, h);
        s.integrateToTime(t1);
        for (int j = 0; j < N; j++) state[(size_t)j * nVars + v] = s.y[j];
    }
    return 0;
}


static bool adjoint_options_ok(const orc_options* o) {
    return o->integrator == 0 && o->qss_itermax == 1 && o->qss_stability_check == 0;
}

static void prepare_interp(Mech& m, const orc_options* o) {
    if (o->kinetics == 1) {
        Gas g(m, 1, o->pressure);
        std::vector<double> Y0(m.K, 1.0 / m.K);
        g.setMassFractions(Y0.data());
        g.setState_TP(300.0);
        g.build_interp();
    }
}

int orc_qss_adjoint(void* mech, const orc_options* o, int nPoints, const double* state_in,
                    const double* split_const, const double* qloss, double dt, const double* Qbar,
                    const double* ybar_final, double* state_out, double* Q, double* grad_y0,
                    double* grad_lnk, int* status, int nthreads) {
    Mech& m = *static_cast<Mech*>(mech);
    if (!adjoint_options_ok(o)) return -6;
    if (!m.ratefn.empty()) return -8;  // the ln-multiplier parameters need per-reaction multipliers
    if (o->ignition_energy != 0.0) return -9;
    for (const Rxn& r : m.rxn)
        if (r.type >= 4 || r.hasOrd) return -10;  // PLOG / Chebyshev / SRI / orders: no adjoint
    const int N = m.K + 2;
    prepare_interp(m, o);
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        Gas gas(m, o->kinetics, o->pressure);
        SourceQSS qss(gas, *o);
        QssAdjoint adj(gas, qss);
        std::vector<QssStepRecord> tape;
        std::vector<double> yb(N), thb(m.I);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int j = 0; j < nPoints; j++) {
            const double* yin = state_in + (size_t)j * N;
            for (int i = 0; i < N; i++) qss.splitConst[i] = split_const ? split_const[(size_t)j * N + i] : 0.0;
            qss.qLoss = qloss ? qloss[j] : 0.0;
            qss.failed_rates = false;
            tape.clear();
            qss.integrator.tape = &tape;
            qss.setState(yin);
            int err = qss.integrateToTime(dt);
            qss.integrator.tape = nullptr;
            double* yout = state_out ? state_out + (size_t)j * N : nullptr;
            if (yout) {
                if (err >= 0) std::copy(qss.integrator.y.begin(), qss.integrator.y.end(), yout);
                else std::copy(yin, yin + N, yout);
            }
            if (Q) Q[j] = qss.integrator.qint;
            for (int i = 0; i < N; i++) yb[i] = ybar_final ? ybar_final[(size_t)j * N + i] : 0.0;
            std::fill(thb.begin(), thb.end(), 0.0);
            if (err >= 0 && !adj.reverse(tape, yin, Qbar ? Qbar[j] : 1.0, yb.data(), thb.data())) err = -2;
            if (err < 0) {
                std::fill(yb.begin(), yb.end(), 0.0);
                std::fill(thb.begin(), thb.end(), 0.0);
            }
            if (grad_y0) std::copy(yb.begin(), yb.end(), grad_y0 + (size_t)j * N);
            if (grad_lnk) std::copy(thb.begin(), thb.end(), grad_lnk + (size_t)j * m.I);
            if (status) status[j] = err;
        }
    }
    return 0;
}

int orc_qss_fd_gradient(void* mech, const orc_options* o, int nPoints, const double* state_in,
                        const double* split_const, const double* qloss, double dt, const double* Qbar,
                        const double* ybar_final, double h, int frozen, double* grad_y0,
                        double* grad_lnk, int nthreads) {
    Mech& m = *static_cast<Mech*>(mech);
    if (!adjoint_options_ok(o)) return -6;
    if (!m.ratefn.empty()) return -8;  // the ln-multiplier parameters need per-reaction multipliers
    if (o->ignition_energy != 0.0) return -9;
    for (const Rxn& r : m.rxn)
        if (r.type >= 4 || r.hasOrd) return -10;  // PLOG / Chebyshev / SRI / orders: no adjoint
    const int N = m.K + 2;
    prepare_interp(m, o);
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        Gas gas(m, o->kinetics, o->pressure);
        SourceQSS qss(gas, *o);
        QssAdjoint adj(gas, qss);
        std::vector<QssStepRecord> tape;
        std::vector<double> y(N), pert(m.perturb);
        const double* yb = nullptr;
        double qb = 1.0;
        auto run = [&](const double* y0) -> double {
            double Qv;
            if (frozen) {
                std::copy(y0, y0 + N, y.begin());
                Qv = adj.replay(tape, y.data());
            } else {
                qss.failed_rates = false;
                qss.setState(y0);
                qss.integrateToTime(dt);
                Qv = qss.integrator.qint;
                std::copy(qss.integrator.y.begin(), qss.integrator.y.end(), y.begin());
            }
            double J = qb * Qv;
            if (yb)
                for (int i = 0; i < N; i++) J += yb[i] * y[i];
            return J;
        };
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int j = 0; j < nPoints; j++) {
            const double* yin = state_in + (size_t)j * N;
            for (int i = 0; i < N; i++) qss.splitConst[i] = split_const ? split_const[(size_t)j * N + i] : 0.0;
            qss.qLoss = qloss ? qloss[j] : 0.0;
            qss.failed_rates = false;
            tape.clear();
            qss.integrator.tape = &tape;
            qss.setState(yin);
            qss.integrateToTime(dt);
            qss.integrator.tape = nullptr;
            qb = Qbar ? Qbar[j] : 1.0;
            yb = ybar_final ? ybar_final + (size_t)j * N : nullptr;
            std::vector<double> yp(yin, yin + N);
            for (int i = 0; i < N; i++) {
                double g = 0.0;
                const double hi = h * std::abs(yin[i]);
                if (i >= 1 && hi > 0.0) {
                    yp[i] = yin[i] + hi;
                    const double Qp = run(yp.data());
                    yp[i] = yin[i] - hi;
                    const double Qm = run(yp.data());
                    yp[i] = yin[i];
                    g = (Qp - Qm) / (2.0 * hi);
                }
                if (grad_y0) grad_y0[(size_t)j * N + i] = g;
            }
            if (grad_lnk) {
                for (int r = 0; r < m.I; r++) {
                    const double base = pert[r];
                    pert[r] = base * std::exp(h);
                    gas.perturb = pert.data();
                    const double Qp = run(yin);
                    pert[r] = base * std::exp(-h);
                    const double Qm = run(yin);
                    pert[r] = base;
                    gas.perturb = nullptr;
                    grad_lnk[(size_t)j * m.I + r] = (Qp - Qm) / (2.0 * h);
                }
            }
        }
    }
    return 0;
}

}  // extern "C"

#include "convection.inc"
