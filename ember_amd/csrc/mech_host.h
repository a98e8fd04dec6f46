// Host-side mechanism tables: reads a compiled ``.emxm`` table and lays it out for
// the wavefront kernels (device slot order, efficiency groups, species-sum chunks,
// interpolation tables).
#pragma once

#include <string>
#include <vector>

#include "emx_common.h"

namespace emx {

struct HostRxn {
    int type = 0, rev = 1, nr = 0, np = 0;
    int r[3] = {-1, -1, -1}, p[3] = {-1, -1, -1};
    double A = 0, b = 0, EaR = 0, A0 = 0, b0 = 0, EaR0 = 0;
    double troeA = 0, T3 = 0, T1 = 0, T2 = 0;
    int hasT2 = 0;
    double defeff = 1.0;
    std::vector<std::pair<int, double>> eff;
    // type 4 (PLOG): (P [Pa], A, b, Ea/R) per rate, in file order
    std::vector<double> plog;
    // type 5 (Chebyshev): nT x nP coefficients of log10 k, row-major, and the ranges
    int chebT = 0, chebP = 0;
    double chebTmin = 0, chebTmax = 0, chebPmin = 0, chebPmax = 0;
    std::vector<double> cheb;
    // type 6 (SRI falloff): a, b, c, d, e
    double sri[5] = {0, 0, 0, 1, 0};
    // explicit forward orders of reactant positions 0..2 (ORD record)
    int hasOrd = 0;
    double ord[3] = {0, 0, 0};
};

struct HostMech {
    std::string name;
    int K = 0, I = 0;
    double Pref = 101325.0;
    std::vector<std::string> species;
    std::vector<double> W, Tmid, nasa;  // nasa: K x 14
    std::vector<HostRxn> rxn;
    // transport fits (TRANSPORT records), five coefficients of a polynomial in ln T each:
    // visc K x 5 (sqrt(mu_k)/T^(1/4)), cond K x 5 (lambda_k/sqrt(T)), diff K(K+1)/2 x 5
    // (D_jk(T, 1 Pa)/T^(3/2), pairs j <= k row by row); empty when the table has none
    double trTmin = 0, trTmax = 0;
    std::vector<double> visc, cond, diff;
    bool has_transport() const { return !visc.empty(); }
};

// index of the pair (j, k) in the upper triangle j <= k, row by row (MixTransport's ic)
inline int pair_index(int j, int k, int K) {
    if (k < j) { int t = j; j = k; k = t; }
    return K * j - (j - 1) * j / 2 + k - j;
}

// Returns an error message, empty on success.
std::string read_table(const std::string& path, HostMech& out);

// Flat arrays in device order, ready to upload into one allocation.
struct DevLayout {
    int nRounds = 0, nFallRounds = 0, nGroups = 0, nChunks = 0, nChunkRounds = 0;
    std::vector<int> slot_of_rxn;  // reaction -> slot
    std::vector<int> orig;         // slot -> reaction (-1 padding)
    std::vector<double> A, b, EaR, dn, A0, b0, EaR0, troeA, rT3, rT1, T2;
    std::vector<uint32_t> reac, prod;
    std::vector<int32_t> info;
    std::vector<double> ord;  // [3][S]: orders of reactant positions (kOrdBit slots)
    std::vector<int32_t> grp_off, grp_sp;
    std::vector<double> grp_eff, grp_def;
    std::vector<uint32_t> entry;     // nChunkRounds * 64 * kChunkLen
    std::vector<int32_t> sp_chunk0, sp_nchunk;
    std::vector<int32_t> grp_slot_off, grp_slot, sp_grp_off, sp_grp, chunk_sp;
    std::vector<double> sp_grp_eff;
    std::vector<double> rW;
};

std::string build_layout(const HostMech& m, DevLayout& L);

// Pressure-dependent rates at the context pressure P, per slot: aux[s] = offset into
// auxd (-1: none).  PLOG (Cantera PlogRate at fixed P): [wP, n1, n2, (A, ln A, b, Ea/R) x
// (n1 + n2)] for the rates at the bracketing pressures P1 <= P < P2 (the end level
// twice outside the listed range), wP = (ln P - ln P1) / (ln P2 - ln P1) (0 for one
// level).  Chebyshev (ChebyshevRate at fixed P): [nT, -1/Tmin - 1/Tmax,
// 1 / (1/Tmax - 1/Tmin), d_0 .. d_{nT-1}] with d_t = sum_p a_tp T_p(Pr).  SRI falloff:
// [a, b, c, d, e].
void build_pressure_tables(const HostMech& m, const DevLayout& L, double P, std::vector<int32_t>& aux,
                           std::vector<double>& auxd);

// InterpKinetics / MultiArrheniusInterp tables at pressure P, [nT][nRounds*64] over nT
// equally spaced temperatures in [Tmin, Tmax] (src/chemistry0d.cpp:291-331, 415-436)
void build_interp_tables(const HostMech& m, const DevLayout& L, double P, int nT, double Tmin,
                         double Tmax, std::vector<double>& kf_c, std::vector<double>& kf_s,
                         std::vector<double>& kc_c, std::vector<double>& kc_s);

}  // namespace emx
