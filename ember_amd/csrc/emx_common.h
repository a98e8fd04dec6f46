// Shared host/device definitions for the ember_amd HIP library.
#pragma once

#include <cstdint>

namespace emx {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kMaxRounds = 8;      // reactions per lane: up to 8 * 64 = 512 reactions
constexpr int kMaxFallRounds = 2;  // falloff reactions occupy the first rounds (<= 128)
constexpr int kMaxVars = 64;       // nSpecies + 2 <= 64: one lane per state variable
constexpr int kMaxChunkRounds = 4; // species-sum chunks: <= 256 chunks
constexpr int kChunkLen = 8;       // entries per species-sum chunk (padded with zero entries)
constexpr int kPadSpecies = 63;    // stoichiometry padding index: C[63] = 1, mu[63] = 0
constexpr int kInterpN = 200;      // InterpKinetics default table size (src/chemistry0d.h:70-72)
constexpr double kInterpTmin = 273.0;  // and its default range; a context can grow it
constexpr double kInterpTmax = 3500.0;

constexpr double kGasConstant = 6.02214076e26 * 1.380649e-23;  // J/kmol/K
constexpr double kSmallNumber = 1e-300;
constexpr double kBigNumber = 1e300;
constexpr double kLn10 = 2.302585092994045684;  // ln 10: Troe's 10^x evaluated as exp(x ln 10)
constexpr double kInvLn10 = 0.43429448190325182765;  // 1 / ln 10: log10(x) = ln(x) / ln 10

// reaction kinds in device order (falloff first, then three-body, then elementary with the
// pressure-dependent PLOG and Chebyshev rates)
enum RxnKind : int { kElementary = 0, kThreeBody = 1, kLindemann = 2, kTroe = 3, kNone = 4,
                     kPlog = 5, kCheb = 6, kSri = 7 };

// info word of a reaction slot
__host__ __device__ inline int info_kind(int32_t w) { return w & 7; }
__host__ __device__ inline int info_rev(int32_t w) { return (w >> 3) & 1; }
__host__ __device__ inline int info_group(int32_t w) { return (w >> 8) & 0xff; }   // 255 = none
__host__ __device__ inline int info_hasT2(int32_t w) { return (w >> 4) & 1; }
// explicit forward orders (Cantera C_AnyN): the concentration product is the ordered one
constexpr int32_t kOrdBit = 32;
__host__ __device__ inline int info_ord(int32_t w) { return (w >> 5) & 1; }

// packed stoichiometry: byte0 = molecule count, bytes 1..3 = species indices
// (unused index bytes hold kPadSpecies)
__host__ __device__ inline int st_n(uint32_t s) { return s & 0xff; }
__host__ __device__ inline int st_k(uint32_t s, int j) { return (s >> (8 * (j + 1))) & 0xff; }

// species-sum entry over the rates of progress held as one array rop = [ropf (nRounds*64
// slots), zero slot, ropr (nRounds*64 slots)]: index added to creation (bits 0..11), index
// added to destruction (12..23), multiplicity (24..31).  A product entry of slot s reads
// (s, S1 + s), a reactant entry (S1 + s, s), S1 = nRounds*64 + 1; padding entries have
// multiplicity 0 and point at the zero slot nRounds*64 twice.
__host__ __device__ inline uint32_t en_pack(uint32_t q, uint32_t d, uint32_t nu) {
    return q | (d << 12) | (nu << 24);
}
__host__ __device__ inline int en_q(uint32_t e) { return e & 0xfff; }
__host__ __device__ inline int en_d(uint32_t e) { return (e >> 12) & 0xfff; }
__host__ __device__ inline int en_nu(uint32_t e) { return e >> 24; }
__host__ __device__ inline int en_slot(uint32_t e) { return en_q(e) < en_d(e) ? en_q(e) : en_d(e); }
__host__ __device__ inline int en_prod(uint32_t e) { return en_q(e) < en_d(e); }

// Device-side mechanism description: pointers into one device allocation.
struct DevMech {
    int K, I, nVars;       // species, reactions, K + 2
    int nRounds;           // reaction rounds of 64 slots
    int nFallRounds;       // rounds holding falloff reactions
    int nGroups;           // distinct third-body efficiency sets
    int nChunks;           // species-sum chunks
    int nChunkRounds;
    int nEntries;          // species-sum entries, nChunkRounds * 64 * kChunkLen
    int nGrpEntries;       // third-body efficiency entries
    int grpMaxLen;         // entries of the largest efficiency group
    int interp;            // 1: InterpKinetics tables
    double P, Pref;
    double lnRPref;  // ln(R / Pref): standard-state term of the 1/Kc factors
    // species [K]
    const double* W;
    const double* rW;
    const double* Tmid;
    const double* nasa;    // [K][14]: low a0..a6, high a0..a6
    const double* exp2t;   // [16] 2^(j/16) for emx_exp
    // reactions in device order [nRounds * 64]
    const double* A;
    const double* b;
    const double* EaR;
    const double* dn;
    const double* A0;
    const double* b0;
    const double* EaR0;
    const double* troeA;
    const double* rT3;
    const double* rT1;
    const double* T2;
    const uint32_t* reac;
    const uint32_t* prod;
    const int32_t* info;
    const int32_t* orig;   // slot -> reaction index in the mechanism file (-1 padding)
    // third-body groups
    const int32_t* grp_off;   // [nGroups + 1]
    const int32_t* grp_sp;
    const double* grp_eff;    // efficiency - default
    const double* grp_def;    // [nGroups]
    // species sums: chunk c holds entries [c * kChunkLen, (c + 1) * kChunkLen)
    const uint32_t* entry;
    const int32_t* sp_chunk0; // [K]
    const int32_t* sp_nchunk; // [K]
    // adjoint tables: group -> slots using it, species -> (group, efficiency - default),
    // chunk -> species
    const int32_t* grp_slot_off;  // [nGroups + 1]
    const int32_t* grp_slot;
    const int32_t* sp_grp_off;    // [K + 1]
    const int32_t* sp_grp;
    const double* sp_grp_eff;
    const int32_t* chunk_sp;      // [nChunkRounds * 64]
    int nGrpSlots;                // entries of grp_slot (>= 1)
    int nSpGrp;                   // entries of sp_grp and sp_grp_eff (>= 1)
    int grpSlotMax, spGrpMax;     // longest group slot list, longest species group list
    int nTbRounds;                // rounds holding slots with a third-body group (>= 1)
    // pressure-dependent rates (PLOG, Chebyshev) at the context pressure: per slot offset
    // into auxd (-1: none); layouts in build_pressure_tables (mech_host.h)
    const int32_t* aux;
    const double* auxd;
    int has_pdep;  // 1 if the mechanism has PLOG or Chebyshev reactions
    // explicit forward orders: ord[j * nRounds*64 + s] = order of reactant position j of
    // slot s (slots flagged kOrdBit; 0 elsewhere)
    const double* ord;
    int has_orders;
    // transport fits (polynomials in ln T, 5 coefficients each): sqrt(mu_k)/T^(1/4) and
    // lambda_k/sqrt(T) per species, D_jk(T, 1 Pa)/T^(3/2) per pair j <= k (row by row);
    // NULL without transport data
    const double* tr_visc;
    const double* tr_cond;
    const double* tr_diff;
    const double* tr_wilke;  // [K * K]: sqrt(W_j) / sqrt(8 (W_j + W_k)) at j * K + k
    // interp tables [interpN][nRounds*64] over [interpTmin, interpTmax]
    int interpN;
    double interpTmin, interpTmax;
    const double* kf_c;
    const double* kf_s;
    const double* kc_c;
    const double* kc_s;
};

constexpr int kMaxRateFnCoeffs = 32;  // Chebyshev coefficients of the rate multiplier function

// Integrator options as seen by the kernels
struct DevOpts {
    double qss_epsmin, qss_epsmax, qss_dtmin, qss_dtmax;
    int qss_itermax;
    double qss_abstol, qss_minval;
    int qss_stability;
    double rtol, atol_momentum, atol_energy, atol_species, min_step;
    int max_steps;
    // Ember's rate multiplier function (a ChebyshevFunction): [T0, T1, c_0 .. c_{n-1}] in
    // device memory, NULL when none is set; the right-hand sides of a stage evaluate it at
    // t_start plus their time from the stage start
    const double* ratefn;
    int ratefn_n;
    double t_start;
    // igniter (SourceSystem::getQdotIgniter): per-point amplitudes in device memory (NULL:
    // off), applied for absolute times in [ign_t0, ign_t1)
    const double* qign;
    double ign_t0, ign_t1;
};

}  // namespace emx
