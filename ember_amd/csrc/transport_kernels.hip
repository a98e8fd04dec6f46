// Transport properties of FlameSolver::updateChemicalProperties (src/flameSolver.cpp:678-692)
// with Ember's default transport model, ApproxMixTransport (src/chemistry0d.cpp:12-261):
// the mixture viscosity, thermal conductivity and mass-weighted mixture diffusion
// coefficients rho D_k of every point of a point-major state.
//
// One wavefront per point at a time, lane k = species k.  The pure-species viscosities and
// conductivities and the binary diffusion coefficients come from the mechanism table's
// temperature fits, evaluated as Cantera's MixTransport evaluates them (update_T: sqrt(T),
// T^(1/4) and [1, ln T, ..., ln^4 T]; the products summed left to right as dot5):
//   sqrt(mu_k) = T^(1/4) dot5(p, visc_k),  lambda_k = sqrt(T) dot5(p, cond_k),
//   D_jk(T, 1 Pa) = T sqrt(T) dot5(p, diff_jk).
// Mole fractions: Phase::setMassFractions (clip, normalise), then MixTransport::update_C's
// floor X_k >= 1e-20; the species with X_k >= threshold are ApproxMixTransport's majors.
//   mu      = sum_{k major} X_k mu_k / sum_{j major} X_j Phi_kj   (Wilke, with Phi_jk for
//             j > k formed from Phi_kj as ApproxMixTransport::updateViscosity_T forms it)
//   lambda  = (sum_k X_k lambda_k + 1 / sum_k X_k / lambda_k) / 2
//   rho D_k = rho / (P sum_i X_i / D_ki + P X_k / (W - W_k X_k) sum_i X_i W_i / D_ki),
//             i != k over all species for a major k, over the majors for a minor one
//             (ApproxMixTransport::getMixDiffCoeffsMass, CanteraGas::getWeighted...Mass).
#include <hip/hip_runtime.h>

#include "chem_device.h"
#include "kernels.h"

namespace emx {

namespace {

__device__ __forceinline__ double dot5(const double (&p)[5], const double* c) {
    return p[0] * c[0] + p[1] * c[1] + p[2] * c[2] + p[3] * c[3] + p[4] * c[4];
}

// slot of the pair (j, k), j <= k, in the row-by-row upper triangle of the binary fits:
// K j - (j - 1) j / 2 + k - j.  For a fixed k and i = 0, 1, ... the slot of the pair (i, k)
// starts at k and grows by K - 1 - i while i < k, by 1 from there on (the diagonal slot
// included).

// Wilke's Phi_kj (ApproxMixTransport::updateViscosity_T: the direct form
// (1 + (sqrt(mu_k) / sqrt(mu_j)) (W_j / W_k)^(1/4))^2 / (sqrt(8) sqrt(1 + W_k / W_j)) for
// j <= k, Phi_jk / (mu_j / mu_k W_k / W_j) for j > k, the same quantity) is evaluated as
// (1 + a_k / a_j)^2 g_jk with a_k = sqrt(mu_k) W_k^(-1/4) per species and point and the
// mechanism constant g_jk = sqrt(W_j) / sqrt(8 (W_j + W_k)) (DevMech::tr_wilke): the same
// quantity again, with two multiplications per pair instead of square roots and divisions.

// 1 / x by Newton-Raphson from the hardware estimate (v_rcp_f64, relative error below
// 2^-22): each step r <- r + r (1 - x r) squares the relative error, so two steps reach the
// rounding level.  For the positive, finite binary diffusion coefficients.
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// A workgroup of kWavesPerBlock waves takes kWavesPerBlock * kPointsPerWave consecutive
// points; the species' fits and the pair fits (5 K(K+1)/2 doubles: 57 KB for GRI) are
// staged in LDS once per workgroup, so the pair terms read LDS instead of gathering their
// coefficients from L2 for every point (112 KB per GRI point).
constexpr int kWavesPerBlock = 8;
constexpr int kPointsPerWave = 16;

__host__ __device__ inline int transport_lds_doubles(int K) {
    return 5 * (K * (K + 1) / 2) + 10 * K + K + 3 * 64 * kWavesPerBlock;
}

__global__ void __launch_bounds__(64 * kWavesPerBlock) transport_props_kernel(
    DevMech m, const double* __restrict__ state, int n, double threshold, double* __restrict__ mu_out,
    double* __restrict__ lambda_out, int ml_stride, double* __restrict__ rhoD_out, int rd_stride) {
    extern __shared__ double lds[];
    const int K = m.K, nv = m.nVars;
    const int npair = K * (K + 1) / 2;
    double* tdiff = lds;                // D_jk fits, pair slot order (pair_slot)
    double* tvisc = tdiff + 5 * npair;  // sqrt(mu_k) / T^(1/4) fits
    double* tcond = tvisc + 5 * K;      // lambda_k / sqrt(T) fits
    double* tW = tcond + 5 * K;         // molecular weights
    for (int i = threadIdx.x; i < 5 * npair; i += blockDim.x) tdiff[i] = m.tr_diff[i];
    for (int i = threadIdx.x; i < 5 * K; i += blockDim.x) {
        tvisc[i] = m.tr_visc[i];
        tcond[i] = m.tr_cond[i];
    }
    for (int i = threadIdx.x; i < K; i += blockDim.x) tW[i] = m.W[i];
    __syncthreads();
    const int lane = lane_id();
    const int wave = wave_in_block();
    double* sX = tW + K + wave * 192;  // per wave: X_k, 1 / a_k and X_k W_k of the current point
    double* sIa = sX + 64;
    double* sXW = sIa + 64;
    const bool sp = lane < K;
    const double Wk = sp ? tW[lane] : 1.0;
    const double rW4 = sp ? 1.0 / sqrt(sqrt(Wk)) : 0.0;  // W_k^(-1/4)
    const double* gk = m.tr_wilke + (sp ? lane : 0);   // g_jk at gk[j * K]
    for (int it = 0; it < kPointsPerWave; it++) {
        const int p = (blockIdx.x * kPointsPerWave + it) * kWavesPerBlock + wave;
        if (p >= n) break;  // uniform over the wave
        wave_lds_sync();  // the previous point's reads of sX, sIa, sXW are done
        const double* s = state + (size_t)p * nv;
        const double T = s[1];
        const double y = sp ? s[2 + lane] : 0.0;
        // Phase::setMassFractions: clip negatives, normalise; mean molecular weight
        const double yc = fmax(y, 0.0);
        const double norm = wave_sum(yc);
        const double ym = sp ? (yc * (1.0 / norm)) * (1.0 / Wk) : 0.0;
        const double mmw = 1.0 / wave_sum(ym);
        const double dens = m.P * mmw / (T * kGasConstant);
        const double X = sp ? fmax(1e-20, ym * mmw) : 0.0;
        const bool major = sp && X >= threshold;
        // MixTransport::update_T
        const double logt = log(T), sqrt_t = sqrt(T), t14 = sqrt(sqrt_t);
        const double pt[5] = {1.0, logt, logt * logt, logt * logt * logt, logt * logt * logt * logt};
        double sqv = 0.0, visc = 0.0, cond = 1.0;
        if (sp) {
            sqv = t14 * dot5(pt, tvisc + 5 * lane);
            visc = sqv * sqv;
            cond = sqrt_t * dot5(pt, tcond + 5 * lane);
        }
        // MixTransport::thermalConductivity
        const double lam = 0.5 * (wave_sum(sp ? X * cond : 0.0) + 1.0 / wave_sum(sp ? X / cond : 0.0));
        const double ak = sqv * rW4;
        sX[lane] = X;  // species lanes; 0 elsewhere
        sIa[lane] = sp ? 1.0 / ak : 0.0;
        sXW[lane] = X * Wk;
        wave_lds_sync();
        // ApproxMixTransport::viscosity over the major species (the loop's skips are uniform:
        // sX[j] is the same for every lane)
        double contrib = 0.0;
        if (major) {
            double sum = 0.0;
            double gj = gk[0];
            for (int j = 0; j < K; j++) {
                const double gn = j + 1 < K ? gk[(j + 1) * K] : 0.0;
                const double Xj = sX[j];
                if (Xj >= threshold) {
                    const double t = ak * sIa[j] + 1.0;
                    sum += (Xj * gj) * (t * t);
                }
                gj = gn;
            }
            contrib = X * visc / sum;
        }
        const double mu = wave_sum(contrib);
        // ApproxMixTransport::getMixDiffCoeffsMass (pairs with a major member, as
        // ApproxMixTransport::updateDiff_T fills them), times rho; X_i / D and
        // X_i W_i / D share one reciprocal of D
        if (sp) {
            const double tts = T * sqrt_t;
            double s1 = 0.0, s2 = 0.0;
            int slot = lane;  // the pair (0, lane)
            for (int i = 0; i < K; i++) {
                const int here = slot;
                slot += i < lane ? K - 1 - i : 1;
                if (i == lane) continue;
                const double Xi = sX[i];
                if (!major && !(Xi >= threshold)) continue;
                const double rb = rcp_nr(tts * dot5(pt, tdiff + 5 * here));
                s1 += Xi * rb;
                s2 += sXW[i] * rb;
            }
            s1 *= m.P;
            s2 *= m.P * X / (mmw - Wk * X);
            rhoD_out[(size_t)p * rd_stride + lane] = dens * (1.0 / (s1 + s2));
        }
        if (lane == 0) {
            mu_out[(size_t)p * ml_stride] = mu;
            lambda_out[(size_t)p * ml_stride] = lam;
        }
    }
}

}  // namespace

int launch_transport_props(const DevMech& m, const double* state, int n, double threshold, double* mu,
                           double* lambda, int ml_stride, double* rhoD, int rd_stride, hipStream_t stream) {
    const int per_block = kWavesPerBlock * kPointsPerWave;
    const int blocks = (int)(((long)n + per_block - 1) / per_block);
    const size_t lds = (size_t)transport_lds_doubles(m.K) * sizeof(double);
    hipLaunchKernelGGL(transport_props_kernel, dim3(blocks), dim3(64 * kWavesPerBlock), lds, stream, m, state,
                       n, threshold, mu, lambda, ml_stride, rhoD, rd_stride);
    return (int)hipGetLastError();
}

}  // namespace emx
