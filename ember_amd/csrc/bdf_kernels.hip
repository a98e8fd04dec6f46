// Batched chemistry production stage with a variable-order BDF integrator:
// SourceSystemCVODE semantics (src/sourceSystem.cpp:104-164, 268-338) on one
// wavefront per grid point.
//
// Per point: the ODE is SourceSystemCVODE::f (net production rates with
// setMassFractions_NoNorm), the Jacobian is Ember's finite-difference
// fdJacobian (src/sourceSystem.cpp:268-293), tolerances are Ember's
// integratorRelTol / *AbsTol (src/sourceSystem.cpp:116-126).  The BDF itself is
// the variable-order, variable-step method of Shampine & Reichelt (SIAM J. Sci.
// Comput. 18, 1997; NDF family with kappa = 0) in difference-array form with
// modified Newton iterations, the same algorithm as the CPU oracle's, so results
// agree step for step.  The step, order and Newton logic restate scipy's
// scipy.integrate._ivp.bdf (SciPy 1.15, BSD-3-Clause: NEWTON_MAXITER = 4, the
// 0.2 / 10 step-factor bounds, solve_bdf_system's convergence test, change_D /
// compute_R, select_initial_step's 0.01 d0/d1 rule, min_step = 10 ulp(t)); the
// per-step failure bounds (10 convergence, 7 error-test failures) are CVODE's
// defaults.  CVODE's own step/order heuristics are not reproduced.
//
// Per-point status: 0 ok, -2 rates failed (interpolation table range), -3 step size
// below the minimum, -4 integrator_max_steps accepted steps, -5 singular Newton
// matrix, -6 non-finite error estimate, -7 / -8 too many error-test / convergence
// failures in one step, -9 invalid input state (non-finite, or T <= 0).
//
// Layout: lane v holds state variable v (and its column of the difference array);
// the Jacobian lives in a per-wave global scratch slot (column j contiguous over
// lanes); the Newton matrix is factored in LDS, row i owned by lane i, partial
// pivoting without row exchanges (each lane remembers the step at which it was
// the pivot row).
#include <hip/hip_runtime.h>

#include "chem_device.h"
#include "kernels.h"

namespace emx {

namespace {

constexpr int kMaxOrder = 5;
constexpr int kNewtonMaxIter = 4;
// failed attempts allowed per step (CVODE's defaults maxncf and maxnef)
constexpr int kMaxConvFail = 10;
constexpr int kMaxErrFail = 7;

struct BdfSystem {
    double W, rW, sc, qloss, atol;
    double qign;  // igniter amplitude at this point's position (0: none)
    double qdot;
    int nfe;
};

// SourceSystemCVODE::f for this lane's variable.  Returns false if rates failed.
// Every device function below is forced inline: outlined, its LDS pointers lose their
// address space (flat loads) and the rate-constant registers go to the stack.
template <int NR, bool INTERP>
__device__ __forceinline__ bool bdf_f(const DevMech& m, int lane, double yv, BdfSystem& S, RateRegs<NR>& R,
                             const double* perturb, double fa, double qi, const WaveLds& L,
                             double& ydot) {
    S.nfe++;
    const double T = bcast(yv, 1);
    GasState g = set_state(m, lane, yv, S.rW, T, true, L);
    double hRT, cpR;
    bool ok = update_T<NR, INTERP>(m, lane, g, R, L, hRT, cpR);
    rates_of_progress<NR>(m, lane, g, R, perturb, fa, L);
    double wQ, wD;
    species_rates(m, lane, L, wQ, wD);
    const bool sp = is_species_lane(lane, m);
    const double wdot = wQ - wD;
    const double hk = sp ? g.T * kGasConstant * hRT : 0.0;
    const double rho = g.dens;
    const double s = wave_sum(g.ym * cpR);
    const double cp = (kGasConstant * (g.mmw * s)) / g.mmw;
    double qdot = -wave_sum(sp ? wdot * hk : 0.0) + qi;  // + getQdotIgniter(t)
    qdot -= S.qloss;
    S.qdot = qdot;
    if (lane == 0) ydot = S.sc;
    else if (lane == 1) ydot = qdot / (rho * cp) + S.sc;
    else if (sp) ydot = wdot * S.W / rho + S.sc;
    else ydot = 0.0;
    return ok;
}

__device__ __forceinline__ double rms(double v, double scale, bool active, int n) {
    double r = active ? v / scale : 0.0;
    return sqrt(wave_sum(r * r) / n);
}

__device__ __forceinline__ double gamma_k(int k) {
    double g = 0.0;
    for (int i = 1; i <= k; i++) g += 1.0 / i;
    return g;
}

// Difference-array rescaling D[0..order] <- (R U)^T D[0..order]
template <int ND>
__device__ __forceinline__ void change_D(double (&D)[ND], int order, double factor, double* ru /*LDS 36*/) {
    const int m1 = order + 1;
    const int lane = lane_id();
    // lane (i, j) computes RU[i][j] = sum_l R[i][l] U[l][j]
    if (lane < m1 * m1) {
        const int i = lane / m1, j = lane % m1;
        double s = 0.0;
        for (int l = 0; l < m1; l++) {
            // R(f)[i][l] = prod_{t=1..i} M[t][l], M[t][l] = (t - 1 - f l)/t, M[0][*] = 1
            double r = 1.0;
            for (int t = 1; t <= i; t++) r *= (l >= 1) ? (t - 1 - factor * l) / (double)t : 0.0;
            double u = 1.0;
            for (int t = 1; t <= l; t++) u *= (j >= 1) ? (t - 1 - 1.0 * j) / (double)t : 0.0;
            s += r * u;
        }
        ru[i * m1 + j] = s;
    }
    wave_lds_sync();
    double Dn[kMaxOrder + 1];
#pragma unroll
    for (int j = 0; j <= kMaxOrder; j++) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i <= kMaxOrder; i++)
            if (i < m1 && j < m1) s += ru[i * m1 + j] * D[i];
        Dn[j] = s;
    }
#pragma unroll
    for (int j = 0; j <= kMaxOrder; j++)
        if (j < m1) D[j] = Dn[j];
    wave_lds_sync();
}

// LU of M = I - c J in LDS (M[j*64 + lane] = row `lane`, column j), row-per-lane,
// partial pivoting: at step k the pivot is the not-yet-pivoted lane with the largest
// |M[lane][k]| (lowest lane on ties).  step_of = k for the pivot lane of step k.
__device__ __forceinline__ bool lu_factor(double* M, const double* J, double c, int n, int lane, int& step_of) {
    const bool row = lane < n;
    for (int j = 0; j < n; j++)
        M[j * 64 + lane] = row ? ((lane == j ? 1.0 : 0.0) - c * J[j * 64 + lane]) : 0.0;
    wave_lds_sync();
    step_of = row ? 1 << 30 : -1;
    bool used = !row;
    for (int k = 0; k < n; k++) {
        const double akk = M[k * 64 + lane];
        double cand = used ? -1.0 : fabs(akk);
        double best = wave_max(cand);
        if (!(best > 0.0)) return false;
        unsigned long long ball = __ballot(!used && cand == best);
        const int p = __ffsll(ball) - 1;
        if (lane == p) {
            used = true;
            step_of = k;
        }
        const double inv = 1.0 / bcast(akk, p);
        const double l = akk * inv;
        if (!used) M[k * 64 + lane] = l;
        for (int j = k + 1; j < n; j++) {
            const double pj = M[j * 64 + p];
            if (!used) M[j * 64 + lane] -= l * pj;
        }
        wave_lds_sync();
    }
    return true;
}

// Solve M x = b; b is this lane's row value (lane = row); returns x for lane = variable.
__device__ __forceinline__ double lu_solve(const double* M, int n, int lane, int step_of, double b) {
    // forward: rows in pivot order
    for (int k = 0; k < n; k++) {
        const int p = __ffsll(__ballot(step_of == k)) - 1;
        const double zk = bcast(b, p);
        if (step_of > k) b -= M[k * 64 + lane] * zk;
    }
    double x = 0.0;
    for (int k = n - 1; k >= 0; k--) {
        const int p = __ffsll(__ballot(step_of == k)) - 1;
        double xk = 0.0;
        if (lane == p) xk = b / M[k * 64 + lane];
        xk = bcast(xk, p);
        if (lane == k) x = xk;
        if (step_of < k && step_of >= 0) b -= M[k * 64 + lane] * xk;
    }
    return x;
}

// SourceSystemCVODE::fdJacobian: column i written to J[i*64 + lane]
template <int NR, bool INTERP>
__device__ __forceinline__ bool fd_jacobian(const DevMech& m, int lane, int n, double yv, double ydot, BdfSystem& S,
                            RateRegs<NR>& R, const double* perturb, double fa, double qi,
                            const WaveLds& L, double* J) {
    const double eps = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON)
    const double atl = 2.220446049250313e-16;   // DBL_EPSILON
    bool ok = true;
    for (int i = 0; i < n; i++) {
        const double yi = bcast(yv, i);
        const double dy = (fabs(yi) > atl) ? fabs(yi) * eps : fabs(yi) * eps + atl;
        double yp = yv;
        if (lane == i) yp += dy;
        double yd2;
        ok = bdf_f<NR, INTERP>(m, lane, yp, S, R, perturb, fa, qi, L, yd2) && ok;
        J[i * 64 + lane] = (yd2 - ydot) / dy;
    }
    return ok;
}

template <int NR, bool INTERP>
__device__ __forceinline__ int bdf_integrate(const DevMech& m, const DevOpts& o, int lane, bool active, double& y,
                             BdfSystem& S, RateRegs<NR>& R, const double* perturb, const WaveLds& L,
                             double* M, double* ru, double* J, double tf, int& nsteps,
                             double* qint_out) {
    const int n = m.nVars;
    const double rtol = o.rtol;
    const double EPS = 2.220446049250313e-16;
    const double newton_tol = fmax(10 * EPS / rtol, fmin(0.03, sqrt(rtol)));
    nsteps = 0;
    bool ok = true;
    double f0;
    // the right-hand sides see t_start + t (CVODE's t0 = tInitial, src/sourceSystem.cpp:298)
    ok = bdf_f<NR, INTERP>(m, lane, y, S, R, perturb, fn_mult(o, 0.0), igniter(o, S.qign, 0.0), L, f0) && ok;
    double qprev = S.qdot, qint = 0.0;
    double scale = S.atol + fabs(y) * rtol;
    const double d0 = rms(y, scale, active, n), d1 = rms(f0, scale, active, n);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    h0 = fmin(h0, tf);
    double f1;
    ok = bdf_f<NR, INTERP>(m, lane, y + h0 * f0, S, R, perturb, fn_mult(o, h0), igniter(o, S.qign, h0), L,
                           f1) && ok;
    const double d2 = rms(f1 - f0, scale, active, n) / h0;
    const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3)
                                                   : pow(0.01 / fmax(d1, d2), 1.0 / 2.0);
    double h_abs = fmin(100 * h0, h1);
    h_abs = fmin(h_abs, tf);
    int order = 1, n_equal = 0;
    double D[kMaxOrder + 3];
#pragma unroll
    for (int i = 0; i < kMaxOrder + 3; i++) D[i] = 0.0;
    D[0] = y;
    D[1] = f0 * h_abs;
    bool lu_valid = false;
    int step_of = -1;
    ok = fd_jacobian<NR, INTERP>(m, lane, n, y, f0, S, R, perturb, fn_mult(o, 0.0), igniter(o, S.qign, 0.0),
                                 L, J) && ok;
    if (!ok) return -2;
    double t = 0.0;
    double error_norm = 0.0, safety = 0.0;
    double dvec = 0.0, ynew = 0.0;
    while (t < tf) {
        double min_step = 10 * fabs(nextafter(t, __builtin_huge_val()) - t);
        min_step = fmax(min_step, o.min_step);
        if (h_abs < min_step) h_abs = min_step;
        bool accepted = false, current_jac = false;
        int n_iter = 0;
        double t_new = 0.0;
        // failed attempts of this step, bounded as CVODE bounds them
        int n_conv_fail = 0, n_err_fail = 0;
        while (!accepted) {
            if (h_abs < min_step) return -3;
            if (nsteps >= o.max_steps) return -4;
            double h = h_abs;
            t_new = t + h;
            if (t_new - tf > 0) {
                t_new = tf;
                change_D(D, order, fabs(t_new - t) / h_abs, ru);
                n_equal = 0;
                lu_valid = false;
            }
            h = t_new - t;
            h_abs = fabs(h);
            double ypred = 0.0, psi = 0.0;
#pragma unroll
            for (int i = 0; i <= kMaxOrder; i++)
                if (i <= order) ypred += D[i];
#pragma unroll
            for (int i = 1; i <= kMaxOrder; i++)
                if (i <= order) psi += D[i] * gamma_k(i);
            const double alpha_o = gamma_k(order);
            psi = psi / alpha_o;
            scale = S.atol + rtol * fabs(ypred);
            const double c = h / alpha_o;
            bool converged = false;
            while (!converged) {
                if (!lu_valid) {
                    if (!lu_factor(M, J, c, n, lane, step_of)) return -5;
                    lu_valid = true;
                }
                dvec = 0.0;
                ynew = ypred;
                double dy_norm_old = -1.0;
                int k;
                for (k = 0; k < kNewtonMaxIter; k++) {
                    double fy;
                    ok = bdf_f<NR, INTERP>(m, lane, ynew, S, R, perturb, fn_mult(o, t_new),
                                           igniter(o, S.qign, t_new), L, fy) && ok;
                    const int bad = wave_or(active && !isfinite(fy) ? 1 : 0);
                    if (bad) break;
                    double dy = lu_solve(M, n, lane, step_of, c * fy - psi - dvec);
                    const double dy_norm = rms(dy, scale, active, n);
                    const double rate = (dy_norm_old < 0) ? -1.0 : dy_norm / dy_norm_old;
                    if (rate >= 0 && (rate >= 1 || pow(rate, kNewtonMaxIter - k) / (1 - rate) * dy_norm > newton_tol))
                        break;
                    ynew += dy;
                    dvec += dy;
                    if (dy_norm == 0 || (rate >= 0 && rate / (1 - rate) * dy_norm < newton_tol)) {
                        converged = true;
                        break;
                    }
                    dy_norm_old = dy_norm;
                }
                n_iter = k + 1;
                if (!converged) {
                    if (current_jac) break;
                    double fp;
                    const double fa = fn_mult(o, t_new), qi = igniter(o, S.qign, t_new);
                    ok = bdf_f<NR, INTERP>(m, lane, ypred, S, R, perturb, fa, qi, L, fp) && ok;
                    ok = fd_jacobian<NR, INTERP>(m, lane, n, ypred, fp, S, R, perturb, fa, qi, L, J) && ok;
                    lu_valid = false;
                    current_jac = true;
                }
            }
            if (!ok) return -2;
            if (!converged) {
                if (++n_conv_fail >= kMaxConvFail) return -8;
                const double factor = 0.5;
                h_abs *= factor;
                change_D(D, order, factor, ru);
                n_equal = 0;
                lu_valid = false;
                continue;
            }
            safety = 0.9 * (2 * kNewtonMaxIter + 1) / (2 * kNewtonMaxIter + n_iter);
            scale = S.atol + rtol * fabs(ynew);
            error_norm = rms((1.0 / (order + 1)) * dvec, scale, active, n);
            // a non-finite error estimate fails the point (it must not pass as a small one)
            if (!isfinite(error_norm)) return -6;
            if (error_norm > 1) {
                if (++n_err_fail >= kMaxErrFail) return -7;
                const double factor = fmax(0.2, safety * pow(error_norm, -1.0 / (order + 1)));
                h_abs *= factor;
                change_D(D, order, factor, ru);
                n_equal = 0;
            } else {
                accepted = true;
            }
        }
        nsteps++;
        n_equal++;
        const double hstep = t_new - t;
        t = t_new;
        y = ynew;
        if (qint_out) {
            double ftmp;
            ok = bdf_f<NR, INTERP>(m, lane, y, S, R, perturb, fn_mult(o, t), igniter(o, S.qign, t), L, ftmp) && ok;
            S.nfe--;
            qint += 0.5 * hstep * (qprev + S.qdot);
            qprev = S.qdot;
        }
        // D[order+2] = d - D[order+1]; D[order+1] = d; D[i] += D[i+1] for i = order..0
#pragma unroll
        for (int i = 0; i <= kMaxOrder; i++) {
            if (i == order) {
                D[i + 2] = dvec - D[i + 1];
                D[i + 1] = dvec;
            }
        }
#pragma unroll
        for (int i = kMaxOrder; i >= 0; i--)
            if (i <= order) D[i] += D[i + 1];
        if (n_equal < order + 1) continue;
        double em = __builtin_huge_val(), ep = __builtin_huge_val();
        double Do = 0.0, Dop2 = 0.0;
#pragma unroll
        for (int i = 0; i < kMaxOrder + 3; i++) {
            if (i == order) Do = D[i];
            if (i == order + 2) Dop2 = D[i];
        }
        if (order > 1) em = rms((1.0 / order) * Do, scale, active, n);
        if (order < kMaxOrder) ep = rms((1.0 / (order + 2)) * Dop2, scale, active, n);
        double f[3];
        f[0] = pow(em, -1.0 / order);
        f[1] = pow(error_norm, -1.0 / (order + 1));
        f[2] = pow(ep, -1.0 / (order + 2));
        int best = 0;
        for (int j = 0; j < 3; j++)
            if (isnan(f[j])) f[j] = 0;
        for (int j = 1; j < 3; j++)
            if (f[j] > f[best]) best = j;
        order = max(1, order + best - 1);
        const double factor = fmin(10.0, safety * f[best]);
        h_abs *= factor;
        change_D(D, order, factor, ru);
        n_equal = 0;
        lu_valid = false;
    }
    if (qint_out && lane == 0) *qint_out = qint;
    return 0;
}

template <int NR, bool INTERP>
__global__ void __launch_bounds__(256) bdf_stage_kernel(
    DevMech m, DevOpts o, const double* __restrict__ state_in, const double* __restrict__ split_const,
    const double* __restrict__ qloss, const double* __restrict__ perturb, double dt, int nPoints,
    const int* __restrict__ order, unsigned int* __restrict__ queue, double* __restrict__ state_out,
    int* __restrict__ status, int* __restrict__ stat1, int* __restrict__ stat2,
    int* __restrict__ cost, double* __restrict__ qint_out, double* __restrict__ jac_scratch) {
    extern __shared__ double lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x / 64;
    const int wpb = blockDim.x / 64;
    const int nv = m.nVars;
    const int chem = wave_lds_doubles(m.nRounds, m.nChunkRounds);
    const int per_wave = chem + nv * 64 + 64;
    double* base = lds + wave * per_wave;
    WaveLds L = carve_lds(base, m.nRounds, m.nChunkRounds);
    init_wave_lds(L, lane, m.nRounds, m.nChunkRounds);
    double* M = base + chem;
    double* ru = M + nv * 64;
    double* J = jac_scratch + (size_t)(blockIdx.x * wpb + wave) * 64 * 64;
    const bool active = lane < nv;
    BdfSystem S;
    S.W = is_species_lane(lane, m) ? m.W[lane - 2] : 0.0;
    S.rW = is_species_lane(lane, m) ? m.rW[lane - 2] : 0.0;
    S.atol = lane == 0 ? o.atol_momentum : lane == 1 ? o.atol_energy : o.atol_species;
    RateRegs<NR> R;
    while (true) {
        unsigned int t = 0;
        if (lane == 0) t = atomicAdd(queue, 1u);
        t = (unsigned int)bcast_i((int)t, 0);
        if (t >= (unsigned int)nPoints) break;
        const int p = order ? order[t] : (int)t;
        double y0 = active ? state_in[(size_t)p * nv + lane] : 0.0;
        S.sc = (active && split_const) ? split_const[(size_t)p * nv + lane] : 0.0;
        S.qloss = qloss ? qloss[p] : 0.0;
        S.qign = o.qign ? o.qign[p] : 0.0;
        S.nfe = 0;
        R.T = -1.0;
        const unsigned long long t_point = point_clock();
        double y = y0;
        int nsteps = 0;
        int err = kInvalidState;
        if (valid_input_state(lane, active, y0))
            err = bdf_integrate<NR, INTERP>(m, o, lane, active, y, S, R, perturb, L, M, ru, J, dt, nsteps,
                                            qint_out ? qint_out + p : nullptr);
        if (active) state_out[(size_t)p * nv + lane] = (err >= 0) ? y : y0;
        if (lane == 0) {
            if (status) status[p] = err;
            if (stat1) stat1[p] = nsteps;
            if (stat2) stat2[p] = S.nfe;
            if (cost) cost[p] = point_cost(t_point);
        }
    }
}

__global__ void reset_queue_kernel(unsigned int* q) {
    if (threadIdx.x == 0) *q = 0u;
}

}  // namespace

size_t bdf_lds_bytes(const DevMech& m, int waves_per_block) {
    return (size_t)waves_per_block *
           (wave_lds_doubles(m.nRounds, m.nChunkRounds) + m.nVars * 64 + 64) * sizeof(double);
}

int launch_bdf_stage(const DevMech& m, const DevOpts& o, const double* state_in,
                     const double* split_const, const double* qloss, const double* perturb,
                     double dt, int nPoints, const int* order, unsigned int* queue,
                     double* state_out, int* status, int* stat1, int* stat2, int* cost, double* qint,
                     int grid_blocks, int waves_per_block, double* jac_scratch, hipStream_t stream) {
    const size_t lds = bdf_lds_bytes(m, waves_per_block);
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, stream, queue);
#define EMX_BDF_LAUNCH(NRV)                                                                     \
    if (m.interp)                                                                               \
        hipLaunchKernelGGL((bdf_stage_kernel<NRV, true>), dim3(grid_blocks),                    \
                           dim3(64 * waves_per_block), lds, stream, m, o, state_in, split_const, \
                           qloss, perturb, dt, nPoints, order, queue, state_out, status, stat1,  \
                           stat2, cost, qint, jac_scratch);                                           \
    else                                                                                        \
        hipLaunchKernelGGL((bdf_stage_kernel<NRV, false>), dim3(grid_blocks),                   \
                           dim3(64 * waves_per_block), lds, stream, m, o, state_in, split_const, \
                           qloss, perturb, dt, nPoints, order, queue, state_out, status, stat1,  \
                           stat2, cost, qint, jac_scratch);
    if (m.nRounds <= 1) { EMX_BDF_LAUNCH(1) }
    else if (m.nRounds <= 2) { EMX_BDF_LAUNCH(2) }
    else if (m.nRounds <= 4) { EMX_BDF_LAUNCH(4) }
    else if (m.nRounds <= 6) { EMX_BDF_LAUNCH(6) }
    else { EMX_BDF_LAUNCH(8) }
#undef EMX_BDF_LAUNCH
    return (int)hipGetLastError();
}

}  // namespace emx
