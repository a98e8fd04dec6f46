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
template <int NR, bool INTERP, bool ORD>
__device__ __forceinline__ bool bdf_f(const DevMech& m, int lane, double yv, BdfSystem& S, RateRegs<NR>& R,
                             const double* perturb, double fa, double qi, const WaveLds& L,
                             double& ydot) {
    S.nfe++;
    const double T = bcast(yv, 1);
    GasState g = set_state(m, lane, yv, S.rW, T, true, L);
    double hRT, cpR;
    bool ok = update_T<NR, INTERP>(m, lane, g, R, L, hRT, cpR);
    rates_of_progress<NR, ORD>(m, lane, g, R, perturb, fa, L);
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

// Factors of the Newton matrix M = I - c J, kept between right-hand sides.
//   M (LDS): row i, column j at M[(j / 2) * 128 + 2 i + (j % 2)]: a lane's two entries of a
//            column pair are adjacent, so one 16-byte access moves both.  Factored in place
//            without row exchanges (L multipliers in the rows not yet pivoted at their step,
//            U in the pivot rows).
//   src:     lane k holds the row that was the pivot of step k (lanes >= n: k itself)
//   u, r:    lane k holds U of step k's pivot, U[src_k][k], and 1 / U[src_k][k]
struct LuFactors {
    int src;
    double u, r;
};

__device__ __forceinline__ int lu_index(int row, int col) { return (col >> 1) * 128 + 2 * row + (col & 1); }

// Pivot of one elimination step over column values `a` (this lane's entry): the
// not-yet-pivoted row with the largest |a| (lowest lane on ties).  Returns the pivot lane
// and 1 / a[pivot] (IEEE division), or -1 if every candidate is zero (singular).
__device__ __forceinline__ int lu_pivot(double a, bool used, double& inv) {
    const double cand = used ? -1.0 : fabs(a);
    const double best = wave_max(cand);
    if (!(best > 0.0)) return -1;
    const int p = __ffsll(__ballot(!used && cand == best)) - 1;
    inv = 1.0 / bcast(a, p);
    return p;
}

// LU with partial pivoting, two elimination steps per pass over the trailing columns.
// Step k: pivot p1 on column k, l1 = M[i][k] * (1 / M[p1][k]); step k + 1: pivot p2 on
// column k + 1 after step k, l2 likewise.  A trailing entry becomes
// v1 = M[i][j] - l1 M[p1][j], then v2 = v1 - l2 v1(p2) (one FMA each, in that order: the
// same operations as two single steps), written once: v2 for rows pivoted after k + 1,
// v1 for p2, nothing for p1 and earlier pivots.  Column pairs are read and written 16
// bytes at a time (the pivot rows' pairs as broadcast reads; v_readlane costs ~10-15
// cycles of issue on gfx950, an LDS read of 16 bytes a few), and a block's loads are
// issued before the previous block's stores.
__device__ __forceinline__ bool lu_factor(double* M, const double* J, double c, int n, int lane,
                                          LuFactors& lu) {
    const bool row = lane < n;
    const int npair = (n + 1) >> 1;
    double2* Mp = reinterpret_cast<double2*>(M) + lane;  // pair q of this row: Mp[q * 64]
    // the per-wave Jacobian slot holds 64 columns (those >= n unused): 16 loads in flight
    for (int q0 = 0; q0 < npair; q0 += 8) {
        double jv[16];
#pragma unroll
        for (int t = 0; t < 16; t++) jv[t] = J[(2 * q0 + t) * 64 + lane];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int j = 2 * (q0 + t);
            double2 v;
            v.x = (row && j < n) ? ((lane == j ? 1.0 : 0.0) - c * jv[2 * t]) : 0.0;
            v.y = (row && j + 1 < n) ? ((lane == j + 1 ? 1.0 : 0.0) - c * jv[2 * t + 1]) : 0.0;
            if (q0 + t < npair) Mp[(q0 + t) * 64] = v;
        }
    }
    wave_lds_sync();
    bool used = !row;
    int step = lane;
    double rd = 0.0;
    for (int k = 0; k < n; k += 2) {
        const int q = k >> 1;
        const double2 col = Mp[q * 64];
        double inv1;
        const int p1 = lu_pivot(col.x, used, inv1);
        if (p1 < 0) return false;
        const bool before = used;  // pivoted at an earlier pass
        if (lane == p1) {
            used = true;
            step = k;
            rd = inv1;
        }
        const double l1 = col.x * inv1;
        if (k + 1 == n) {  // odd n: the last step has no column after it
            if (!used) M[lu_index(lane, k)] = l1;
            break;
        }
        // column k + 1 after step k, then step k + 1
        const double a1 = col.y - l1 * bcast(col.y, p1);
        double inv2;
        const int p2 = lu_pivot(a1, used, inv2);
        if (p2 < 0) return false;
        const bool is_p2 = lane == p2;
        if (is_p2) {
            used = true;
            step = k + 1;
            rd = inv2;
        }
        const double l2 = a1 * inv2;
        // rows written this pass: p2 and those still not pivoted
        const bool write = !before && lane != p1;
        if (write) Mp[q * 64] = make_double2(l1, is_p2 ? a1 : l2);
        // Trailing pairs q + 1 .. npair - 1, 2 pairs per block (loads past the last pair
        // clamp to it).  The pivot rows' pairs are broadcast LDS reads; row p2 after step k
        // is formed in every lane as M[p2][j] - l1(p2) M[p1][j], the FMA lane p2 itself does.
        const double l1p2 = bcast(l1, p2);
        const double2* R1 = reinterpret_cast<const double2*>(M) + p1;
        const double2* R2 = reinterpret_cast<const double2*>(M) + p2;
        const int qlast = npair - 1;
        double2 own[2], r1[2], r2[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int qq = min(q + 1 + t, qlast) * 64;
            own[t] = Mp[qq];
            r1[t] = R1[qq];
            r2[t] = R2[qq];
        }
        for (int q0 = q + 1; q0 <= qlast; q0 += 2) {
            double2 nown[2], nr1[2], nr2[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int qq = min(q0 + 2 + t, qlast) * 64;
                nown[t] = Mp[qq];
                nr1[t] = R1[qq];
                nr2[t] = R2[qq];
            }
            double2 res[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const double x1 = own[t].x - l1 * r1[t].x;
                const double y1 = own[t].y - l1 * r1[t].y;
                const double ux = r2[t].x - l1p2 * r1[t].x;
                const double uy = r2[t].y - l1p2 * r1[t].y;
                const double x2 = x1 - l2 * ux;
                const double y2 = y1 - l2 * uy;
                res[t] = is_p2 ? make_double2(x1, y1) : make_double2(x2, y2);
            }
            if (write) {
                Mp[q0 * 64] = res[0];
                if (q0 + 1 <= qlast) Mp[(q0 + 1) * 64] = res[1];
            }
#pragma unroll
            for (int t = 0; t < 2; t++) {
                own[t] = nown[t];
                r1[t] = nr1[t];
                r2[t] = nr2[t];
            }
        }
        wave_lds_sync();
    }
    wave_lds_sync();
    // src: lane step_i receives i (ds_permute writes lane i's value into lane addr / 4)
    lu.src = __builtin_amdgcn_ds_permute(step * 4, lane);
    const uint64_t rb = __builtin_bit_cast(uint64_t, rd);
    const int rlo = __builtin_amdgcn_ds_permute(step * 4, (int)(uint32_t)rb);
    const int rhi = __builtin_amdgcn_ds_permute(step * 4, (int)(uint32_t)(rb >> 32));
    lu.r = __builtin_bit_cast(double, ((uint64_t)(uint32_t)rhi << 32) | (uint32_t)rlo);
    lu.u = 1.0;
    if (row) lu.u = M[lu_index(lu.src, lane)];
    return true;
}

__device__ __forceinline__ double lane_gather(double v, int src) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_ds_bpermute(src * 4, (int)(uint32_t)b);
    const int hi = __builtin_amdgcn_ds_bpermute(src * 4, (int)(uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Solve M x = b; b is this lane's row value (lane = row); returns x for lane = variable.
// Lane k works on the pivot row of step k (the right-hand side gathered into that
// order): forward substitution, then back substitution, in which the division by the
// pivot is the reciprocal times the value with one correction (r = 1/u rounded to
// nearest, q0 = z r, q = q0 + (z - q0 u) r: the correctly rounded quotient by Markstein's
// theorem).  The LDS reads do not depend on the running values and are issued ahead.
__device__ __forceinline__ double lu_solve(const double* M, int n, int lane, const LuFactors& lu, double b) {
    // this lane's row in pivot order, column pair q at Mr[q * 64] (16-byte reads)
    const double2* Mr = reinterpret_cast<const double2*>(M) + lu.src;
    const int qlast = (n - 1) >> 1;
    double z = lane_gather(b, lu.src);
    // 8 steps (4 column pairs) per block, unrolled by hand (the loop holds v_readlane,
    // which the compiler does not unroll at run-time trip counts); steps past the end read
    // a clamped pair and change nothing
    for (int q0 = 0; q0 <= qlast; q0 += 4) {
        double2 m[4];
#pragma unroll
        for (int t = 0; t < 4; t++) m[t] = Mr[min(q0 + t, qlast) * 64];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int k = 2 * q0 + t;
            const double mk = (t & 1) ? m[t >> 1].y : m[t >> 1].x;
            const double zk = bcast(z, min(k, 63));
            z -= (lane > k && k < n ? mk : 0.0) * zk;
        }
    }
    double x = 0.0;
    for (int q0 = qlast; q0 >= 0; q0 -= 4) {
        double2 m[4];
#pragma unroll
        for (int t = 0; t < 4; t++) m[t] = Mr[max(q0 - t, 0) * 64];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int k = 2 * q0 + 1 - t;  // columns 2 q0 + 1, 2 q0, 2 q0 - 1, ...
            const double mk = (t & 1) ? m[t >> 1].x : m[t >> 1].y;
            const bool live = k >= 0 && k < n;
            const double q0v = z * lu.r;
            const double q = fma(fma(-q0v, lu.u, z), lu.r, q0v);
            const double xk = bcast(q, live ? k : 0);
            if (live && lane == k) x = xk;
            z -= (live && lane < k ? mk : 0.0) * xk;
        }
    }
    return x;
}

// SourceSystemCVODE::fdJacobian: column i written to J[i*64 + lane]
template <int NR, bool INTERP, bool ORD>
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
        ok = bdf_f<NR, INTERP, ORD>(m, lane, yp, S, R, perturb, fa, qi, L, yd2) && ok;
        J[i * 64 + lane] = (yd2 - ydot) / dy;
    }
    return ok;
}

template <int NR, bool INTERP, bool ORD>
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
    ok = bdf_f<NR, INTERP, ORD>(m, lane, y, S, R, perturb, fn_mult(o, 0.0), igniter(o, S.qign, 0.0), L, f0) && ok;
    double qprev = S.qdot, qint = 0.0;
    double scale = S.atol + fabs(y) * rtol;
    const double d0 = rms(y, scale, active, n), d1 = rms(f0, scale, active, n);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    h0 = fmin(h0, tf);
    double f1;
    ok = bdf_f<NR, INTERP, ORD>(m, lane, y + h0 * f0, S, R, perturb, fn_mult(o, h0), igniter(o, S.qign, h0), L,
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
    LuFactors lu;
    ok = fd_jacobian<NR, INTERP, ORD>(m, lane, n, y, f0, S, R, perturb, fn_mult(o, 0.0), igniter(o, S.qign, 0.0),
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
                    if (!lu_factor(M, J, c, n, lane, lu)) return -5;
                    lu_valid = true;
                }
                dvec = 0.0;
                ynew = ypred;
                double dy_norm_old = -1.0;
                int k;
                for (k = 0; k < kNewtonMaxIter; k++) {
                    double fy;
                    ok = bdf_f<NR, INTERP, ORD>(m, lane, ynew, S, R, perturb, fn_mult(o, t_new),
                                           igniter(o, S.qign, t_new), L, fy) && ok;
                    const int bad = wave_or(active && !isfinite(fy) ? 1 : 0);
                    if (bad) break;
                    double dy = lu_solve(M, n, lane, lu, c * fy - psi - dvec);
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
                    ok = bdf_f<NR, INTERP, ORD>(m, lane, ypred, S, R, perturb, fa, qi, L, fp) && ok;
                    ok = fd_jacobian<NR, INTERP, ORD>(m, lane, n, ypred, fp, S, R, perturb, fa, qi, L, J) && ok;
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
            ok = bdf_f<NR, INTERP, ORD>(m, lane, y, S, R, perturb, fn_mult(o, t), igniter(o, S.qign, t), L, ftmp) && ok;
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

template <int NR, bool INTERP, bool ORD>
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
    const int per_wave = chem + ((nv + 1) / 2) * 128 + 64;
    double* base = lds + wave * per_wave;
    WaveLds L = carve_lds(base, m.nRounds, m.nChunkRounds);
    init_wave_lds(L, lane, m.nRounds, m.nChunkRounds);
    double* M = base + chem;
    double* ru = M + ((nv + 1) / 2) * 128;
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
            err = bdf_integrate<NR, INTERP, ORD>(m, o, lane, active, y, S, R, perturb, L, M, ru, J, dt, nsteps,
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
           (wave_lds_doubles(m.nRounds, m.nChunkRounds) + ((m.nVars + 1) / 2) * 128 + 64) * sizeof(double);
}

int launch_bdf_stage(const DevMech& m, const DevOpts& o, const double* state_in,
                     const double* split_const, const double* qloss, const double* perturb,
                     double dt, int nPoints, const int* order, unsigned int* queue,
                     double* state_out, int* status, int* stat1, int* stat2, int* cost, double* qint,
                     int grid_blocks, int waves_per_block, double* jac_scratch, hipStream_t stream) {
    const size_t lds = bdf_lds_bytes(m, waves_per_block);
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, stream, queue);
// ORD: the mechanism has reactions with explicit forward orders
#define EMX_BDF_LAUNCH_O(NRV, INTERP, ORD)                                                       \
    hipLaunchKernelGGL((bdf_stage_kernel<NRV, INTERP, ORD>), dim3(grid_blocks),                  \
                       dim3(64 * waves_per_block), lds, stream, m, o, state_in, split_const,    \
                       qloss, perturb, dt, nPoints, order, queue, state_out, status, stat1,     \
                       stat2, cost, qint, jac_scratch)
#define EMX_BDF_LAUNCH(NRV)                                                                     \
    if (m.has_orders) {                                                                         \
        if (m.interp) EMX_BDF_LAUNCH_O(NRV, true, true);                                        \
        else EMX_BDF_LAUNCH_O(NRV, false, true);                                                \
    } else {                                                                                    \
        if (m.interp) EMX_BDF_LAUNCH_O(NRV, true, false);                                       \
        else EMX_BDF_LAUNCH_O(NRV, false, false);                                               \
    }
    if (m.nRounds <= 1) { EMX_BDF_LAUNCH(1) }
    else if (m.nRounds <= 2) { EMX_BDF_LAUNCH(2) }
    else if (m.nRounds <= 4) { EMX_BDF_LAUNCH(4) }
    else if (m.nRounds <= 6) { EMX_BDF_LAUNCH(6) }
    else { EMX_BDF_LAUNCH(8) }
#undef EMX_BDF_LAUNCH_O
#undef EMX_BDF_LAUNCH
    return (int)hipGetLastError();
}

}  // namespace emx
