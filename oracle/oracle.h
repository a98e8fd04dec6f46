/*
 * CPU FP64 oracle for the Ember split-step hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++17 restatement of Ember's chemistry substep (QssIntegrator,
 * SourceSystemQSS, SourceSystemCVODE semantics), of the Cantera kinetics and
 * thermo it calls, and of the diffusion substep (DiffusionSystem,
 * TridiagonalIntegrator).  Only the tests, the smoke check and the CPU-baseline
 * leg of the benchmark may load it; the product library never links it.
 *
 * Layouts follow Ember: a state is an nVars x nPoints column-major matrix
 * (one point's U, T, Y_0..Y_{K-1} contiguous), src/splitSolver.h:73.
 */
#ifndef EMBER_ORACLE_H
#define EMBER_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_options {
    int integrator;        /* 0 = QSS (SourceSystemQSS), 1 = BDF (SourceSystemCVODE semantics) */
    int kinetics;          /* 0 = standard (Cantera BulkKinetics), 1 = interp (InterpKinetics) */
    double pressure;       /* Pa */
    /* QSS tolerances, python/ember/input.py:748-776 */
    double qss_epsmin, qss_epsmax, qss_dtmin, qss_dtmax;
    int qss_itermax;
    double qss_abstol, qss_minval;
    int qss_stability_check;
    /* BDF tolerances, python/ember/input.py:729-746 */
    double rtol, atol_momentum, atol_energy, atol_species, min_step;
    int max_steps;
    /* absolute time of the stage start (SourceSystem::setState tInitial); the right-hand
     * sides of a stage see t_start + t (the rate multiplier function and the igniter depend
     * on it) */
    double t_start;
    /* igniter, ConfigOptions ignition_* (src/readConfig.h:135-139, python/ember/input.py
     * class Ignition): SourceSystem::getQdotIgniter (src/sourceSystem.cpp:33-45) adds
     * energy / (stddev sqrt(2 pi) duration) exp(-(x - center)^2 / (2 stddev^2)) to qDot for
     * t in [t_start_ign, t_start_ign + duration).  positions: nPoints grid positions x
     * (SourceSystem::setPosition); NULL = 0 for every point. */
    double ignition_t_start, ignition_duration, ignition_energy, ignition_center, ignition_stddev;
    const double* positions;
} orc_options;

void orc_default_options(orc_options* o);

void* orc_mech_load(const char* path);
void orc_mech_free(void* mech);
int orc_mech_nspecies(void* mech);
int orc_mech_nreactions(void* mech);
/* Per-reaction rate multipliers (Cantera setMultiplier / m_perturb). */
void orc_mech_set_multipliers(void* mech, const double* m);
/* InterpKinetics table grid: n equally spaced temperatures over [Tmin, Tmax] (Ember's
 * defaults 200 over [273, 3500], src/chemistry0d.h:70-72).  Rates at a T outside it fail
 * (status -2).  Returns -1 for an invalid grid. */
int orc_mech_set_interp_range(void* mech, int n, double Tmin, double Tmax);
/* Ember's rate multiplier function, FlameSolver::rateMultiplierFunction: a ChebyshevFunction
 * whose coefficients ChebyshevFunction::setCoefficients(n, data) sets (data = [T0, T1,
 * c_0, ...], src/scalarFunction.cpp:33-51).  While set, every right-hand side multiplies
 * all reactions by a(t) at its own time, in place of the per-reaction multipliers
 * (src/sourceSystem.cpp:133-135, :453-455).  n = 0 or data NULL removes it; -1 for n = 1. */
int orc_mech_set_rate_function(void* mech, int n, const double* data);
/* a(t) of the function set on mech (1 when none is set) */
double orc_rate_function_eval(void* mech, double t);

/* Rates at n states.  nonorm=0: setState_TPY (clip + normalise);
 * nonorm=1: setMassFractions_NoNorm + setState_TP.  Any output may be NULL. */
int orc_rates(void* mech, const orc_options* o, int n, const double* T, const double* Y,
              int nonorm, double* wdotQ, double* wdotD, double* wdot_net, double* hk,
              double* rho, double* cp, double* ropf, double* ropr);

/* Thermodynamic properties of FlameSolver::updateChemicalProperties for nPoints points:
 * state point-major (nSpecies + 2 per point: U, T, Y); rho, Wmx, cp per point; cpSpec and
 * hk (J/kmol/K, J/kmol) nPoints x nSpecies; Wmx, cpSpec, hk nullable. */
int orc_thermo_properties(void* mech, double pressure, int nPoints, const double* state,
                          double* rho, double* Wmx, double* cp, double* cpSpec, double* hk,
                          int nthreads);

/* FlameSolver::prepareIntegrators' density derivative (src/flameSolver.cpp:203-213): with
 * ddt = ddt_conv + ddt_diff + ddt_prod (+ ddt_cross when balanced; nullable = zeros), all
 * point-major nPoints x (nSpecies + 2), drhodt[j] = -rho[j] (ddt_T / T + Wmx[j] sum_k
 * ddt_Yk / W_k) at the point's temperature in state. */
int orc_density_derivative(void* mech, int nPoints, const double* state, const double* ddt_conv,
                           const double* ddt_diff, const double* ddt_prod, const double* ddt_cross,
                           int balanced, const double* rho, const double* Wmx, double* drhodt);

/* Transport properties of FlameSolver::updateChemicalProperties with Ember's default
 * transport model ("Approx": ApproxMixTransport with mole-fraction threshold `threshold`)
 * for nPoints points of a point-major state: mu [Pa s], lambda [W/m/K] per point, rhoD
 * [kg/m/s] nPoints x nSpecies.  -1 when the table has no transport fits. */
int orc_transport_properties(void* mech, double pressure, double threshold, int nPoints,
                             const double* state, double* mu, double* lambda, double* rhoD,
                             int nthreads);

/* Convection stage (ConvectionSystemSplit::integrateToTime, src/convectionSystem.cpp)
 * for continuity_bc 0 (fixedLeft).  Boundary codes as DiffusionSystem's; strain rate
 * Ember's LinearFunction (strain_initial until strain_t0, linear over strain_dt, then
 * strain_final); tolerances as ConvectionSystemSplit::setTolerances (the W atol is 20 x
 * atol_species). */
typedef struct orc_conv_params {
    int alpha, beta, left_bc, right_bc, continuity_bc;
    double pressure, rVzero, Tleft, rhou;
    double strain_initial, strain_final, strain_t0, strain_dt;
    double rtol, atol_momentum, atol_energy, atol_species;
    int max_steps;
} orc_conv_params;

/* state, split_const: N x nVars point-major; drhodt: N (nullable = 0); Yleft: nSpecies
 * (nullable = 0).  state_out receives U, T and Y_k after [t0, t1].  steps_out (nullable,
 * 1 + nSpecies): accepted steps of the UTW system and of each species system.  Returns 0,
 * or < 0 on an integration failure (-1 step limit, -3 step size, -4 non-finite). */
int orc_convection_integrate(void* mech, const orc_conv_params* p, int N, const double* x,
                             const double* state, const double* split_const, const double* drhodt,
                             const double* Yleft, double t0, double t1, double* state_out,
                             int* steps_out, int nthreads);

/* One production stage for nPoints points (FlameSolver::integrateProductionTerms).
 * status: 0 ok, <0 failure (the point keeps its input state).
 * stat1/stat2: QSS gcount/rcount, BDF nsteps/nfe.  qint: integral of qdot over the
 * stage on accepted steps (nullable).  nthreads <= 0: OpenMP default. */
int orc_source_integrate(void* mech, const orc_options* o, int nPoints,
                         const double* state_in, const double* split_const,
                         const double* qloss, double dt, double* state_out,
                         int* status, int* stat1, int* stat2, double* qint, int nthreads);

/* Work counts of the BDF stage per point (zero split constants and heat loss), for the
 * benchmark's FLOP count: work[4 j ..] = right-hand-side evaluations (including the
 * finite-difference Jacobian columns; -1 for a failed point), Jacobians, LU
 * factorisations, linear solves.  Standard kinetics only (-1 otherwise). */
int orc_bdf_work(void* mech, const orc_options* o, int nPoints, const double* state_in, double dt,
                 int* work, int nthreads);

/* QssIntegrator on a test ODE.  problem 0: CHEMEQ2 cesium problem (n = 6);
 * problem 1: y' = qc - kc*y (n components).  y in/out.  Returns the integrator status. */
int orc_qss_test_problem(int problem, int n, const double* qc, const double* kc, double epsmin,
                         double epsmax, double dtmin, double dtmax, int itermax, double abstol,
                         double ymin, int stability, double tf, double* y, int* gcount, int* rcount);

/* SourceSystemCVODE::f at one state y (nVars), for independent ODE solvers. */
int orc_source_rhs(void* mech, const orc_options* o, const double* y, const double* split_const,
                   double qloss, double* ydot);

/* TridiagonalIntegrator with a constant (a, b, c, k) system
 * (test/tridiagonalIntegrator.cpp ConstantTridiagonalODE).
 * mode 0: take nsteps single steps of size h from t0.
 * mode 1: integrateToTime(t1) from t0 with initial step h.  y is in/out. */
int orc_tridiag_integrate(int N, const double* a, const double* b, const double* c,
                          const double* k, double t0, double h, double t1, int mode,
                          int nsteps, double* y, double* t_out, double* h_out);

/* Diffusion stage for nVars variables on one grid (FlameSolver::integrateDiffusionTerms).
 * B: nB x N rows (row-contiguous) with B_row[nVars]; D, split_const, state: nVars x N
 * column-major.  h_per_var NULL -> Ember's rule min(mult*min dlj^2/(B D), dtGlobal).
 * BC codes: 0 FixedValue, 1 ZeroGradient, 2 ControlVolume, 3 WallFlux; yInf and wallConst
 * apply to variable wallVar only (-1: none). */
int orc_diffusion_integrate(int nVars, int N, const double* B, int nB, const int* B_row,
                            const double* D, const double* split_const, const double* x,
                            int alpha, int leftBC, int rightBC, double yInf, double wallConst,
                            int wallVar, double multiplier, double dt_global,
                            const double* h_per_var, double t0, double t1, double* state,
                            int nthreads);

/* Discrete adjoint of the QSS stage (qss_itermax 1, no stability check; -6 otherwise).
 * Per point: the stage as orc_source_integrate (state_out, Q = heat release integrated
 * over accepted steps) and the gradients of J = Qbar*Q + ybar_final . y_final with the
 * step sequence frozen: grad_y0 (nPoints x nVars) w.r.t. the initial state and grad_lnk
 * (nPoints x nReactions) w.r.t. the logs of the rate multipliers.  Qbar NULL = ones,
 * ybar_final NULL = zeros.  status as orc_source_integrate.  -6: unsupported options,
 * -8: a rate multiplier function is set, -9: the igniter is on, -10: the mechanism has PLOG,
 * Chebyshev or SRI reactions. */
int orc_qss_adjoint(void* mech, const orc_options* o, int nPoints, const double* state_in,
                    const double* split_const, const double* qloss, double dt, const double* Qbar,
                    const double* ybar_final, double* state_out, double* Q, double* grad_y0,
                    double* grad_lnk, int* status, int nthreads);

/* Central finite differences of J = Qbar*Q + ybar_final . y_final for the same quantities.
 * frozen = 1: the perturbed runs replay the unperturbed step sequence; 0: they rerun the
 * integrator.  State steps are h*|y_i| (0 where y_i = 0, and for U), ln-multiplier steps h. */
int orc_qss_fd_gradient(void* mech, const orc_options* o, int nPoints, const double* state_in,
                        const double* split_const, const double* qloss, double dt, const double* Qbar,
                        const double* ybar_final, double h, int frozen, double* grad_y0,
                        double* grad_lnk, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
