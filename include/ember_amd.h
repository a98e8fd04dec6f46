/*
 * ember_amd: MI355X-native batched split-step kernels for Ember.
 *
 * C ABI (plain pointers and sizes).  Each entry point replaces one per-point or
 * per-variable loop of Ember's FlameSolver with one batched call:
 *
 *   chemistry  FlameSolver::integrateProductionTerms()  src/flameSolver.cpp:754-761
 *              tbb::parallel_for over points             src/flameSolver.cpp:758-759
 *              -> SourceSystem{QSS,CVODE}::setState / integrateToTime / unroll_y
 *                 (src/sourceSystem.h:26-41, src/sourceSystem.cpp:295-338, 434-526)
 *   diffusion  FlameSolver::integrateDiffusionTerms()   src/flameSolver.cpp:838-854
 *              tbb::parallel_for over variables          src/flameSolver.cpp:842-843
 *              -> setDiffusionSolverState (src/flameSolver.cpp:702-718) +
 *                 TridiagonalIntegrator::integrateToTime (src/integrator.cpp:37-47,133-204)
 *                 over DiffusionSystem::get_A / get_k   (src/diffusionSystem.cpp:10-76)
 *
 * Layouts are Ember's: a state is the nVars x nPoints column-major matrix of
 * SplitSolver::state (src/splitSolver.h:73), i.e. the nVars values of one point
 * are contiguous, rows U, T, Y_0..Y_{K-1} (src/readConfig.h:7-9).
 *
 * Return codes: 0 = success, < 0 = API or HIP error (text in emx_last_error()).
 * Per-point integration failures are reported in `status` (0 ok, < 0 failed) and the
 * failed point keeps its input state, as FlameSolver does (src/flameSolver.cpp:794-834).
 *
 * Calls on one context are not re-entrant; use one context per host thread.
 */
#ifndef EMBER_AMD_H
#define EMBER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emx_context emx_context;

enum { EMX_QSS = 0, EMX_BDF = 1 };                 /* chemistryIntegrator "qss" / "cvode" */
enum { EMX_KIN_STANDARD = 0, EMX_KIN_INTERP = 1 };  /* kineticsModel "standard" / "interp" */
enum { EMX_BC_FIXED_VALUE = 0, EMX_BC_ZERO_GRADIENT = 1, EMX_BC_CONTROL_VOLUME = 2,
       EMX_BC_WALL_FLUX = 3 };                      /* BoundaryCondition, src/grid.h:7-25 */

/* Hot-path subset of ConfigOptions (src/readConfig.h:29-217); defaults are
 * python/ember/input.py:371, 686-776. */
typedef struct emx_options {
    int32_t chemistry_integrator;
    int32_t kinetics_model;
    double qss_epsmin, qss_epsmax, qss_dtmin, qss_dtmax;
    int32_t qss_iteration_count;
    double qss_abstol, qss_minval;
    int32_t qss_stability_check;
    double integrator_rel_tol, integrator_momentum_abs_tol, integrator_energy_abs_tol,
        integrator_species_abs_tol, integrator_min_timestep;
    int32_t integrator_max_steps;
    double diffusion_timestep_multiplier;
    double global_timestep;
    /* igniter, ConfigOptions ignition_* (src/readConfig.h:135-139; python/ember/input.py
     * class Ignition): SourceSystem::getQdotIgniter (src/sourceSystem.cpp:33-45) adds
     * energy / (stddev sqrt(2 pi) duration) exp(-(x - center)^2 / (2 stddev^2)) [W/m^3] to
     * the production stages' qDot for absolute times in [t_start, t_start + duration).
     * Off while energy = 0 (the default); on, it needs emx_set_point_positions. */
    double ignition_t_start, ignition_duration, ignition_energy, ignition_center,
        ignition_stddev;
    /* ConfigOptions transportThreshold (python/ember/input.py class Chemistry, threshold):
     * ApproxMixTransport's mole-fraction threshold of the major species. */
    double transport_threshold;
} emx_options;

/* One-dimensional grid for the diffusion systems (OneDimGrid, src/grid.h:120-126). */
typedef struct emx_grid {
    const double* x;      /* nPoints positions (host pointer for host entry points) */
    int32_t alpha;        /* 0 planar, 1 cylindrical */
    int32_t left_bc, right_bc;
    double y_inf;         /* DiffusionSystem::yInf (WallFlux) */
    double wall_const;    /* DiffusionSystem::wallConst (WallFlux) */
    int32_t wall_var;     /* the variable whose DiffusionSystem carries y_inf and wall_const
                           * (Ember: kEnergy = 1, src/flameSolver.cpp:79-82); the others use
                           * 0 and 0, DiffusionSystem's defaults.  -1: none. */
} emx_grid;

/* Fills `o` with Ember's defaults. */
void emx_default_options(emx_options* o);

/* Creates a context on the current HIP device.  `mech_table_path` is a compiled
 * mechanism table (ember_amd.mechanism.table_path); replaces CanteraGas::setOptions +
 * initialize (src/chemistry0d.cpp:445-485).  pressure in Pa. */
int emx_create(const char* mech_table_path, double pressure, const emx_options* opts,
               emx_context** out);
void emx_destroy(emx_context* ctx);
const char* emx_last_error(void);
int emx_num_species(const emx_context* ctx);
int emx_num_reactions(const emx_context* ctx);
int emx_set_options(emx_context* ctx, const emx_options* opts);

/* Per-reaction rate multipliers, Cantera Kinetics::setMultiplier as used by
 * CanteraGas::setRateMultiplier (src/chemistry0d.cpp:634-642).  NULL resets to 1. */
int emx_set_rate_multipliers(emx_context* ctx, const double* multipliers);

/* Ember's rate multiplier function (FlameSolver::rateMultiplierFunction, a
 * ChebyshevFunction whose coefficients the Python layer sets before every step,
 * python/ember/_ember.pyx:533-543): data = [T0, T1, c_0, ..., c_{n-3}] with the meaning of
 * ChebyshevFunction::setCoefficients(n, data) (src/scalarFunction.cpp:33-51), at most 32
 * coefficients.  While set, every right-hand side of the production stages multiplies all
 * reactions by a(t) = sum c_i T_i(2 (t - T0)/(T1 - T0) - 1) at its absolute time t
 * (t_stage_start plus the integrator's time, src/sourceSystem.cpp:133-135, :453-455), in
 * place of the per-reaction multipliers; emx_heat_release_device uses a(t_now)
 * (src/flameSolver.cpp:929-931).  n = 0 or data NULL removes it.  Waits for the context
 * stream before replacing the coefficients (a captured graph keeps the t_stage_start of
 * its capture but reads the current coefficients).  QSS stages with a function run at
 * most 8 waves per workgroup.  The adjoint entry points refuse to run while a function is
 * set, and the plain kernel variant does not support it. */
int emx_set_rate_multiplier_function(emx_context* ctx, int n, const double* data);

/* InterpKinetics table grid (kinetics_model EMX_KIN_INTERP): n equally spaced
 * temperatures over [Tmin, Tmax] for the interpolated forward rate constants and 1/Kc
 * (InterpKinetics / MultiArrheniusInterp, src/chemistry0d.h:66-72, 147-149,
 * src/chemistry0d.cpp:291-331, 392-397, 415-436); Ember's defaults 200 over [273, 3500] K.
 * A right-hand side at a T outside the grid fails its point with status -2.  Ember instead
 * rebuilds its tables (src/chemistry0d.cpp:279-291): emx_source_integrate does the same for
 * its failed points (10 more nodes and 100 K more above; 2 more nodes and a lower bound of
 * max(T - 20, 10) K below; then those points are integrated again from their input states),
 * and the grown grid stays for later calls.  The device entry points report -2 and leave
 * the growth to the caller through this function.  Waits for the context stream. */
int emx_set_interp_range(emx_context* ctx, int n, double Tmin, double Tmax);
int emx_get_interp_range(const emx_context* ctx, int* n, double* Tmin, double* Tmax);

/* Grid positions x of the points of the production stages (SourceSystem::setPosition,
 * src/sourceSystem.cpp:73, called from src/flameSolver.cpp:501, :539), for the igniter.
 * Host array of nPoints; NULL clears.  The stages' nPoints must match while the igniter is
 * on. */
int emx_set_point_positions(emx_context* ctx, size_t nPoints, const double* x);

/* HIP stream (hipStream_t) for the *_device entry points; NULL = default stream. */
int emx_set_stream(emx_context* ctx, void* stream);

/* Launch shape for the chemistry kernels: waves per workgroup and workgroups per CU
 * (0 = built-in default: 4 waves per workgroup, which the BDF lowers to the most that fit in
 * a CU's 160 KB of LDS; a shape set here that does not fit is an error).  The autotuner in
 * ember_amd.autotune picks these. */
int emx_set_launch(emx_context* ctx, int waves_per_block, int blocks_per_cu);

/* Chemistry kernel variant for the *_integrate entry points.  EMX_VARIANT_WAVE (default):
 * one wavefront per point, persistent work queue.  EMX_VARIANT_PLAIN: one thread per point
 * running the serial QssIntegrator loop (QSS only).  EMX_VARIANT_WAVE_PER_POINT (QSS only):
 * the wavefront kernel launched with one wave per point and no work queue.  The last two
 * are comparison baselines for benchmarks; none has a counterpart in the reference. */
enum { EMX_VARIANT_WAVE = 0, EMX_VARIANT_PLAIN = 1, EMX_VARIANT_WAVE_PER_POINT = 2 };
int emx_set_kernel_variant(emx_context* ctx, int variant);

/* Work-queue order of the chemistry kernels.  EMX_SCHED_COST (default): when the
 * previous stage ran over the same number of points, the points are taken in
 * descending order of their measured integration times in that stage (most expensive
 * first); otherwise, and with EMX_SCHED_FIFO, in index order.  Results do not depend on
 * the order.  No counterpart in the reference: TBB's dynamic grain-1 scheduling over
 * points (src/flameSolver.cpp:758-759) is the analogue. */
enum { EMX_SCHED_FIFO = 0, EMX_SCHED_COST = 1 };
int emx_set_schedule(emx_context* ctx, int mode);

/* Chemistry production stage over [t_stage_start, t_stage_start + dt] for all points:
 * FlameSolver::integrateProductionTerms (src/flameSolver.cpp:754-836) with
 * SourceSystemQSS or SourceSystemCVODE per point.
 *   state_in / state_out   nVars x nPoints (nVars = nSpecies + 2), host memory
 *   split_const            nVars x nPoints, SplitSolver::splitConstProd (NULL = zeros)
 *   qloss                  nPoints heat loss [W/m^3] evaluated at setState (NULL = zeros)
 *   status                 nPoints, 0 ok / <0 failed (may be NULL): -1 QSS step below
 *                          qss_dtmin, -2 T outside the interpolation tables (after the
 *                          growth described at emx_set_interp_range), -3 step size below
 *                          the minimum / step budget, -4 non-finite QSS error or BDF
 *                          integrator_max_steps, -5 singular BDF Newton matrix, -6
 *                          non-finite BDF error estimate, -7 / -8 too many BDF error-test
 *                          / convergence failures in one step, -9 invalid input state
 *                          (a NaN or infinite value, or T <= 0)
 *   stat1, stat2           nPoints: QSS gcount/rcount, BDF steps/RHS evaluations (NULL ok)
 * Synchronous. */
int emx_source_integrate(emx_context* ctx, size_t nPoints, size_t nVars, const double* state_in,
                         const double* split_const, const double* qloss, double t_stage_start,
                         double dt, double* state_out, int32_t* status, int32_t* stat1,
                         int32_t* stat2);

/* Same with device pointers; asynchronous on the context stream.  qint (nullable): the
 * heat release integrated over the stage on accepted internal steps [J/m^3]. */
int emx_source_integrate_device(emx_context* ctx, size_t nPoints, size_t nVars,
                                const double* state_in, const double* split_const,
                                const double* qloss, double t_stage_start, double dt,
                                double* state_out, int32_t* status, int32_t* stat1,
                                int32_t* stat2, double* qint);

/* The production stage of emx_source_integrate_device without any device allocation.
 * No counterpart in the reference: for callers that cannot allocate device memory during
 * a stage (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as emx_source_integrate_device (bit for bit),
 * QSS integrator only; the work-queue counter, per-point costs and queue order live in a
 * fixed library scratch of 2 MiB + 256 B per GPU (held until the process ends), so one
 * call takes at most 262144 points (-1 beyond, and for the BDF or the plain variant).
 * With the igniter on, one emx_source_integrate_device stage over the same points must
 * come first.  Every call on one GPU shares that GPU's scratch: calls that overlap in
 * time on one GPU (from two host threads, or on two streams) race; the costs left by the
 * last call order the next one over the same number of points, which changes only the
 * order in which points are integrated, never their results. */
int emx_source_integrate_device_noalloc(emx_context* ctx, size_t nPoints, size_t nVars,
                                        const double* state_in, const double* split_const,
                                        const double* qloss, double t_stage_start, double dt,
                                        double* state_out, int32_t* status, int32_t* stat1,
                                        int32_t* stat2, double* qint);

/* Discrete adjoint of the QSS production stage (QSS integrator with
 * qss_iteration_count = 1 and no stability check).  No counterpart in the reference: it
 * differentiates the stage that SourceSystemQSS::integrateToTime runs per point
 * (src/sourceSystem.cpp:434-443, src/qssintegrator.cpp:98-265).
 *
 * emx_source_adjoint_forward: the stage as emx_source_integrate_device (device pointers,
 * asynchronous on the context stream), plus heat_release (nPoints, nullable) = the heat
 * release rate integrated over the accepted internal steps [J/m^3], and checkpoints of the
 * integrator state every 16 accepted steps kept in the context (status -5 for a point
 * that needs more checkpoints than emx_set_adjoint_capacity allows).
 *
 * emx_source_adjoint_backward: for the points of the preceding adjoint forward call
 * (same state_in, split_const, qloss, dt), the gradients of
 *     J_j = Qbar_j * heat_release_j + ybar_final_j . state_out_j
 * with the internal step sequence held fixed: grad_state (nPoints x nVars) w.r.t. the
 * input state and grad_lnk (nPoints x nReactions, nullable) w.r.t. the natural logs of the
 * reaction rate multipliers (emx_set_rate_multipliers), per point.  Qbar NULL = ones,
 * ybar_final NULL = zeros.  status: 0 ok, < 0 failed point (its gradients are zero). */
int emx_set_adjoint_capacity(emx_context* ctx, int max_checkpoints);
int emx_source_adjoint_forward(emx_context* ctx, size_t nPoints, size_t nVars,
                               const double* state_in, const double* split_const,
                               const double* qloss, double dt, double* state_out,
                               int32_t* status, double* heat_release);
int emx_source_adjoint_backward(emx_context* ctx, size_t nPoints, size_t nVars,
                                const double* state_in, const double* split_const,
                                const double* qloss, double dt, const double* Qbar,
                                const double* ybar_final, double* grad_state, double* grad_lnk,
                                int32_t* status);

/* Creation, destruction and net molar production rates [kmol/m^3/s] at n states
 * (CanteraGas::getCreationRates / getDestructionRates / getReactionRates,
 * src/chemistry0d.cpp:644-672).  T: n, Y: n x nSpecies, outputs n x nSpecies.
 * nonorm = 0: setState_TPY (QSS path); 1: setMassFractions_NoNorm (CVODE path).
 * status (nullable): 0 ok, -2 if T is outside the interpolation table (interp only). */
int emx_rates(emx_context* ctx, size_t n, const double* T, const double* Y, int nonorm,
              double* wdot_creation, double* wdot_destruction, double* wdot_net,
              int32_t* status);

/* Diffusion stage over [t0, t1] for nVars variables on one grid:
 * FlameSolver::setDiffusionSolverState + integrateDiffusionTerms.
 *   B         nB x nPoints, row b contiguous; B_row[v] selects the row used by variable v
 *             (NULL: nB == nVars, row v).  DiffusionSystem::B (1/rho, 1/(rho cp), ...)
 *   D         nVars x nPoints (Ember layout): mu, lambda, rhoD_k
 *   split_const  nVars x nPoints, splitConstDiff (NULL = zeros)
 *   h_per_var nVars initial step sizes (NULL: Ember's rule
 *             min(diffusion_timestep_multiplier * min_j dlj^2/(B D), global_timestep))
 *   state     nVars x nPoints, in/out.
 * Synchronous (host pointers). */
int emx_diffusion_integrate(emx_context* ctx, size_t nVars, size_t nPoints, const double* B,
                            size_t nB, const int32_t* B_row, const double* D,
                            const double* split_const, const emx_grid* grid,
                            const double* h_per_var, double t0, double t1, double* state);

/* Same with device pointers (grid->x and B_row also device; h_per_var stays a host array,
 * as the h of TridiagonalIntegrator::initialize, src/integrator.cpp:16-20, 111-116).  Asynchronous
 * on the context stream when h_per_var is given; with NULL the call reads the grid minimum
 * back to the host once (one stream synchronisation) to apply Ember's step rule. */
int emx_diffusion_integrate_device(emx_context* ctx, size_t nVars, size_t nPoints,
                                   const double* B, size_t nB, const int32_t* B_row,
                                   const double* D, const double* split_const,
                                   const emx_grid* grid, const double* h_per_var, double t0,
                                   double t1, double* state);

/* Device-resident split step: SplitSolver::step (src/splitSolver.cpp:34-58), i.e. the
 * split constants (:38-44), integrateSplitTerms' stage sequence D1/4 C1/2 D1/4 P D1/4 C1/2
 * D1/4 with its delta bookkeeping (:60-121) and calculateTimeDerivatives (:27-32), with
 * FlameSolver's production and diffusion stages (flameSolver.cpp:702-734, 754-854) run
 * on the GPU and the state resident in device memory throughout.  The convection stages
 * are empty here (emx_split_step_conv_device runs them).  All matrices are nVars x nPoints,
 * point-contiguous, device pointers owned by the caller:
 *   state             in/out
 *   ddt_conv/diff/prod  in: the previous step's time derivatives (zeros initially);
 *                     out: this step's
 *   ddt_cross         nullable (zeros): the cross-diffusion derivative
 *   delta_diff/prod   out: the step's diffusion and production increments
 *   work              scratch of 3 x nVars x nPoints doubles
 * B, B_row, D, grid as emx_diffusion_integrate_device (device pointers, fixed over the
 * step); qloss (nPoints, nullable) as emx_source_integrate_device; status (nPoints,
 * device) receives the production stage's per-point status.  balanced: 1 = Ember's
 * balanced splitting, 0 = Strang with splitConstDiff = ddt_cross.  The diffusion step
 * sizes follow Ember's rule (one read of the grid minimum per step; the four diffusion
 * stages share it since B and D are fixed) unless h_per_var (host, nVars) is given, in
 * which case the call is fully asynchronous on the context stream. */
typedef struct emx_split_arrays {
    double* state;
    double* ddt_conv;
    double* ddt_diff;
    double* ddt_prod;
    const double* ddt_cross;
    double* delta_diff;
    double* delta_prod;
    double* work;
} emx_split_arrays;

int emx_split_step_device(emx_context* ctx, size_t nVars, size_t nPoints,
                          const emx_split_arrays* arrays, const double* B, size_t nB,
                          const int32_t* B_row, const double* D, const emx_grid* grid,
                          const double* qloss, const double* h_per_var, double t, double dt,
                          int balanced, int32_t* status);

/* Convection stage: ConvectionSystemSplit::integrateToTime (src/convectionSystem.cpp:588-627)
 * as FlameSolver::integrateConvectionTerms runs it (src/flameSolver.cpp:720-750): the
 * (U, T, Wmx) system of all points (ConvectionSystemUTW, continuity condition fixedLeft),
 * then every species' advection by the recorded v = V / rho (ConvectionSystemY), with
 * Ember's tolerances (integrator_rel_tol, momentum / energy abs tol, 20 x species abs tol
 * for Wmx, species abs tol).  CVODE's Adams method is replaced by the explicit
 * Dormand-Prince 5(4) pair with error control (no counterpart in the reference; the step
 * sequence differs from CVODE's, the tolerances are Ember's).  3 to 4096 grid points. */
typedef struct emx_convection_params {
    int32_t alpha;          /* OneDimGrid::alpha: 0 planar, 1 cylindrical */
    int32_t beta;           /* OneDimGrid::beta: 1, or 2 for disc (axisymmetric jet) flames */
    int32_t left_bc, right_bc;  /* EMX_BC_* of grid.leftBC / rightBC */
    int32_t continuity_bc;  /* 0: fixedLeft (ContinuityBoundaryCondition::Left); others refused */
    double rVzero;          /* ConvectionSystemUTW::rVzero */
    double Tleft;           /* setLeftBC: inflow temperature (ControlVolume / WallFlux left end) */
    const double* Yleft;    /* setLeftBC: nSpecies inflow mass fractions, host (NULL: zeros) */
    double rhou;            /* FlameSolver::rhou, unburned density of the strain term */
    double strain_initial, strain_final, strain_t0, strain_dt;  /* LinearFunction strain rate */
    int32_t max_steps;      /* per system (0: CVODE's 1000000 of ConvectionSystemSplit::setState) */
} emx_convection_params;

/* Device pointers: x (nPoints), state_in / state_out (nVars x nPoints; may be the same
 * array), split_const (splitConstConv, nullable = zeros), drhodt (nPoints, nullable =
 * zeros; FlameSolver::drhodt).  steps (device, 1 + nSpecies, nullable): accepted steps of
 * the UTW system and of each species system, < 0 for a failed system (its variables are
 * not written).  Asynchronous on the context stream. */
int emx_convection_integrate_device(emx_context* ctx, size_t nVars, size_t nPoints,
                                    const emx_convection_params* params, const double* x,
                                    const double* state_in, const double* split_const,
                                    const double* drhodt, double t0, double t1, double* state_out,
                                    int32_t* steps);

/* The same on host arrays, state in/out; synchronous; returns -3 when a system failed
 * (steps, host, nullable, as above). */
int emx_convection_integrate(emx_context* ctx, size_t nVars, size_t nPoints,
                             const emx_convection_params* params, const double* x,
                             const double* split_const, const double* drhodt, double t0, double t1,
                             double* state, int32_t* steps);

/* The device-resident split step with its convection stages: emx_split_step_device plus
 * FlameSolver::integrateConvectionTerms (src/flameSolver.cpp:736-751) in the two stages
 * C1/2 of integrateSplitTerms (src/splitSolver.cpp:68, :72) through
 * emx_convection_integrate_device, the balanced splitting's convection constants
 * splitConstConv (:40; zeros for Strang), deltaConv and ddtConv (:29), and the density
 * derivative of FlameSolver::prepareIntegrators (src/flameSolver.cpp:203-213) as
 * emx_density_derivative_device forms it.  arrays as emx_split_step_device, with work of
 * 4 x nVars x nPoints + nPoints doubles.  conv:
 *   params      the convection stage's parameters (Yleft is uploaded when it changes, so a
 *               step captured in a graph after one call records no host copy)
 *   rho, Wmx    nPoints, device: the density and mean molecular weight at the step's start
 *               (updateChemicalProperties); both NULL: drhodt = 0
 *   delta_conv  nVars x nPoints, device, out: the step's convection increment
 *   steps       device, 2 x (1 + nSpecies), nullable: each convection stage's accepted
 *               steps as emx_convection_integrate_device reports them
 * 3 to 4096 grid points (the convection stage's range). */
typedef struct emx_split_convection {
    const emx_convection_params* params;
    const double* rho;
    const double* Wmx;
    double* delta_conv;
    int32_t* steps;
} emx_split_convection;

int emx_split_step_conv_device(emx_context* ctx, size_t nVars, size_t nPoints,
                               const emx_split_arrays* arrays, const emx_split_convection* conv,
                               const double* B, size_t nB, const int32_t* B_row, const double* D,
                               const emx_grid* grid, const double* qloss, const double* h_per_var,
                               double t, double dt, int balanced, int32_t* status);

/* FlameSolver::prepareIntegrators' density derivative (src/flameSolver.cpp:203-213) for the
 * convection stage: with ddt = ddt_conv + ddt_diff + ddt_prod (+ ddt_cross when balanced,
 * nullable = zeros), drhodt = -rho (ddt_T / T + Wmx sum_k ddt_Yk / W_k) at each point of the
 * device state matrix.  Device pointers: the ddt matrices nVars x nPoints, rho, Wmx and
 * drhodt nPoints.  Asynchronous on the context stream. */
int emx_density_derivative_device(emx_context* ctx, size_t nVars, size_t nPoints,
                                  const double* state, const double* ddt_conv,
                                  const double* ddt_diff, const double* ddt_prod,
                                  const double* ddt_cross, int balanced, const double* rho,
                                  const double* Wmx, double* drhodt);

/* FlameSolver::correctMassFractions (src/flameSolver.cpp:942-949): per point, negative
 * mass fractions clipped to zero and the rest scaled to sum to one (Cantera
 * Phase::setMassFractions).  state: nVars x nPoints device matrix, in place, async. */
int emx_correct_mass_fractions_device(emx_context* ctx, size_t nVars, size_t nPoints,
                                      double* state);

/* FlameSolver::calculateQdot (src/flameSolver.cpp:926-939): qdot[j] = -sum_k wdot_k h_k
 * [W/m^3] at each point's (T, Y) of the device state matrix; qdot (nPoints) device.
 * t_now: the solver time, for the rate multiplier function (ignored without one).
 * status (nPoints, device, nullable): 0, or -2 if T is outside the interpolation table. */
int emx_heat_release_device(emx_context* ctx, size_t nVars, size_t nPoints, const double* state,
                            double t_now, double* qdot, int32_t* status);

/* FlameSolver::updateChemicalProperties, thermodynamic block (src/flameSolver.cpp:667-673):
 * CanteraGas::setStateMass(Y_j, T_j) (chemistry0d.cpp:492-495), then getDensity,
 * getMixtureMolecularWeight, getSpecificHeatCapacity, getSpecificHeatCapacities and
 * getEnthalpies (chemistry0d.cpp:527-535, 608-631) for every point of the device state
 * matrix.  Device pointers: rho [kg/m^3], Wmx [kg/kmol], cp [J/kg/K] (nPoints each);
 * cpSpec [J/kmol/K] and hk [J/kmol] (nPoints x nSpecies, point-major as Ember's
 * column-major cpSpec/hk matrices).  Wmx, cpSpec and hk may be NULL. */
int emx_thermo_properties_device(emx_context* ctx, size_t nVars, size_t nPoints,
                                 const double* state, double* rho, double* Wmx, double* cp,
                                 double* cpSpec, double* hk);

/* FlameSolver::updateChemicalProperties, transport block (src/flameSolver.cpp:678-692)
 * with Ember's default transport model "Approx" (ApproxMixTransport,
 * src/chemistry0d.cpp:12-261; ConfigOptions transportThreshold = transport_threshold):
 * setStateMass(Y_j, T_j), then getViscosity, getThermalConductivity and
 * getWeightedDiffusionCoefficientsMass (chemistry0d.cpp:549-596) for every point of the
 * device state matrix.  Device pointers: mu [Pa s], lambda [W/m/K] (nPoints each), rhoD
 * [kg/m/s] (nPoints x nSpecies, point-major as Ember's rhoD matrix).  The pure-species
 * and binary properties come from the mechanism table's temperature fits (the species'
 * `transport` blocks; -1 without them).  Soret coefficients (soretDiffusion) are not
 * computed. */
int emx_transport_properties_device(emx_context* ctx, size_t nVars, size_t nPoints,
                                    const double* state, double* mu, double* lambda, double* rhoD);

/* The same on host arrays (synchronous). */
int emx_transport_properties(emx_context* ctx, size_t nVars, size_t nPoints, const double* state,
                             double* mu, double* lambda, double* rhoD);

/* The diffusion coefficients of the split step from the device state: the property update
 * (updateChemicalProperties, thermodynamic and transport blocks) followed by
 * FlameSolver::prepareIntegrators' diffusion inputs (src/flameSolver.cpp:160-172):
 *   B  2 x nPoints, row 0 = 1/rho (momentum and species), row 1 = 1/(rho cp) (energy),
 *      for B_row = {0, 1, 0, 0, ...};
 *   D  nVars x nPoints (point-major): mu, lambda, rho D_0 .. rho D_{K-1}.
 * Device pointers, asynchronous on the context stream; the inputs of
 * emx_split_step_device / emx_diffusion_integrate_device. */
int emx_diffusion_coefficients_device(emx_context* ctx, size_t nVars, size_t nPoints,
                                      const double* state, double* B, double* D);

/* The same on host arrays, the drop-in for the TBB loop over points (copied through the
 * context's staging buffers; returns after the results are on the host). */
int emx_thermo_properties(emx_context* ctx, size_t nVars, size_t nPoints, const double* state,
                          double* rho, double* Wmx, double* cp, double* cpSpec, double* hk);

/* TridiagonalIntegrator::integrateToTime for nSys constant systems y' = A y + k
 * (test/tridiagonalIntegrator.cpp ConstantTridiagonalODE).  a, b, c, k, y: nSys x N,
 * row-major (system s contiguous).  h: initial step per system.  Host pointers. */
int emx_tridiagonal_integrate(emx_context* ctx, size_t nSys, size_t N, const double* a,
                              const double* b, const double* c, const double* k,
                              const double* h, double t0, double t1, double* y);

#ifdef __cplusplus
}
#endif
#endif /* EMBER_AMD_H */
