// Kernel launchers (device code in *.hip), called by the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include "emx_common.h"

namespace emx {

int launch_qss_stage(const DevMech& m, const DevOpts& o, const double* state_in,
                     const double* split_const, const double* qloss, const double* perturb,
                     double dt, int nPoints, const int* order, unsigned int* queue,
                     double* state_out, int* status, int* stat1, int* stat2, int* cost, double* qint,
                     int grid_blocks, int waves_per_block, hipStream_t stream);

int launch_rates(const DevMech& m, const double* T, const double* Y, int nonorm,
                 const double* perturb, int n, double* wQ, double* wD, double* wnet, int* status,
                 hipStream_t stream);

// device-resident split step (split_kernels.hip)
int launch_split_constants(size_t n, int balanced, double* ddt_conv, const double* ddt_diff,
                           const double* ddt_prod, const double* ddt_cross, double* sc_diff,
                           double* sc_prod, hipStream_t st);
int launch_split_constants_conv(size_t n, int balanced, const double* ddt_conv, const double* ddt_diff,
                                const double* ddt_prod, const double* ddt_cross, double* sc_diff,
                                double* sc_prod, double* sc_conv, hipStream_t st);
int launch_density_derivative(const DevMech& m, int N, int balanced, const double* state,
                              const double* ddt_conv, const double* ddt_diff, const double* ddt_prod,
                              const double* ddt_cross, const double* rho, const double* Wmx,
                              double* drhodt, hipStream_t st);
int launch_conv_time_derivative(size_t n, double dt, const double* delta_conv, const double* sc_conv,
                                double* ddt_conv, hipStream_t st);
int launch_fill(size_t n, double v, double* a, double* b, hipStream_t st);
int launch_copy(size_t n, const double* src, double* dst, hipStream_t st);
int launch_accumulate(size_t n, const double* state, const double* start, double* delta,
                      hipStream_t st);
int launch_time_derivatives(size_t n, double dt, const double* delta_diff, const double* delta_prod,
                            const double* sc_diff, const double* sc_prod, double* ddt_diff,
                            double* ddt_prod, hipStream_t st);
int launch_correct_mass_fractions(int nVars, int nPoints, double* state, hipStream_t st);
int launch_heat_release(const DevMech& m, const DevOpts& o, const double* state, const double* perturb, int n,
                        double* qdot, int* status, hipStream_t stream);
// Convection stage (convection_kernels.hip)
constexpr int kConvMaxPoints = 4096;  // grid points of one convection system (LDS rows)
struct ConvParams {
    int N, K, nv, alpha, beta, leftBC, rightBC;
    double P, rVzero, Tleft, Wleft, rhou;
    double aInitial, aFinal, T0, Dt;  // Ember's LinearFunction strain rate
    double rtol, atolU, atolT, atolW, atolY;
    double t0, t1;
    int maxSteps, maxRec;
};
struct ConvArrays {
    const double* state;     // N x nv input
    const double* sc;        // N x nv split constants (nullable)
    const double* x;         // N
    const double* drhodt;    // N
    const double* Wmx;       // N mean molecular weights of the input state
    const double* scU;       // N
    const double* scT;       // N
    const double* scW;       // N: sum_k splitConst_Yk / W_k
    const double* Yleft;     // K (nullable)
    const double* rW;        // K: 1 / W_k (register path: Wmx and scW formed in the kernel)
    double* state_out;       // N x nv
    double* rec_t;           // maxRec
    double* rec_v;           // maxRec x N
    int* nrec;               // 1
    int* steps;              // 1 + K
    double* work;            // convection_work_doubles
};
size_t convection_work_doubles(int N, int K);
// grids of at most kConvRegMax points take the register path (one point per thread; the
// kernels form Wmx and the UTW split constants themselves: Wmx, scU, scT, scW unused)
constexpr int kConvRegMax = 1024;
int launch_convection(const ConvParams& p, const ConvArrays& a, hipStream_t stream);
// split constants of the UTW system from the convection split constants (N x nv):
// scU, scT rows and scW = sum_k sc_Yk / W_k
int launch_conv_split_constants(const DevMech& m, int N, const double* sc, double* scU, double* scT,
                                double* scW, hipStream_t stream);

// ApproxMixTransport properties (mu, lambda, rho D_k) of n points; needs m.tr_* fits.
// mu and lambda of point p at [p * ml_stride], rho D_k at rhoD[p * rd_stride + k].
int launch_transport_props(const DevMech& m, const double* state, int n, double threshold, double* mu,
                           double* lambda, int ml_stride, double* rhoD, int rd_stride, hipStream_t stream);
// B rows of the diffusion systems from rho (row 0) and cp (row 1) in place:
// 1 / rho and 1 / (rho cp) (FlameSolver::prepareIntegrators)
int launch_diffusion_B(int n, double* B, hipStream_t stream);

int launch_thermo_props(const DevMech& m, const double* state, int n, double* rho, double* Wmx,
                        double* cp, double* cpSpec, double* hk, hipStream_t stream);

}  // namespace emx

namespace emx {

struct DiffArgs {
    const double* B;
    const int32_t* B_row;
    const double* D;
    const double* sc;
    const double* x;
    int N, nV, alpha, left, right;
    double yInf, wallConst;
    int wallVar;
};

struct DiffWork {
    double *own1, *nxt1, *X1, *own2, *nxt2, *X2, *cp, *dp, *cmin, *yprev, *g1, *g2, *edge, *coef;
};

struct DiffSteps {
    double h[kMaxVars];
    int ns[kMaxVars];
};
int launch_set_steps(const DiffSteps& s, int nV, double* h, int* ns, hipStream_t st);
size_t diffusion_work_doubles(int nV, int N);
void diffusion_carve(double* base, int nV, int N, DiffWork& w);
int launch_diffusion_hmin(const DiffArgs& a, DiffWork& w, double* hmin, hipStream_t st);
int launch_diffusion_solve(const DiffArgs& a, double* y, const double* h, const int* nsteps, int q,
                           DiffWork& w, hipStream_t st);
// all nsolve solves of a stage (fused path when the grid allows it)
int launch_diffusion_stage(const DiffArgs& a, double* y, const DiffSteps& steps, double* h, int* nsteps,
                           int nsolve, DiffWork& w, hipStream_t st);
int launch_explicit_solve(const double* A, const double* B, const double* C, const double* K,
                          int nSys, int N, double* y, const double* h, const int* nsteps, int q,
                          DiffWork& w, hipStream_t st);

}  // namespace emx

namespace emx {

int launch_bdf_stage(const DevMech& m, const DevOpts& o, const double* state_in,
                     const double* split_const, const double* qloss, const double* perturb,
                     double dt, int nPoints, const int* order, unsigned int* queue,
                     double* state_out, int* status, int* stat1, int* stat2, int* cost, double* qint,
                     int grid_blocks, int waves_per_block, double* jac_scratch, hipStream_t stream);
size_t bdf_lds_bytes(const DevMech& m, int waves_per_block);
size_t qss_lds_bytes(const DevMech& m, int waves_per_block);
// discrete adjoint of the QSS stage
size_t adjoint_record_doubles_per_wave();
int adjoint_checkpoint_interval();
size_t adjoint_lds_bytes(const DevMech& m, int waves_per_block, bool backward);
int launch_qss_adjoint_forward(const DevMech& m, const DevOpts& o, const double* state_in,
                               const double* split_const, const double* qloss,
                               const double* perturb, double dt, int nPoints, const int* order,
                               unsigned int* queue, double* state_out, int* status, double* Q,
                               int* cost, double* ck_y, double* ck_s, int* nckpt, int cap,
                               int grid_blocks, int waves_per_block, hipStream_t stream);
int launch_qss_adjoint_backward(const DevMech& m, const DevOpts& o, const double* state_in,
                                const double* split_const, const double* qloss,
                                const double* perturb, double dt, int nPoints, const int* order,
                                unsigned int* queue, const double* ck_y, const double* ck_s,
                                const int* nckpt, int cap, const double* Qbar, const double* ybar,
                                double* record, double* grad_state, double* grad_lnk, int* status,
                                int grid_blocks, int waves_per_block, hipStream_t stream);
// order[0..n) = point indices by descending cost (single workgroup)
int launch_cost_order(const int* cost, int n, int* order, hipStream_t stream);

}  // namespace emx

namespace emx {

size_t plain_workspace_doubles(const DevMech& m, int nPoints);
int launch_qss_plain(const DevMech& m, const DevOpts& o, const double* state_in,
                     const double* split_const, const double* qloss, const double* perturb,
                     double dt, int nPoints, double* state_out, int* status, int* stat1, int* stat2,
                     double* ws, hipStream_t stream);

}  // namespace emx
