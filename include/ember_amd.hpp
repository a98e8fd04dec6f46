// C++ host layer over the ember_amd C ABI, shaped like Ember's split driver.
//
// ember_amd::SplitSolver mirrors SplitSolver (src/splitSolver.h, src/splitSolver.cpp):
// the balanced/Strang stage sequence of integrateSplitTerms (splitSolver.cpp:60-75),
// the per-stage delta bookkeeping (:77-121), the balanced split constants
// (:38-44) and calculateTimeDerivatives (:27-32).  ember_amd::FlameStages fills the
// production and diffusion hooks the way FlameSolver does
// (setProductionSolverState + integrateProductionTerms, flameSolver.cpp:727-836;
// setDiffusionSolverState + integrateDiffusionTerms, :702-718, :838-854), each as one
// batched call into the GPU library instead of a tbb::parallel_for.
// Matrices are Ember's column-major nVars x nPoints (point-contiguous) vectors.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "ember_amd.h"

namespace ember_amd {

using Matrix = std::vector<double>;  // nVars x nPoints, column-major

inline void check(int rc) {
    if (rc != 0) throw std::runtime_error(std::string("ember_amd: ") + emx_last_error());
}

// RAII owner of one emx_context (the analogue of a per-thread CanteraGas + the
// per-point SourceSystem / per-variable DiffusionSystem objects).
class Context {
public:
    Context(const std::string& mech_table, double pressure, const emx_options& opts) {
        check(emx_create(mech_table.c_str(), pressure, &opts, &ctx_));
    }
    ~Context() { emx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    emx_context* get() const { return ctx_; }
    int nSpec() const { return emx_num_species(ctx_); }

private:
    emx_context* ctx_ = nullptr;
};

// SplitSolver: stage sequence and splitting bookkeeping.
class SplitSolver {
public:
    virtual ~SplitSolver() {}

    void resize(size_t nRows, size_t nCols) {
        nVars = nRows;
        nPoints = nCols;
        size_t n = nRows * nCols;
        state.assign(n, 0.0);
        startState.assign(n, 0.0);
        for (Matrix* m : {&deltaConv, &deltaDiff, &deltaProd, &ddtConv, &ddtDiff, &ddtProd,
                          &ddtCross, &splitConstConv, &splitConstDiff, &splitConstProd})
            m->assign(n, 0.0);
    }

    // SplitSolver::step (splitSolver.cpp:34-58) without output, regrid and finishStep
    void step(double t, double dt_) {
        dt = dt_;
        if (useBalancedSplitting) {
            for (size_t i = 0; i < state.size(); i++) {
                splitConstDiff[i] = 0.25 * (ddtProd[i] + ddtConv[i] + ddtCross[i] - 3 * ddtDiff[i]);
                splitConstConv[i] = 0.25 * (ddtProd[i] + ddtDiff[i] + ddtCross[i] - 3 * ddtConv[i]);
                splitConstProd[i] = 0.5 * (ddtConv[i] + ddtDiff[i] + ddtCross[i] - ddtProd[i]);
            }
        } else {
            splitConstDiff = ddtCross;
        }
        integrateSplitTerms(t, dt);
        calculateTimeDerivatives(dt);
    }

    // SplitSolver::integrateSplitTerms (splitSolver.cpp:60-75)
    void integrateSplitTerms(double t, double dt_) {
        std::fill(deltaConv.begin(), deltaConv.end(), 0.0);
        std::fill(deltaProd.begin(), deltaProd.end(), 0.0);
        std::fill(deltaDiff.begin(), deltaDiff.end(), 0.0);
        stage(deltaDiff, t, t + 0.25 * dt_, &SplitSolver::integrateDiffusionTerms);
        stage(deltaConv, t, t + 0.5 * dt_, &SplitSolver::integrateConvectionTerms);
        stage(deltaDiff, t + 0.25 * dt_, t + 0.5 * dt_, &SplitSolver::integrateDiffusionTerms);
        stage(deltaProd, t, t + dt_, &SplitSolver::integrateProductionTerms);
        stage(deltaDiff, t + 0.5 * dt_, t + 0.75 * dt_, &SplitSolver::integrateDiffusionTerms);
        stage(deltaConv, t + 0.5 * dt_, t + dt_, &SplitSolver::integrateConvectionTerms);
        stage(deltaDiff, t + 0.75 * dt_, t + dt_, &SplitSolver::integrateDiffusionTerms);
    }

    // SplitSolver::calculateTimeDerivatives (splitSolver.cpp:27-32)
    void calculateTimeDerivatives(double dt_) {
        for (size_t i = 0; i < state.size(); i++) {
            ddtConv[i] = deltaConv[i] / dt_ - splitConstConv[i];
            ddtDiff[i] = deltaDiff[i] / dt_ - splitConstDiff[i];
            ddtProd[i] = deltaProd[i] / dt_ - splitConstProd[i];
        }
    }

    Matrix state, startState, deltaConv, deltaDiff, deltaProd, ddtConv, ddtDiff, ddtProd, ddtCross,
        splitConstConv, splitConstDiff, splitConstProd;
    size_t nVars = 0, nPoints = 0;
    double tStageStart = 0, tStageEnd = 0, dt = 0;
    bool useBalancedSplitting = true;

protected:
    virtual void integrateConvectionTerms() {}
    virtual void integrateProductionTerms() = 0;
    virtual void integrateDiffusionTerms() = 0;

private:
    void stage(Matrix& delta, double t0, double t1, void (SplitSolver::*fn)()) {
        tStageStart = t0;
        tStageEnd = t1;
        startState = state;
        (this->*fn)();
        for (size_t i = 0; i < state.size(); i++) delta[i] += state[i] - startState[i];
    }
};

// Ember's BoundaryCondition::BC (src/grid.h:7-25: FixedValue 0, ZeroGradient 1,
// ControlVolume 2, WallFlux 3, Floating 4) as the diffusion systems use it:
// DiffusionSystem::get_A treats every left condition other than the first four, and every
// right condition other than FixedValue, as ZeroGradient (src/diffusionSystem.cpp:24-58).
inline int diffusion_left_bc(int emberBC) {
    return (emberBC >= EMX_BC_FIXED_VALUE && emberBC <= EMX_BC_WALL_FLUX) ? emberBC : EMX_BC_ZERO_GRADIENT;
}
inline int diffusion_right_bc(int emberBC) {
    return emberBC == EMX_BC_FIXED_VALUE ? EMX_BC_FIXED_VALUE : EMX_BC_ZERO_GRADIENT;
}

// Production, diffusion and (once set up) convection stages on the GPU (FlameSolver's
// parallel_for loops and its convection system).
class FlameStages : public SplitSolver {
public:
    FlameStages(Context& ctx, const std::vector<double>& x, int alpha = 0,
                int leftBC = EMX_BC_FIXED_VALUE, int rightBC = EMX_BC_FIXED_VALUE)
        : ctx_(ctx), x_(x) {
        grid_.x = x_.data();
        grid_.alpha = alpha;
        grid_.left_bc = diffusion_left_bc(leftBC);
        grid_.right_bc = diffusion_right_bc(rightBC);
        // DiffusionSystem's defaults (yInf = wallConst = 0) until setWallFlux
        grid_.y_inf = 0.0;
        grid_.wall_const = 0.0;
        grid_.wall_var = 1;  // kEnergy
        // SourceSystem::setPosition for every point (flameSolver.cpp:501, :539): the
        // igniter's positions
        check(emx_set_point_positions(ctx_.get(), x_.size(), x_.data()));
    }

    // OneDimGrid::leftBC / rightBC as FlameSolver::updateBC sets them (Ember's enum values).
    void setBoundaryConditions(int leftBC, int rightBC) {
        grid_.left_bc = diffusion_left_bc(leftBC);
        grid_.right_bc = diffusion_right_bc(rightBC);
    }

    // ConfigOptions wallFlux / Tinf / Kwall (src/readConfig.h:130-132): FlameSolver gives the
    // energy equation's DiffusionSystem yInf = Tinf and wallConst = Kwall when wallFlux is
    // set (src/flameSolver.cpp:79-82); the other variables keep 0 and 0.
    void setWallFlux(bool wallFlux, double Tinf, double Kwall) {
        grid_.y_inf = wallFlux ? Tinf : 0.0;
        grid_.wall_const = wallFlux ? Kwall : 0.0;
        grid_.wall_var = 1;  // kEnergy
    }
    const emx_grid& grid() const { return grid_; }

    // The convection stages (FlameSolver::integrateConvectionTerms, flameSolver.cpp:736-751):
    // ConvectionSystemSplit's parameters (rVzero, the inflow state, the strain function,
    // the grid's boundary conditions) and drhodt (FlameSolver::drhodt, nPoints; empty:
    // zeros).  Until setConvection the convection stages change nothing.
    void setConvection(const emx_convection_params& p) {
        conv_ = p;
        convOn_ = true;
    }
    std::vector<double> drhodt;

    // Diffusion coefficients as FlameSolver::prepareIntegrators sets them
    // (flameSolver.cpp:160-172): B rows (1/rho, 1/(rho cp)) with B_row per variable,
    // D nVars x nPoints (mu, lambda, rhoD_k).
    Matrix B, D;
    std::vector<int32_t> B_row;
    std::vector<double> qloss;     // heat loss per point (optional)
    std::vector<int32_t> status;   // per-point status of the last production stage
    std::vector<int32_t> stat1, stat2;

protected:
    void integrateConvectionTerms() override {
        if (!convOn_) return;
        check(emx_convection_integrate(ctx_.get(), nVars, nPoints, &conv_, x_.data(), splitConstConv.data(),
                                       drhodt.empty() ? nullptr : drhodt.data(), tStageStart, tStageEnd,
                                       state.data(), nullptr));
    }
    void integrateProductionTerms() override {
        status.assign(nPoints, 0);
        stat1.assign(nPoints, 0);
        stat2.assign(nPoints, 0);
        Matrix out(state.size());
        check(emx_source_integrate(ctx_.get(), nPoints, nVars, state.data(), splitConstProd.data(),
                                   qloss.empty() ? nullptr : qloss.data(), tStageStart,
                                   tStageEnd - tStageStart, out.data(), status.data(),
                                   stat1.data(), stat2.data()));
        state.swap(out);  // failed points already carry their old state
    }
    void integrateDiffusionTerms() override {
        check(emx_diffusion_integrate(ctx_.get(), nVars, nPoints, B.data(), B.size() / nPoints,
                                      B_row.empty() ? nullptr : B_row.data(), D.data(),
                                      splitConstDiff.data(), &grid_, nullptr, tStageStart, tStageEnd,
                                      state.data()));
    }

private:
    Context& ctx_;
    std::vector<double> x_;
    emx_grid grid_;
    emx_convection_params conv_{};
    bool convOn_ = false;
};

}  // namespace ember_amd
