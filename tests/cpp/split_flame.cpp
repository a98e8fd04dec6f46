// Split-step sequence on a flame-like profile (H2/O2 mechanism, ~150 points):
// ember_amd::FlameStages (GPU production + diffusion through the host-pointer C ABI)
// against the same SplitSolver driven by the CPU oracle, T and Y compared after every
// step; and the device-resident split step (emx_split_step_device, state kept in device
// memory) on the same inputs.
//
// usage: split_flame <mech_table> <input.bin> <output.bin> <nsteps> <dt>
//                    [leftBC Tinf Kwall [rVzero strain rhou]]
// (leftBC: Ember's BoundaryCondition value for the left end; with leftBC = 3 (WallFlux)
// the energy equation carries yInf = Tinf and wallConst = Kwall, flameSolver.cpp:79-82.
// rVzero, strain, rhou: the convection stages run too, planar, fixedLeft continuity, a
// constant strain rate, drhodt = 0; the device-resident step is then
// emx_split_step_conv_device with the same convection parameters and drhodt = 0)
// input.bin: int64 nVars, nPoints, nB; doubles x[N], state[N*nV], B[nB*N], D[N*nV];
//            int32 B_row[nV]
// output.bin: doubles gpu[nsteps][N*nV], oracle[nsteps][N*nV], device[nsteps][N*nV];
//             int32 failures[3]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ember_amd.hpp"
#include "../../oracle/oracle.h"

namespace {

class OracleStages : public ember_amd::SplitSolver {
public:
    OracleStages(void* mech, const orc_options& o, const std::vector<double>& x)
        : mech_(mech), o_(o), x_(x) {}
    ember_amd::Matrix B, D;
    std::vector<int32_t> B_row;
    int failures = 0;
    int leftBC = 0;
    double yInf = 0.0, wallConst = 0.0;
    orc_conv_params conv{};
    bool convOn = false;

protected:
    void integrateProductionTerms() override {
        std::vector<double> out(state.size());
        std::vector<int> status(nPoints);
        orc_source_integrate(mech_, &o_, (int)nPoints, state.data(), splitConstProd.data(), nullptr,
                             tStageEnd - tStageStart, out.data(), status.data(), nullptr, nullptr,
                             nullptr, 0);
        for (int s : status) failures += (s != 0);
        state.swap(out);
    }
    void integrateConvectionTerms() override {
        if (!convOn) return;
        std::vector<double> out(state.size());
        std::vector<int> steps(nVars - 1);
        const int rc = orc_convection_integrate(mech_, &conv, (int)nPoints, x_.data(), state.data(),
                                                splitConstConv.data(), nullptr, nullptr, tStageStart,
                                                tStageEnd, out.data(), steps.data(), 0);
        if (rc != 0) failures++;
        state.swap(out);
    }
    void integrateDiffusionTerms() override {
        std::vector<int> brow(B_row.begin(), B_row.end());
        orc_diffusion_integrate((int)nVars, (int)nPoints, B.data(), (int)(B.size() / nPoints), brow.data(),
                                D.data(), splitConstDiff.data(), x_.data(), 0, leftBC, 0, yInf, wallConst, 1,
                                10.0, dt,
                                nullptr, tStageStart, tStageEnd, state.data(), 0);
    }

private:
    void* mech_;
    orc_options o_;
    std::vector<double> x_;
};

template <class T>
bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6 && argc != 9 && argc != 12) {
        fprintf(stderr, "usage: %s table input output nsteps dt [leftBC Tinf Kwall [rVzero strain rhou]]\n",
                argv[0]);
        return 2;
    }
    const bool conv = argc == 12;
    const int nsteps = atoi(argv[4]);
    const double dt = atof(argv[5]);
    const int leftBC = argc == 9 ? atoi(argv[6]) : EMX_BC_FIXED_VALUE;
    const bool wallFlux = leftBC == EMX_BC_WALL_FLUX;
    const double Tinf = argc == 9 ? atof(argv[7]) : 0.0;
    const double Kwall = argc == 9 ? atof(argv[8]) : 0.0;
    FILE* f = fopen(argv[2], "rb");
    if (!f) return 3;
    int64_t hdr[3];
    if (!rd(f, hdr, 3)) return 3;
    const size_t nV = hdr[0], N = hdr[1], nB = hdr[2];
    std::vector<double> x(N), st(N * nV), B(nB * N), D(N * nV);
    std::vector<int32_t> B_row(nV);
    if (!rd(f, x.data(), N) || !rd(f, st.data(), N * nV) || !rd(f, B.data(), nB * N) ||
        !rd(f, D.data(), N * nV) || !rd(f, B_row.data(), nV))
        return 3;
    fclose(f);

    emx_options eo;
    emx_default_options(&eo);
    eo.global_timestep = dt;
    ember_amd::Context ctx(argv[1], 101325.0, eo);
    ember_amd::FlameStages gpu(ctx, x, 0, leftBC, EMX_BC_FIXED_VALUE);
    gpu.setWallFlux(wallFlux, Tinf, Kwall);
    gpu.resize(nV, N);
    gpu.state = st;
    gpu.B = B;
    gpu.D = D;
    gpu.B_row = B_row;

    orc_options oo;
    orc_default_options(&oo);
    void* mech = orc_mech_load(argv[1]);
    if (!mech) return 4;
    OracleStages cpu(mech, oo, x);
    cpu.leftBC = leftBC;
    cpu.yInf = wallFlux ? Tinf : 0.0;
    cpu.wallConst = wallFlux ? Kwall : 0.0;
    cpu.resize(nV, N);
    cpu.state = st;
    cpu.B = B;
    cpu.D = D;
    cpu.B_row = B_row;
    emx_convection_params ep{};
    if (conv) {
        const double rVzero = atof(argv[9]), strain = atof(argv[10]), rhou = atof(argv[11]);
        ep.alpha = 0;
        ep.beta = 1;
        ep.left_bc = leftBC;
        ep.right_bc = EMX_BC_FIXED_VALUE;
        ep.continuity_bc = 0;
        ep.rVzero = rVzero;
        ep.Tleft = st[1];
        ep.Yleft = &st[2];
        ep.rhou = rhou;
        ep.strain_initial = ep.strain_final = strain;
        ep.strain_t0 = 0.0;
        ep.strain_dt = 1.0;
        gpu.setConvection(ep);
        orc_conv_params& cp = cpu.conv;
        cp.alpha = 0;
        cp.beta = 1;
        cp.left_bc = leftBC;
        cp.right_bc = EMX_BC_FIXED_VALUE;
        cp.continuity_bc = 0;
        cp.pressure = 101325.0;
        cp.rVzero = rVzero;
        cp.Tleft = st[1];
        cp.rhou = rhou;
        cp.strain_initial = cp.strain_final = strain;
        cp.strain_t0 = 0.0;
        cp.strain_dt = 1.0;
        cp.rtol = eo.integrator_rel_tol;
        cp.atol_momentum = eo.integrator_momentum_abs_tol;
        cp.atol_energy = eo.integrator_energy_abs_tol;
        cp.atol_species = eo.integrator_species_abs_tol;
        cp.max_steps = 1000000;
        cpu.convOn = true;
    }

    FILE* o = fopen(argv[3], "wb");
    if (!o) return 5;
    std::vector<double> traj_g, traj_c;
    int gfail = 0;
    double t = 0.0;
    for (int s = 0; s < nsteps; s++) {
        gpu.step(t, dt);
        cpu.step(t, dt);
        for (int32_t v : gpu.status) gfail += (v != 0);
        traj_g.insert(traj_g.end(), gpu.state.begin(), gpu.state.end());
        traj_c.insert(traj_c.end(), cpu.state.begin(), cpu.state.end());
        t += dt;
        if (s % 10 == 0) {
            printf("[split_flame] step %d done\n", s);
            fflush(stdout);
        }
    }
    // device-resident split step on the same inputs
    std::vector<double> traj_d;
    int dfail = 0;
    {
        const size_t n = N * nV;
        double *d_state, *d_ddt, *d_delta, *d_work, *d_B, *d_D, *d_x;
        int32_t *d_Brow, *d_status;
        // with convection: delta_conv after delta_diff and delta_prod, and work for
        // splitConstConv and drhodt
        bool ok = hipMalloc(&d_state, n * 8) == hipSuccess && hipMalloc(&d_ddt, 3 * n * 8) == hipSuccess &&
                  hipMalloc(&d_delta, 3 * n * 8) == hipSuccess &&
                  hipMalloc(&d_work, (4 * n + N) * 8) == hipSuccess &&
                  hipMalloc(&d_B, nB * N * 8) == hipSuccess && hipMalloc(&d_D, n * 8) == hipSuccess &&
                  hipMalloc(&d_x, N * 8) == hipSuccess && hipMalloc(&d_Brow, nV * 4) == hipSuccess &&
                  hipMalloc(&d_status, N * 4) == hipSuccess;
        ok = ok && hipMemcpy(d_state, st.data(), n * 8, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemset(d_ddt, 0, 3 * n * 8) == hipSuccess &&
             hipMemcpy(d_B, B.data(), nB * N * 8, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d_D, D.data(), n * 8, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d_x, x.data(), N * 8, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d_Brow, B_row.data(), nV * 4, hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) return 6;
        emx_split_arrays sa{d_state, d_ddt, d_ddt + n, d_ddt + 2 * n, nullptr, d_delta, d_delta + n, d_work};
        emx_grid g = gpu.grid();
        g.x = d_x;
        std::vector<double> host(n);
        std::vector<int32_t> hst(N);
        t = 0.0;
        for (int s = 0; s < nsteps; s++) {
            if (conv) {
                emx_split_convection sc{&ep, nullptr, nullptr, d_delta + 2 * n, nullptr};
                ember_amd::check(emx_split_step_conv_device(ctx.get(), nV, N, &sa, &sc, d_B, nB, d_Brow, d_D,
                                                            &g, nullptr, nullptr, t, dt, 1, d_status));
            } else {
                ember_amd::check(emx_split_step_device(ctx.get(), nV, N, &sa, d_B, nB, d_Brow, d_D, &g,
                                                       nullptr, nullptr, t, dt, 1, d_status));
            }
            if (hipMemcpy(host.data(), d_state, n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(hst.data(), d_status, N * 4, hipMemcpyDeviceToHost) != hipSuccess)
                return 7;
            for (int32_t v : hst) dfail += (v != 0);
            traj_d.insert(traj_d.end(), host.begin(), host.end());
            t += dt;
            if (s % 10 == 0) {
                printf("[split_flame] device-resident step %d done\n", s);
                fflush(stdout);
            }
        }
        for (void* p : {(void*)d_state, (void*)d_ddt, (void*)d_delta, (void*)d_work, (void*)d_B,
                        (void*)d_D, (void*)d_x, (void*)d_Brow, (void*)d_status})
            (void)hipFree(p);
    }
    fwrite(traj_g.data(), sizeof(double), traj_g.size(), o);
    fwrite(traj_c.data(), sizeof(double), traj_c.size(), o);
    fwrite(traj_d.data(), sizeof(double), traj_d.size(), o);
    int32_t fails[3] = {gfail, cpu.failures, dfail};
    fwrite(fails, sizeof(int32_t), 3, o);
    fclose(o);
    orc_mech_free(mech);
    printf("[split_flame] ok\n");
    return 0;
}
