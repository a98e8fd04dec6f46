"""Edge cases of the C ABI on the GPU against the oracle: empty and ragged point counts
(around the 64-lane wavefront and the 4-wave workgroup), the smallest grids around the
16-point diffusion chunks, the widest variable count, and the argument checks.  Bars as
the parity tests: stages |dT| <= 1e-6 T, |dY| <= 1e-6 |Y| + 1e-12; diffusion 1e-10."""
import ctypes as C

import numpy as np
import pytest

from ember_amd import capi, mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,integrator", [(0, 0), (1, 0), (3, 0), (63, 0), (65, 0), (257, 0),
                                          (1, 1), (5, 1)])
def test_stage_point_counts(gpu, oracle_mod, n, integrator):
    print(f"[edge] stage n={n} integrator={integrator}", flush=True)
    st = W.tanh_progress_states(max(n, 1), seed=11, name="h2o2")[:n]
    ctx = capi.Context("h2o2", chemistry_integrator=integrator)
    out, status, *_ = ctx.source_integrate(st, 2e-5)
    assert out.shape == st.shape
    if n == 0:
        return
    om = oracle_mod.Mech(mechanism.table_path("h2o2"))
    ref, rstat, *_ = om.integrate(st, 2e-5, oracle_mod.options(integrator=integrator))
    assert (status == rstat).all() and (status == 0).all()
    assert (np.abs(out[:, 1] - ref[:, 1]) <= 1e-6 * ref[:, 1]).all()
    assert (np.abs(out[:, 2:] - ref[:, 2:]) <= 1e-6 * np.abs(ref[:, 2:]) + 1e-12).all()


def test_stage_argument_checks(gpu):
    print("[edge] stage argument checks", flush=True)
    ctx = capi.Context("h2o2")
    st = np.ascontiguousarray(W.tanh_progress_states(4, seed=11, name="h2o2"))
    out = np.empty_like(st)
    dp = C.POINTER(C.c_double)
    p = lambda a: a.ctypes.data_as(dp)
    # nVars must be nSpecies + 2
    rc = capi.lib().emx_source_integrate(ctx.h, 4, ctx.n_vars + 1, p(st), None, None, 0.0, 2e-5,
                                         p(out), None, None, None)
    assert rc != 0
    # NULL state
    rc = capi.lib().emx_source_integrate(ctx.h, 4, ctx.n_vars, None, None, None, 0.0, 2e-5, p(out),
                                         None, None, None)
    assert rc != 0


@pytest.mark.parametrize("N,nV", [(4, 3), (5, 3), (16, 3), (17, 3), (31, 2), (32, 2), (33, 5),
                                  (64, 1), (65, 7), (200, 64)])
def test_diffusion_small_grids(gpu, oracle_mod, N, nV):
    print(f"[edge] diffusion N={N} nV={nV}", flush=True)
    rng = np.random.default_rng(N * 100 + nV)
    if N >= 5:
        case = W.diffusion_case(n_vars=nV, n_points=N, seed=7 * N + 1)
    else:  # too few spacings for the smooth random grid: a uniform one
        x = np.linspace(0.0, 0.02, N)
        D = rng.uniform(1e-5, 1e-4, (N, nV))
        case = W.DiffusionCase(x, np.full((1, N), 4.0), np.zeros(nV, np.int32), D,
                               rng.uniform(0.0, 1.0, (N, nV)), np.zeros((N, nV)))
    sc = rng.normal(0, 1.0, (N, nV))
    B = np.atleast_2d(case.B[0])
    B_row = np.zeros(nV, np.int32)
    ctx = capi.Context("h2o2")
    t0, t1 = 1e-3, 1e-3 + 5e-6
    got = ctx.diffusion_integrate(case.y0, B, case.D, case.x, t0, t1, B_row=B_row, split_const=sc)
    ref = oracle_mod.diffusion_integrate(B, B_row, case.D, sc, case.x, case.y0, t0, t1,
                                         multiplier=10.0, dt_global=2e-5)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * scale)


def test_diffusion_argument_checks(gpu):
    print("[edge] diffusion argument checks", flush=True)
    ctx = capi.Context("h2o2")

    def case(N, nV=2):
        x = np.linspace(0.0, 0.02, N)
        return x, np.full((1, N), 4.0), np.full((N, nV), 1e-4), np.full((N, nV), 0.5)

    x, B, D, y = case(3)
    with pytest.raises(Exception):  # fewer than 4 points
        ctx.diffusion_integrate(y, B, D, x, 0.0, 1e-6, B_row=np.zeros(2, np.int32))
    x, B, D, y = case(8)
    with pytest.raises(Exception):  # right boundary condition out of range
        ctx.diffusion_integrate(y, B, D, x, 0.0, 1e-6, B_row=np.zeros(2, np.int32), right_bc=3)
    with pytest.raises(Exception):  # B_row entry beyond the B rows
        ctx.diffusion_integrate(y, B, D, x, 0.0, 1e-6, B_row=np.array([0, 1], np.int32))
    out = ctx.diffusion_integrate(y, B, D, x, 0.0, 1e-6, B_row=np.zeros(2, np.int32))
    np.testing.assert_allclose(out, 0.5, rtol=1e-14)  # a constant profile stays constant


def test_device_tensor_checks(gpu):
    """The device wrappers refuse tensors too small for the call or of the wrong dtype
    before anything reaches a kernel."""
    import torch
    print("[edge] device tensor checks", flush=True)
    ctx = capi.Context("h2o2")
    st = torch.from_numpy(W.tanh_progress_states(8, seed=11, name="h2o2")).to(gpu)
    with pytest.raises(capi.EmberAmdError):  # output one row short
        ctx.source_integrate_device(st, torch.empty_like(st)[:7], 2e-5)
    with pytest.raises(capi.EmberAmdError):  # status of the wrong dtype
        ctx.source_integrate_device(st, torch.empty_like(st), 2e-5,
                                    status=torch.zeros(8, dtype=torch.int64, device=gpu))
    with pytest.raises(capi.EmberAmdError):  # cp_k of one point too few
        rho = torch.empty(8, dtype=torch.float64, device=gpu)
        ctx.thermo_properties_device(st, rho, rho.clone(),
                                     cpSpec=torch.empty((7, ctx.n_species), dtype=torch.float64, device=gpu))
    with pytest.raises(capi.EmberAmdError):  # heat release of float32
        ctx.heat_release_device(st, torch.empty(8, dtype=torch.float32, device=gpu))
