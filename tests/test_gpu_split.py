"""Split-step sequence (diffusion 1/4 x 4 + production, balanced splitting) on a
150-point H2/air flame-like profile for 100 steps of dt = 1e-5 s: the C++ SplitSolver
mirror over the GPU library against the same driver over the CPU oracle, T and Y
compared after every step (rtol 1e-6); and the device-resident split step
(emx_split_step_device) against the host-driven one, bit for bit.  With convection: the
full Strang sequence D C D P D C D, the convection stages through
emx_convection_integrate (GPU) and orc_convection_integrate (oracle), and the
device-resident step with its convection stages (emx_split_step_conv_device) against the
host-driven one, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from ember_amd import mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "split_flame")


@pytest.mark.parametrize("wall,conv", [(None, False), ((350.0, 200.0), False), (None, True)],
                         ids=["fixed", "wallflux", "convection"])
def test_split_flame_100_steps(gpu, tmp_path, wall, conv):
    """wall = (Tinf, Kwall): the left end is a WallFlux boundary whose energy equation
    carries yInf = Tinf and wallConst = Kwall (src/flameSolver.cpp:79-82).  conv: the
    convection stages run too, with inflow rV(0) = 0.2 rho_left (kg/m^2/s) from the left
    and a strain rate of 100 1/s (rhou = rho_left)."""
    case = W.flame_profile(150, seed=1, name="h2o2")
    N, nV = case.state.shape
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([nV, N, case.B.shape[0]], np.int64).tofile(f)
        for a in (case.x, case.state, case.B, case.D):
            np.ascontiguousarray(a, np.float64).tofile(f)
        case.B_row.astype(np.int32).tofile(f)
    out = tmp_path / "out.bin"
    nsteps, dt = 100, 1e-5
    print(f"[split] running 100 split steps, wall={wall}, convection={conv}", flush=True)
    extra = [] if wall is None else ["3", repr(wall[0]), repr(wall[1])]
    if conv:
        m = mechanism.load("h2o2")
        Y0 = case.state[0, 2:]
        rho0 = 101325.0 / (mechanism.GAS_CONSTANT * case.state[0, 1] * (Y0 / m.W).sum())
        extra = ["0", "0", "0", repr(float(0.2 * rho0)), "100.0", repr(float(rho0))]
    r = subprocess.run([EXE, mechanism.table_path("h2o2"), str(inp), str(out), str(nsteps), str(dt)]
                       + extra, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    data = np.fromfile(out, dtype=np.float64, count=3 * nsteps * N * nV).reshape(3, nsteps, N, nV)
    fails = np.fromfile(out, dtype=np.int32, offset=3 * nsteps * N * nV * 8)
    assert fails.tolist() == [0, 0, 0]
    g, c, d = data
    # the device-resident step (with convection: emx_split_step_conv_device) runs the same
    # kernels with the same arithmetic
    print(f"[split] device-resident vs host-driven max |diff| {np.abs(d - g).max():.3e}")
    np.testing.assert_array_equal(d, g)
    if conv:
        # the strain term drove U up from 0 and the inflow advected the profile
        assert g[-1, :, 0].max() > 1.0
        dU = np.abs(g[:, :, 0] - c[:, :, 0]) / (np.abs(c[:, :, 0]) + 1e-3)
        print(f"[split] max rel dU {dU.max():.2e}")
        assert dU.max() <= 1e-6
    dT = np.abs(g[:, :, 1] - c[:, :, 1]) / c[:, :, 1]
    dY = np.abs(g[:, :, 2:] - c[:, :, 2:]) - 1e-6 * np.abs(c[:, :, 2:])
    print(f"[split] max rel dT {dT.max():.2e}, max Y excess {dY.max():.2e}")
    assert dT.max() <= 1e-6
    assert dY.max() <= 1e-12
    # the flame evolved: chemistry and diffusion both acted
    assert np.abs(g[-1, :, 1] - case.state[:, 1]).max() > 1.0
    if wall is not None:
        # the wall pulls the left-end temperature towards Tinf
        T0 = case.state[0, 1]
        assert abs(g[-1, 0, 1] - wall[0]) < abs(T0 - wall[0]), (T0, g[-1, 0, 1])


def test_correct_mass_fractions_device(gpu):
    """FlameSolver::correctMassFractions: clip negative Y to zero, scale to unit sum."""
    import torch
    from ember_amd import capi
    print("[split] correct mass fractions", flush=True)
    ctx = capi.Context("gri30")
    st = W.tanh_progress_states(200, seed=7, name="gri30")
    rng = np.random.default_rng(7)
    st[:, 2:] += rng.normal(0.0, 1e-4, st[:, 2:].shape)  # some negative, sums off one
    d = torch.from_numpy(st.copy()).to(gpu)
    ctx.correct_mass_fractions_device(d)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    y = np.maximum(st[:, 2:], 0.0)
    ref = y * (1.0 / y.sum(axis=1, keepdims=True))
    np.testing.assert_allclose(got[:, 2:], ref, rtol=1e-15, atol=0)
    np.testing.assert_array_equal(got[:, :2], st[:, :2])
    assert (got[:, 2:] >= 0).all()


@pytest.mark.parametrize("name", ["h2o2", "gri30"])
def test_heat_release_device(gpu, oracle_mod, name):
    """FlameSolver::calculateQdot against the oracle's rates and enthalpies."""
    import torch
    from ember_amd import capi
    print(f"[split] heat release {name}", flush=True)
    ctx = capi.Context(name)
    st = W.tanh_progress_states(96, seed=3, name=name)
    d = torch.from_numpy(st).to(gpu)
    q = torch.zeros(st.shape[0], dtype=torch.float64, device=gpu)
    status = torch.zeros(st.shape[0], dtype=torch.int32, device=gpu)
    ctx.heat_release_device(d, q, status)
    torch.cuda.synchronize()
    m = oracle_mod.Mech(mechanism.table_path(name))
    r = m.rates(st[:, 1], st[:, 2:])
    ref = -(r["wdot"] * r["hk"]).sum(axis=1)
    assert (status.cpu().numpy() == 0).all()
    np.testing.assert_allclose(q.cpu().numpy(), ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
