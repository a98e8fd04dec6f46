"""Split-step sequence (diffusion 1/4 x 4 + production, balanced splitting) on a
150-point H2/air flame-like profile for 100 steps of dt = 1e-5 s: the C++ SplitSolver
mirror over the GPU library against the same driver over the CPU oracle, T and Y
compared after every step (rtol 1e-6).  Convection is out of scope (no-op stage)."""
import os
import subprocess

import numpy as np
import pytest

from ember_amd import mechanism
from ember_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "split_flame")


def test_split_flame_100_steps(gpu, tmp_path):
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
    print("[split] running 100 split steps", flush=True)
    r = subprocess.run([EXE, mechanism.table_path("h2o2"), str(inp), str(out), str(nsteps), str(dt)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    data = np.fromfile(out, dtype=np.float64, count=2 * nsteps * N * nV).reshape(2, nsteps, N, nV)
    fails = np.fromfile(out, dtype=np.int32, offset=2 * nsteps * N * nV * 8)
    assert fails.tolist() == [0, 0]
    g, c = data
    dT = np.abs(g[:, :, 1] - c[:, :, 1]) / c[:, :, 1]
    dY = np.abs(g[:, :, 2:] - c[:, :, 2:]) - 1e-6 * np.abs(c[:, :, 2:])
    print(f"[split] max rel dT {dT.max():.2e}, max Y excess {dY.max():.2e}")
    assert dT.max() <= 1e-6
    assert dY.max() <= 1e-12
    # the flame evolved: chemistry and diffusion both acted
    assert np.abs(g[-1, :, 1] - case.state[:, 1]).max() > 1.0
