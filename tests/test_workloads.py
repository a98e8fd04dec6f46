"""The synthetic workloads are reproducible: regenerated arrays match the committed
SHA-256 hashes and the committed H2/O2 fixtures (tests/golden/make_workload_hashes.py)."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_workload_hashes as MWH  # noqa: E402


@pytest.fixture(scope="module")
def generated():
    return MWH.arrays()


def test_workload_hashes(generated):
    want = json.load(open(os.path.join(HERE, "golden", "workload_hashes.json")))
    assert set(want) == set(generated)
    bad = [k for k, v in generated.items() if MWH.sha(v) != want[k]]
    assert not bad, bad


def test_h2o2_fixtures(generated):
    fl = np.load(os.path.join(HERE, "golden", "h2o2_flame_profile_150.npz"))
    for f in ("x", "state", "B", "B_row", "D"):
        np.testing.assert_array_equal(fl[f], generated[f"h2o2_flame_profile_150_seed1_{f}"])
    st = np.load(os.path.join(HERE, "golden", "h2o2_tanh_states_4096.npy"))
    np.testing.assert_array_equal(st, generated["h2o2_tanh_states_4096_seed2"])


def test_workload_shapes(generated):
    assert generated["gri30_tanh_states_4096_seed2"].shape == (4096, 55)
    assert generated["gri30_uneven_states_65536_seed3"].shape == (65536, 55)
    assert generated["diffusion_53x65536_seed4_D"].shape == (65536, 53)
    assert generated["h2o2_tanh_states_4096_seed2"].shape == (4096, 12)
    # mass fractions sum to one, temperatures in the flame range
    for k in ("gri30_tanh_states_4096_seed2", "gri30_flame_front_states_4096_seed5",
              "h2o2_tanh_states_4096_seed2"):
        s = generated[k]
        np.testing.assert_allclose(s[:, 2:].sum(axis=1), 1.0, rtol=1e-12)
        assert s[:, 1].min() >= 290.0 and s[:, 1].max() <= 2500.0
