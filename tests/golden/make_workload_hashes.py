"""Writes tests/golden/workload_hashes.json (SHA-256 of every synthetic input array the
benchmark and parity workloads generate) and the small H2/O2 fixtures
tests/golden/h2o2_flame_profile_150.npz and tests/golden/h2o2_tanh_states_4096.npy.

The generators are ember_amd/workloads.py with numpy.random.default_rng(seed):
  tanh progress states (4096 GRI points, seed 2), flame-front states (4096, seed 5),
  uneven-stiffness states (65536, seed 3), the diffusion case (53 x 65536, seed 4),
  the H2/air flame profile (150 points, seed 1) and H2/O2 tanh states (4096, seed 2).
tests/test_workloads.py regenerates them and checks the hashes and fixtures.

Run with `python tests/golden/make_workload_hashes.py`.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ember_amd import workloads as W  # noqa: E402


def sha(a) -> str:
    a = np.ascontiguousarray(a)
    h = hashlib.sha256()
    h.update(str(a.dtype).encode())
    h.update(str(a.shape).encode())
    h.update(a.tobytes())
    return h.hexdigest()


def arrays():
    out = {}
    out["gri30_tanh_states_4096_seed2"] = W.tanh_progress_states(4096, seed=2, name="gri30")
    out["gri30_flame_front_states_4096_seed5"] = W.flame_front_states(4096, seed=5, name="gri30")
    st, kind = W.uneven_stiffness_states(65536, seed=3, name="gri30")
    out["gri30_uneven_states_65536_seed3"] = st
    out["gri30_uneven_kind_65536_seed3"] = kind
    d = W.diffusion_case(53, 65536, seed=4)
    for f in ("x", "B", "B_row", "D", "y0"):
        out[f"diffusion_53x65536_seed4_{f}"] = getattr(d, f)
    fl = W.flame_profile(150, seed=1, name="h2o2")
    for f in ("x", "state", "B", "B_row", "D"):
        out[f"h2o2_flame_profile_150_seed1_{f}"] = getattr(fl, f)
    out["h2o2_tanh_states_4096_seed2"] = W.tanh_progress_states(4096, seed=2, name="h2o2")
    return out


def main():
    arr = arrays()
    with open(os.path.join(HERE, "workload_hashes.json"), "w") as f:
        json.dump({k: sha(v) for k, v in arr.items()}, f, indent=1, sort_keys=True)
    fl = {k.split("_seed1_")[1]: v for k, v in arr.items() if k.startswith("h2o2_flame_profile")}
    np.savez(os.path.join(HERE, "h2o2_flame_profile_150.npz"), **fl)
    np.save(os.path.join(HERE, "h2o2_tanh_states_4096.npy"), arr["h2o2_tanh_states_4096_seed2"])


if __name__ == "__main__":
    main()
