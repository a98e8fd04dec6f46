"""Writes ember_amd/data/<mech>_burned.json: the adiabatic constant-pressure
equilibrium ("HP equilibrium") of the stoichiometric fuel/air mixtures used by the
synthetic workloads.

Method: complete-combustion products (CO2, H2O, N2 for CH4/air; H2O, N2 for H2/air)
at the temperature T* that conserves the unburned enthalpy, then integrated with the
oracle's BDF (adiabatic, constant P, rtol 1e-10) until dT/dt and dY/dt vanish.  No
external equilibrium solver is involved.

Run once with `python tests/golden/make_burned_states.py`; the JSON files are data.
"""
import json
import os
import sys

import numpy as np
from scipy.optimize import brentq

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ember_amd import mechanism as M  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {
    "gri30": ({"CH4": 1.0, "O2": 2.0, "N2": 7.52}, {"CO2": 1.0, "H2O": 2.0, "N2": 7.52}),
    "h2o2": ({"H2": 2.0, "O2": 1.0, "N2": 3.76}, {"H2O": 2.0, "N2": 3.76}),
}


def mass_fractions(m, moles):
    X = np.zeros(m.n_species)
    for k, v in moles.items():
        X[m.species_index(k)] = v
    return m.mole_to_mass(X / X.sum())


def main():
    for name, (reac, prod) in CASES.items():
        m = M.load(name)
        Yu = mass_fractions(m, reac)
        Tu = 300.0
        hu = m.mixture_enthalpy_mass(Tu, Yu)
        Yp = mass_fractions(m, prod)
        Tstar = brentq(lambda T: m.mixture_enthalpy_mass(T, Yp) - hu, 300.0, 5000.0)
        om = O.Mech(M.table_path(name))
        opts = O.options(integrator=1, rtol=1e-10, atol_species=1e-20, atol_energy=1e-10,
                         atol_momentum=1e-10)
        state = np.zeros((1, m.n_species + 2))
        state[0, 1] = Tstar
        state[0, 2:] = Yp
        t = 0.0
        for dt in [1e-6] * 10 + [1e-5] * 10 + [1e-4] * 10 + [1e-3] * 20 + [1e-2] * 20:
            out, status, *_ = om.integrate(state, dt, opts)
            assert status[0] == 0, status
            state = out
            t += dt
        r = om.rates(state[0, 1], state[0, 2:], opts)
        Yb = state[0, 2:] / state[0, 2:].sum()
        Tb = float(state[0, 1])
        hb = m.mixture_enthalpy_mass(Tb, Yb)
        res = {
            "mechanism": name, "pressure": M.ONE_ATM, "T_unburned": Tu,
            "X_unburned": reac, "T_complete_combustion": Tstar,
            "T": Tb, "Y": {s: float(y) for s, y in zip(m.species, Yb)},
            "time_integrated": t,
            "max_abs_wdot_kmol_m3_s": float(np.abs(r["wdot"]).max()),
            "enthalpy_error_rel": float((hb - hu) / abs(hu) if hu != 0 else hb - hu),
        }
        path = os.path.join(ROOT, "ember_amd", "data", f"{name}_burned.json")
        with open(path, "w") as f:
            json.dump(res, f, indent=1)
        print(name, "T* %.2f -> Tb %.2f K" % (Tstar, Tb), "max|wdot| %.3g" % res["max_abs_wdot_kmol_m3_s"])


if __name__ == "__main__":
    main()
