"""Seeded synthetic inputs for the hot-path workloads.

Chemistry states are Ember state columns (U, T, Y_0..Y_{K-1}) at P = 1 atm; the
diffusion workload is a set of smooth per-variable coefficient profiles on a
non-uniform grid.  Random numbers come from ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass

import numpy as np

from . import mechanism

DATA = mechanism.DATA_DIR

FUELS = {
    "gri30": {"CH4": 1.0, "O2": 2.0, "N2": 7.52},
    "h2o2": {"H2": 2.0, "O2": 1.0, "N2": 3.76},
}


def unburned(mech: mechanism.Mechanism, name: str):
    X = np.zeros(mech.n_species)
    for k, v in FUELS[name].items():
        X[mech.species_index(k)] = v
    return 300.0, mech.mole_to_mass(X / X.sum())


def burned(mech: mechanism.Mechanism, name: str):
    """Adiabatic equilibrium of the stoichiometric mixture (data file)."""
    with open(os.path.join(DATA, f"{name}_burned.json")) as f:
        d = json.load(f)
    Y = np.array([d["Y"][s] for s in mech.species])
    return float(d["T"]), Y


def blend(mech, name, c, rng, noise=0.01, T_override=None):
    """States on the unburned -> burned line at progress c, with log-normal noise
    on minor species (Y < 1e-3), renormalised."""
    Tu, Yu = unburned(mech, name)
    Tb, Yb = burned(mech, name)
    c = np.asarray(c, dtype=np.float64)
    n = c.shape[0]
    T = (1 - c) * Tu + c * Tb
    if T_override is not None:
        T = np.where(np.isnan(T_override), T, T_override)
    Y = (1 - c)[:, None] * Yu[None, :] + c[:, None] * Yb[None, :]
    minor = Y < 1e-3
    Y = np.where(minor, Y * np.exp(noise * rng.standard_normal(Y.shape)), Y)
    Y = Y / Y.sum(axis=1, keepdims=True)
    state = np.zeros((n, mech.n_species + 2))
    state[:, 1] = T
    state[:, 2:] = Y
    return state


def tanh_progress_states(n: int, seed: int = 2, name: str = "gri30"):
    """Progress variable c = (1 + tanh z)/2, z ~ U(-3, 3)."""
    mech = mechanism.load(name)
    rng = np.random.default_rng(seed)
    z = rng.uniform(-3.0, 3.0, n)
    c = 0.5 * (1.0 + np.tanh(z))
    return blend(mech, name, c, rng)


def flame_front_states(n: int, seed: int = 5, name: str = "gri30"):
    mech = mechanism.load(name)
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.3, 0.8, n)
    return blend(mech, name, c, rng)


def uneven_stiffness_states(n: int, seed: int = 3, name: str = "gri30"):
    """60% near-equilibrium burned, 30% flame front, 10% pre-ignition (900-1200 K)."""
    mech = mechanism.load(name)
    rng = np.random.default_rng(seed)
    nb = int(round(0.6 * n))
    nf = int(round(0.3 * n))
    npi = n - nb - nf
    c = np.concatenate([rng.uniform(0.97, 1.0, nb), rng.uniform(0.3, 0.8, nf), np.zeros(npi)])
    Tov = np.concatenate([np.full(nb + nf, np.nan), rng.uniform(900.0, 1200.0, npi)])
    kind = np.concatenate([np.zeros(nb, int), np.ones(nf, int), np.full(npi, 2)])
    perm = rng.permutation(n)
    st = blend(mech, name, c[perm], rng, T_override=Tov[perm])
    return st, kind[perm]


@dataclass
class DiffusionCase:
    x: np.ndarray        # (N,)
    B: np.ndarray        # (1, N) shared row 1/rho
    B_row: np.ndarray    # (nVars,) all zero
    D: np.ndarray        # (N, nVars) rho * Dk
    y0: np.ndarray       # (N, nVars)
    split_const: np.ndarray  # (N, nVars) zeros


def _smooth(rng, n, npts, lo, hi):
    """Smooth random profile on [0, 1] with values in [lo, hi]."""
    knots = rng.uniform(lo, hi, npts)
    xs = np.linspace(0.0, 1.0, npts)
    return np.interp(np.linspace(0.0, 1.0, n), xs, knots)


def diffusion_case(n_vars: int = 53, n_points: int = 65536, seed: int = 4,
                   length: float = 0.02) -> DiffusionCase:
    """Species diffusion systems on a smooth non-uniform grid: log(spacing) follows a
    seeded smooth profile spanning a factor of 8, so the ratio of neighbouring spacings
    stays within [1/1.5, 1.5] (in fact within 1 +- 1e-3 for large N)."""
    rng = np.random.default_rng(seed)
    log_h = _smooth(rng, n_points - 1, int(np.clip(n_points // 16, 2, 12)), np.log(1.0 / 8.0), 0.0)
    hh = np.exp(log_h)
    hh = hh / hh.sum() * length
    assert (hh[1:] / hh[:-1]).max() <= 1.5 and (hh[1:] / hh[:-1]).min() >= 1 / 1.5
    x = np.concatenate([[0.0], np.cumsum(hh)])
    s = (x - 0.5 * length) / (0.05 * length)
    T = 300.0 + 1900.0 * 0.5 * (1 + np.tanh(s))
    rho = mechanism.ONE_ATM * 28.0 / (mechanism.GAS_CONSTANT * T)
    B = (1.0 / rho)[None, :]
    D = np.empty((n_points, n_vars))
    y0 = np.empty((n_points, n_vars))
    for k in range(n_vars):
        logd = _smooth(rng, n_points, 9, -5.0, -3.0)
        D[:, k] = rho * 10.0 ** logd
        amp = rng.uniform(0.01, 1.0)
        w = rng.uniform(0.02, 0.2)
        y0[:, k] = amp * 0.5 * (1 + np.tanh((x / length - rng.uniform(0.3, 0.7)) / w))
    return DiffusionCase(x, B, np.zeros(n_vars, np.int32), D, y0, np.zeros_like(y0))


@dataclass
class FlameCase:
    x: np.ndarray        # (N,)
    state: np.ndarray    # (N, nVars)
    B: np.ndarray        # (2, N): 1/rho, 1/(rho cp)
    B_row: np.ndarray    # (nVars,)
    D: np.ndarray        # (N, nVars): mu, lambda, rho D_k


def flame_profile(n_points: int = 150, seed: int = 1, name: str = "h2o2",
                  width: float = 0.01, delta: float = 5e-4) -> FlameCase:
    """Premixed-flame-like profile (tanh in T between the unburned mixture at 300 K and
    its adiabatic equilibrium) with frozen seeded transport coefficients."""
    mech = mechanism.load(name)
    rng = np.random.default_rng(seed)
    u = np.linspace(-1.0, 1.0, n_points)
    x = 0.5 * width * np.sinh(2.5 * u) / np.sinh(2.5)   # finer spacing at the centre
    Tu, Yu = unburned(mech, name)
    Tb, Yb = burned(mech, name)
    c = 0.5 * (1 + np.tanh(x / delta))
    T = Tu + c * (Tb - Tu)
    Y = (1 - c)[:, None] * Yu + c[:, None] * Yb
    K = mech.n_species
    state = np.zeros((n_points, K + 2))
    state[:, 0] = 0.0
    state[:, 1] = T
    state[:, 2:] = Y
    Wmix = 1.0 / (Y / mech.W).sum(axis=1)
    rho = mechanism.ONE_ATM * Wmix / (mechanism.GAS_CONSTANT * T)
    cp = np.array([(mech.cp_R(t) * y / mech.W).sum() * mechanism.GAS_CONSTANT for t, y in zip(T, Y)])
    B = np.vstack([1.0 / rho, 1.0 / (rho * cp)])
    B_row = np.zeros(K + 2, np.int32)
    B_row[1] = 1
    D = np.empty((n_points, K + 2))
    D[:, 0] = 1.8e-5 * (T / 300.0) ** 0.7
    D[:, 1] = 0.026 * (T / 300.0) ** 0.8
    D0 = rng.uniform(1e-5, 1e-4, K)
    for k, s in enumerate(mech.species):
        if s in ("H", "H2"):
            D0[k] *= 4.0
        D[:, 2 + k] = rho * D0[k] * (T / 300.0) ** 1.75
    return FlameCase(x, state, B, B_row, D)
