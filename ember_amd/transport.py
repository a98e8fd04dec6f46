"""Pure-species and binary transport properties from molecular parameters, and their
temperature fits.

Ember's default transport model (ApproxMixTransport, src/chemistry0d.cpp:12-261) is
Cantera's MixTransport with a mole-fraction threshold.  MixTransport does not evaluate
kinetic theory per call: Cantera's GasTransport fits, once per mechanism, polynomials of
degree 4 in ln T to

    sqrt(mu_k) / T^(1/4),   lambda_k / sqrt(T),   D_jk(T, P = 1 Pa) / T^(3/2)

and the mixture rules evaluate those polynomials at the current T.  This module
computes the same three families of fits from the species' transport parameters
(Cantera YAML ``transport`` blocks: geometry, diameter, well-depth, dipole,
polarizability, rotational-relaxation), by Chapman-Enskog theory:

- collision integrals Omega(1,1)* and Omega(2,2)* of the Lennard-Jones 12-6 potential by
  the fits of Neufeld, Janzen and Aziz (J. Chem. Phys. 57, 1100, 1972), with Brokaw's
  correction for polar pairs (Ind. Eng. Chem. Process Des. Dev. 8, 240, 1969):
  + 0.19 delta*^2 / T* and + 0.2 delta*^2 / T*;
- combining rules sigma_jk = (sigma_j + sigma_k) / 2, eps_jk = sqrt(eps_j eps_k), with the
  polar / non-polar correction xi = 1 + alpha*_n mu*_p^2 sqrt(eps_p / eps_n) / 4
  (sigma xi^(-1/6), eps xi^2) and delta*_jk = mu_j mu_k / (8 pi eps0 eps_jk sigma_jk^3);
- viscosity mu = 5/16 sqrt(pi m k T) / (pi sigma^2 Omega(2,2)*), binary diffusion
  D = 3/16 sqrt(2 pi (k T)^3 / m_jk) / (P pi sigma_jk^2 Omega(1,1)*);
- thermal conductivity by the Warnatz model with Parker's rotational-relaxation
  temperature dependence, as in Kee, Coltrin and Glarborg, "Chemically Reacting Flow"
  (2003), section 12.3: translational, rotational and vibrational parts weighted by
  rho D_kk / mu_k and the rotational collision number Z_rot(T).

The fits use 50 temperatures equally spaced over the range where every species' NASA7
polynomials hold, least squares relative to the fitted quantity.  Outside that range the
polynomials extrapolate, as Cantera's do.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

BOLTZMANN = 1.380649e-23          # J/K
AVOGADRO = 6.02214076e26          # 1/kmol
GAS_CONSTANT = AVOGADRO * BOLTZMANN
EPSILON_0 = 8.8541878128e-12      # F/m
LIGHT_SPEED = 299792458.0         # m/s
DEBYE = 1e-21 / LIGHT_SPEED       # C m
ANGSTROM = 1e-10
N_FIT = 50
DEGREE = 4

GEOMETRY_CV_ROT = {"atom": 0.0, "linear": 1.0, "nonlinear": 1.5}


@dataclass
class SpeciesTransport:
    geometry: str
    diameter: float          # m
    well_depth: float        # J
    dipole: float = 0.0      # C m
    polarizability: float = 0.0   # m^3
    zrot: float = 0.0        # rotational relaxation collision number at 298 K

    @classmethod
    def from_yaml(cls, t: dict) -> "SpeciesTransport":
        if t.get("model", "gas") != "gas":
            raise ValueError(f"transport model {t.get('model')!r} is not supported")
        if t["geometry"] not in GEOMETRY_CV_ROT:
            raise ValueError(f"unknown geometry {t['geometry']!r}")
        unknown = set(t) - {"model", "geometry", "diameter", "well-depth", "dipole",
                            "polarizability", "rotational-relaxation", "note",
                            "acentric-factor", "dispersion-coefficient", "quadrupole-polarizability"}
        if unknown:
            raise ValueError(f"unsupported transport fields {sorted(unknown)}")
        return cls(geometry=t["geometry"], diameter=float(t["diameter"]) * ANGSTROM,
                   well_depth=float(t["well-depth"]) * BOLTZMANN,
                   dipole=float(t.get("dipole", 0.0)) * DEBYE,
                   polarizability=float(t.get("polarizability", 0.0)) * ANGSTROM ** 3,
                   zrot=float(t.get("rotational-relaxation", 0.0)))


def omega11_lj(tstar):
    """Omega(1,1)* of the Lennard-Jones potential (Neufeld, Janzen and Aziz 1972)."""
    t = np.asarray(tstar, dtype=np.float64)
    return (1.06036 / t ** 0.15610 + 0.19300 / np.exp(0.47635 * t) + 1.03587 / np.exp(1.52996 * t)
            + 1.76474 / np.exp(3.89411 * t))


def omega22_lj(tstar):
    """Omega(2,2)* of the Lennard-Jones potential (Neufeld, Janzen and Aziz 1972)."""
    t = np.asarray(tstar, dtype=np.float64)
    return 1.16145 / t ** 0.14874 + 0.52487 / np.exp(0.77320 * t) + 2.16178 / np.exp(2.43787 * t)


def omega11(tstar, delta):
    return omega11_lj(tstar) + 0.19 * np.asarray(delta) ** 2 / np.asarray(tstar)


def omega22(tstar, delta):
    return omega22_lj(tstar) + 0.2 * np.asarray(delta) ** 2 / np.asarray(tstar)


def parker(eps_over_kT):
    """Parker's temperature dependence of the rotational collision number, F(T)."""
    e = eps_over_kT
    return 1.0 + np.pi ** 1.5 / 2.0 * np.sqrt(e) + (np.pi ** 2 / 4.0 + 2.0) * e + np.pi ** 1.5 * e ** 1.5


def pair_parameters(a: SpeciesTransport, b: SpeciesTransport):
    """(sigma_jk, eps_jk, delta*_jk) with the polar / non-polar correction."""
    pa, pb = a.dipole > 0, b.dipole > 0
    f_sigma = f_eps = 1.0
    if pa != pb:
        p, n = (a, b) if pa else (b, a)
        alpha_star = n.polarizability / n.diameter ** 3
        mu_star = p.dipole / np.sqrt(4 * np.pi * EPSILON_0 * p.diameter ** 3 * p.well_depth)
        xi = 1.0 + 0.25 * alpha_star * mu_star ** 2 * np.sqrt(p.well_depth / n.well_depth)
        f_sigma = xi ** (-1.0 / 6.0)
        f_eps = xi * xi
    sigma = 0.5 * (a.diameter + b.diameter) * f_sigma
    eps = np.sqrt(a.well_depth * b.well_depth) * f_eps
    mu = np.sqrt(a.dipole * b.dipole)
    delta = 0.5 * mu * mu / (4 * np.pi * EPSILON_0 * eps * sigma ** 3)
    return sigma, eps, delta


def viscosity(s: SpeciesTransport, W: float, T):
    """Pure-species viscosity [Pa s] at T [K]; W in kg/kmol."""
    T = np.asarray(T, dtype=np.float64)
    _, eps, delta = pair_parameters(s, s)
    tstar = BOLTZMANN * T / eps
    m = W / AVOGADRO
    return 5.0 / 16.0 * np.sqrt(np.pi * m * BOLTZMANN * T) / (np.pi * s.diameter ** 2 * omega22(tstar, delta))


def binary_diffusion(a: SpeciesTransport, Wa: float, b: SpeciesTransport, Wb: float, T, P=1.0):
    """Binary diffusion coefficient D_ab [m^2/s] at T [K] and P [Pa]."""
    T = np.asarray(T, dtype=np.float64)
    sigma, eps, delta = pair_parameters(a, b)
    tstar = BOLTZMANN * T / eps
    m = (Wa * Wb / (Wa + Wb)) / AVOGADRO
    return 3.0 / 16.0 * np.sqrt(2.0 * np.pi * (BOLTZMANN * T) ** 3 / m) / \
        (P * np.pi * sigma ** 2 * omega11(tstar, delta))


def conductivity(s: SpeciesTransport, W: float, T, cp_R):
    """Pure-species thermal conductivity [W/m/K] at T, with cp/R of the species at T
    (Warnatz model with Parker's Z_rot(T))."""
    T = np.asarray(T, dtype=np.float64)
    mu = viscosity(s, W, T)
    D = binary_diffusion(s, W, s, W, T, P=1.0)   # self-diffusion at 1 Pa
    rho = W * 1.0 / (GAS_CONSTANT * T)            # density at 1 Pa
    f = rho * D / mu                              # rho D_kk / mu_k
    cv_rot = GEOMETRY_CV_ROT[s.geometry]
    cv_vib = np.asarray(cp_R) - 1.0 - 1.5 - cv_rot
    if s.geometry == "atom":
        f_trans = 2.5 * np.ones_like(T)
        f_rot = np.zeros_like(T)
    else:
        e = s.well_depth / BOLTZMANN
        zrot = s.zrot * parker(e / 298.0) / parker(e / T)
        A = 2.5 - f
        B = zrot + 2.0 / np.pi * (5.0 / 3.0 * cv_rot + f)
        f_trans = 2.5 * (1.0 - 2.0 / np.pi * cv_rot / 1.5 * A / B)
        f_rot = f * (1.0 + 2.0 / np.pi * A / B)
    f_vib = f
    return mu / W * GAS_CONSTANT * (f_trans * 1.5 + f_rot * cv_rot + f_vib * cv_vib)


def _fit(x, y):
    """Degree-4 least-squares fit of y(x), residuals relative to y; coefficients c0..c4
    of c0 + c1 x + ... + c4 x^4."""
    c = np.polyfit(x, y, DEGREE, w=1.0 / np.abs(y))
    return c[::-1].copy()


@dataclass
class TransportFits:
    Tmin: float
    Tmax: float
    visc: np.ndarray    # (K, 5): sqrt(mu_k) / T^(1/4) in ln T
    cond: np.ndarray    # (K, 5): lambda_k / sqrt(T) in ln T
    diff: np.ndarray    # (K (K + 1) / 2, 5): D_jk(T, 1 Pa) / T^(3/2) in ln T, j <= k


def pair_index(j: int, k: int, K: int) -> int:
    """Index of the pair (j, k) in the upper triangle j <= k, row by row (MixTransport's
    ic = K j - (j - 1) j / 2 + k - j)."""
    if k < j:
        j, k = k, j
    return K * j - (j - 1) * j // 2 + k - j


def fit(species: Sequence[SpeciesTransport], W: np.ndarray, cp_R_of_T, Tmin: float,
        Tmax: float) -> TransportFits:
    """cp_R_of_T(T) -> (K,) array of cp_k/R at T (the mechanism's NASA7 thermo)."""
    K = len(species)
    T = np.linspace(Tmin, Tmax, N_FIT)
    x = np.log(T)
    cp = np.array([cp_R_of_T(t) for t in T])  # (N_FIT, K)
    visc = np.zeros((K, 5))
    cond = np.zeros((K, 5))
    for k, s in enumerate(species):
        mu = viscosity(s, W[k], T)
        visc[k] = _fit(x, np.sqrt(mu) / T ** 0.25)
        cond[k] = _fit(x, conductivity(s, W[k], T, cp[:, k]) / np.sqrt(T))
    diff = np.zeros((K * (K + 1) // 2, 5))
    for j in range(K):
        for k in range(j, K):
            D = binary_diffusion(species[j], W[j], species[k], W[k], T, P=1.0)
            diff[pair_index(j, k, K)] = _fit(x, D / T ** 1.5)
    return TransportFits(Tmin, Tmax, visc, cond, diff)


def poly(c, logT):
    """c0 + c1 ln T + ... + c4 ln^4 T (MixTransport's dot5 with [1, ln T, ..., ln^4 T])."""
    lt = np.asarray(logT, dtype=np.float64)
    return c[0] + c[1] * lt + c[2] * lt ** 2 + c[3] * lt ** 3 + c[4] * lt ** 4


def fitted_viscosity(fits: TransportFits, k: int, T):
    return (T ** 0.25 * poly(fits.visc[k], np.log(T))) ** 2


def fitted_conductivity(fits: TransportFits, k: int, T):
    return np.sqrt(T) * poly(fits.cond[k], np.log(T))


def fitted_diffusion(fits: TransportFits, j: int, k: int, K: int, T, P=1.0):
    return T ** 1.5 * poly(fits.diff[pair_index(j, k, K)], np.log(T)) / P
