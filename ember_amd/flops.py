"""Static FP64 operation counts of the chemistry right-hand side.

One operation per FP64 add/sub/mul/div/min/max, 2 per fused multiply-add, and 1 per
exp/log/log10/pow call (so fractions of peak built on these counts are lower
bounds).  Counted from the arithmetic of the restated Cantera/Ember pipeline:

  full RHS      = state set-up + NASA7 thermo + T-dependent rate constants and 1/Kc
                  + concentration-dependent rates of progress + species sums + q/d
  corrector RHS = the same without the T-dependent parts (temperature frozen)
"""
from __future__ import annotations

from dataclasses import dataclass

from .mechanism import TYPE_ELEMENTARY, TYPE_LINDEMANN, TYPE_THREE_BODY, TYPE_TROE, Mechanism


@dataclass
class RhsFlops:
    full: int
    corrector: int
    integrator_per_step: int


def rhs_flops(m: Mechanism) -> RhsFlops:
    K, I = m.n_species, m.n_reactions
    nv = K + 2
    # Phase::setMassFractions (clip, normalise, Y/W, mean W) + density + concentrations
    state = K * 5 + 6 + K
    # NASA7 per species: T powers (shared 4), cp 4 mul + 4 add, h 5 mul + 5 add,
    # s 5 mul + 5 add, g = h - s, mu = RT g + tmp (2)
    thermo = 4 + 2 + K * (8 + 10 + 10 + 1 + 2) + 3 * K  # + h_k, cp_mass sum
    tdep = 0
    conc = 0
    for r in m.reactions:
        nr, npd = len(r.reactants), len(r.products)
        if r.rtype in (TYPE_ELEMENTARY, TYPE_THREE_BODY):
            tdep += 5            # A exp(b logT - EaR/T)
        if r.reversible:
            tdep += (nr - 1) + (npd - 1) + 1 + 3 + 1 + 1   # sums, diff, scale, exp, min
        if r.rtype in (TYPE_LINDEMANN, TYPE_TROE):
            tdep += 10           # k_low, k_high
            conc += 4            # Pr, Pr/(1+Pr), kf
            if r.rtype == TYPE_TROE:
                tdep += 9        # Fcent: 3 exp, 4 mul/add, max, log10
                conc += 14       # log10, max, cc, nn, f1, lf, pow
        if r.rtype != TYPE_ELEMENTARY:
            conc += 2 * len(r.efficiencies) + 2
        conc += 1 + (1 if r.rtype == TYPE_THREE_BODY else 0) + nr
        if r.reversible:
            conc += 1 + npd
        # species sums: creation and destruction contributions (2 FMA each)
        conc += 4 * (len(set(r.reactants)) + len(set(r.products)))
    assemble = 6 * K + 3 * K + 6   # q/d per species, qdot
    full = state + thermo + tdep + conc + assemble
    corr = state + conc + assemble
    return RhsFlops(full=full, corrector=corr, integrator_per_step=40 * nv)


def qss_point_flops(f: RhsFlops, gcount, rcount, itermax: int = 1):
    """FLOPs of one QSS stage per point from its gcount/rcount (numpy arrays ok).

    With itermax corrector iterations each attempt evaluates the RHS once as a
    predictor (new step only) and itermax times as corrector:
    gcount = steps + itermax * (steps + rcount)."""
    steps = (gcount - itermax * rcount) / (1 + itermax)
    corr = gcount - steps
    return steps * f.full + corr * f.corrector + (steps + rcount) * f.integrator_per_step
