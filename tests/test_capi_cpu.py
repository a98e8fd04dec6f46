"""CPU-side checks of the product library: it loads, exports every symbol the public
header declares, and rejects bad input without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from ember_amd import capi, mechanism

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "ember_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(emx_[a-z_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    L = ctypes.CDLL(capi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 16
    for s in syms:
        assert hasattr(L, s), s
    bound = {name for name, _, _ in capi.SIGNATURES}
    assert set(syms) == bound


def test_default_options_match_ember():
    o = capi.default_options()
    # python/ember/input.py:729-776 defaults
    assert o.qss_epsmin == 2e-2 and o.qss_epsmax == 10.0 and o.qss_dtmin == 1e-16
    assert o.qss_dtmax == 1e-6 and o.qss_iteration_count == 1 and o.qss_abstol == 1e-11
    assert o.qss_minval == 1e-60 and o.qss_stability_check == 0
    assert o.integrator_rel_tol == 1e-6 and o.integrator_species_abs_tol == 1e-13
    assert o.global_timestep == 2e-5 and o.diffusion_timestep_multiplier == 10.0


def test_create_rejects_missing_table():
    h = ctypes.c_void_p()
    rc = capi.lib().emx_create(b"/nonexistent.emxm", 101325.0, None, ctypes.byref(h))
    assert rc < 0
    assert b"cannot open" in capi.lib().emx_last_error()


def test_table_roundtrip(tmp_path):
    m = mechanism.load("gri30")
    p = tmp_path / "g.emxm"
    m.write_table(str(p))
    text = p.read_text().splitlines()
    assert text[0] == "EMXM 1"
    assert sum(1 for l in text if l.startswith("S ")) == 53
    assert sum(1 for l in text if l.startswith("R ")) == 325
    # SI conversion of A: elementary bimolecular cm^3/mol/s -> m^3/kmol/s is x 1e-3
    r = next(r for r in m.reactions if r.equation.replace(" ", "") in ("O+H2<=>H+OH", "H2+O<=>H+OH"))
    assert abs(r.A - 3.87e4 * 1e-3) < 1e-9 and abs(r.b - 2.7) < 1e-15
    assert abs(r.EaR - 6260.0 * 4184.0 / mechanism.GAS_CONSTANT) < 1e-9
