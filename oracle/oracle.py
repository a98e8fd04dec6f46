"""ctypes wrapper for the CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py``; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Options(C.Structure):
    _fields_ = [("integrator", C.c_int), ("kinetics", C.c_int), ("pressure", C.c_double),
                ("qss_epsmin", C.c_double), ("qss_epsmax", C.c_double), ("qss_dtmin", C.c_double),
                ("qss_dtmax", C.c_double), ("qss_itermax", C.c_int), ("qss_abstol", C.c_double),
                ("qss_minval", C.c_double), ("qss_stability_check", C.c_int),
                ("rtol", C.c_double), ("atol_momentum", C.c_double), ("atol_energy", C.c_double),
                ("atol_species", C.c_double), ("min_step", C.c_double), ("max_steps", C.c_int),
                ("t_start", C.c_double), ("ignition_t_start", C.c_double),
                ("ignition_duration", C.c_double), ("ignition_energy", C.c_double),
                ("ignition_center", C.c_double), ("ignition_stddev", C.c_double),
                ("positions", C.c_void_p)]


class ConvParams(C.Structure):
    _fields_ = [("alpha", C.c_int), ("beta", C.c_int), ("left_bc", C.c_int), ("right_bc", C.c_int),
                ("continuity_bc", C.c_int), ("pressure", C.c_double), ("rVzero", C.c_double),
                ("Tleft", C.c_double), ("rhou", C.c_double), ("strain_initial", C.c_double),
                ("strain_final", C.c_double), ("strain_t0", C.c_double), ("strain_dt", C.c_double),
                ("rtol", C.c_double), ("atol_momentum", C.c_double), ("atol_energy", C.c_double),
                ("atol_species", C.c_double), ("max_steps", C.c_int)]


def conv_params(**kw) -> ConvParams:
    """Convection-stage parameters with Ember's defaults (planar, FixedValue ends, fixedLeft
    continuity, constant strain rate, input.py tolerances)."""
    d = dict(alpha=0, beta=1, left_bc=0, right_bc=0, continuity_bc=0, pressure=101325.0, rVzero=0.0,
             Tleft=300.0, rhou=1.0, strain_initial=0.0, strain_final=0.0, strain_t0=0.0, strain_dt=1.0,
             rtol=1e-6, atol_momentum=1e-7, atol_energy=1e-8, atol_species=1e-13, max_steps=1000000)
    d.update(kw)
    return ConvParams(**d)


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_default_options.argtypes = [C.POINTER(Options)]
        L.orc_mech_load.restype = C.c_void_p
        L.orc_mech_load.argtypes = [C.c_char_p]
        L.orc_mech_free.argtypes = [C.c_void_p]
        L.orc_mech_nspecies.argtypes = [C.c_void_p]
        L.orc_mech_nreactions.argtypes = [C.c_void_p]
        L.orc_mech_set_multipliers.argtypes = [C.c_void_p, _dp]
        L.orc_mech_set_interp_range.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_mech_set_rate_function.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_rate_function_eval.restype = C.c_double
        L.orc_rate_function_eval.argtypes = [C.c_void_p, C.c_double]
        L.orc_rates.argtypes = [C.c_void_p, C.POINTER(Options), C.c_int, _dp, _dp, C.c_int,
                                _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_thermo_properties.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _dp, _dp, _dp,
                                            _dp, _dp, C.c_int]
        L.orc_source_integrate.argtypes = [C.c_void_p, C.POINTER(Options), C.c_int, _dp, _dp, _dp,
                                           C.c_double, _dp, _ip, _ip, _ip, _dp, C.c_int]
        L.orc_density_derivative.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp,
                                             _dp, _dp]
        L.orc_transport_properties.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                               _dp, _dp, C.c_int]
        L.orc_convection_integrate.argtypes = [C.c_void_p, C.POINTER(ConvParams), C.c_int, _dp, _dp, _dp,
                                               _dp, _dp, C.c_double, C.c_double, _dp, _ip, C.c_int]
        L.orc_bdf_work.argtypes = [C.c_void_p, C.POINTER(Options), C.c_int, _dp, C.c_double, _ip,
                                   C.c_int]
        L.orc_qss_test_problem.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double,
                                           C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                           C.c_int, C.c_double, _dp, _ip, _ip]
        L.orc_source_rhs.argtypes = [C.c_void_p, C.POINTER(Options), _dp, _dp, C.c_double, _dp]
        L.orc_tridiag_integrate.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                            C.c_double, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_diffusion_integrate.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, _dp, _dp, _dp,
                                              C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                              C.c_int, C.c_double, C.c_double, _dp, C.c_double,
                                              C.c_double, _dp, C.c_int]
        L.orc_qss_adjoint.argtypes = [C.c_void_p, C.POINTER(Options), C.c_int, _dp, _dp, _dp,
                                      C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _ip, C.c_int]
        L.orc_qss_fd_gradient.argtypes = [C.c_void_p, C.POINTER(Options), C.c_int, _dp, _dp, _dp,
                                          C.c_double, _dp, _dp, C.c_double, C.c_int, _dp, _dp,
                                          C.c_int]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp if a.dtype == np.float64 else _ip)


def options(**kw) -> Options:
    o = Options()
    lib().orc_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Mech:
    def __init__(self, table_path: str):
        self.h = lib().orc_mech_load(table_path.encode())
        if not self.h:
            raise RuntimeError(f"oracle could not load {table_path}")
        self.K = lib().orc_mech_nspecies(self.h)
        self.I = lib().orc_mech_nreactions(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mech_free(self.h)
            self.h = None

    def set_multipliers(self, m):
        m = None if m is None else np.ascontiguousarray(m, dtype=np.float64)
        lib().orc_mech_set_multipliers(self.h, _p(m))

    def set_interp_range(self, n, Tmin, Tmax):
        if lib().orc_mech_set_interp_range(self.h, int(n), float(Tmin), float(Tmax)):
            raise ValueError("invalid interpolation grid")

    def set_rate_function(self, data):
        """ChebyshevFunction::setCoefficients(len(data), data), data = [T0, T1, c0, ...];
        None removes the function."""
        if data is None:
            rc = lib().orc_mech_set_rate_function(self.h, 0, None)
        else:
            d = np.ascontiguousarray(data, dtype=np.float64)
            rc = lib().orc_mech_set_rate_function(self.h, len(d), _p(d))
        if rc:
            raise ValueError(f"orc_mech_set_rate_function failed: {rc}")

    def rate_function(self, t):
        return lib().orc_rate_function_eval(self.h, float(t))

    def rates(self, T, Y, opts=None, nonorm=False):
        opts = opts or options()
        T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
        Y = np.ascontiguousarray(np.atleast_2d(Y), dtype=np.float64)
        n = T.shape[0]
        out = {k: np.zeros((n, self.K)) for k in ("wdotQ", "wdotD", "wdot", "hk")}
        out["rho"] = np.zeros(n)
        out["cp"] = np.zeros(n)
        out["ropf"] = np.zeros((n, self.I))
        out["ropr"] = np.zeros((n, self.I))
        rc = lib().orc_rates(self.h, C.byref(opts), n, _p(T), _p(Y), int(nonorm), _p(out["wdotQ"]),
                             _p(out["wdotD"]), _p(out["wdot"]), _p(out["hk"]), _p(out["rho"]),
                             _p(out["cp"]), _p(out["ropf"]), _p(out["ropr"]))
        if rc:
            raise RuntimeError(f"orc_rates failed: {rc}")
        return out

    def thermo_properties(self, state, pressure=101325.0, nthreads=0):
        """rho, Wmx, cp (N) and cpSpec, hk (N, K) of point-major states (N, K + 2)."""
        state = np.ascontiguousarray(np.atleast_2d(state), dtype=np.float64)
        n = state.shape[0]
        out = {k: np.zeros(n) for k in ("rho", "Wmx", "cp")}
        out["cpSpec"] = np.zeros((n, self.K))
        out["hk"] = np.zeros((n, self.K))
        rc = lib().orc_thermo_properties(self.h, float(pressure), n, _p(state), _p(out["rho"]),
                                         _p(out["Wmx"]), _p(out["cp"]), _p(out["cpSpec"]),
                                         _p(out["hk"]), int(nthreads))
        if rc:
            raise RuntimeError(f"orc_thermo_properties failed: {rc}")
        return out

    def density_derivative(self, state, ddt_conv, ddt_diff, ddt_prod, rho, Wmx, ddt_cross=None,
                           balanced=True):
        """drhodt (N) of FlameSolver::prepareIntegrators from point-major (N, K + 2) arrays."""
        a = [np.ascontiguousarray(np.atleast_2d(v), dtype=np.float64)
             for v in (state, ddt_conv, ddt_diff, ddt_prod)]
        n = a[0].shape[0]
        cross = None if ddt_cross is None else np.ascontiguousarray(ddt_cross, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        Wmx = np.ascontiguousarray(Wmx, dtype=np.float64)
        out = np.zeros(n)
        rc = lib().orc_density_derivative(self.h, n, *(_p(v) for v in a), _p(cross), 1 if balanced else 0,
                                          _p(rho), _p(Wmx), _p(out))
        if rc:
            raise RuntimeError(f"orc_density_derivative failed: {rc}")
        return out

    def rhs(self, y, opts=None, split_const=None, qloss=0.0):
        """SourceSystemCVODE::f at state y (nVars)."""
        opts = opts or options()
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros_like(y)
        sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
        rc = lib().orc_source_rhs(self.h, C.byref(opts), _p(y), _p(sc), qloss, _p(out))
        if rc:
            raise RuntimeError(f"orc_source_rhs failed: {rc}")
        return out

    def integrate(self, state, dt, opts=None, split_const=None, qloss=None, nthreads=0):
        """One production stage; state is (nPoints, nVars) (Ember column-major)."""
        opts = opts or options()
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[0]
        out = np.empty_like(state)
        status = np.zeros(n, dtype=np.int32)
        s1 = np.zeros(n, dtype=np.int32)
        s2 = np.zeros(n, dtype=np.int32)
        qint = np.zeros(n)
        sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
        ql = None if qloss is None else np.ascontiguousarray(qloss, dtype=np.float64)
        lib().orc_source_integrate(self.h, C.byref(opts), n, _p(state), _p(sc), _p(ql), dt, _p(out),
                                   _p(status), _p(s1), _p(s2), _p(qint), nthreads)
        return out, status, s1, s2, qint

    def transport_properties(self, state, pressure=101325.0, threshold=1e-5, nthreads=0):
        """(mu, lambda, rhoD) of ApproxMixTransport at each point (orc_transport_properties)."""
        state = np.ascontiguousarray(np.atleast_2d(state), dtype=np.float64)
        n = state.shape[0]
        mu, lam = np.zeros(n), np.zeros(n)
        rhoD = np.zeros((n, self.K))
        rc = lib().orc_transport_properties(self.h, pressure, threshold, n, _p(state), _p(mu), _p(lam),
                                            _p(rhoD), nthreads)
        if rc:
            raise RuntimeError(f"orc_transport_properties failed: {rc}")
        return mu, lam, rhoD

    def convection_integrate(self, x, state, t0, t1, params=None, split_const=None, drhodt=None,
                             Yleft=None, nthreads=0):
        """Convection stage over [t0, t1]: (state_out, steps (1 + K: UTW, species))."""
        params = params or conv_params()
        state = np.ascontiguousarray(state, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = state.shape[0]
        out = state.copy()
        steps = np.zeros(1 + self.K, dtype=np.int32)
        sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
        dr = None if drhodt is None else np.ascontiguousarray(drhodt, dtype=np.float64)
        yl = None if Yleft is None else np.ascontiguousarray(Yleft, dtype=np.float64)
        rc = lib().orc_convection_integrate(self.h, C.byref(params), n, _p(x), _p(state), _p(sc), _p(dr), _p(yl),
                                            t0, t1, _p(out), steps.ctypes.data_as(_ip), nthreads)
        if rc:
            raise RuntimeError(f"orc_convection_integrate failed: {rc}")
        return out, steps

    def bdf_work(self, state, dt, opts=None, nthreads=0):
        """BDF work per point: (n_rhs, n_jac, n_lu, n_solve) columns (orc_bdf_work)."""
        opts = opts or options(integrator=1)
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[0]
        w = np.zeros((n, 4), dtype=np.int32)
        rc = lib().orc_bdf_work(self.h, C.byref(opts), n, _p(state), dt,
                                w.ctypes.data_as(_ip), nthreads)
        if rc:
            raise RuntimeError(f"orc_bdf_work failed: {rc}")
        return w

    def adjoint(self, state, dt, opts=None, Qbar=None, ybar=None, split_const=None, qloss=None,
                nthreads=0):
        """Discrete adjoint of the QSS stage: returns (state_out, Q, grad_y0, grad_lnk, status)
        for J = Qbar*Q + ybar . y_final (frozen step sequence)."""
        opts = opts or options()
        state = np.ascontiguousarray(state, dtype=np.float64)
        n, nv = state.shape
        nr = lib().orc_mech_nreactions(self.h)
        out = np.empty_like(state)
        Q = np.zeros(n)
        gy = np.zeros((n, nv))
        gk = np.zeros((n, nr))
        status = np.zeros(n, dtype=np.int32)
        qb = None if Qbar is None else np.ascontiguousarray(np.broadcast_to(Qbar, (n,)), dtype=np.float64)
        yb = None if ybar is None else np.ascontiguousarray(ybar, dtype=np.float64)
        sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
        ql = None if qloss is None else np.ascontiguousarray(qloss, dtype=np.float64)
        rc = lib().orc_qss_adjoint(self.h, C.byref(opts), n, _p(state), _p(sc), _p(ql), dt, _p(qb),
                                   _p(yb), _p(out), _p(Q), _p(gy), _p(gk), _p(status), nthreads)
        if rc:
            raise RuntimeError(f"orc_qss_adjoint failed: {rc}")
        return out, Q, gy, gk, status

    def fd_gradient(self, state, dt, opts=None, h=1e-6, frozen=True, split_const=None, qloss=None,
                    with_lnk=True, Qbar=None, ybar=None, nthreads=0):
        """Central finite differences of J = Qbar*Q + ybar . y_final (orc_qss_fd_gradient)."""
        opts = opts or options()
        state = np.ascontiguousarray(state, dtype=np.float64)
        n, nv = state.shape
        nr = lib().orc_mech_nreactions(self.h)
        gy = np.zeros((n, nv))
        gk = np.zeros((n, nr)) if with_lnk else None
        sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
        ql = None if qloss is None else np.ascontiguousarray(qloss, dtype=np.float64)
        qb = None if Qbar is None else np.ascontiguousarray(np.broadcast_to(Qbar, (n,)), dtype=np.float64)
        yb = None if ybar is None else np.ascontiguousarray(ybar, dtype=np.float64)
        rc = lib().orc_qss_fd_gradient(self.h, C.byref(opts), n, _p(state), _p(sc), _p(ql), dt, _p(qb),
                                       _p(yb), h, 1 if frozen else 0, _p(gy), _p(gk), nthreads)
        if rc:
            raise RuntimeError(f"orc_qss_fd_gradient failed: {rc}")
        return gy, gk


def qss_test_problem(problem, y0, tf, epsmin=1e-2, epsmax=20.0, dtmin=1e-15, dtmax=1e-6,
                     itermax=2, abstol=1e-8, ymin=1e-20, stability=True, qc=None, kc=None):
    """QssIntegrator on the cesium problem (problem=0) or y' = qc - kc*y (problem=1)."""
    y = np.array(y0, dtype=np.float64)
    qc = None if qc is None else np.ascontiguousarray(qc, dtype=np.float64)
    kc = None if kc is None else np.ascontiguousarray(kc, dtype=np.float64)
    g = np.zeros(1, np.int32)
    r = np.zeros(1, np.int32)
    err = lib().orc_qss_test_problem(problem, len(y), _p(qc), _p(kc), epsmin, epsmax, dtmin, dtmax,
                                     itermax, abstol, ymin, int(stability), tf, _p(y), _p(g), _p(r))
    return y, int(err), int(g[0]), int(r[0])


def tridiag_integrate(a, b, c, k, y0, t0, h, t1=0.0, mode=1, nsteps=0):
    a, b, c, k = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, k))
    y = np.array(y0, dtype=np.float64)
    t = np.zeros(1)
    hh = np.zeros(1)
    lib().orc_tridiag_integrate(len(y), _p(a), _p(b), _p(c), _p(k), t0, h, t1, mode, nsteps, _p(y),
                                _p(t), _p(hh))
    return y, float(t[0]), float(hh[0])


def diffusion_integrate(B, B_row, D, split_const, x, state, t0, t1, alpha=0, leftBC=0, rightBC=0,
                        yInf=0.0, wallConst=0.0, wallVar=-1, multiplier=10.0, dt_global=2e-5,
                        h_per_var=None, nthreads=0):
    """B: (nB, N); D, split_const, state: (N, nVars) (Ember column-major). Returns new state."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    D = np.ascontiguousarray(D, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    st = np.array(state, dtype=np.float64, order="C")
    N, nVars = st.shape
    Br = np.ascontiguousarray(B_row, dtype=np.int32)
    sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
    hv = None if h_per_var is None else np.ascontiguousarray(h_per_var, dtype=np.float64)
    lib().orc_diffusion_integrate(nVars, N, _p(B), B.shape[0], _p(Br), _p(D), _p(sc), _p(x), alpha,
                                  leftBC, rightBC, yInf, wallConst, wallVar, multiplier, dt_global,
                                  _p(hv),
                                  t0, t1, _p(st), nthreads)
    return st
