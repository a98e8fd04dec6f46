"""ctypes binding of the ember_amd C ABI (include/ember_amd.h).

The binding loads the in-tree ``libember_amd.so``.  There is no fallback: if the
library is missing or cannot find a GPU, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import mechanism

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libember_amd.so")

EMX_QSS, EMX_BDF = 0, 1
EMX_KIN_STANDARD, EMX_KIN_INTERP = 0, 1
EMX_VARIANT_WAVE, EMX_VARIANT_PLAIN, EMX_VARIANT_WAVE_PER_POINT = 0, 1, 2
EMX_SCHED_FIFO, EMX_SCHED_COST = 0, 1
BC_FIXED_VALUE, BC_ZERO_GRADIENT, BC_CONTROL_VOLUME, BC_WALL_FLUX = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p
_sz = C.c_size_t


class Options(C.Structure):
    _fields_ = [
        ("chemistry_integrator", C.c_int32), ("kinetics_model", C.c_int32),
        ("qss_epsmin", C.c_double), ("qss_epsmax", C.c_double), ("qss_dtmin", C.c_double),
        ("qss_dtmax", C.c_double), ("qss_iteration_count", C.c_int32),
        ("qss_abstol", C.c_double), ("qss_minval", C.c_double),
        ("qss_stability_check", C.c_int32),
        ("integrator_rel_tol", C.c_double), ("integrator_momentum_abs_tol", C.c_double),
        ("integrator_energy_abs_tol", C.c_double), ("integrator_species_abs_tol", C.c_double),
        ("integrator_min_timestep", C.c_double), ("integrator_max_steps", C.c_int32),
        ("diffusion_timestep_multiplier", C.c_double), ("global_timestep", C.c_double),
        ("ignition_t_start", C.c_double), ("ignition_duration", C.c_double),
        ("ignition_energy", C.c_double), ("ignition_center", C.c_double),
        ("ignition_stddev", C.c_double), ("transport_threshold", C.c_double),
    ]


class Grid(C.Structure):
    _fields_ = [("x", _dp), ("alpha", C.c_int32), ("left_bc", C.c_int32),
                ("right_bc", C.c_int32), ("y_inf", C.c_double), ("wall_const", C.c_double),
                ("wall_var", C.c_int32)]


class ConvectionParams(C.Structure):
    _fields_ = [("alpha", C.c_int32), ("beta", C.c_int32), ("left_bc", C.c_int32), ("right_bc", C.c_int32),
                ("continuity_bc", C.c_int32), ("rVzero", C.c_double), ("Tleft", C.c_double),
                ("Yleft", C.c_void_p), ("rhou", C.c_double), ("strain_initial", C.c_double),
                ("strain_final", C.c_double), ("strain_t0", C.c_double), ("strain_dt", C.c_double),
                ("max_steps", C.c_int32)]


class SplitArrays(C.Structure):
    _fields_ = [("state", _vp), ("ddt_conv", _vp), ("ddt_diff", _vp), ("ddt_prod", _vp),
                ("ddt_cross", _vp), ("delta_diff", _vp), ("delta_prod", _vp), ("work", _vp)]


class SplitConvection(C.Structure):
    _fields_ = [("params", C.c_void_p), ("rho", _vp), ("Wmx", _vp), ("delta_conv", _vp), ("steps", _vp)]


# (name, restype, argtypes) for every entry point of include/ember_amd.h
SIGNATURES = [
    ("emx_default_options", None, [C.POINTER(Options)]),
    ("emx_create", C.c_int, [C.c_char_p, C.c_double, C.POINTER(Options), C.POINTER(_vp)]),
    ("emx_destroy", None, [_vp]),
    ("emx_last_error", C.c_char_p, []),
    ("emx_num_species", C.c_int, [_vp]),
    ("emx_num_reactions", C.c_int, [_vp]),
    ("emx_set_options", C.c_int, [_vp, C.POINTER(Options)]),
    ("emx_set_rate_multipliers", C.c_int, [_vp, _dp]),
    ("emx_set_rate_multiplier_function", C.c_int, [_vp, C.c_int, _dp]),
    ("emx_set_interp_range", C.c_int, [_vp, C.c_int, C.c_double, C.c_double]),
    ("emx_get_interp_range", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
    ("emx_set_point_positions", C.c_int, [_vp, _sz, _dp]),
    ("emx_set_stream", C.c_int, [_vp, _vp]),
    ("emx_set_launch", C.c_int, [_vp, C.c_int, C.c_int]),
    ("emx_set_kernel_variant", C.c_int, [_vp, C.c_int]),
    ("emx_set_schedule", C.c_int, [_vp, C.c_int]),
    ("emx_source_integrate", C.c_int,
     [_vp, _sz, _sz, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _ip, _ip, _ip]),
    ("emx_source_integrate_device", C.c_int,
     [_vp, _sz, _sz, _vp, _vp, _vp, C.c_double, C.c_double, _vp, _vp, _vp, _vp, _vp]),
    ("emx_source_integrate_device_noalloc", C.c_int,
     [_vp, _sz, _sz, _vp, _vp, _vp, C.c_double, C.c_double, _vp, _vp, _vp, _vp, _vp]),
    ("emx_set_adjoint_capacity", C.c_int, [_vp, C.c_int]),
    ("emx_source_adjoint_forward", C.c_int,
     [_vp, _sz, _sz, _vp, _vp, _vp, C.c_double, _vp, _vp, _vp]),
    ("emx_source_adjoint_backward", C.c_int,
     [_vp, _sz, _sz, _vp, _vp, _vp, C.c_double, _vp, _vp, _vp, _vp, _vp]),
    ("emx_rates", C.c_int, [_vp, _sz, _dp, _dp, C.c_int, _dp, _dp, _dp, _ip]),
    ("emx_diffusion_integrate", C.c_int,
     [_vp, _sz, _sz, _dp, _sz, _ip, _dp, _dp, C.POINTER(Grid), _dp, C.c_double, C.c_double, _dp]),
    ("emx_diffusion_integrate_device", C.c_int,
     [_vp, _sz, _sz, _vp, _sz, _vp, _vp, _vp, C.POINTER(Grid), _dp, C.c_double, C.c_double, _vp]),
    ("emx_tridiagonal_integrate", C.c_int,
     [_vp, _sz, _sz, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp]),
    ("emx_split_step_device", C.c_int,
     [_vp, _sz, _sz, C.POINTER(SplitArrays), _vp, _sz, _vp, _vp, C.POINTER(Grid), _vp, _dp,
      C.c_double, C.c_double, C.c_int, _vp]),
    ("emx_split_step_conv_device", C.c_int,
     [_vp, _sz, _sz, C.POINTER(SplitArrays), C.POINTER(SplitConvection), _vp, _sz, _vp, _vp,
      C.POINTER(Grid), _vp, _dp, C.c_double, C.c_double, C.c_int, _vp]),
    ("emx_density_derivative_device", C.c_int,
     [_vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp]),
    ("emx_convection_integrate_device", C.c_int,
     [_vp, _sz, _sz, C.c_void_p, _vp, _vp, _vp, _vp, C.c_double, C.c_double, _vp, _vp]),
    ("emx_convection_integrate", C.c_int,
     [_vp, _sz, _sz, C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _ip]),
    ("emx_correct_mass_fractions_device", C.c_int, [_vp, _sz, _sz, _vp]),
    ("emx_heat_release_device", C.c_int, [_vp, _sz, _sz, _vp, C.c_double, _vp, _vp]),
    ("emx_thermo_properties", C.c_int, [_vp, _sz, _sz, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("emx_transport_properties", C.c_int, [_vp, _sz, _sz, _dp, _dp, _dp, _dp]),
    ("emx_transport_properties_device", C.c_int, [_vp, _sz, _sz, _vp, _vp, _vp, _vp]),
    ("emx_diffusion_coefficients_device", C.c_int, [_vp, _sz, _sz, _vp, _vp, _vp]),
    ("emx_thermo_properties_device", C.c_int, [_vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp]),
]

_lib = None


class EmberAmdError(RuntimeError):
    pass


def lib():
    """Load libember_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EmberAmdError(f"{LIB_PATH} is missing; build it with `python -m ember_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise EmberAmdError(f"ember_amd error {rc}: {lib().emx_last_error().decode()}")


def _p(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def default_options(**kw) -> Options:
    o = Options()
    lib().emx_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


_F64, _I32 = "torch.float64", "torch.int32"


def _tptr(t, count=None, dtype=_F64):
    """Device pointer of a torch tensor (None passes through).  count: elements the call
    reads or writes through it; the tensor must hold at least that many, of dtype."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise EmberAmdError("device entry points need contiguous GPU tensors")
    if count is not None:
        if str(t.dtype) != dtype:
            raise EmberAmdError(f"expected a {dtype} tensor, got {t.dtype}")
        if t.numel() < count:
            raise EmberAmdError(f"tensor holds {t.numel()} elements, the call uses {count}")
    return C.c_void_p(t.data_ptr())


class Context:
    """One emx_context: mechanism tables on the current GPU plus options."""

    def __init__(self, mech: str = "gri30", pressure: float = mechanism.ONE_ATM,
                 options: Optional[Options] = None, **kw):
        self.table = mechanism.table_path(mech)
        self.options = options or default_options(**kw)
        h = _vp()
        _check(lib().emx_create(self.table.encode(), pressure, C.byref(self.options), C.byref(h)))
        self.h = h
        self.n_species = lib().emx_num_species(h)
        self.n_reactions = lib().emx_num_reactions(h)
        self.n_vars = self.n_species + 2
        self.rate_multipliers = None  # per-reaction multipliers last set (None: all 1)

    def close(self):
        if getattr(self, "h", None):
            lib().emx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, **kw):
        for k, v in kw.items():
            setattr(self.options, k, v)
        _check(lib().emx_set_options(self.h, C.byref(self.options)))

    def set_rate_multipliers(self, m):
        m = None if m is None else np.ascontiguousarray(m, dtype=np.float64)
        _check(lib().emx_set_rate_multipliers(self.h, _p(m)))
        self.rate_multipliers = None if m is None else m.copy()

    def set_rate_multiplier_function(self, data):
        """Ember's rate multiplier function: data = [T0, T1, c_0, ...] as
        ChebyshevFunction::setCoefficients takes them (src/scalarFunction.cpp:33-51);
        None removes it."""
        if data is None:
            _check(lib().emx_set_rate_multiplier_function(self.h, 0, None))
            return
        d = np.ascontiguousarray(data, dtype=np.float64)
        _check(lib().emx_set_rate_multiplier_function(self.h, len(d), _p(d)))

    def set_interp_range(self, n: int, Tmin: float, Tmax: float):
        """InterpKinetics table grid: n temperatures over [Tmin, Tmax] (emx_set_interp_range)."""
        _check(lib().emx_set_interp_range(self.h, int(n), float(Tmin), float(Tmax)))

    def interp_range(self):
        n = C.c_int()
        lo, hi = C.c_double(), C.c_double()
        _check(lib().emx_get_interp_range(self.h, C.byref(n), C.byref(lo), C.byref(hi)))
        return n.value, lo.value, hi.value

    def set_point_positions(self, x):
        """Grid positions of the stage's points (SourceSystem::setPosition), for the
        igniter; None clears."""
        if x is None:
            _check(lib().emx_set_point_positions(self.h, 0, None))
            return
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().emx_set_point_positions(self.h, len(x), _p(x)))

    def set_stream(self, stream_handle: int):
        _check(lib().emx_set_stream(self.h, C.c_void_p(stream_handle)))

    def set_launch(self, waves_per_block: int, blocks_per_cu: int):
        _check(lib().emx_set_launch(self.h, waves_per_block, blocks_per_cu))

    def set_kernel_variant(self, variant: int):
        """0: wavefront-per-point kernel with work queue (default), 1: plain thread-per-point
        baseline, 2: wavefront kernel with one wave per point and no queue."""
        _check(lib().emx_set_kernel_variant(self.h, variant))

    def set_schedule(self, mode: int):
        """EMX_SCHED_COST (default): most expensive points of the previous stage first;
        EMX_SCHED_FIFO: index order."""
        _check(lib().emx_set_schedule(self.h, mode))

    # ---- chemistry
    def source_integrate(self, state, dt, split_const=None, qloss=None, t_start=0.0):
        """Host arrays: state (nPoints, nVars).  Returns (state_out, status, stat1, stat2)."""
        st = np.ascontiguousarray(state, dtype=np.float64)
        n = st.shape[0]
        if st.shape[1] != self.n_vars:
            raise ValueError("state must be (nPoints, nSpecies + 2)")
        out = np.empty_like(st)
        status = np.zeros(n, np.int32)
        s1 = np.zeros(n, np.int32)
        s2 = np.zeros(n, np.int32)
        sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
        ql = None if qloss is None else np.ascontiguousarray(qloss, dtype=np.float64)
        _check(lib().emx_source_integrate(self.h, n, self.n_vars, _p(st), _p(sc), _p(ql), t_start, dt,
                                          _p(out), _p(status, _ip), _p(s1, _ip), _p(s2, _ip)))
        return out, status, s1, s2

    def source_integrate_device(self, state_in, state_out, dt, split_const=None, qloss=None,
                                status=None, stat1=None, stat2=None, qint=None, t_start=0.0):
        """torch CUDA tensors, asynchronous on the context stream."""
        n, nv = state_in.shape[0], self.n_vars
        _check(lib().emx_source_integrate_device(
            self.h, n, nv, _tptr(state_in, n * nv), _tptr(split_const, n * nv), _tptr(qloss, n), t_start,
            dt, _tptr(state_out, n * nv), _tptr(status, n, _I32), _tptr(stat1, n, _I32),
            _tptr(stat2, n, _I32), _tptr(qint, n)))

    # ---- discrete adjoint of the QSS stage (torch CUDA tensors, context stream)
    def set_adjoint_capacity(self, max_checkpoints: int):
        _check(lib().emx_set_adjoint_capacity(self.h, max_checkpoints))

    def adjoint_forward(self, state_in, state_out, dt, heat_release=None, status=None,
                        split_const=None, qloss=None):
        n, nv = state_in.shape[0], self.n_vars
        _check(lib().emx_source_adjoint_forward(
            self.h, n, nv, _tptr(state_in, n * nv), _tptr(split_const, n * nv), _tptr(qloss, n), dt,
            _tptr(state_out, n * nv), _tptr(status, n, _I32), _tptr(heat_release, n)))

    def adjoint_backward(self, state_in, dt, grad_state, grad_lnk=None, Qbar=None, ybar=None,
                         status=None, split_const=None, qloss=None):
        n, nv = state_in.shape[0], self.n_vars
        _check(lib().emx_source_adjoint_backward(
            self.h, n, nv, _tptr(state_in, n * nv), _tptr(split_const, n * nv), _tptr(qloss, n), dt,
            _tptr(Qbar, n), _tptr(ybar, n * nv), _tptr(grad_state, n * nv),
            _tptr(grad_lnk, n * self.n_reactions), _tptr(status, n, _I32)))

    def rates(self, T, Y, nonorm=False):
        T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
        Y = np.ascontiguousarray(np.atleast_2d(Y), dtype=np.float64)
        n = T.shape[0]
        wQ = np.zeros((n, self.n_species))
        wD = np.zeros_like(wQ)
        wN = np.zeros_like(wQ)
        status = np.zeros(n, np.int32)
        _check(lib().emx_rates(self.h, n, _p(T), _p(Y), int(nonorm), _p(wQ), _p(wD), _p(wN),
                               _p(status, _ip)))
        return wQ, wD, wN, status

    # ---- diffusion
    def diffusion_integrate(self, state, B, D, x, t0, t1, B_row=None, split_const=None,
                            alpha=0, left_bc=0, right_bc=0, y_inf=0.0, wall_const=0.0, wall_var=-1,
                            h_per_var=None):
        """state, D, split_const: (nPoints, nVars); B: (nB, nPoints).  Returns new state."""
        st = np.array(state, dtype=np.float64, order="C")
        N, nV = st.shape
        B = np.ascontiguousarray(np.atleast_2d(B), dtype=np.float64)
        D = np.ascontiguousarray(D, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        Br = None if B_row is None else np.ascontiguousarray(B_row, dtype=np.int32)
        sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
        hv = None if h_per_var is None else np.ascontiguousarray(h_per_var, dtype=np.float64)
        g = Grid(_p(x), alpha, left_bc, right_bc, y_inf, wall_const, wall_var)
        _check(lib().emx_diffusion_integrate(self.h, nV, N, _p(B), B.shape[0], _p(Br, _ip), _p(D),
                                             _p(sc), C.byref(g), _p(hv), t0, t1, _p(st)))
        return st

    def diffusion_integrate_device(self, state, B, D, x, t0, t1, B_row=None, split_const=None,
                                   alpha=0, left_bc=0, right_bc=0, y_inf=0.0, wall_const=0.0, wall_var=-1,
                                   h_per_var=None):
        """torch CUDA tensors; state is updated in place."""
        N, nV = state.shape
        g = Grid(C.cast(_tptr(x, N), _dp), alpha, left_bc, right_bc, y_inf, wall_const, wall_var)
        _check(lib().emx_diffusion_integrate_device(
            self.h, nV, N, _tptr(B, B.shape[0] * N), B.shape[0], _tptr(B_row, nV, _I32), _tptr(D, N * nV),
            _tptr(split_const, N * nV),
            C.byref(g), _p(None if h_per_var is None else np.ascontiguousarray(h_per_var, dtype=np.float64)),
            t0, t1, _tptr(state, N * nV)))

    def split_step_device(self, arrays: "SplitState", B, D, x, t, dt, B_row=None, qloss=None,
                          alpha=0, left_bc=0, right_bc=0, y_inf=0.0, wall_const=0.0, wall_var=1,
                          h_per_var=None, balanced=True, status=None):
        """One device-resident split step (SplitSolver::step) on arrays.state, in place."""
        N, nV = arrays.state.shape
        g = Grid(C.cast(_tptr(x, N), _dp), alpha, left_bc, right_bc, y_inf, wall_const, wall_var)
        f = arrays.fields()
        sa = SplitArrays(*(_tptr(t_, (3 if i == len(f) - 1 else 1) * N * nV) for i, t_ in enumerate(f)))
        hv = None if h_per_var is None else np.ascontiguousarray(h_per_var, dtype=np.float64)
        _check(lib().emx_split_step_device(
            self.h, nV, N, C.byref(sa), _tptr(B, B.shape[0] * N), B.shape[0], _tptr(B_row, nV, _I32),
            _tptr(D, N * nV), C.byref(g), _tptr(qloss, N), _p(hv), t, dt, 1 if balanced else 0,
            _tptr(status, N, _I32)))

    def split_step_conv_device(self, arrays: "SplitState", conv_params, B, D, x, t, dt, rho=None, Wmx=None,
                               steps=None, B_row=None, qloss=None, alpha=0, left_bc=0, right_bc=0,
                               y_inf=0.0, wall_const=0.0, wall_var=1, h_per_var=None, balanced=True,
                               status=None):
        """One device-resident split step with its convection stages (SplitSolver::step,
        D1/4 C1/2 D1/4 P D1/4 C1/2 D1/4) on arrays.state, in place; arrays made with
        SplitState(..., convection=True); conv_params from convection_params()."""
        N, nV = arrays.state.shape
        if arrays.delta_conv is None:
            raise ValueError("the split state needs convection=True")
        g = Grid(C.cast(_tptr(x, N), _dp), alpha, left_bc, right_bc, y_inf, wall_const, wall_var)
        f = arrays.fields()
        sa = SplitArrays(*(_tptr(t_, (4 * nV + 1) * N if i == len(f) - 1 else N * nV)
                           for i, t_ in enumerate(f)))
        sc = SplitConvection(C.cast(C.pointer(conv_params), C.c_void_p), _tptr(rho, N), _tptr(Wmx, N),
                             _tptr(arrays.delta_conv, N * nV), _tptr(steps, 2 * (nV - 1), _I32))
        hv = None if h_per_var is None else np.ascontiguousarray(h_per_var, dtype=np.float64)
        _check(lib().emx_split_step_conv_device(
            self.h, nV, N, C.byref(sa), C.byref(sc), _tptr(B, B.shape[0] * N), B.shape[0],
            _tptr(B_row, nV, _I32), _tptr(D, N * nV), C.byref(g), _tptr(qloss, N), _p(hv), t, dt,
            1 if balanced else 0, _tptr(status, N, _I32)))

    def density_derivative_device(self, state, ddt_conv, ddt_diff, ddt_prod, rho, Wmx, drhodt,
                                  ddt_cross=None, balanced=True):
        """FlameSolver::prepareIntegrators' drhodt (nPoints) from device matrices."""
        N, nV = state.shape
        _check(lib().emx_density_derivative_device(
            self.h, nV, N, _tptr(state, N * nV), _tptr(ddt_conv, N * nV), _tptr(ddt_diff, N * nV),
            _tptr(ddt_prod, N * nV), _tptr(ddt_cross, N * nV), 1 if balanced else 0, _tptr(rho, N),
            _tptr(Wmx, N), _tptr(drhodt, N)))

    def convection_params(self, Yleft=None, **kw):
        """emx_convection_params with Ember's defaults (planar, FixedValue ends, fixedLeft
        continuity, zero strain); Yleft is kept alive on the returned object."""
        d = dict(alpha=0, beta=1, left_bc=0, right_bc=0, continuity_bc=0, rVzero=0.0, Tleft=300.0,
                 rhou=1.0, strain_initial=0.0, strain_final=0.0, strain_t0=0.0, strain_dt=1.0, max_steps=0)
        d.update(kw)
        p = ConvectionParams(**d)
        if Yleft is not None:
            p._yleft = np.ascontiguousarray(Yleft, dtype=np.float64)
            p.Yleft = p._yleft.ctypes.data
        return p

    def convection_integrate(self, state, x, t0, t1, params=None, split_const=None, drhodt=None):
        """Convection stage on host arrays: (state_out, steps (1 + K))."""
        params = params or self.convection_params()
        st = np.ascontiguousarray(state, dtype=np.float64).copy()
        n = st.shape[0]
        steps = np.zeros(1 + self.n_species, dtype=np.int32)
        sc = None if split_const is None else np.ascontiguousarray(split_const, dtype=np.float64)
        dr = None if drhodt is None else np.ascontiguousarray(drhodt, dtype=np.float64)
        _check(lib().emx_convection_integrate(self.h, self.n_vars, n, C.byref(params),
                                              _p(np.ascontiguousarray(x, dtype=np.float64)), _p(sc), _p(dr),
                                              t0, t1, _p(st), _p(steps, _ip)))
        return st, steps

    def convection_integrate_device(self, state_in, state_out, x, t0, t1, params=None, split_const=None,
                                    drhodt=None, steps=None):
        params = params or self.convection_params()
        n = state_in.shape[0]
        _check(lib().emx_convection_integrate_device(
            self.h, self.n_vars, n, C.byref(params), _tptr(x, n), _tptr(state_in, n * self.n_vars),
            _tptr(split_const, n * self.n_vars), _tptr(drhodt, n), t0, t1, _tptr(state_out, n * self.n_vars),
            _tptr(steps, 1 + self.n_species, _I32)))

    def correct_mass_fractions_device(self, state):
        N, nV = state.shape
        _check(lib().emx_correct_mass_fractions_device(self.h, nV, N, _tptr(state, N * nV)))

    def heat_release_device(self, state, qdot, status=None, t_now=0.0):
        N, nV = state.shape
        _check(lib().emx_heat_release_device(self.h, nV, N, _tptr(state, N * nV), float(t_now),
                                             _tptr(qdot, N), _tptr(status, N, _I32)))

    def thermo_properties(self, state):
        """Host arrays: rho, Wmx, cp (N) and cpSpec, hk (N, K) of states (N, nVars)."""
        st = np.ascontiguousarray(np.atleast_2d(state), dtype=np.float64)
        N, nV = st.shape
        out = {k: np.zeros(N) for k in ("rho", "Wmx", "cp")}
        out["cpSpec"] = np.zeros((N, nV - 2))
        out["hk"] = np.zeros((N, nV - 2))
        _check(lib().emx_thermo_properties(self.h, nV, N, _p(st), _p(out["rho"]), _p(out["Wmx"]),
                                           _p(out["cp"]), _p(out["cpSpec"]), _p(out["hk"])))
        return out

    def transport_properties(self, state):
        """(mu, lambda, rhoD) at each point of a host state (emx_transport_properties)."""
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[0]
        mu, lam = np.zeros(n), np.zeros(n)
        rhoD = np.zeros((n, self.n_species))
        _check(lib().emx_transport_properties(self.h, self.n_vars, n, _p(state), _p(mu), _p(lam), _p(rhoD)))
        return mu, lam, rhoD

    def transport_properties_device(self, state, mu, lam, rhoD):
        n = state.shape[0]
        _check(lib().emx_transport_properties_device(
            self.h, self.n_vars, n, _tptr(state, n * self.n_vars), _tptr(mu, n), _tptr(lam, n),
            _tptr(rhoD, n * self.n_species)))

    def diffusion_coefficients_device(self, state, B, D):
        """B (2 x N: 1/rho, 1/(rho cp)) and D (N x nVars: mu, lambda, rho D_k) of the split
        step from the device state (emx_diffusion_coefficients_device)."""
        n = state.shape[0]
        _check(lib().emx_diffusion_coefficients_device(
            self.h, self.n_vars, n, _tptr(state, n * self.n_vars), _tptr(B, 2 * n), _tptr(D, n * self.n_vars)))

    def thermo_properties_device(self, state, rho, cp, Wmx=None, cpSpec=None, hk=None):
        """Density, mean molecular weight, cp and the species' partial molar cp and
        enthalpies at every point of the device state (N, nVars)."""
        N, nV = state.shape
        K = nV - 2
        _check(lib().emx_thermo_properties_device(self.h, nV, N, _tptr(state, N * nV), _tptr(rho, N),
                                                  _tptr(Wmx, N), _tptr(cp, N), _tptr(cpSpec, N * K),
                                                  _tptr(hk, N * K)))

    def tridiagonal_integrate(self, a, b, c, k, y, h, t0, t1):
        """Constant systems, (nSys, N) arrays.  Returns y after integrateToTime(t1)."""
        a, b, c, k = (np.ascontiguousarray(np.atleast_2d(v), dtype=np.float64) for v in (a, b, c, k))
        yy = np.array(np.atleast_2d(y), dtype=np.float64, order="C")
        nS, N = yy.shape
        h = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (nS,)))
        _check(lib().emx_tridiagonal_integrate(self.h, nS, N, _p(a), _p(b), _p(c), _p(k), _p(h), t0,
                                               t1, _p(yy)))
        return yy


class SplitState:
    """Device matrices of SplitSolver's bookkeeping for emx_split_step_device
    (torch CUDA tensors, (nPoints, nVars) = Ember's point-contiguous layout)."""

    def __init__(self, state, ddt_cross=None, convection=False):
        import torch
        self.state = state
        z = lambda: torch.zeros_like(state)
        self.ddt_conv, self.ddt_diff, self.ddt_prod = z(), z(), z()
        self.ddt_cross = ddt_cross
        self.delta_diff, self.delta_prod = z(), z()
        # emx_split_step_conv_device: delta_conv, and work for splitConstConv and drhodt
        self.delta_conv = z() if convection else None
        N, nV = state.shape
        self.work = torch.zeros((4 * nV + 1) * N if convection else 3 * N * nV, dtype=state.dtype,
                                device=state.device)

    def fields(self):
        return (self.state, self.ddt_conv, self.ddt_diff, self.ddt_prod, self.ddt_cross,
                self.delta_diff, self.delta_prod, self.work)
