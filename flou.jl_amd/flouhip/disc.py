"""MultielementDisc on the GPU and the entry points that replace Flou's hot path.

    MultielementDisc(mesh, std, equation, operator, bcs; precision)  -> flou_hip_create
    rhs_(dQ, Q, EquationConfig(disc, eq), t)   (FlouCommon.rhs!)        -> flou_hip_rhs
    timeintegrate(Q0, disc, eq, solver, tf; dt, cfl)                   -> flou_hip_timeintegrate
    get_max_dt(Q, disc, eq, cfl)                                       -> flou_hip_max_dt
    integrate(Q, disc, ivar)                                           -> flou_hip_integrate
    get_limiter(disc, eq, "zhang_shu", minval) / solver stage_limiter  -> flou_hip_limit,
                                                                          flou_hip_set_limiter
    get_monitor(disc, eq, name) / get_monitor_callback                 -> flou_hip_monitor

State buffers are torch tensors on the GPU holding Flou's (ndofs, nvars) column-major
matrix, i.e. a contiguous (nvars, ndofs) tensor (or its transpose view).  torch is only the
device allocator and stream provider here; all arithmetic runs in libflouhip.
"""
import ctypes
import time

import numpy as np

from . import _capi as C
from .equations import EulerEquation, LinearAdvection, _surface_codes
from .mesh import curved_geometry
from .stdregions import StdHex, StdQuad, StdSegment

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the runtime image
    torch = None


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class EquationConfig:
    def __init__(self, disc, equation):
        self.disc, self.equation = disc, equation


class MultielementDisc:
    """GPU MultielementDisc.  `precision` is "fp64" (plain form) or "fp32" (FP32 state and
    bulk arithmetic; dt and integrals stay FP64).  `source` is the reference's source-term
    argument (MultielementDiscontinuous.jl:29-36, applied per node by apply_sourceterm!,
    139-146): an arbitrary host function per node has no device form, so any source other
    than None is refused."""

    def __init__(self, mesh, std, equation, operator, bcs=None, source=None, precision="fp64"):
        if source is not None:
            raise NotImplementedError("source terms (apply_sourceterm!) are not available on the GPU path")
        bcs = dict(bcs or {})
        nd = mesh.nd
        if std.ND != nd or equation.nd != nd:
            raise ValueError("mesh, standard element and equation dimensions differ")
        if len(bcs) != mesh.nboundaries():
            raise ValueError("The number of BCs does not match the number of boundaries.")
        self.mesh, self.std, self.equation, self.operator = mesh, std, equation, operator
        self.precision = precision
        self.dtype = {"fp64": np.float64, "fp32": np.float32}[precision]
        self.nd, self.n = nd, std.n
        self.nvars = equation.nvariables
        self.nelements = mesh.nelements
        self.npe = std.n ** nd
        self.ndofs = self.nelements * self.npe

        d = C.Desc()
        d.nd, d.n, d.nvars = nd, std.n, self.nvars
        d.precision = C.FP64 if precision == "fp64" else C.FP32
        d.nelements, d.nfaces = mesh.nelements, mesh.nfaces
        d.nboundaries = mesh.nboundaries()
        if isinstance(equation, EulerEquation):
            d.equation, d.gamma = C.EULER, equation.γ
        elif isinstance(equation, LinearAdvection):
            d.equation = C.ADVECTION
            for k, a in enumerate(equation.a):
                d.velocity[k] = a
        else:
            raise ValueError("unsupported equation")
        d.volume_op = operator.code
        d.tp_flux = operator.tpflux.tp_code if operator.code in (C.SPLIT, C.HYBRID) else C.TP_STDAVG
        d.surface_avg, d.surface_diss, d.intensity = _surface_codes(operator.numflux)
        d.blend = getattr(operator, "blend", 1.0)

        keep = []  # arrays referenced by the descriptor must outlive create()

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a

        D = arr(std.D, np.float64)
        d.D = _ptr(D, ctypes.c_double)
        d.Ds = _ptr(arr(std.Ds, np.float64), ctypes.c_double)
        d.Dsharp = _ptr(arr(std.Dsharp, np.float64), ctypes.c_double)
        d.l_minus = _ptr(arr(std.l[0], np.float64), ctypes.c_double)
        d.l_plus = _ptr(arr(std.l[1], np.float64), ctypes.c_double)
        d.dg_minus = _ptr(arr(std.dg[0], np.float64), ctypes.c_double)
        d.dg_plus = _ptr(arr(std.dg[1], np.float64), ctypes.c_double)
        d.omega = _ptr(arr(std.ω1, np.float64), ctypes.c_double)
        d.elem_faceinds = _ptr(arr(mesh.elem_faces, np.int64), ctypes.c_int64)
        d.elem_facepos = _ptr(arr(mesh.elem_side, np.int8), ctypes.c_int8)
        d.face_eleminds = _ptr(arr(mesh.face_elems, np.int64), ctypes.c_int64)
        d.face_elempos = _ptr(arr(mesh.face_pos, np.int8), ctypes.c_int8)
        forient = getattr(mesh, "face_orientation", None)  # mesh.faces[i].orientation; None: all 0
        d.face_orientation = None if forient is None else _ptr(arr(forient, np.uint8), ctypes.c_uint8)
        d.face_boundary = _ptr(arr(mesh.face_boundary, np.int32), ctypes.c_int32)
        if getattr(mesh, "mapping", None) is None:
            ejac, emet, evol = mesh.element_geometry(std)
            ffr, fjac = mesh.face_geometry()
            d.geometry = C.GEOM_AFFINE
        else:
            ejac, emet, evol, ffr, fjac = curved_geometry(mesh, std)
            d.geometry = C.GEOM_CURVED
        d.elem_jac = _ptr(arr(ejac, np.float64), ctypes.c_double)
        d.elem_metric = _ptr(arr(emet, np.float64), ctypes.c_double)
        d.elem_volume = _ptr(arr(evol, np.float64), ctypes.c_double)
        d.face_frames = _ptr(arr(ffr, np.float64), ctypes.c_double)
        d.face_jac = _ptr(arr(fjac, np.float64), ctypes.c_double)
        self.elem_jac, self.elem_volume = ejac, evol

        h = ctypes.c_void_p()
        C.check(C.lib.flou_hip_create(ctypes.byref(d), ctypes.byref(h)), None, "MultielementDisc")
        self.h = h
        self._set_bcs(bcs)
        self._sync_stream()

    # ------------------------------------------------------------------ lifecycle
    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and C is not None and C.lib is not None:
            C.lib.flou_hip_destroy(h)
            self.h = None

    def _sync_stream(self):
        if torch is not None and torch.cuda.is_available():
            s = torch.cuda.current_stream().cuda_stream
            C.check(C.lib.flou_hip_set_stream(self.h, ctypes.c_void_p(s)), self.h)

    def set_option(self, which, value):
        C.check(C.lib.flou_hip_set_option(self.h, which, value), self.h)

    def uses_fast_kernel(self):
        """True when this (FP32) discretisation runs the fast fused kernel."""
        return C.lib.flou_hip_set_option(self.h, 2, 0) == 1

    def uses_fast_vjp(self):
        """True when this (FP32) discretisation's adjoint runs the fast fused kernel."""
        return C.lib.flou_hip_set_option(self.h, 3, 0) == 1

    def _set_bcs(self, bcs):
        mesh = self.mesh
        for name, bc in bcs.items():
            orig = int(name) - 1
            if orig not in mesh.bdalive:
                raise ValueError(f"unknown boundary {name}")
            ib = mesh.bdalive.index(orig)
            if bc.code == C.BC_DIRICHLET:
                fx = mesh.face_node_coords(self.std)
                npf = fx.shape[1]
                data = np.zeros((mesh.nfaces, npf, self.nvars))
                for f in np.nonzero(mesh.face_boundary == ib)[0]:
                    for i in range(npf):
                        data[f, i] = np.asarray(bc.Qext(None, fx[f, i], None, 0.0, self.equation),
                                                dtype=np.float64)
                C.check(C.lib.flou_hip_set_bc(self.h, ib, bc.code, _ptr(np.ascontiguousarray(data), ctypes.c_double)), self.h)
            elif bc.code == C.BC_INFLOW:
                q = np.ascontiguousarray(bc.Qext, dtype=np.float64)
                C.check(C.lib.flou_hip_set_bc(self.h, ib, bc.code, _ptr(q, ctypes.c_double)), self.h)
            else:
                C.check(C.lib.flou_hip_set_bc(self.h, ib, bc.code, None), self.h)

    # ------------------------------------------------------------------ state helpers
    def torch_dtype(self):
        return torch.float64 if self.precision == "fp64" else torch.float32

    def new_state(self):
        """Uninitialised device state, a (nvars, ndofs) tensor (Flou's Q.data transposed)."""
        return torch.empty((self.nvars, self.ndofs), dtype=self.torch_dtype(), device="cuda")

    def to_device(self, Q):
        """Host (ndofs, nvars) array -> device state tensor."""
        Q = np.asarray(Q)
        if Q.shape != (self.ndofs, self.nvars):
            raise ValueError(f"state must be ({self.ndofs}, {self.nvars})")
        return torch.from_numpy(np.ascontiguousarray(Q.T.astype(self.dtype))).to("cuda")

    def to_host(self, Qd):
        torch.cuda.synchronize()
        return Qd.detach().to("cpu").numpy().T.copy()

    def _dptr(self, t):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError("state buffers must be CUDA (HIP) torch tensors")
        if t.dtype != self.torch_dtype():
            raise TypeError(f"state dtype must be {self.torch_dtype()}")
        if t.numel() != self.ndofs * self.nvars:
            raise ValueError("state has the wrong size")
        if t.shape == (self.nvars, self.ndofs) and t.is_contiguous():
            return ctypes.c_void_p(t.data_ptr())
        if t.shape == (self.ndofs, self.nvars) and t.stride() == (1, self.ndofs):
            return ctypes.c_void_p(t.data_ptr())
        raise ValueError("state must be a contiguous (nvars, ndofs) tensor or its transpose")

    def check(self):
        """Synchronise and raise DomainError if a kernel met a non-physical state."""
        C.check(C.lib.flou_hip_check(self.h), self.h)


def rhs_(dQ, Q, config, t=0.0):
    """FlouCommon.rhs!(dQ, Q, EquationConfig(disc, equation), t) on the GPU."""
    disc = config.disc
    disc._sync_stream()
    C.check(C.lib.flou_hip_rhs(disc.h, disc._dptr(dQ), disc._dptr(Q), float(t)), disc.h, "rhs!")
    return None


def rk_stage(disc, scheme, stage, dt, u_in, u_out, u_base=None, reg=None, t=0.0):
    disc._sync_stream()
    C.check(C.lib.flou_hip_rk_stage(disc.h, scheme, stage, float(dt), float(t), disc._dptr(u_in),
                                    disc._dptr(u_base) if u_base is not None else None,
                                    disc._dptr(reg) if reg is not None else None,
                                    disc._dptr(u_out)), disc.h, "rk_stage")


class SSPRK33:
    """OrdinaryDiffEq's SSPRK33; `stage_limiter` is what get_limiter_callback returns."""
    code = C.SSPRK33
    nstages = 3

    def __init__(self, stage_limiter=None):
        self.stage_limiter = stage_limiter


class CarpenterKennedy2N54:
    code = C.CK2N54
    nstages = 5

    def __init__(self, stage_limiter=None, williamson_condition=False):
        self.stage_limiter = stage_limiter
        self.williamson_condition = williamson_condition


class ORK256:
    code = C.ORK256
    nstages = 5

    def __init__(self, stage_limiter=None, williamson_condition=False):
        self.stage_limiter = stage_limiter
        self.williamson_condition = williamson_condition


def _solver(solver):
    return solver() if isinstance(solver, type) else solver


def _solver_code(solver):
    return _solver(solver).code


# -- limiters and monitors (Equations/Euler.jl:541-654, FlouTime.jl:107-148) ------------

def list_limiters(disc, equation):
    return ("zhang_shu",) if isinstance(equation, EulerEquation) else ()


class ZhangShuLimiter:
    """get_limiter(disc, eq, :zhang_shu, minval) (Euler.jl:612-654): called as
    limiter(Q, disc, equation), it limits the device state Q in place."""

    def __init__(self, minval):
        if minval is None:
            raise ValueError("The minimum value must be specified when using the limiter of Zhang & Shu.")
        self.minval = float(minval)

    def __call__(self, Q, disc, equation=None):
        disc._sync_stream()
        C.check(C.lib.flou_hip_limit(disc.h, disc._dptr(Q), self.minval), disc.h, "zhang_shu limiter")


def get_limiter(disc, equation, name, p=None):
    if name != "zhang_shu":
        raise ValueError(f"Unknown limiter '{name}'.")
    if not isinstance(equation, EulerEquation):
        raise ValueError("the Zhang & Shu limiter is defined for the Euler equations")
    return ZhangShuLimiter(p)


def get_limiter_callback(disc, equation, name, p=None):
    """get_limiter_callback (FlouTime.jl:136-148): the stage limiter handed to a solver,
    e.g. ORK256(stage_limiter=get_limiter_callback(disc, eq, "zhang_shu", 1e-10)); the
    library then applies it after every stage."""
    return get_limiter(disc, equation, name, p)


_MONITORS = {"kinetic_energy": C.MONITOR_KINETIC_ENERGY, "energy": C.MONITOR_KINETIC_ENERGY,
             "entropy": C.MONITOR_ENTROPY}


def list_monitors(disc, equation):
    return ("kinetic_energy", "entropy") if isinstance(equation, EulerEquation) else ()


class _Monitor:
    def __init__(self, kind):
        self.kind = kind

    def __call__(self, Q, disc, equation=None):
        disc._sync_stream()
        out = ctypes.c_double()
        C.check(C.lib.flou_hip_monitor(disc.h, disc._dptr(Q), self.kind, ctypes.byref(out)), disc.h, "monitor")
        return out.value


def get_monitor(disc, equation, name, p=None):
    """get_monitor(disc, eq, name) (Euler.jl:545-590): a function (Q, disc, eq) -> Σ ∫ s(Q),
    s the kinetic energy ("kinetic_energy", or the reference's own key "energy") or the
    mathematical entropy ("entropy"), accumulated in FP64."""
    if not isinstance(equation, EulerEquation) or name not in _MONITORS:
        raise ValueError(f"Unknown monitor '{name}'.")
    return _Monitor(_MONITORS[name])


class MonitorOutput:
    """MonitorOutput (FlouTime.jl:107-111): time, iteration and value after each selected step."""

    def __init__(self):
        self.time, self.iter, self.value = [], [], []


class MonitorCallback:
    def __init__(self, func, output, iter=True):
        self.func, self.output, self.iter = func, output, iter

    def selected(self, it):
        return self.iter is True or it in self.iter

    def __call__(self, Q, disc, equation, t, it):
        if self.selected(it):
            self.output.time.append(t)
            self.output.iter.append(it)
            self.output.value.append(self.func(Q, disc, equation))


def get_monitor_callback(timetype, valuetype, disc, equation, name, p=None, iter=True):
    """get_monitor_callback (FlouTime.jl:113-134) -> (callback, MonitorOutput); pass the
    callback to timeintegrate(..., callback=...)."""
    out = MonitorOutput()
    return MonitorCallback(get_monitor(disc, equation, name, p), out, iter), out


def make_callback_list(*callbacks):
    if not callbacks:
        return None
    return callbacks[0] if len(callbacks) == 1 else list(callbacks)


def timeintegrate(Q0, disc, equation, solver, tfinal, dt, cfl=None, t0=0.0, callback=None):
    """timeintegrate(Q0, disc, equation, solver, tf; dt, adaptive=false) -- FlouTime.jl:34-56.

    Integrates in place on the device tensor Q0 with fixed steps of `dt` (the last step is
    shortened to land on tfinal, as OrdinaryDiffEq does).  With `cfl`, every step uses
    min(get_max_dt(Q, cfl), dt) (get_cfl_callback(cfl, dt), FlouTime.jl:92-104).  A
    solver's stage_limiter is applied after every stage; monitor callbacks run after every
    step they select.  Returns (Q0, elapsed seconds).  Raises DomainError on a
    non-physical state (the reference logs "Simulation crashed!")."""
    solver = _solver(solver)
    code = solver.code
    limiter = getattr(solver, "stage_limiter", None)
    cbs = [] if callback is None else (list(callback) if isinstance(callback, (list, tuple)) else [callback])
    disc._sync_stream()
    if limiter is not None:
        C.check(C.lib.flou_hip_set_limiter(disc.h, C.LIMITER_ZHANG_SHU, limiter.minval), disc.h, "stage limiter")
    try:
        if cbs:
            return _timeintegrate_steps(Q0, disc, equation, code, tfinal, dt, cfl, t0, cbs)
        return _timeintegrate(Q0, disc, code, tfinal, dt, cfl, t0)
    finally:
        if limiter is not None:
            C.lib.flou_hip_set_limiter(disc.h, C.LIMITER_NONE, 0.0)


def _timeintegrate_steps(Q0, disc, equation, code, tfinal, dt, cfl, t0, cbs):
    ptr = disc._dptr(Q0)
    t = ctypes.c_double(float(t0))
    tic = time.perf_counter()
    it = 0
    for cb in cbs:  # DiscreteCallback's `initialize` (the save callback writes iteration 0)
        init = getattr(cb, "initialize", None)
        if init is not None:
            init(Q0, disc, equation, t.value, it)
    while t.value < float(tfinal) * (1 - 1e-14):
        step = min(float(dt), float(tfinal) - t.value)
        C.check(C.lib.flou_hip_timeintegrate(disc.h, code, ptr, step, 1, float(cfl or 0.0),
                                             ctypes.byref(t)), disc.h, "timeintegrate")
        it += 1
        for cb in cbs:
            cb(Q0, disc, equation, t.value, it)
    torch.cuda.synchronize()
    return Q0, time.perf_counter() - tic


def _timeintegrate(Q0, disc, code, tfinal, dt, cfl, t0):
    ptr = disc._dptr(Q0)
    t = ctypes.c_double(float(t0))
    tic = time.perf_counter()
    if cfl is None:
        span = float(tfinal) - float(t0)
        nfull = int(np.floor(span / dt * (1 + 1e-12)))
        rest = span - nfull * dt
        C.check(C.lib.flou_hip_timeintegrate(disc.h, code, ptr, float(dt), nfull, 0.0,
                                             ctypes.byref(t)), disc.h, "timeintegrate")
        if rest > 1e-12 * max(1.0, abs(float(tfinal))):
            C.check(C.lib.flou_hip_timeintegrate(disc.h, code, ptr, float(rest), 1, 0.0,
                                                 ctypes.byref(t)), disc.h, "timeintegrate")
    else:
        while t.value < float(tfinal) * (1 - 1e-14):
            step = min(float(dt), float(tfinal) - t.value)
            C.check(C.lib.flou_hip_timeintegrate(disc.h, code, ptr, step, 1, float(cfl),
                                                 ctypes.byref(t)), disc.h, "timeintegrate")
    torch.cuda.synchronize()
    return Q0, time.perf_counter() - tic


def get_max_dt(Q, disc, equation, cfl):
    """get_max_dt(q, disc, eq, cfl) (MultielementDiscontinuous.jl:162-178), FP64."""
    disc._sync_stream()
    out = ctypes.c_double()
    C.check(C.lib.flou_hip_max_dt(disc.h, disc._dptr(Q), float(cfl), ctypes.byref(out)), disc.h, "get_max_dt")
    return out.value


def integrate(Q, disc, ivar=0):
    """integrate(f, disc, ivar) (MultielementDiscontinuous.jl:148-160), FP64; ivar 0-based."""
    disc._sync_stream()
    out = ctypes.c_double()
    C.check(C.lib.flou_hip_integrate(disc.h, disc._dptr(Q), int(ivar), ctypes.byref(out)), disc.h, "integrate")
    return out.value


def std_for(nd):
    return {1: StdSegment, 2: StdQuad, 3: StdHex}[nd]
