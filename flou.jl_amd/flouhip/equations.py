"""Equations, numerical fluxes, operators and boundary conditions (host-side types).

These mirror the Flou types a user composes (src/FlouCommon/Euler.jl:16-31,
src/FlouCommon/LinearAdvection.jl:16-31, src/FlouSpatial/Interfaces.jl:16-23,
src/FlouSpatial/Equations/Euler.jl:69-94,167,228-306, src/FlouSpatial/Equations/
OpDivergence.jl:105-107,184-194, src/FlouSpatial/FlouSpatial.jl:85-91) and only carry
parameters: the arithmetic runs in libflouhip.  The point-wise helpers at the end
(pressure, prim2cons, ...) are host conveniences for building initial states.
"""
import numpy as np

from . import _capi as C


class LinearAdvection:
    """LinearAdvection(a...) -- nvariables = 1."""

    def __init__(self, *velocity):
        if not 1 <= len(velocity) <= 3:
            raise ValueError("Linear advection is implemented in 1D, 2D and 3D.")
        self.a = tuple(float(v) for v in velocity)
        self.nd = len(velocity)
        self.nvariables = 1

    def variablenames(self, unicode=False):
        """FlouCommon/LinearAdvection.jl:34-40."""
        return ("u",)


class EulerEquation:
    """EulerEquation{ND}(γ) -- nvariables = ND + 2."""

    def __init__(self, nd, gamma=1.4):
        if not 1 <= nd <= 3:
            raise ValueError("The Euler equations are only implemented in 1D, 2D and 3D.")
        self.nd = nd
        self.γ = float(gamma)
        self.gamma = self.γ
        self.nvariables = nd + 2

    def variablenames(self, unicode=False):
        """FlouCommon/Euler.jl:33-50."""
        vel = ("u", "v", "w")[:self.nd]
        if unicode:
            return ("ρ",) + tuple("ρ" + c for c in vel) + ("ρe",)
        return ("rho",) + tuple("rho" + c for c in vel) + ("rhoe",)


def nvariables(eq):
    return eq.nvariables


def variablenames(eq, unicode=False):
    """variablenames(equation; unicode=false) (FlouCommon/Equations.jl:59-66)."""
    return eq.variablenames(unicode=unicode)


# ---------------------------------------------------------------- numerical fluxes
class StdAverage:
    code = C.AVG_STD
    tp_code = C.TP_STDAVG


class ChandrasekharAverage:
    code = C.AVG_CHANDRASEKHAR
    tp_code = C.TP_CHANDRASEKHAR


class KennedyGruber:
    """Kennedy-Gruber split two-point flux (Kennedy & Gruber, JCP 227, 2008).  Has no
    counterpart in the reference; usable as the two-point flux of SplitDivOperator."""
    code = None
    tp_code = C.TP_KENNEDY_GRUBER


class LxF:
    """LxF(avg, intensity): avg + λ (Ql - Qr)/2 * intensity, λ = max(|u|+c) (Rusanov)."""

    def __init__(self, avg, intensity=1.0):
        self.avg, self.intensity = avg, float(intensity)
        self.diss = C.DISS_LXF


class MatrixDissipation:
    """MatrixDissipation(avg, intensity) (Equations/Euler.jl:303-380)."""

    def __init__(self, avg, intensity=1.0):
        self.avg, self.intensity = avg, float(intensity)
        self.diss = C.DISS_MATRIX


class ScalarDissipation:
    """ScalarDissipation(avg, intensity) (Equations/Euler.jl:228-301): avg - λ/2 D intensity,
    D the entropy-variable jump mapped to conservative variables with scalar λ."""

    def __init__(self, avg, intensity=1.0):
        self.avg, self.intensity = avg, float(intensity)
        self.diss = C.DISS_SCALAR


def _surface_codes(numflux):
    if isinstance(numflux, (StdAverage, ChandrasekharAverage)):
        return numflux.code, C.DISS_NONE, 1.0
    if isinstance(numflux, (LxF, MatrixDissipation, ScalarDissipation)):
        if numflux.avg.code is None:
            raise ValueError("this average cannot be used as a surface flux")
        return numflux.avg.code, numflux.diss, numflux.intensity
    raise ValueError(f"unsupported numerical flux {type(numflux).__name__}")


# ---------------------------------------------------------------- divergence operators
class StrongDivOperator:
    def __init__(self, numflux):
        self.numflux = numflux
        self.code = C.STRONG
        self.tpflux = None


class SplitDivOperator:
    """SplitDivOperator([tpflux=numflux.avg], numflux)."""

    def __init__(self, tpflux, numflux=None):
        if numflux is None:
            numflux = tpflux
            tpflux = numflux.avg
        self.tpflux, self.numflux = tpflux, numflux
        self.code = C.SPLIT


class HybridDivOperator:
    """HybridDivOperator([tpflux=numflux.avg, [fvflux=numflux]], numflux, blend)
    (Equations/OpDivergence.jl:441-470): the split form in telescopic subcell form, each
    subcell interface flux blended with a finite-volume flux by the entropy production of
    the high-order one (δ = max((√(b²+c) - b)/√(b²+c), 1/2), c = blend).  The GPU path
    takes fvflux = numflux, as the two- and three-argument constructors do."""

    def __init__(self, *args):
        if len(args) == 2:
            numflux, blend = args
            tpflux, fvflux = numflux.avg, numflux
        elif len(args) == 3:
            tpflux, numflux, blend = args
            fvflux = numflux
        elif len(args) == 4:
            tpflux, fvflux, numflux, blend = args
            if fvflux is not numflux and _surface_codes(fvflux) != _surface_codes(numflux):
                raise NotImplementedError("HybridDivOperator: a finite-volume flux other than the "
                                          "surface flux is not on the GPU path")
        else:
            raise TypeError("HybridDivOperator([tpflux, [fvflux]], numflux, blend)")
        self.tpflux, self.fvflux, self.numflux = tpflux, fvflux, numflux
        self.blend = float(blend)
        self.code = C.HYBRID


# ---------------------------------------------------------------- boundary conditions
class GenericBC:
    """GenericBC(Qext) with Qext(Qin, x, frame, time, eq) -> state.  The GPU path evaluates
    Qext once per boundary face node at set-up (time-independent exterior states)."""

    def __init__(self, Qext):
        self.Qext = Qext
        self.code = C.BC_DIRICHLET


class EulerInflowBC:
    def __init__(self, Qext):
        q = np.asarray(Qext, dtype=np.float64)
        if not 3 <= q.size <= 5:
            raise ValueError("`Qext` must have a length of 3, 4 or 5.")
        self.Qext = q
        self.code = C.BC_INFLOW


class EulerOutflowBC:
    code = C.BC_OUTFLOW


class EulerSlipBC:
    code = C.BC_SLIP


# ---------------------------------------------------------------- point-wise helpers
def pressure(Q, eq):
    """Pressure from conservative variables; Q has the variables on the last axis."""
    Q = np.asarray(Q, dtype=np.float64)
    nd = eq.nd
    s = Q[..., 1] * Q[..., 1]
    for k in range(1, nd):
        s = s + Q[..., 1 + k] * Q[..., 1 + k]
    return (eq.γ - 1) * (Q[..., nd + 1] - s / (2 * Q[..., 0]))


def vars_prim2cons(P, eq):
    """(ρ, u.., p) -> (ρ, ρu.., ρe) (FlouCommon/Euler.jl:182-195, 255-271)."""
    P = np.asarray(P, dtype=np.float64)
    nd = eq.nd
    rho = P[..., 0]
    u2 = P[..., 1] * P[..., 1]
    for k in range(1, nd):
        u2 = u2 + P[..., 1 + k] * P[..., 1 + k]
    out = np.empty(P.shape, dtype=np.float64)
    out[..., 0] = rho
    for k in range(nd):
        out[..., 1 + k] = rho * P[..., 1 + k]
    out[..., nd + 1] = P[..., nd + 1] / (eq.γ - 1) + rho * u2 / 2
    return out


def vars_cons2prim(Q, eq):
    Q = np.asarray(Q, dtype=np.float64)
    nd = eq.nd
    out = np.empty(Q.shape, dtype=np.float64)
    out[..., 0] = Q[..., 0]
    for k in range(nd):
        out[..., 1 + k] = Q[..., 1 + k] / Q[..., 0]
    out[..., nd + 1] = pressure(Q, eq)
    return out


def soundvelocity(rho, p, eq):
    return np.sqrt(eq.γ * p / rho)


def normal_shockwave(rho0, u0, p0, eq):
    """(ρ1, u1, p1) downstream of a normal shock (FlouCommon/Euler.jl:346-359)."""
    g = eq.γ
    M0 = u0 / soundvelocity(rho0, p0, eq)
    rho1 = rho0 * M0 ** 2 * (g + 1) / ((g - 1) * M0 ** 2 + 2)
    p1 = p0 * (2 * g * M0 ** 2 - (g - 1)) / (g + 1)
    M1 = np.sqrt(((g - 1) * M0 ** 2 + 2) / (2 * g * M0 ** 2 - (g - 1)))
    return np.array([rho1, M1 * soundvelocity(rho1, p1, eq), p1])
