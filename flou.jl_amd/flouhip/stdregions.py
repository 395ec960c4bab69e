"""Nodal bases and standard elements (host side, setup only).

Mirrors Flou's LagrangeBasis / DGSEMrec / StdSegment / StdQuad / StdHex
(src/FlouSpatial/StdRegions/ApproximationBases.jl:60-91, Reconstruction.jl:121-145,
StdSegment.jl:35-121, StdQuad.jl:35-93, StdHex.jl:36-112).  Nodes, weights, the Lagrange
derivative matrix and the face interpolation vectors are computed in extended precision
with the barycentric formulas (Flou fits monomials, ApproximationBases.jl:154-186, which
adds round-off of ~1e-15..1e-12 to the same quantities); the reconstruction vectors and the
Ds / D# combinations are then formed in double exactly as the reference forms them.
"""
from fractions import Fraction

import numpy as np

LD = np.longdouble


def _legendre(N, x):
    """P_N(x), P_N'(x) in extended precision (x scalar longdouble)."""
    p0, p1 = LD(1), x
    if N == 0:
        return LD(1), LD(0)
    for k in range(2, N + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    if abs(x) == 1:
        dP = LD(0.5) * N * (N + 1) * (1 if (x > 0 or N % 2 == 1) else -1)
    else:
        dP = N * (x * p1 - p0) / (x * x - 1)
    return p1, dP


def _gll(n):
    N = n - 1
    x = np.zeros(n, dtype=LD)
    x[0], x[-1] = -1, 1
    for i in range(1, N):
        xi = -np.cos(LD(np.pi) * i / N)
        for _ in range(100):
            P, dP = _legendre(N, xi)
            d2P = (2 * xi * dP - N * (N + 1) * P) / (1 - xi * xi)
            dx = dP / d2P
            xi = xi - dx
            if abs(dx) < LD(1e-22):
                break
        x[i] = xi
    w = np.array([2 / (N * (N + 1) * _legendre(N, xi)[0] ** 2) for xi in x], dtype=LD)
    return x, w


def _gl(n):
    x = np.zeros(n, dtype=LD)
    w = np.zeros(n, dtype=LD)
    for i in range(n):
        xi = -np.cos(LD(np.pi) * (i + LD(0.75)) / (n + LD(0.5)))
        for _ in range(100):
            P, dP = _legendre(n, xi)
            dx = P / dP
            xi = xi - dx
            if abs(dx) < LD(1e-22):
                break
        x[i] = xi
        P, dP = _legendre(n, xi)
        w[i] = 2 / ((1 - xi * xi) * dP * dP)
    return x, w


def _symmetrize(x, w):
    n = len(x)
    for i in range(n // 2):
        m = (x[n - 1 - i] - x[i]) / 2
        x[i], x[n - 1 - i] = -m, m
        ww = (w[i] + w[n - 1 - i]) / 2
        w[i] = w[n - 1 - i] = ww
    if n % 2:
        x[n // 2] = 0
    return x, w


class LagrangeBasis:
    """LagrangeBasis(nodetype, n): n nodes of type "GLL" (Gauss-Lobatto) or "GL" (Gauss)."""

    def __init__(self, nodetype, n):
        if nodetype in ("GLL", "GaussLobatto"):
            x, w = _gll(n)
            self.hasboundaries = True
            self.nname = "Gauss-Lobatto"
        elif nodetype in ("GL", "Gauss"):
            x, w = _gl(n)
            self.hasboundaries = False
            self.nname = "Gauss"
        else:
            raise ValueError(f"Nodes of type {nodetype} cannot be used in Lagrange bases.")
        if n < 2:
            raise ValueError("a Lagrange basis needs at least 2 nodes")
        x, w = _symmetrize(x, w)
        self._x = x
        self.ξ = x.astype(np.float64)
        self.ω = w.astype(np.float64)
        self.n = n
        bw = np.ones(n, dtype=LD)
        for j in range(n):
            for k in range(n):
                if k != j:
                    bw[j] /= (x[j] - x[k])
        self._bw = bw

    def derivative_matrix(self):
        """D_ij = l_j'(ξ_i) (barycentric, extended precision)."""
        x, bw, n = self._x, self._bw, self.n
        D = np.zeros((n, n), dtype=LD)
        for i in range(n):
            for j in range(n):
                if i != j:
                    D[i, j] = (bw[j] / bw[i]) / (x[i] - x[j])
            D[i, i] = -sum(D[i, j] for j in range(n) if j != i)
        return D.astype(np.float64)

    def interp_vector(self, xe):
        """(l_1(xe), ..., l_n(xe))."""
        x, bw, n = self._x, self._bw, self.n
        xe = LD(xe)
        hit = np.nonzero(x == xe)[0]
        if len(hit):
            v = np.zeros(n)
            v[hit[0]] = 1.0
            return v
        t = bw / (xe - x)
        return (t / t.sum()).astype(np.float64)


class DGSEMrec:
    """DGSEM reconstruction (Reconstruction.jl:121-134): ∂g = (-l(-1)./ω, +l(+1)./ω)."""

    def __init__(self, basis):
        self.basis = basis

    def reconstruction(self):
        b = self.basis
        return (-b.interp_vector(-1.0) / b.ω, b.interp_vector(1.0) / b.ω)


def _equispaced_1d(m):
    """range(-1, 1, m) as StdSegment.jl:56 forms it: the double nearest to -1 + 2k/(m-1) for
    each k (Julia's float ranges are exact up to the final rounding); [0.0] for one point."""
    if m < 1:
        raise ValueError("nequispaced must be positive")
    if m == 1:
        return np.zeros(1)
    return np.array([float(Fraction(-1) + Fraction(2 * k, m - 1)) for k in range(m)])


class _StdBase:
    ND = 0
    VTK_TYPE = 0

    def __init__(self, basis, rec, nvars=1, nequispaced=None):
        if rec.basis is not basis:
            raise ValueError("over-integration (different flux basis) is not supported")
        n = basis.n
        self.basis = basis
        self.solbasis = basis
        self.reconstruction = rec
        self.nvars = nvars
        self.n = n
        # StdSegment.jl:76-89
        self.l = (basis.interp_vector(-1.0), basis.interp_vector(1.0))
        self.D = basis.derivative_matrix()
        g = rec.reconstruction()
        self.dg = (-g[0], g[1])
        B = np.outer(self.dg[1], self.l[1]) - np.outer(self.dg[0], self.l[0])
        self.Ds = self.D - B
        self.Dsharp = 2 * self.D - B
        w1 = basis.ω
        if self.ND == 1:
            self.ω = w1.copy()
        elif self.ND == 2:
            self.ω = np.array([wx * wy for wy in w1 for wx in w1])
        else:
            self.ω = np.array([(wx * wy) * wz for wz in w1 for wy in w1 for wx in w1])
        self.ω1 = w1
        # equispaced output points (StdSegment.jl:55-58, StdQuad.jl:51-53, StdHex.jl:56-61):
        # per direction ξe and node2eq = interp_matrix(ξe, basis); the ND-D points x-fastest
        m = n if nequispaced is None else int(nequispaced)
        self.ξe1 = _equispaced_1d(m)
        self.node2eq1 = np.array([basis.interp_vector(x) for x in self.ξe1])
        grids = np.meshgrid(*([self.ξe1] * self.ND), indexing="ij")
        self.ξe = np.stack([g.ravel(order="F") for g in grids], axis=1)

    # Flou's D♯ and ∂g are `Dsharp` and `dg` here (♯ and ∂ are not valid in Python names)

    def ndofs(self):
        return self.n ** self.ND

    def nequispaced(self):
        """Number of equispaced output points of the region (StdRegions.jl:136-138)."""
        return self.ξe1.size ** self.ND

    def equisize(self):
        return (self.ξe1.size,) * self.ND

    def vtk_type(self):
        """VTK cell type: Lagrange curve 68, quadrilateral 70, hexahedron 72."""
        return self.VTK_TYPE

    def vtk_connectivities(self):
        """The cell's points in VTK Lagrange order (vertices, edges, faces, interior), 0-based
        indices into the x-fastest lattice of equispaced points -- the order of
        StdSegment.jl:171-175, StdQuad.jl:185-196 and StdHex.jl:174-199.  The reference
        lists the n-node solution lattice while writing the equispaced points; the two agree
        for the default nequispaced = n, and here the listing always follows the points."""
        m = self.ξe1.size
        if m < 2:
            raise ValueError("a VTK Lagrange cell needs at least two points per direction")
        return _vtk_lagrange_order(self.ND, m)


def _vtk_lagrange_order(nd, m):
    L = np.arange(m ** nd).reshape((m,) * nd, order="F")  # L[i, j, k] = i + m j + m² k
    inner = slice(1, m - 1)

    def F(a):  # column-major flattening (Julia's vec)
        return np.asarray(a).ravel(order="F")

    if nd == 1:
        parts = [[L[0], L[-1]], L[inner]]
    elif nd == 2:
        corners = [L[0, 0], L[-1, 0], L[-1, -1], L[0, -1]]
        edges = [L[inner, 0], L[-1, inner], L[inner, -1], L[0, inner]]
        parts = [corners] + edges + [F(L[inner, inner])]
    else:
        corners = [L[0, 0, 0], L[-1, 0, 0], L[-1, -1, 0], L[0, -1, 0],
                   L[0, 0, -1], L[-1, 0, -1], L[-1, -1, -1], L[0, -1, -1]]
        edges = []
        for z in (0, -1):
            edges += [L[inner, 0, z], L[-1, inner, z], L[inner, -1, z], L[0, inner, z]]
        edges += [L[0, 0, inner], L[-1, 0, inner], L[-1, -1, inner], L[0, -1, inner]]
        faces = [F(L[0, inner, inner]), F(L[-1, inner, inner]), F(L[inner, 0, inner]),
                 F(L[inner, -1, inner]), F(L[inner, inner, 0]), F(L[inner, inner, -1])]
        parts = [corners] + edges + faces + [F(L[inner, inner, inner])]
    return np.concatenate([np.asarray(p, dtype=np.int64).ravel() for p in parts])


class StdSegment(_StdBase):
    ND = 1
    VTK_TYPE = 68


class StdQuad(_StdBase):
    ND = 2
    VTK_TYPE = 70


class StdHex(_StdBase):
    ND = 3
    VTK_TYPE = 72
