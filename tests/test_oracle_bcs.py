"""Pins of the oracle's Euler boundary conditions (EulerInflowBC, EulerOutflowBC,
EulerSlipBC; src/FlouSpatial/Equations/Euler.jl:69-94, applied by applyBCs!,
src/FlouSpatial/Interfaces.jl:25-49).

The reference holds no known answer for these three boundaries, so each is fixed by what
the mathematics of the boundary forces, chosen so that a flipped sign, a swapped momentum
component or a wrong source of the exterior state fails a check:
  * point form on random rotated frames: a slip wall keeps ρ and ρe, keeps the tangential
    momentum and reverses the normal one (m_ext = m - 2(m·n)n, a vector formula that does
    not go through the face rotation); outflow returns the trace; inflow its constant;
  * mirror images: a box with slip walls at y = 0 and 1 has the rhs of the upper half of
    the periodic box [0,1] x [-1,1] holding the mirrored state, where the walls are
    interior faces between mirror-image elements;
  * equivalences: outflow equals GenericBC fed the interior trace (1-D) and inflow equals
    GenericBC fed the same constant (2-D); the GenericBC branch is pinned by the
    reference's Sod tube answer;
  * balance: Σ Jω dQ equals minus the boundary integral of the numerical flux between the
    trace and the exterior state computed here in numpy (interior faces cancel);
  * wavy slip walls on a curved mesh: no mass or energy crosses a slip wall in any frame,
    so both totals are conserved while the momentum is not.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from cases import GAMMA, oracle_for, random_euler_state

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "flou.jl_amd"))
from flouhip import mesh as M  # noqa: E402
from flouhip import stdregions as S  # noqa: E402


def _euler(nd, n, nel, finish, periodic, tp=O.TP_CHANDRASEKHAR, sdiss=O.DISS_LXF):
    return oracle_for(nd, n, nel, [0.0] * nd, finish, periodic, O.EQ_EULER, [GAMMA], O.VOL_SPLIT,
                      tp, O.AVG_STD, sdiss, 1.0)


def _random_frame(rng, nd):
    A, _ = np.linalg.qr(rng.standard_normal((nd, nd)))
    return A.T.copy()  # rows n, t, (b): orthonormal


def _state(rng, nd):
    rho = rng.uniform(0.5, 2.0)
    u = rng.uniform(-1, 1, nd)
    p = rng.uniform(0.5, 2.0)
    return np.concatenate([[rho], rho * u, [p / (GAMMA - 1) + rho * (u @ u) / 2]])


# ------------------------------------------------------------------ point form
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_slip_reflects_normal_momentum_in_any_frame(nd):
    d = _euler(nd, 3, [2] * nd, [1.0] * nd, [0] + [1] * (nd - 1))
    d.set_bc(0, O.BC_SLIP)
    rng = np.random.default_rng(10 + nd)
    for _ in range(20):
        fr = _random_frame(rng, nd)
        frame = np.zeros((3, nd))
        frame[:nd] = fr
        Q = _state(rng, nd)
        X = d.bc_point(0, Q, frame.reshape(-1))
        nrm = fr[0]
        m = Q[1:nd + 1]
        assert X[0] == Q[0] and X[nd + 1] == Q[nd + 1]
        np.testing.assert_allclose(X[1:nd + 1], m - 2 * (m @ nrm) * nrm, rtol=0, atol=4e-16 * np.abs(m).max() * 4)
        # tangential components kept, normal reversed
        for k in range(1, nd):
            assert abs(X[1:nd + 1] @ fr[k] - m @ fr[k]) <= 1e-15 * np.abs(m).max() * 4
        assert abs(X[1:nd + 1] @ nrm + m @ nrm) <= 1e-15 * np.abs(m).max() * 4


@pytest.mark.parametrize("nd", [2, 3])
def test_outflow_returns_trace_and_inflow_its_constant(nd):
    d = _euler(nd, 3, [2] * nd, [1.0] * nd, [0] + [1] * (nd - 1))
    qin = np.array([1.2, 0.3, -0.2, 0.1, 2.5])[:nd + 2]
    qin[nd + 1] = 2.5
    d.set_bc(0, O.BC_INFLOW, qin)
    d.set_bc(1, O.BC_OUTFLOW)
    rng = np.random.default_rng(3)
    for _ in range(10):
        frame = np.zeros((3, nd))
        frame[:nd] = _random_frame(rng, nd)
        Q = _state(rng, nd)
        assert np.array_equal(d.bc_point(1, Q, frame.reshape(-1)), Q)
        assert np.array_equal(d.bc_point(0, Q, frame.reshape(-1)), qin)


# ------------------------------------------------------------------ mirror images
def _mirror_y(Qh, nd, n, nel):
    """State of the lower half [-1, 0] mirroring Qh of the upper half [0, 1] about y = 0.
    Elements are x-fastest then y (then z); nodes likewise inside an element."""
    nx, ny = nel[0], nel[1]
    nz = nel[2] if nd == 3 else 1
    npe = n ** nd
    nv = nd + 2
    Q4 = Qh.reshape(nz, ny, nx, npe, nv)  # element (k, j, i), node, var  (C order of rows)
    Q4 = Q4.reshape(nz, ny, nx, *(([n] * (nd - 2)) + [n, n]), nv)  # node (.., iy, ix)
    M_ = Q4[:, ::-1]                       # element row j -> ny-1-j
    M_ = np.flip(M_, axis=3 + (nd - 2))    # node iy -> n-1-iy
    M_ = M_.copy()
    M_[..., 2] = -M_[..., 2]               # ρv -> -ρv
    return M_.reshape(-1, nv)


@pytest.mark.parametrize("nd,n,tp,sdiss", [(2, 4, O.TP_CHANDRASEKHAR, O.DISS_LXF),
                                           (2, 5, O.TP_KENNEDY_GRUBER, O.DISS_LXF),
                                           (3, 4, O.TP_CHANDRASEKHAR, O.DISS_LXF),
                                           (3, 3, O.TP_STDAVG, O.DISS_NONE)])
def test_slip_walls_equal_mirror_image(nd, n, tp, sdiss):
    nel = [3, 2, 2][:nd]
    walls = _euler(nd, n, nel, [1.0] * nd, [1, 0, 1][:nd], tp, sdiss)
    walls.set_bc(0, O.BC_SLIP)
    walls.set_bc(1, O.BC_SLIP)
    X, _ = walls.coords()
    Qh = random_euler_state(X, nd, seed=4, amp=0.2)

    neld = list(nel)
    neld[1] *= 2
    start = [0.0, -1.0, 0.0][:nd]
    dbl = oracle_for(nd, n, neld, start, [1.0] * nd, [1] * nd, O.EQ_EULER, [GAMMA], O.VOL_SPLIT,
                     tp, O.AVG_STD, sdiss, 1.0)
    # rows of the doubled mesh: element layers y in [-1, 0] first, then the upper half
    npe, nv = n ** nd, nd + 2
    lower = _mirror_y(Qh, nd, n, nel)
    nz = nel[2] if nd == 3 else 1
    up = Qh.reshape(nz, nel[1] * nel[0] * npe, nv)
    lo = lower.reshape(nz, nel[1] * nel[0] * npe, nv)
    Qd = np.asfortranarray(np.concatenate([lo, up], axis=1).reshape(-1, nv))
    Xd, _ = dbl.coords()
    Xm = Xd.reshape(nz, 2, -1, nd)[:, 1].reshape(-1, nd)
    assert np.abs(Xm - X).max() < 1e-14  # the upper half of the doubled mesh is the walled box

    r_walls = walls.rhs(Qh)
    r_dbl = np.asarray(dbl.rhs(Qd)).reshape(nz, 2, -1, nv)[:, 1].reshape(-1, nv)
    scale = np.abs(r_walls).max()
    assert np.abs(r_walls - r_dbl).max() <= 1e-13 * scale
    # the rhs of the mirror half is the mirror of the rhs (a check of the construction)
    r_lo = np.asarray(dbl.rhs(Qd)).reshape(nz, 2, -1, nv)[:, 0].reshape(-1, nv)
    assert np.abs(r_lo - _mirror_y(r_walls, nd, n, nel)).max() <= 1e-13 * scale


# ------------------------------------------------------------------ equivalences
def test_outflow_equals_genericbc_fed_the_trace_1d():
    n, ne = 5, 6
    a = _euler(1, n, [ne], [1.0], [0])
    b = _euler(1, n, [ne], [1.0], [0])
    X, FX = a.coords()
    Q = random_euler_state(X, 1, seed=5, amp=0.2)
    a.set_bc(0, O.BC_OUTFLOW)
    a.set_bc(1, O.BC_OUTFLOW)
    # Dirichlet data per face node: the interior trace at x = 0 and x = 1 (GLL end nodes)
    bc = np.zeros((b.nf, b.npf, 3))
    for f in range(b.nf):
        x = FX[f, 0, 0]
        if abs(x) < 1e-12:
            bc[f, 0] = Q[0]
        elif abs(x - 1.0) < 1e-12:
            bc[f, 0] = Q[-1]
    b.set_bc(0, O.BC_DIRICHLET, bc)
    b.set_bc(1, O.BC_DIRICHLET, bc)
    assert np.array_equal(a.rhs(Q), b.rhs(Q))


def test_inflow_equals_genericbc_fed_the_constant_2d():
    n, nel = 4, [3, 2]
    a = _euler(2, n, nel, [1.0, 1.0], [1, 0])
    b = _euler(2, n, nel, [1.0, 1.0], [1, 0])
    X, _ = a.coords()
    Q = random_euler_state(X, 2, seed=6, amp=0.2)
    qin = np.array([1.1, 0.2, 0.3, 2.6])
    a.set_bc(0, O.BC_INFLOW, qin)
    a.set_bc(1, O.BC_OUTFLOW)
    b.set_bc(0, O.BC_DIRICHLET, np.broadcast_to(qin, (b.nf, b.npf, 4)).copy())
    b.set_bc(1, O.BC_OUTFLOW)
    assert np.array_equal(a.rhs(Q), b.rhs(Q))
    # a free stream equal to the inflow state is steady
    Qc = np.asfortranarray(np.broadcast_to(qin, Q.shape).copy())
    r = a.rhs(Qc)
    assert np.abs(r).max() <= 1e-13 * np.abs(qin).max()


# ------------------------------------------------------------------ boundary flux balance
def _reflect(Q, nrm):
    m = Q[1:-1]
    out = Q.copy()
    out[1:-1] = m - 2 * (m @ nrm) * nrm
    return out


@pytest.mark.parametrize("kinds", [("slip", "slip"), ("inflow", "outflow"), ("outflow", "slip")])
def test_boundary_flux_balance_2d(kinds):
    """Σ_e Σ_i Jω dQ = -Σ_boundary nodes ω jac F*(Q_in, Q_ext)·n, with Q_ext formed here."""
    n, nel, L = 5, [3, 4], [1.5, 1.0]
    d = _euler(2, n, nel, L, [1, 0])
    qin = np.array([1.1, 0.2, 0.3, 2.6])
    codes = {"slip": O.BC_SLIP, "inflow": O.BC_INFLOW, "outflow": O.BC_OUTFLOW}
    for ib, k in enumerate(kinds):
        d.set_bc(ib, codes[k], qin if k == "inflow" else None)
    X, _ = d.coords()
    Q = random_euler_state(X, 2, seed=7, amp=0.2)
    r = d.rhs(Q)
    w = O.build_ops(n, "GLL")["w"]
    dx, dy = L[0] / nel[0], L[1] / nel[1]
    J = dx * dy / 4
    wn = np.outer(w, w).reshape(-1)  # node weight, x fastest
    total = np.array([np.sum(J * np.tile(wn, nel[0] * nel[1]) * r[:, v]) for v in range(4)])

    expect = np.zeros(4)
    npe = n * n
    for j, (row, iy, nrm) in enumerate([(0, 0, np.array([0.0, -1.0])), (nel[1] - 1, n - 1, np.array([0.0, 1.0]))]):
        kind = kinds[j]
        frame = np.array([nrm, [-nrm[1], nrm[0]], [0.0, 0.0]]).reshape(-1)
        for ex in range(nel[0]):
            e = row * nel[0] + ex
            for ix in range(n):
                Qn = Q[e * npe + iy * n + ix]
                Qe = _reflect(Qn, nrm) if kind == "slip" else (qin if kind == "inflow" else Qn)
                Fn = d.numflux(Qn, Qe, frame)
                expect -= w[ix] * (dx / 2) * Fn
    scale = np.abs(expect).max() + np.abs(total).max()
    assert np.abs(total - expect).max() <= 1e-13 * scale
    if "slip" == kinds[0] == kinds[1]:
        # nothing but the wall pressure acts: mass, x momentum and energy are conserved
        wr = np.abs(J * np.tile(wn, nel[0] * nel[1])[:, None] * r).sum(axis=0)
        for v in (0, 1, 3):
            assert abs(total[v]) <= 1e-13 * wr[v]
        assert abs(total[2]) > 1e-3 * wr[2]


# ------------------------------------------------------------------ curved slip walls
def test_wavy_slip_walls_conserve_mass_and_energy():
    nd, n, ne = 2, 5, 4
    b = S.LagrangeBasis("GLL", n)
    std = S.StdQuad(b, S.DGSEMrec(b), nd + 2)
    m = M.CartesianMesh(nd, (0.0, 0.0), (1.0, 1.0), (ne, ne))
    m.apply_periodicBCs(("1", "2"))

    def wavy(X):
        Y = X.copy()
        Y[:, 1] = X[:, 1] + 0.08 * np.sin(2 * np.pi * X[:, 0])
        return Y

    m.apply_mapping(wavy)
    J, met, ev, fr, fj = M.curved_geometry(m, std)
    d = O.OracleDisc(nd, n, [ne, ne], [0.0, 0.0], [1.0, 1.0], [1, 0], O.build_ops(n, "GLL"),
                     O.EQ_EULER, [GAMMA], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF, 1.0)
    d.set_curved(met, J, fr.reshape(-1), fj, ev)
    d.set_bc(0, O.BC_SLIP)
    d.set_bc(1, O.BC_SLIP)
    # the wall frames really are rotated
    frames = fr.reshape(-1, 3, nd)
    assert np.abs(frames[:, 0, 0]).max() > 0.2
    X = m.node_coords(std)
    Q = random_euler_state(X, nd, seed=8, amp=0.2)
    r = d.rhs(Q)
    w1 = O.build_ops(n, "GLL")["w"]
    w = np.outer(w1, w1).reshape(-1)
    jw = J * np.tile(w, ne * ne)
    tot = [float(np.sum(jw * r[:, v])) for v in range(4)]
    scale = float(np.sum(jw * np.abs(r[:, 0])))
    assert abs(tot[0]) <= 1e-13 * scale
    assert abs(tot[3]) <= 1e-13 * float(np.sum(jw * np.abs(r[:, 3])))
    # the wavy walls push on the fluid: momentum changes
    assert max(abs(tot[1]), abs(tot[2])) > 1e-6 * float(np.sum(jw * np.abs(r[:, 2])))
