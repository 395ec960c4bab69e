"""Pins of the CPU oracle (the reference the HIP path is compared with).

Each check is chosen so that a plausible mistake (a dropped term, a wrong sign, a
transposed operand, a wrong index) fails it:
  * the reference's own known answers (test/runtests.jl:26,30,37-38);
  * closed forms (GLL nodes, exactness of D, the SBP identity, RK order conditions);
  * invariants of the discretisation (free-stream preservation, discrete conservation,
    semi-discrete entropy conservation of the EC split form, Tadmor's condition);
  * convergence order on the exact solution of examples/src/Convergence.jl:31-42.
"""
import numpy as np
import pytest

import oracle_lib as O
from cases import GAMMA, isapprox, oracle_for, random_euler_state, reference_answer


# ------------------------------------------------------------------ reference known answers
def _sod():
    n, g = 4, GAMMA
    d = oracle_for(1, n, [20], [0.0], [1.0], [0], O.EQ_EULER, [g], O.VOL_SPLIT,
                   O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1.0)
    X, FX = d.coords()

    def qext(x):
        rho, u, p = (1.0, 0.0, 100.0) if x < 0.5 else (0.125, 0.0, 10.0)
        return np.array([rho, rho * u, p / (g - 1) + rho * (u * u) / 2])

    Q = np.asfortranarray(np.array([qext(x) for x in X[:, 0]]))
    bc = np.zeros((d.nf, d.npf, 3))
    for f in range(d.nf):
        bc[f, 0] = qext(FX[f, 0, 0])
    d.set_bc(0, O.BC_DIRICHLET, bc)
    d.set_bc(1, O.BC_DIRICHLET, bc)
    return d, Q


def test_sod_tube_golden():
    """SodTube1D (test/tests.jl:90-134) -> test/runtests.jl:37-38 at rtol 1e-7."""
    ref = reference_answer("sod_tube_1d")
    d, Q = _sod()
    u = d.rk_run(O.RK_ORK256, Q, 1e-4, 180)
    assert isapprox(u.min(), ref["min"], ref["rtol"]) and isapprox(u.max(), ref["max"], ref["rtol"])
    # far tighter than the reference asks: only round-off separates the two codes
    assert abs(u.min() - ref["min"]) / abs(ref["min"]) < 1e-12
    assert abs(u.max() - ref["max"]) / abs(ref["max"]) < 1e-13


def _bump(X, x0, s):
    return np.exp(sum(-(X[:, k] - x0[k]) ** 2 / (2 * s[k] ** 2) for k in range(X.shape[1])))


def test_advection_1d_full_period():
    """Advection1D (test/tests.jl:16-51) -> test/runtests.jl:26."""
    d = oracle_for(1, 5, [20], [0.0], [1.0], [1], O.EQ_ADVECTION, [2.0], O.VOL_STRONG, 0, 0,
                   O.DISS_LXF, 1.0, nodes="GL")
    X, _ = d.coords()
    Q = np.asfortranarray(_bump(X, [0.5], [0.1])[:, None])
    u = d.rk_run(O.RK_ORK256, Q, 1e-3, 500)
    rtol = reference_answer("advection_1d")["rtol"]
    assert np.linalg.norm(u - Q) <= rtol * max(np.linalg.norm(u), np.linalg.norm(Q))


def test_advection_2d_full_period():
    """Advection2D (test/tests.jl:53-88) -> test/runtests.jl:30."""
    d = oracle_for(2, 5, [20, 10], [0.0, 0.0], [1.5, 2.0], [1, 1], O.EQ_ADVECTION, [3.0, 4.0],
                   O.VOL_STRONG, 0, 0, O.DISS_LXF, 1.0, nodes="GL")
    X, _ = d.coords()
    Q = np.asfortranarray(_bump(X, [0.75, 1.0], [0.2, 0.2])[:, None])
    u = d.rk_run(O.RK_ORK256, Q, 1e-3, 500)
    rtol = reference_answer("advection_2d")["rtol"]
    assert np.linalg.norm(u - Q) <= rtol * max(np.linalg.norm(u), np.linalg.norm(Q))


# ------------------------------------------------------------------ operators
def test_gll_nodes_closed_form():
    assert np.allclose(O.build_ops(3, "GLL")["xi"], [-1, 0, 1], atol=0, rtol=0)
    x4 = O.build_ops(4, "GLL")["xi"]
    assert np.allclose(x4, [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], rtol=0, atol=2e-16)
    x5 = O.build_ops(5, "GLL")
    assert np.allclose(x5["xi"], [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], rtol=0, atol=2e-16)
    assert np.allclose(x5["w"], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], rtol=0, atol=3e-16)
    xg = O.build_ops(3, "GL")
    assert np.allclose(xg["xi"], [-np.sqrt(3 / 5), 0, np.sqrt(3 / 5)], rtol=0, atol=2e-16)
    assert np.allclose(xg["w"], [5 / 9, 8 / 9, 5 / 9], rtol=0, atol=2e-16)


@pytest.mark.parametrize("nodes", ["GLL", "GL"])
@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8])
def test_operators_exactness_and_sbp(nodes, n):
    o = O.build_ops(n, nodes)
    x, w, D = o["xi"], o["w"], o["D"]
    # D differentiates every polynomial of degree <= n-1 exactly (row order i, column j)
    for k in range(n):
        assert np.allclose(D @ x ** k, k * x ** max(k - 1, 0) * (k > 0), rtol=0, atol=1e-12 * n * n)
    # quadrature: exact to degree 2n-3 (GLL) / 2n-1 (GL)
    deg = 2 * n - 3 if nodes == "GLL" else 2 * n - 1
    for k in range(deg + 1):
        assert abs(w @ x ** k - (2 / (k + 1) if k % 2 == 0 else 0)) < 1e-14
    # face interpolation reproduces x^k at ±1
    for k in range(n):
        assert abs(o["lm"] @ x ** k - (-1.0) ** k) < 1e-13
        assert abs(o["lp"] @ x ** k - 1.0) < 1e-13
    # SBP: W D + (W D)^T = l_+ l_+^T - l_- l_-^T (summation-by-parts)
    W = np.diag(w)
    S = W @ D + (W @ D).T
    B = np.outer(o["lp"], o["lp"]) - np.outer(o["lm"], o["lm"])
    assert np.allclose(S, B, rtol=0, atol=1e-13 * n)
    # ∂g = l./ω with the signs of StdSegment.jl:84-85, Ds = D - B, D♯ = 2D - B
    assert np.array_equal(o["gm"], o["lm"] / w)
    assert np.array_equal(o["gp"], o["lp"] / w)
    Bg = np.outer(o["gp"], o["lp"]) - np.outer(o["gm"], o["lm"])
    assert np.array_equal(o["Ds"], D - Bg)
    assert np.array_equal(o["Dsh"], 2 * D - Bg)
    # rows of D♯ sum to the boundary term (constants are preserved by flux differencing)
    assert np.allclose(o["Dsh"].sum(axis=1), -Bg.sum(axis=1), atol=1e-12 * n)


# ------------------------------------------------------------------ point-wise fluxes
def _disc1(nd, tp, savg=O.AVG_STD, sdiss=O.DISS_LXF):
    return oracle_for(nd, 3, [2] * nd, [0.0] * nd, [1.0] * nd, [1] * nd, O.EQ_EULER, [GAMMA],
                      O.VOL_SPLIT, tp, savg, sdiss)


def _prim2cons(P, nd):
    P = np.asarray(P, dtype=np.float64)
    Q = np.zeros(nd + 2)
    Q[0] = P[0]
    Q[1:nd + 1] = P[0] * P[1:nd + 1]
    Q[nd + 1] = P[nd + 1] / (GAMMA - 1) + P[0] * np.sum(P[1:nd + 1] ** 2) / 2
    return Q


def _euler_flux(Q, nd, nvec):
    rho, m = Q[0], Q[1:nd + 1]
    u = m / rho
    p = (GAMMA - 1) * (Q[nd + 1] - np.dot(m, m) / (2 * rho))
    un = u @ nvec
    F = np.zeros(nd + 2)
    F[0] = rho * un
    F[1:nd + 1] = m * un + p * nvec
    F[nd + 1] = (Q[nd + 1] + p) * un
    return F


def test_logarithmic_mean_both_branches():
    for a, b in [(1.0, 1.0001), (1.0, 1.1), (1.0, 1.2), (1.0, 1.25), (1.0, 3.0), (0.2, 7.0)]:
        exact = (b - a) / (np.log(b) - np.log(a))
        assert O.logmean(a, b) == pytest.approx(exact, rel=2e-9 if abs(b / a - 1) < 0.3 else 1e-14)
        assert O.logmean(b, a) == pytest.approx(exact, rel=2e-9)
    assert O.logmean(2.0, 2.0) == 2.0


@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("tp", [O.TP_STDAVG, O.TP_CHANDRASEKHAR, O.TP_KENNEDY_GRUBER])
def test_twopoint_consistency_and_symmetry(nd, tp):
    d = _disc1(nd, tp)
    rng = np.random.default_rng(7 + nd + 10 * tp)
    for _ in range(5):
        P1 = np.concatenate([[rng.uniform(0.5, 2)], rng.uniform(-1, 1, nd), [rng.uniform(0.5, 2)]])
        P2 = np.concatenate([[rng.uniform(0.5, 2)], rng.uniform(-1, 1, nd), [rng.uniform(0.5, 2)]])
        Q1, Q2 = _prim2cons(P1, nd), _prim2cons(P2, nd)
        Ja = rng.uniform(-1, 1, nd)
        # consistency: F(Q, Q) = f(Q)·n
        assert np.allclose(d.twopoint(Q1, Q1, Ja, Ja), _euler_flux(Q1, nd, Ja), rtol=1e-12, atol=1e-13)
        # symmetry: F(Q1, Q2) = F(Q2, Q1)
        assert np.allclose(d.twopoint(Q1, Q2, Ja, Ja), d.twopoint(Q2, Q1, Ja, Ja), rtol=1e-12, atol=1e-13)
        # averaged metric: F(Q1, Q2; Ja1, Ja2) = F(Q1, Q2; n̄, n̄)
        Jb = rng.uniform(-1, 1, nd)
        assert np.allclose(d.twopoint(Q1, Q2, Ja, Jb), d.twopoint(Q1, Q2, (Ja + Jb) / 2, (Ja + Jb) / 2),
                           rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_chandrasekhar_tadmor_condition(nd):
    """Entropy conservation: (w1 - w2)·F(1,2) = ψ1 - ψ2, ψ = ρ u·n (math_entropy potential)."""
    d = _disc1(nd, O.TP_CHANDRASEKHAR)
    rng = np.random.default_rng(3 + nd)
    for _ in range(10):
        P1 = np.concatenate([[rng.uniform(0.5, 2)], rng.uniform(-1, 1, nd), [rng.uniform(0.5, 2)]])
        P2 = np.concatenate([[rng.uniform(0.5, 2)], rng.uniform(-1, 1, nd), [rng.uniform(0.5, 2)]])
        Q1, Q2 = _prim2cons(P1, nd), _prim2cons(P2, nd)
        nv = rng.uniform(-1, 1, nd)
        F = d.twopoint(Q1, Q2, nv, nv)
        w1, w2 = O.cons2entropy(nd, GAMMA, Q1), O.cons2entropy(nd, GAMMA, Q2)
        psi1, psi2 = Q1[1:nd + 1] @ nv, Q2[1:nd + 1] @ nv
        # exact up to the truncation of the log-mean series (u^4/9 at u < 0.01, Utilities.jl:39)
        assert (w1 - w2) @ F == pytest.approx(psi1 - psi2, abs=1e-8)


def test_kennedy_gruber_closed_form():
    """KG: ρ̄ (ū·n)(1, ū, ē) + p̄ (0, n, ū·n) with arithmetic means of ρ, u, p, e = ρe/ρ."""
    nd = 3
    d = _disc1(nd, O.TP_KENNEDY_GRUBER)
    Q1 = _prim2cons([1.1, 0.2, -0.3, 0.1, 0.9], nd)
    Q2 = _prim2cons([0.8, -0.1, 0.4, 0.3, 1.3], nd)
    nv = np.array([0.3, -0.5, 0.7])
    r = lambda Q: Q[0]
    u1, u2 = Q1[1:4] / Q1[0], Q2[1:4] / Q2[0]
    p = lambda Q: (GAMMA - 1) * (Q[4] - Q[1:4] @ Q[1:4] / (2 * Q[0]))
    rb, ub, pb = (r(Q1) + r(Q2)) / 2, (u1 + u2) / 2, (p(Q1) + p(Q2)) / 2
    eb = (Q1[4] / Q1[0] + Q2[4] / Q2[0]) / 2
    un = ub @ nv
    ref = np.concatenate([[rb * un], rb * ub * un + pb * nv, [rb * un * eb + pb * un]])
    assert np.allclose(d.twopoint(Q1, Q2, nv, nv), ref, rtol=1e-14, atol=1e-15)


def test_lxf_surface_flux_closed_form():
    """LxF(StdAverage): (f(Ql) + f(Qr))·n/2 + λ/2 (Ql - Qr), λ = max(|u_n| + c)."""
    nd = 2
    d = oracle_for(nd, 3, [2, 2], [0, 0], [1, 1], [1, 1], O.EQ_EULER, [GAMMA], O.VOL_STRONG, 0,
                   O.AVG_STD, O.DISS_LXF, 0.7)
    Ql = _prim2cons([1.1, 0.2, -0.3, 0.9], nd)
    Qr = _prim2cons([0.8, -0.1, 0.4, 1.3], nd)
    frame = np.array([0.0, -1.0, 1.0, 0.0, 0.0, 0.0])  # n = -y (the pos-3 frame)
    nv = frame[:2]
    F = d.numflux(Ql, Qr, frame)
    c = lambda Q: np.sqrt(GAMMA * (GAMMA - 1) * (Q[3] - Q[1:3] @ Q[1:3] / (2 * Q[0])) / Q[0])
    lam = max(abs(Ql[1:3] @ nv / Ql[0]) + c(Ql), abs(Qr[1:3] @ nv / Qr[0]) + c(Qr))
    ref = (_euler_flux(Ql, nd, nv) + _euler_flux(Qr, nd, nv)) / 2 + lam * (Ql - Qr) / 2 * 0.7
    assert np.allclose(F, ref, rtol=1e-14, atol=1e-15)


# ------------------------------------------------------------------ invariants of rhs!
CONFIGS = [
    # nd, n, nel, volop, tp, savg, sdiss
    (1, 4, [6], O.VOL_STRONG, 0, O.AVG_STD, O.DISS_LXF),
    (1, 5, [5], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF),
    (2, 4, [3, 4], O.VOL_STRONG, 0, O.AVG_STD, O.DISS_LXF),
    (2, 5, [3, 3], O.VOL_SPLIT, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF),
    (2, 4, [3, 2], O.VOL_SPLIT, O.TP_STDAVG, O.AVG_STD, O.DISS_NONE),
    (3, 3, [2, 3, 2], O.VOL_STRONG, 0, O.AVG_STD, O.DISS_LXF),
    (3, 4, [2, 2, 3], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF),
    (3, 3, [2, 2, 2], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX),
]


def _euler_disc(cfg, finish=None):
    nd, n, nel, vol, tp, savg, sdiss = cfg
    finish = finish or [1.0, 1.5, 2.0][:nd]
    return oracle_for(nd, n, nel, [0.0] * nd, finish, [1] * nd, O.EQ_EULER, [GAMMA], vol, tp,
                      savg, sdiss)


@pytest.mark.parametrize("cfg", CONFIGS)
def test_free_stream_preservation(cfg):
    d = _euler_disc(cfg)
    nd = cfg[0]
    q = _prim2cons([1.2] + [0.3, -0.2, 0.1][:nd] + [0.8], nd)
    Q = np.asfortranarray(np.tile(q, (d.ndofs, 1)))
    dQ = d.rhs(Q)
    assert np.max(np.abs(dQ)) < 1e-11


@pytest.mark.parametrize("cfg", CONFIGS)
def test_discrete_conservation(cfg):
    d = _euler_disc(cfg)
    X, _ = d.coords()
    Q = random_euler_state(X, cfg[0], seed=11)
    dQ = d.rhs(Q)
    scale = np.max(np.abs(dQ)) * d.ndofs
    for v in range(d.nv):
        assert abs(d.integrate(dQ, v)) < 1e-13 * scale


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_entropy_conservation_ec_split_form(nd):
    """Chandrasekhar volume flux + Chandrasekhar surface average, no dissipation:
    Σ Jω w·dQ = 0 on a periodic mesh (semi-discrete entropy conservation)."""
    nel = [4, 3, 2][:nd]
    d = oracle_for(nd, 4, nel, [0.0] * nd, [1.0] * nd, [1] * nd, O.EQ_EULER, [GAMMA], O.VOL_SPLIT,
                   O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_NONE)
    d.set_option(0, 0)  # corrected 3-D surface flux (Euler.jl:216 typo off)
    X, _ = d.coords()
    Q = random_euler_state(X, nd, seed=5, amp=0.2)
    dQ = d.rhs(Q)
    W = np.array([O.cons2entropy(nd, GAMMA, Q[i]) for i in range(d.ndofs)])
    dS = sum(d.integrate(np.asfortranarray((W[:, v] * dQ[:, v])[:, None] * np.ones((1, d.nv))), 0)
             for v in range(d.nv))
    scale = sum(abs(d.integrate(np.asfortranarray(np.abs(W[:, v] * dQ[:, v])[:, None] * np.ones((1, d.nv))), 0))
                for v in range(d.nv))
    # exact up to round-off and the log-mean series truncation (Utilities.jl:39)
    assert abs(dS) < 1e-10 * scale
    if nd == 3:
        d.set_option(0, 1)  # the reference's 3-D surface flux breaks entropy conservation
        dQt = d.rhs(Q)
        dSt = sum(d.integrate(np.asfortranarray((W[:, v] * dQt[:, v])[:, None] * np.ones((1, d.nv))), 0)
                  for v in range(d.nv))
        assert abs(dSt) > 1e-6 * scale  # ~5e-5 relative here, against < 1e-10 without it


def _convergence_error(nel, n, nd=1, t_end=0.05):
    """Convergence.jl:31-42 exact solution ρ = 2 + sin(π(x - t)), u = 1, p = 1 on [-1, 1]."""
    d = oracle_for(nd, n, [nel] * nd, [-1.0] * nd, [1.0] * nd, [1] * nd, O.EQ_EULER, [GAMMA],
                   O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF)
    X, _ = d.coords()

    def exact(t):
        rho = 2 + np.sin(np.pi * (X[:, 0] - t))
        P = np.zeros((d.ndofs, nd + 2))
        P[:, 0], P[:, 1], P[:, nd + 1] = rho, 1.0, 1.0
        return np.asfortranarray(np.array([_prim2cons(p, nd) for p in P]))

    dt = 0.2 * (2.0 / nel) / n ** 2
    steps = int(round(t_end / dt))
    dt = t_end / steps
    u = d.rk_run(O.RK_CK2N54, exact(0.0), dt, steps)
    return np.sqrt(np.mean((u - exact(t_end)) ** 2))


def test_convergence_order():
    """Errors fall at high order: the asymptotic rate exceeds the polynomial order N."""
    n = 5
    e = [_convergence_error(nel, n) for nel in (4, 8, 16, 32)]
    rates = np.log2(np.array(e[:-1]) / np.array(e[1:]))
    assert np.all(rates > 3.5), rates
    assert rates[-1] > n - 1, rates


# ------------------------------------------------------------------ time integration
def _butcher_2n(A, B, s):
    """Equivalent Butcher (a, b) of a 2N scheme with stage updates dU = A dU + dt f; U += B dU."""
    # U_i = U_0 + Σ_j a_ij dt f_j ; dU_i = Σ_j g_ij dt f_j
    a = np.zeros((s + 1, s))
    g = np.zeros(s)
    for i in range(s):
        g = A[i] * g
        g[i] += 1.0
        a[i + 1] = a[i] + B[i] * g
    return a[:s], a[s]


@pytest.mark.parametrize("scheme,order", [(O.RK_CK2N54, 4), (O.RK_ORK256, 2)])
def test_2n_tableau_order_conditions(scheme, order):
    A, B, c = O.tableau_2n(scheme)
    a, b = _butcher_2n(A, B, 5)
    cc = a.sum(axis=1)
    assert np.allclose(cc, c, atol=2e-6 if scheme == O.RK_ORK256 else 1e-12)
    assert b.sum() == pytest.approx(1.0, abs=1e-12)
    assert b @ cc == pytest.approx(0.5, abs=1e-5)
    if order >= 4:
        assert b @ cc ** 2 == pytest.approx(1 / 3, abs=1e-10)
        assert b @ (a @ cc) == pytest.approx(1 / 6, abs=1e-10)
        assert b @ cc ** 3 == pytest.approx(1 / 4, abs=1e-10)
        assert b @ (cc * (a @ cc)) == pytest.approx(1 / 8, abs=1e-10)
        assert b @ (a @ cc ** 2) == pytest.approx(1 / 12, abs=1e-10)
        assert b @ (a @ (a @ cc)) == pytest.approx(1 / 24, abs=1e-10)


@pytest.mark.parametrize("scheme,order", [(O.RK_SSPRK33, 3), (O.RK_CK2N54, 4), (O.RK_ORK256, 2)])
def test_rk_temporal_order(scheme, order):
    """Linear advection of sin(2πx) on a spatially converged mesh: error ∝ dt^order."""
    d = oracle_for(1, 8, [8], [0.0], [1.0], [1], O.EQ_ADVECTION, [1.0], O.VOL_STRONG, 0, 0,
                   O.DISS_LXF, 1.0)
    X, _ = d.coords()
    T = 0.25
    errs = []
    for steps in (20, 40):
        u = d.rk_run(scheme, np.asfortranarray(np.sin(2 * np.pi * X)), T / steps, steps)
        errs.append(np.max(np.abs(u[:, 0] - np.sin(2 * np.pi * (X[:, 0] - T)))))
    rate = np.log2(errs[0] / errs[1])
    assert rate > order - 0.3, (errs, rate)


# ------------------------------------------------------------------ dt, integrals, errors
def test_max_dt_brute_force():
    nd = 2
    d = _euler_disc((2, 4, [3, 2], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF))
    X, _ = d.coords()
    Q = random_euler_state(X, nd, seed=3)
    g = d.geometry()
    rho = Q[:, 0]
    u = Q[:, 1:3] / rho[:, None]
    p = (GAMMA - 1) * (Q[:, 3] - (Q[:, 1] ** 2 + Q[:, 2] ** 2) / (2 * rho))
    dx = np.sqrt(np.repeat(g["e_vol"], d.npe) / d.npe)
    ref = np.min(0.3 * dx / (np.sqrt((u ** 2).sum(1)) + np.sqrt(GAMMA * p / rho)))
    assert d.max_dt(Q, 0.3) == pytest.approx(ref, rel=1e-15)


def test_integrate_polynomial():
    d = _euler_disc((3, 4, [2, 3, 2], O.VOL_STRONG, 0, O.AVG_STD, O.DISS_LXF), finish=[1.0, 1.5, 2.0])
    X, _ = d.coords()
    f = X[:, 0] ** 2 * X[:, 1] + X[:, 2]
    Q = np.asfortranarray(np.tile(f[:, None], (1, 5)))
    exact = (1.0 / 3) * (1.5 ** 2 / 2) * 2.0 + 1.0 * 1.5 * 2.0 ** 2 / 2
    assert d.integrate(Q, 0) == pytest.approx(exact, rel=1e-14)


def test_negative_pressure_is_a_domain_error():
    d = _euler_disc((2, 4, [3, 2], O.VOL_STRONG, 0, O.AVG_STD, O.DISS_LXF))
    X, _ = d.coords()
    Q = random_euler_state(X, 2, seed=1)
    Q[0, 3] = -1.0  # ρe < 0 at a face node => p < 0 => sqrt of a negative in LxF
    with pytest.raises(ArithmeticError):
        d.rhs(Q)


# ---------------------------------------------------------------- ScalarDissipation
def _scalar_disc(nd, typo):
    d = O.OracleDisc(nd, 3, [2] * nd, [0.0] * nd, [1.0] * nd, [1] * nd, O.build_ops(3, "GLL"),
                     O.EQ_EULER, [GAMMA], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR,
                     O.DISS_SCALAR, 1.0)
    d.set_option(0, typo)
    avg = O.OracleDisc(nd, 3, [2] * nd, [0.0] * nd, [1.0] * nd, [1] * nd, O.build_ops(3, "GLL"),
                       O.EQ_EULER, [GAMMA], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR,
                       O.DISS_NONE, 1.0)
    avg.set_option(0, typo)
    return d, avg


def _frame(nd):
    fr = np.zeros((3, nd))
    fr[0, 0] = 1.0
    if nd >= 2:
        fr[1, 1] = 1.0
    if nd == 3:
        fr[2, 2] = 1.0
    return fr


def _lam(Ql, Qr, nd):
    def side(Q):
        rho = Q[0]
        u = Q[1:nd + 1] / rho
        p = (GAMMA - 1) * (Q[nd + 1] - 0.5 * rho * np.dot(u, u))
        return abs(u[0]) + np.sqrt(GAMMA * p / rho)
    return max(side(Ql), side(Qr))


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_scalar_dissipation_is_the_linearised_jump(nd):
    """ScalarDissipation (Euler.jl:228-301): Fn = avg - λ/2·D; for Qr = Ql + εδ the
    dissipation vector D/ε tends to δ (D is ∂Q/∂W·[W], energy row included), and vanishes
    for equal states (consistency)."""
    d, avg = _scalar_disc(nd, 0)
    rng = np.random.default_rng(nd)
    fr = _frame(nd).ravel()
    for _ in range(10):
        X = rng.uniform(0, 1, (1, nd))
        Ql = random_euler_state(X, nd, seed=int(rng.integers(1e6)), amp=0.3, smooth=False)[0]
        assert np.allclose(d.numflux(Ql, Ql, fr), avg.numflux(Ql, Ql, fr), rtol=0, atol=1e-14)
        delta = rng.standard_normal(nd + 2)
        eps = 1e-6
        Qr = Ql + eps * delta
        D = (avg.numflux(Ql, Qr, fr) - d.numflux(Ql, Qr, fr)) / (_lam(Ql, Qr, nd) / 2)
        assert np.allclose(D / eps, delta, rtol=0, atol=1e-4 * np.max(np.abs(delta)))


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_scalar_dissipation_dissipates_entropy(nd):
    """(W_r - W_l)·D ≥ 0 for strongly different random states: the dissipation term never
    produces entropy."""
    d, avg = _scalar_disc(nd, 0)
    rng = np.random.default_rng(10 + nd)
    fr = _frame(nd).ravel()
    X = rng.uniform(0, 1, (200, nd))
    QL = random_euler_state(X, nd, seed=1, amp=0.5, smooth=False)
    QR = random_euler_state(X, nd, seed=2, amp=0.5, smooth=False)
    for Ql, Qr in zip(QL, QR):
        D = avg.numflux(Ql, Qr, fr) - d.numflux(Ql, Qr, fr)
        dW = O.cons2entropy(nd, GAMMA, Qr) - O.cons2entropy(nd, GAMMA, Ql)
        assert np.dot(dW, D) >= -1e-14 * np.sum(np.abs(dW * D))


def test_scalar_dissipation_3d_typo_drops_momentum_jumps():
    """Euler.jl:257 takes the 3-D left momenta from Qr: with the typo switch on, the
    momentum rows of D vanish; the density and energy rows are unchanged."""
    good, avg = _scalar_disc(3, 0)
    bad, avg_bad = _scalar_disc(3, 1)  # same typo switch for the average part
    rng = np.random.default_rng(4)
    fr = _frame(3).ravel()
    X = rng.uniform(0, 1, (5, 3))
    QL = random_euler_state(X, 3, seed=3, amp=0.3, smooth=False)
    QR = random_euler_state(X, 3, seed=4, amp=0.3, smooth=False)
    for Ql, Qr in zip(QL, QR):
        Dg = avg.numflux(Ql, Qr, fr) - good.numflux(Ql, Qr, fr)
        Db = avg_bad.numflux(Ql, Qr, fr) - bad.numflux(Ql, Qr, fr)
        assert np.allclose(Db[1:4], 0.0, atol=1e-15)
        assert np.allclose(Db[[0, 4]], Dg[[0, 4]], rtol=1e-12, atol=1e-14)
        assert np.max(np.abs(Dg[1:4])) > 1e-3
