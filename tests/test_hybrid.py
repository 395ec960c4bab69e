"""HybridDivOperator in the oracle (Equations/OpDivergence.jl:441-620, 781-801).

Pins:
* Shockwave2D (test/tests.jl:136-181) -> the reference's own known answers
  (test/runtests.jl:42-43): max at rtol 1e-7, min at rtol 1;
* with blend c -> ∞ every δ is 1 and the telescopic subcell form must equal the split form
  (SBP: ω_i D_ik = -ω_k D_ki off the diagonal, and D♯ has a zero diagonal on GLL nodes), so
  a wrong index, weight or sign in the interface sums fails it;
* δ ∈ [1/2, 2]: the hybrid rhs differs from the split one when c is small;
* conservation Σ Jω dQ = 0 on a periodic mesh for any δ (the interface fluxes telescope),
  and free-stream preservation.
On Gauss nodes (_hybrid_nb_surface_contribution!, OpDivergence.jl:633-779) the subcell
fluxes telescope through the blended values along each line, starting from the face flux.
With c -> ∞ the telescoped sum returns exactly the other face flux (W D♯ is skew on Gauss
nodes and the projected corrections sum to the face fluxes), so the operator must equal
the Gauss split form (_splitdiv_nb_surface_contribution!).  A wrong row of D♯, a dropped
ω, a sign on l·Fl / r·Fr or a missing face term breaks that identity; conservation and free
stream are checked for small c too.
"""
import numpy as np
import pytest

import oracle_lib as O
from cases import GAMMA, isapprox, oracle_for, random_euler_state, reference_answer


def cons(r, u, v, p, g=GAMMA):
    return np.array([r, r * u, r * v, p / (g - 1) + r * (u * u + v * v) / 2])


def normal_shock(r0, u0, p0, g=GAMMA):
    """FlouCommon/Euler.jl:346-359."""
    M0 = u0 / np.sqrt(g * p0 / r0)
    r1 = r0 * M0 ** 2 * (g + 1) / ((g - 1) * M0 ** 2 + 2)
    p1 = p0 * (2 * g * M0 ** 2 - (g - 1)) / (g + 1)
    M1 = np.sqrt(((g - 1) * M0 ** 2 + 2) / (2 * g * M0 ** 2 - (g - 1)))
    return r1, M1 * np.sqrt(g * p1 / r1), p1


def shockwave_case():
    """Shockwave2D: GLL 6, 11x3 elements on [-1,1]x[0,1], periodic in y, GenericBC
    Q(x) = Q0 (x < 0) / Q1 on "1", "2", HybridDivOperator(MatrixDissipation(Chandrasekhar,
    1.0), 1.0), ORK256, dt 1e-2, t = 1."""
    d = oracle_for(2, 6, [11, 3], [-1.0, 0.0], [1.0, 1.0], [0, 1], O.EQ_EULER, [GAMMA], O.VOL_HYBRID,
                   O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1.0)
    d.set_blend(1.0)
    X, FX = d.coords()
    r0, M0, p0 = 1.0, 2.0, 1.0
    u0 = M0 * np.sqrt(GAMMA * p0 / r0)
    Q0, Q1 = cons(r0, u0, 0.0, p0), cons(*normal_shock(r0, u0, p0)[:2], 0.0, normal_shock(r0, u0, p0)[2])

    def qext(x):
        return Q0 if x < 0 else Q1

    Q = np.asfortranarray(np.array([qext(x) for x in X[:, 0]]))
    bc = np.zeros((d.nf, d.npf, 4))
    for f in range(d.nf):
        for k in range(d.npf):
            bc[f, k] = qext(FX[f, k, 0])
    d.set_bc(0, O.BC_DIRICHLET, bc)
    d.set_bc(1, O.BC_DIRICHLET, bc)
    return d, Q


def test_shockwave_golden():
    ref = reference_answer("shockwave_2d")
    d, Q = shockwave_case()
    u = d.rk_run(O.RK_ORK256, Q, 1e-2, 100)
    print("Shockwave2D oracle: min", u.min(), "max", u.max())
    assert isapprox(u.max(), ref["max"], ref["max_rtol"])
    # isapprox(min, -4.68e-13, rtol=1): same sign, round-off size
    assert isapprox(u.min(), ref["min"], ref["min_rtol"])
    # only round-off separates the two codes
    assert abs(u.max() - ref["max"]) / ref["max"] < 1e-11


def _pair(nd, n, nel, tp, savg, sdiss, blend, periodic=True):
    args = (nd, n, [nel] * nd, [0.0] * nd, [1.0, 1.5, 2.0][:nd], [int(periodic)] * nd, O.EQ_EULER, [GAMMA])
    h = oracle_for(*args, O.VOL_HYBRID, tp, savg, sdiss, 1.0)
    h.set_blend(blend)
    s = oracle_for(*args, O.VOL_SPLIT, tp, savg, sdiss, 1.0)
    return h, s


CASES = [(1, 5, 6, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX),
         (2, 4, 3, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF),
         (2, 6, 2, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF),
         (3, 4, 2, O.TP_STDAVG, O.AVG_STD, O.DISS_LXF),
         (3, 3, 2, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-n{c[1]}-tp{c[3]}-d{c[5]}")
def test_large_blend_is_the_split_form(case):
    nd, n, nel, tp, savg, sdiss = case
    h, s = _pair(nd, n, nel, tp, savg, sdiss, 1e300)
    X, _ = h.coords()
    Q = random_euler_state(X, nd, seed=3, amp=0.3, smooth=False)
    a, b = h.rhs(Q), s.rhs(Q)
    err = np.max(np.abs(a - b)) / np.max(np.abs(b))
    print(f"{case}: hybrid(c=1e300) vs split {err:.2e}")
    assert err <= 1e-13


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-n{c[1]}-tp{c[3]}-d{c[5]}")
def test_hybrid_conservation_and_free_stream(case):
    nd, n, nel, tp, savg, sdiss = case
    h, s = _pair(nd, n, nel, tp, savg, sdiss, 1e-2)
    X, _ = h.coords()
    Q = random_euler_state(X, nd, seed=4, amp=0.3, smooth=False)
    dq = h.rhs(Q)
    assert np.max(np.abs(dq - s.rhs(Q))) > 1e-6 * np.max(np.abs(dq))  # the blend acts
    g = h.geometry()
    w = np.ones(1)
    for _ in range(nd):
        w = np.kron(O.build_ops(n, "GLL")["w"], w)
    wq = np.repeat(g["e_jac"], n ** nd) * np.tile(w, h.ne)
    for v in range(nd + 2):
        tot = np.sum(wq * dq[:, v])
        assert abs(tot) <= 1e-13 * np.sum(np.abs(wq * dq[:, v]))
    Q0 = np.zeros_like(Q)
    Q0[:, 0], Q0[:, 1:nd + 1], Q0[:, nd + 1] = 1.2, [0.3, -0.2, 0.1][:nd], 2.5
    assert np.max(np.abs(h.rhs(np.asfortranarray(Q0)))) <= 1e-12


def _pair_gauss(nd, n, nel, tp, savg, sdiss, blend):
    args = (nd, n, [nel] * nd, [0.0] * nd, [1.0, 1.5, 2.0][:nd], [1] * nd, O.EQ_EULER, [GAMMA])
    h = oracle_for(*args, O.VOL_HYBRID, tp, savg, sdiss, 1.0, nodes="GL")
    h.set_blend(blend)
    s = oracle_for(*args, O.VOL_SPLIT, tp, savg, sdiss, 1.0, nodes="GL")
    return h, s


GAUSS_CASES = [(1, 4, 6, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX),
               (1, 6, 4, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF),
               (2, 4, 3, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_SCALAR),
               (2, 5, 2, O.TP_STDAVG, O.AVG_STD, O.DISS_LXF),
               (3, 3, 2, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX),
               (3, 4, 2, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF)]


@pytest.mark.parametrize("case", GAUSS_CASES, ids=lambda c: f"{c[0]}d-n{c[1]}-tp{c[3]}-d{c[5]}")
def test_gauss_large_blend_is_the_gauss_split_form(case):
    nd, n, nel, tp, savg, sdiss = case
    h, s = _pair_gauss(nd, n, nel, tp, savg, sdiss, 1e300)
    X, _ = h.coords()
    Q = random_euler_state(X, nd, seed=3, amp=0.2)
    a, b = h.rhs(Q), s.rhs(Q)
    err = np.max(np.abs(a - b)) / np.max(np.abs(b))
    print(f"Gauss {case}: hybrid(c=1e300) vs Gauss split {err:.2e}")
    assert err <= 1e-13


@pytest.mark.parametrize("case", GAUSS_CASES, ids=lambda c: f"{c[0]}d-n{c[1]}-tp{c[3]}-d{c[5]}")
def test_gauss_hybrid_conservation_and_free_stream(case):
    nd, n, nel, tp, savg, sdiss = case
    h, s = _pair_gauss(nd, n, nel, tp, savg, sdiss, 1e-2)
    X, _ = h.coords()
    Q = random_euler_state(X, nd, seed=4, amp=0.2)
    dq = h.rhs(Q)
    assert np.max(np.abs(dq - s.rhs(Q))) > 1e-5 * np.max(np.abs(dq))  # the blend acts
    w = np.ones(1)
    for _ in range(nd):
        w = np.kron(O.build_ops(n, "GL")["w"], w)
    wq = np.repeat(h.geometry()["e_jac"], n ** nd) * np.tile(w, h.ne)
    for v in range(nd + 2):
        assert abs(np.sum(wq * dq[:, v])) <= 1e-13 * np.sum(np.abs(wq * dq[:, v]))
    Q0 = np.zeros_like(Q)
    Q0[:, 0], Q0[:, 1:nd + 1], Q0[:, nd + 1] = 1.2, [0.3, -0.2, 0.1][:nd], 2.5
    assert np.max(np.abs(h.rhs(np.asfortranarray(Q0)))) <= 1e-12
