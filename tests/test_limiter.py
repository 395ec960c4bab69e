"""Zhang & Shu limiter and the kinetic-energy / entropy monitors in the oracle
(Equations/Euler.jl:541-654), pinned by hand-worked cases and by the invariants the
limiter guarantees, not by re-evaluating its formulas:

* a two-node element worked by hand (density pulled to exactly minval);
* an element whose nodal-pressure mean is below minval while p(Q̄) is not: the
  reference's nodal-mean rule (θ = 0) sends every node to the element mean;
* element means (Jω'Q / vol) unchanged for every variable (the limiter is conservative);
* min ρ' ≥ minval and min p' ≥ minval after limiting a state with negative nodes
  (concavity of p in Q);
* no change at all when every node is above minval;
* the limited RK run (stage limiter after every stage) conserves the integrals;
* monitors: closed forms on constant states, spectral convergence to analytic integrals.
"""
import numpy as np
import pytest

import oracle_lib as O
from cases import GAMMA, random_euler_state


def euler(nd, n, nel, length=1.0, periodic=True, savg=O.AVG_STD):
    return O.OracleDisc(nd, n, [nel] * nd, [0.0] * nd, [length] * nd, [int(periodic)] * nd,
                        O.build_ops(n, "GLL"), O.EQ_EULER, [GAMMA], O.VOL_SPLIT,
                        O.TP_CHANDRASEKHAR, savg, O.DISS_LXF, 1.0)


def pressure(Q, nd):
    return (GAMMA - 1) * (Q[:, nd + 1] - np.sum(Q[:, 1:nd + 1] ** 2, 1) / (2 * Q[:, 0]))


def element_means(d, Q):
    """Jω'Q / vol per element from the oracle's own geometry (Cartesian: J per element)."""
    g = d.geometry()
    w = np.ones(1)
    for _ in range(d.nd):
        w = np.kron(O.build_ops(d.n, "GLL")["w"], w)
    Qe = Q.reshape(d.ne, d.npe, Q.shape[1])
    return np.einsum("e,i,eiv->ev", g["e_jac"], w, Qe) / g["e_vol"][:, None]


def test_two_node_density_case_by_hand():
    # 1-D, n = 2 GLL nodes (ω = 1, 1), one element on [0, 1]: J = 1/2, vol = 1
    d = euler(1, 2, 1)
    Q = np.asfortranarray([[-0.1, 0.0, 2.5], [1.1, 0.0, 2.5]])
    out = d.limit(Q, 1e-3)
    # ρ̄ = 0.5, m = 1e-3, θ = 0.499 / 0.6: ρ' = θ(ρ - ρ̄) + ρ̄ = [1e-3, 0.999]
    assert np.allclose(out[:, 0], [1e-3, 0.999], rtol=0, atol=1e-15)
    # p = 0.4 * 2.5 = 1 at both nodes after the density pull (no momentum): no pressure pull
    assert np.array_equal(out[:, 1:], Q[:, 1:])


def test_nodal_pressure_mean_rule():
    """p̄ is the mean of the nodal pressures (Euler.jl:636-640), not p(Q̄): with p̄ < minval
    the limiter uses θ = 0 and every node becomes Q̄; p(Q̄) = 0.84 would not do that."""
    d = euler(1, 2, 1)
    Q = np.asfortranarray([[1.0, 2.0, 2.05], [1.0, -2.0, 2.15]])  # p_i = 0.4 (E_i - 2) = 0.02, 0.06
    out = d.limit(Q, 0.1)  # p̄ = 0.04 = m, pmin = 0.02: θ = 0
    assert np.allclose(out, [[1.0, 0.0, 2.1], [1.0, 0.0, 2.1]], rtol=0, atol=1e-15)


def _negative_state(d, which, seed=21):
    """Random state with a tenth of the nodes pushed to negative density (which = "rho") or
    to negative pressure at positive density (which = "p")."""
    nd = d.nd
    X, _ = d.coords()
    Q = random_euler_state(X, nd, seed=seed, amp=0.5, smooth=False)
    rng = np.random.default_rng(seed + 1)
    bad = rng.choice(d.ndofs, d.ndofs // 10, replace=False)
    if which == "rho":
        Q[bad, 0] *= -0.2
    else:
        Q[bad, nd + 1] = np.sum(Q[bad, 1:nd + 1] ** 2, 1) / (2 * Q[bad, 0]) - 0.05
    return Q


def _check_common(d, Q, out, minval):
    mq, mo = element_means(d, Q), element_means(d, out)
    assert np.abs(mo - mq).max() <= 1e-13 * np.abs(mq).max()  # conservative
    pq = pressure(Q, d.nd).reshape(d.ne, d.npe)
    clean = (Q[:, 0].reshape(d.ne, d.npe).min(1) > minval) & (pq.min(1) > minval)
    rows = np.repeat(clean, d.npe)
    assert clean.any() and not clean.all()
    assert np.array_equal(out[rows], Q[rows])  # elements with every node above minval
    return mq


LCASES = [(1, 5, 8), (2, 4, 4), (3, 3, 3)]


@pytest.mark.parametrize("nd,n,nel", LCASES)
def test_limiter_density_positivity(nd, n, nel):
    d = euler(nd, n, nel)
    Q = _negative_state(d, "rho")
    minval = 1e-6
    out = d.limit(Q, minval)
    mq = _check_common(d, Q, out, minval)
    ok = mq[:, 0] > minval
    assert Q[:, 0].reshape(d.ne, d.npe)[ok].min() < 0
    assert out[:, 0].reshape(d.ne, d.npe)[ok].min() >= minval * (1 - 1e-9)


@pytest.mark.parametrize("nd,n,nel", LCASES)
def test_limiter_pressure_positivity(nd, n, nel):
    """With ρ > 0 everywhere the density pass is a no-op, and concavity of p(Q) with
    p(Q̄) ≥ p̄ gives min p' ≥ minval in every element whose nodal-mean pressure exceeds it."""
    d = euler(nd, n, nel)
    Q = _negative_state(d, "p")
    minval = 1e-6
    out = d.limit(Q, minval)
    mq = _check_common(d, Q, out, minval)
    g = d.geometry()
    w = np.ones(1)
    for _ in range(nd):
        w = np.kron(O.build_ops(n, "GLL")["w"], w)
    pq = pressure(Q, nd).reshape(d.ne, d.npe)
    pbar = (g["e_jac"][:, None] * w[None] * pq).sum(1) / g["e_vol"]
    ok = pbar > minval
    assert pq[ok].min() < 0
    assert pressure(out, nd).reshape(d.ne, d.npe)[ok].min() >= minval * (1 - 1e-9)
    assert out[:, 0].min() >= Q[:, 0].min()


def test_limiter_noop_above_minval():
    d = euler(2, 5, 4)
    X, _ = d.coords()
    Q = random_euler_state(X, 2, seed=2, amp=0.2)
    assert np.array_equal(d.limit(Q, 1e-10), Q)


def test_limited_rk_run_conserves():
    d = euler(2, 4, 6)
    X, _ = d.coords()
    Q = random_euler_state(X, 2, seed=8, amp=0.3)
    minval = 0.85  # above the smallest nodal density and pressure: the limiter acts
    assert Q[:, 0].min() < minval
    d.set_limiter(True, minval)
    dt = 0.2 * d.max_dt(Q, 1.0)
    u = d.rk_run(O.RK_SSPRK33, Q, dt, 5)
    d.set_limiter(False)
    v = d.rk_run(O.RK_SSPRK33, Q, dt, 5)
    assert np.abs(u - v).max() > 1e-3  # the limiter changed the run
    for k in range(4):
        a, b = d.integrate(Q, k), d.integrate(u, k)
        assert abs(a - b) <= 1e-13 * max(1.0, abs(a))


def test_monitors_on_constant_state():
    d = euler(3, 4, 2, length=2.0)
    Q = np.zeros((d.ndofs, 5))
    Q[:, 0], Q[:, 1:4], Q[:, 4] = 1.3, [0.2, -0.4, 0.5], 3.0
    Q = np.asfortranarray(Q)
    vol = 8.0
    ke = (0.04 + 0.16 + 0.25) / (2 * 1.3)
    p = 0.4 * (3.0 - ke)
    ent = -1.3 * (np.log(p) - GAMMA * np.log(1.3)) / (GAMMA - 1)
    assert abs(d.monitor(Q, O.MON_KINETIC_ENERGY) - vol * ke) <= 1e-14 * vol * ke
    assert abs(d.monitor(Q, O.MON_ENTROPY) - vol * ent) <= 1e-14 * abs(vol * ent)


def test_monitors_converge_to_analytic_integrals():
    """ρ = 2 + sin 2πx, ρu = 1, ρv = 0, p = ρ^γ on the unit square: ∫ KE = ∫ 1/(2ρ) =
    1/(2√3) (∫_0^1 dx/(a + b sin 2πx) = 1/√(a²-b²)), and s = log p - γ log ρ = 0, so the
    entropy monitor vanishes."""
    errs = []
    for n in (3, 5, 7, 9):
        d = euler(2, n, 4)
        X, _ = d.coords()
        rho = 2 + np.sin(2 * np.pi * X[:, 0])
        one = np.ones_like(rho)
        Q = np.asfortranarray(np.stack([rho, one, 0 * one, rho ** GAMMA / (GAMMA - 1) + 1 / (2 * rho)], 1))
        errs.append(abs(d.monitor(Q, O.MON_KINETIC_ENERGY) - 1 / (2 * np.sqrt(3))))
        assert abs(d.monitor(Q, O.MON_ENTROPY)) <= 1e-13
    print("KE monitor errors", errs)
    assert errs[-1] <= 1e-9 and errs[-1] < errs[0] * 1e-4


def test_entropy_monitor_rejects_negative_pressure():
    d = euler(1, 3, 2)
    Q = np.asfortranarray(np.tile([1.0, 0.0, -1.0], (d.ndofs, 1)))
    with pytest.raises(ArithmeticError):
        d.monitor(Q, O.MON_ENTROPY)
