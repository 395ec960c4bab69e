"""Split form on Gauss nodes (no boundary nodes) in the oracle: the entropy-projected surface
terms of _splitdiv_nb_surface_contribution! (Equations/OpDivergence.jl:300-436).

No reference fixture covers this path, so it is pinned by what the construction is for:
* free-stream preservation;
* discrete conservation Σ Jω dQ = 0 on periodic meshes (W D♯ is skew on Gauss nodes and
  the projected corrections telescope to the face fluxes);
* semi-discrete entropy conservation Σ Jω w·dQ = 0 with the entropy-conservative volume
  and surface fluxes (Chandrasekhar, no dissipation) -- the property the entropy
  projection exists for; with LxF the entropy must decrease;
* entropy2cons inverts cons2entropy;
* spectral convergence to the exact divergence of a smooth periodic flow.
A dropped term, a wrong sign on either end or the wrong interpolation vector breaks the
entropy identity.
"""
import numpy as np
import pytest

import oracle_lib as O
from cases import GAMMA, oracle_for, periodic_euler_state, random_euler_state


def gauss(nd, n, nel, savg=O.AVG_CHANDRASEKHAR, sdiss=O.DISS_NONE, tp=O.TP_CHANDRASEKHAR, L=1.0):
    return oracle_for(nd, n, [nel] * nd, [0.0] * nd, [L] * nd, [1] * nd, O.EQ_EULER, [GAMMA], O.VOL_SPLIT,
                      tp, savg, sdiss, 1.0, nodes="GL")


def weights(d, n, nd):
    w = np.ones(1)
    for _ in range(nd):
        w = np.kron(O.build_ops(n, "GL")["w"], w)
    return np.repeat(d.geometry()["e_jac"], n ** nd) * np.tile(w, d.ne)


CASES = [(1, 4, 6), (2, 3, 4), (2, 5, 3), (3, 3, 2), (3, 4, 2)]


@pytest.mark.parametrize("nd,n,nel", CASES)
def test_free_stream_and_conservation(nd, n, nel):
    for tp, savg, sdiss in ((O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_NONE),
                            (O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF),
                            (O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX)):
        d = gauss(nd, n, nel, savg, sdiss, tp)
        X, _ = d.coords()
        Q = random_euler_state(X, nd, seed=2, amp=0.2)
        dq = d.rhs(Q)
        wq = weights(d, n, nd)
        for v in range(nd + 2):
            assert abs(np.sum(wq * dq[:, v])) <= 1e-13 * np.sum(np.abs(wq * dq[:, v]))
        Q0 = np.zeros_like(Q)
        Q0[:, 0], Q0[:, 1:nd + 1], Q0[:, nd + 1] = 1.2, [0.3, -0.2, 0.1][:nd], 2.5
        assert np.max(np.abs(d.rhs(np.asfortranarray(Q0)))) <= 1e-12


def _entropy_identity(nd, n, nel, savg, sdiss):
    """Σ Jω w·dQ of the Gauss split form, and the value the construction predicts for it.

    With an entropy-conservative two-point flux (Tadmor: (w_a - w_b)·f♯ = ψ_a - ψ_b), the
    skew volume operator and the projected corrections leave only face terms: per element,
    direction and line (transverse weight ω_f, metric scale s = surface Jacobian),
        ω_f [ s ψ̃_r - w̃_r·Fn_r + (-s ψ̃_l - w̃_l·Fn_l) ],   ψ = ρ u_dir,
    where w̃ = l·w are the interpolated entropy variables, ψ̃ = ψ(u(w̃)) and Fn the face
    fluxes (outward, times the surface Jacobian) evaluated on the conservative traces.
    u(w̃) is found here by Newton on cons2entropy, independently of entropy2cons."""
    from scipy.optimize import fsolve
    d = gauss(nd, n, nel, savg, sdiss)
    X, FX = d.coords()
    Q = random_euler_state(X, nd, seed=5, amp=0.2)
    dq = d.rhs(Q)
    W = np.array([O.cons2entropy(nd, GAMMA, q) for q in Q])
    lhs = np.sum(weights(d, n, nd)[:, None] * W * dq)
    ops = O.build_ops(n, "GL")
    lm, lp, w1 = ops["lm"], ops["lp"], ops["w"]
    g = d.geometry()
    cn = d.connectivity()
    e_face, f_elem, f_pos = cn["e_face"], cn["f_elem"], cn["f_pos"]
    fr = g["f_frame"]
    pred = 0.0
    for e in range(d.ne):
        for k in range(nd):
            s = g["e_metric"][e][k, k]
            for f in range(n ** (nd - 1)):
                nodes = [e * n ** nd + _line_node(nd, n, k, f, a) for a in range(n)]
                wf = np.prod([w1[(f // n ** j) % n] for j in range(nd - 1)]) if nd > 1 else 1.0
                for side, lv, sign in ((0, lm, -1.0), (1, lp, 1.0)):
                    wt = lv @ W[nodes]
                    ut = fsolve(lambda q: O.cons2entropy(nd, GAMMA, q) - wt, lv @ Q[nodes], xtol=1e-13)
                    psi = ut[1 + k]  # ρ u_k
                    # conservative traces on both sides of the face and the +axis flux
                    face = e_face[e, 2 * k + side]
                    ql, qr = _face_traces(d, Q, face, f, f_elem, f_pos, lm, lp, n, nd)
                    fstar = d.numflux(ql, qr, fr[face]) * g["f_jac"][face]
                    axis_sign = fr[face][0][k]  # master normal along ±axis
                    flux_plus = fstar * axis_sign  # flux in the +axis direction
                    Fn_out = sign * flux_plus
                    pred += wf * (sign * s * psi - wt @ Fn_out)
    return lhs, pred, np.sum(np.abs(weights(d, n, nd)[:, None] * W * dq))


def _line_node(nd, n, k, f, a):
    if nd == 1:
        return a
    if nd == 2:
        return a + n * f if k == 0 else f + n * a
    fi, fj = f % n, f // n
    return [a + n * (fi + n * fj), fi + n * (a + n * fj), fi + n * (fj + n * a)][k]


def _face_traces(d, Q, face, f, f_elem, f_pos, lm, lp, n, nd):
    out = []
    for side in (0, 1):
        e, p = f_elem[face, side], f_pos[face, side]
        k, lr = p // 2, p % 2
        nodes = [e * n ** nd + _line_node(nd, n, k, f, a) for a in range(n)]
        out.append((lp if lr else lm) @ Q[nodes])
    return out


@pytest.mark.filterwarnings("ignore::RuntimeWarning")
@pytest.mark.parametrize("nd,n,nel", [(1, 4, 6), (1, 5, 4), (2, 3, 3), (2, 4, 2)])
@pytest.mark.parametrize("surf", ["ec", "lxf"])
def test_entropy_production_identity(nd, n, nel, surf):
    savg, sdiss = (O.AVG_CHANDRASEKHAR, O.DISS_NONE) if surf == "ec" else (O.AVG_STD, O.DISS_LXF)
    lhs, pred, scale = _entropy_identity(nd, n, nel, savg, sdiss)
    print(nd, n, surf, lhs, pred)
    assert abs(lhs - pred) <= 1e-10 * scale


def test_entropy2cons_inverts_cons2entropy():
    for nd in (1, 2, 3):
        d = gauss(nd, 3, 2)
        X, _ = d.coords()
        Q = random_euler_state(X, nd, seed=1, amp=0.4, smooth=False)
        for q in Q[:20]:
            back = O.entropy2cons(nd, GAMMA, O.cons2entropy(nd, GAMMA, q))
            assert np.max(np.abs(back - q) / np.abs(q).max()) <= 1e-14
