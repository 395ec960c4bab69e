"""Curved (mapped) meshes: host geometry (flouhip.mesh.curved_geometry) and the oracle on it.

The metric terms are pinned by mathematics, not by the reference (Flou has trilinear maps
only): the identity map must give back the Cartesian geometry, the discrete metric
identities Σ_i D_i Ja^i = 0 must hold to round-off, J must converge spectrally to the
analytic Jacobian of the map, and on the periodic curved mesh the split form must preserve
free streams, conserve mass/momentum/energy and -- with the entropy-conserving surface
flux -- conserve entropy.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from cases import GAMMA, periodic_euler_state, random_euler_state

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "flou.jl_amd"))
from flouhip import mesh as M  # noqa: E402
from flouhip import stdregions as S  # noqa: E402

AMP = 0.05


def warp3(X):
    s = np.sin(2 * np.pi * X[:, [0]]) * np.sin(2 * np.pi * X[:, [1]]) * np.sin(2 * np.pi * X[:, [2]])
    return X + AMP * s


def warp2(X):
    s = np.sin(2 * np.pi * X[:, [0]]) * np.sin(2 * np.pi * X[:, [1]])
    return X + AMP * s


def curved(nd, n, nel, mapping=None):
    b = S.LagrangeBasis("GLL", n)
    std = [None, S.StdSegment, S.StdQuad, S.StdHex][nd](b, S.DGSEMrec(b), nd + 2)
    m = M.CartesianMesh(nd, (0.0,) * nd, (1.0,) * nd, (nel,) * nd)
    m.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
    m.apply_mapping(mapping or (warp3 if nd == 3 else warp2))
    return m, std


def oracle_curved(nd, n, nel, m, std, tp=O.TP_CHANDRASEKHAR, savg=O.AVG_STD, sdiss=O.DISS_LXF, vol=O.VOL_SPLIT):
    J, met, ev, fr, fj = M.curved_geometry(m, std)
    d = O.OracleDisc(nd, n, [nel] * nd, [0.0] * nd, [1.0] * nd, [1] * nd, O.build_ops(n, "GLL"),
                     O.EQ_EULER, [GAMMA], vol, tp, savg, sdiss, 1.0)
    d.set_option(0, 0)
    d.set_curved(met, J, fr.reshape(-1), fj, ev)
    return d, J


@pytest.mark.parametrize("nd", [2, 3])
def test_identity_map_gives_cartesian_geometry(nd):
    m, std = curved(nd, 5, 3, mapping=lambda X: X)
    J, met, ev, fr, fj = M.curved_geometry(m, std)
    ej, em, evol = m.element_geometry(std)
    ffr, ffj = m.face_geometry()
    assert np.abs(J - ej[0]).max() <= 1e-15
    assert np.abs(met - em[0]).max() <= 1e-13
    assert np.abs(ev - evol).max() <= 1e-15
    npf = std.n ** (nd - 1)
    assert np.abs(fj.reshape(m.nfaces, npf) - ffj[:, None]).max() <= 1e-13
    assert np.abs(fr.reshape(m.nfaces, npf, 3, nd)[:, :, 0] - ffr[:, None, 0]).max() <= 1e-12


@pytest.mark.parametrize("nd", [2, 3])
def test_discrete_metric_identities(nd):
    n, nel = 6, 3
    m, std = curved(nd, n, nel)
    _, met, _, _, _ = M.curved_geometry(m, std)
    Ja = met.reshape(m.nelements, n ** nd, nd, nd)  # [.., c, r] = (Ja^c)_r
    for r in range(nd):
        s = sum(M._dref(Ja[:, :, c, r], std.D, nd, n, c) for c in range(nd))
        assert np.abs(s).max() <= 1e-12


def test_jacobian_converges_to_the_analytic_one():
    """J of the interpolated map vs det(∂x'/∂x)·(Δx/2)^3 of the analytic map."""
    errs = []
    for n in (3, 5, 7, 9):
        m, std = curved(3, n, 2)
        J, _, ev, _, _ = M.curved_geometry(m, std)
        X = m._cartesian_coords(std.basis.ξ)
        x, y, z = (2 * np.pi * X[:, k] for k in range(3))
        s, c = np.sin, np.cos
        g = AMP * 2 * np.pi * np.stack([c(x) * s(y) * s(z), s(x) * c(y) * s(z), s(x) * s(y) * c(z)], 1)
        A = np.eye(3)[None] + np.ones(3)[None, :, None] * g[:, None, :]
        Jex = np.linalg.det(A) * (0.5 / 2) ** 3
        errs.append(np.abs(J - Jex).max() / Jex.max())
        assert abs(ev.sum() - 1.0) <= 1e-12  # the map preserves the unit cube's volume
    print("J errors", errs)
    assert errs[-1] < 1e-5 and errs[-1] < errs[0] * 1e-2


@pytest.mark.parametrize("nd", [2, 3])
def test_free_stream_and_conservation_on_curved_mesh(nd):
    n, nel = 5, 3
    m, std = curved(nd, n, nel)
    X = m.node_coords(std)
    for vol, tp in ((O.VOL_SPLIT, O.TP_CHANDRASEKHAR), (O.VOL_SPLIT, O.TP_KENNEDY_GRUBER), (O.VOL_STRONG, 0)):
        d, J = oracle_curved(nd, n, nel, m, std, tp=tp, vol=vol)
        Q = periodic_euler_state(X, nd, 1.0, seed=3, amp=0.1)
        dq = d.rhs(Q)
        Q0 = np.zeros_like(Q)
        Q0[:, 0] = 1.2
        Q0[:, 1:nd + 1] = [0.3, -0.2, 0.1][:nd]
        Q0[:, nd + 1] = 2.5
        fs = np.abs(d.rhs(np.asfortranarray(Q0))).max() / np.abs(dq).max()
        wq = J * np.tile(std.ω, nel ** nd)
        cons = max(abs(np.sum(wq * dq[:, v])) / np.sum(np.abs(wq * dq[:, v])) for v in range(nd + 2))
        print(f"nd={nd} vol={vol} tp={tp}: free stream {fs:.2e}, conservation {cons:.2e}")
        assert fs <= 1e-12
        assert cons <= 1e-13


def test_entropy_conservation_on_curved_mesh():
    """EC split form (Chandrasekhar volume and surface, no dissipation) on the curved mesh:
    Σ Jω w·dQ = 0 for a discontinuous state; StdAverage/LxF surfaces do not conserve it."""
    n, nel = 5, 3
    m, std = curved(3, n, nel)
    X = m.node_coords(std)
    Q = random_euler_state(X, 3, seed=5, amp=0.2, smooth=False)
    W = np.array([O.cons2entropy(3, GAMMA, q) for q in Q])
    res = {}
    for key, savg, sdiss in (("ec", O.AVG_CHANDRASEKHAR, O.DISS_NONE), ("avg", O.AVG_STD, O.DISS_NONE),
                             ("lxf", O.AVG_STD, O.DISS_LXF)):
        d, J = oracle_curved(3, n, nel, m, std, savg=savg, sdiss=sdiss)
        dq = d.rhs(Q)
        wq = J * np.tile(std.ω, nel ** 3)
        res[key] = np.sum(wq[:, None] * W * dq) / np.sum(np.abs(wq[:, None] * W * dq))
    print(res)
    assert abs(res["ec"]) <= 1e-12
    assert abs(res["avg"]) > 1e-6
    assert res["lxf"] < -1e-3
