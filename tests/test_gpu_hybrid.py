"""HybridDivOperator on the GPU against the oracle and the reference's Shockwave2D answer
(Equations/OpDivergence.jl:441-620, 781-801; test/runtests.jl:42-43).

FP64 plain kernels: rhs within 1e-12 of the oracle (logs and square roots of the device
library may differ by an ulp), Shockwave2D through the product API at the reference's
tolerance and at 1e-10.  FP32: within 2e-5 of the FP64 oracle.  The same on Gauss nodes
(_hybrid_nb_surface_contribution!, OpDivergence.jl:633-779), where pow/exp of the
entropy projection join the device-library functions.
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA, isapprox, random_euler_state, reference_answer

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TP = {O.TP_STDAVG: F.StdAverage, O.TP_CHANDRASEKHAR: F.ChandrasekharAverage,
      O.TP_KENNEDY_GRUBER: F.KennedyGruber}


def surface(savg, sdiss):
    avg = F.ChandrasekharAverage() if savg == O.AVG_CHANDRASEKHAR else F.StdAverage()
    return {O.DISS_NONE: lambda: avg, O.DISS_LXF: lambda: F.LxF(avg, 1.0),
            O.DISS_MATRIX: lambda: F.MatrixDissipation(avg, 1.0),
            O.DISS_SCALAR: lambda: F.ScalarDissipation(avg, 1.0)}[sdiss]()


def build(nd, n, nel, tp, savg, sdiss, blend, precision="fp64", nodes="GLL"):
    finish = [1.0, 1.5, 2.0][:nd]
    b = F.LagrangeBasis(nodes, n)
    std = [None, F.StdSegment, F.StdQuad, F.StdHex][nd](b, F.DGSEMrec(b), nd + 2)
    mesh = F.CartesianMesh(nd, [0.0] * nd, finish, [nel] * nd)
    mesh.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
    eq = F.EulerEquation(nd, GAMMA)
    op = F.HybridDivOperator(TP[tp](), surface(savg, sdiss), blend)
    disc = F.MultielementDisc(mesh, std, eq, op, precision=precision)
    orc = O.OracleDisc(nd, n, [nel] * nd, [0.0] * nd, finish, [1] * nd, O.build_ops(n, nodes), O.EQ_EULER,
                       [GAMMA], O.VOL_HYBRID, tp, savg, sdiss, 1.0)
    orc.set_blend(blend)
    orc.set_option(0, 0)  # corrected 3-D surface fluxes on both sides
    disc.set_option(0, 0)
    return disc, eq, orc


CASES = [(1, 5, 8, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1.0),
         (2, 6, 3, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1.0),
         (2, 4, 4, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF, 1e-2),
         (2, 8, 2, O.TP_STDAVG, O.AVG_STD, O.DISS_SCALAR, 0.1),
         (3, 4, 2, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF, 1.0),
         (3, 3, 3, O.TP_KENNEDY_GRUBER, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1e-3),
         (3, 6, 2, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1.0)]


def _ids(c):
    return f"{c[0]}d-n{c[1]}-tp{c[3]}-avg{c[4]}-diss{c[5]}-c{c[6]:g}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_hybrid_rhs_fp64_matches_oracle(case):
    nd, n, nel, tp, savg, sdiss, blend = case
    disc, eq, orc = build(*case)
    X, _ = orc.coords()
    Q = random_euler_state(X, nd, seed=12, amp=0.3, smooth=False)
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
    disc.check()
    got, ref = disc.to_host(dQ), orc.rhs(Q)
    err = np.max(np.abs(got - ref) / np.max(np.abs(ref), axis=0))
    print(f"{_ids(case)}: max rel err {err:.2e}, bit-identical {np.mean(got == ref):.4f}")
    assert err <= 1e-12


@pytest.mark.parametrize("case", [CASES[1], CASES[4]], ids=["2d", "3d"])
def test_hybrid_rhs_fp32_against_fp64_oracle(case):
    nd = case[0]
    disc, eq, orc = build(*case, precision="fp32")
    assert not disc.uses_fast_kernel()
    X, _ = orc.coords()
    Q = random_euler_state(X, nd, seed=12, amp=0.3, smooth=False).astype(np.float32).astype(np.float64)
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
    err = np.max(np.abs(disc.to_host(dQ).astype(np.float64) - orc.rhs(Q)) / np.max(np.abs(orc.rhs(Q)), axis=0))
    print(f"FP32 hybrid {_ids(case)}: max rel err {err:.2e}")
    assert err <= 2e-5


def test_hybrid_ork256_steps_match_oracle():
    case = CASES[1]
    disc, eq, orc = build(*case)
    X, _ = orc.coords()
    Q = random_euler_state(X, 2, seed=13, amp=0.2)
    dt = 0.2 * orc.max_dt(Q, 1.0)
    ref = orc.rk_run(O.RK_ORK256, Q, dt, 5)
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, eq, F.ORK256(), 5 * dt, dt=dt)
    err = np.max(np.abs(disc.to_host(Qd) - ref) / np.max(np.abs(ref), axis=0))
    print(f"hybrid ORK256 5 steps: {err:.2e}")
    assert err <= 1e-12


def test_shockwave2d_on_gpu_matches_reference_golden():
    """Shockwave2D (test/tests.jl:136-181) through the product API -> runtests.jl:42-43."""
    eq = F.EulerEquation(2, GAMMA)
    b = F.LagrangeBasis("GLL", 6)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    mesh = F.CartesianMesh(2, (-1.0, 0.0), (1.0, 1.0), (11, 3)).apply_periodicBCs(("3", "4"))
    r0, M0, p0 = 1.0, 2.0, 1.0
    u0 = M0 * F.soundvelocity(r0, p0, eq)
    r1, u1, p1 = F.normal_shockwave(r0, u0, p0, eq)
    Q0 = F.vars_prim2cons(np.array([r0, u0, 0.0, p0]), eq)
    Q1 = F.vars_prim2cons(np.array([r1, u1, 0.0, p1]), eq)

    def Qext(_, xy, __, ___, ____):
        return Q0 if xy[0] < 0 else Q1

    bcs = {"1": F.GenericBC(Qext), "2": F.GenericBC(Qext)}
    op = F.HybridDivOperator(F.MatrixDissipation(F.ChandrasekharAverage(), 1.0), 1.0)
    disc = F.MultielementDisc(mesh, std, eq, op, bcs)
    X = mesh.node_coords(std)
    Q = np.array([Qext(None, x, None, 0.0, eq) for x in X])
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, eq, F.ORK256(williamson_condition=False), 1.0, dt=1e-2)
    u = disc.to_host(Qd)
    ref = reference_answer("shockwave_2d")
    dev = abs(u.max() - ref["max"]) / ref["max"]
    print(f"Shockwave2D GPU: min {u.min():.6e} max {u.max():.15f} (rel dev {dev:.2e})")
    assert isapprox(u.max(), ref["max"], ref["max_rtol"])
    assert isapprox(u.min(), ref["min"], ref["min_rtol"])
    assert dev < 1e-10


def test_hybrid_unsupported_paths():
    disc, eq, orc = build(*CASES[2])
    X, _ = orc.coords()
    Q = random_euler_state(X, 2, seed=1)
    from flouhip import adjoint as A
    Qbar = torch.zeros((4, orc.ndofs), dtype=torch.float64, device="cuda")
    with pytest.raises(NotImplementedError, match="Hybrid"):
        A.rhs_vjp(Qbar, disc.to_device(Q), disc.to_device(Q), disc)
    b = F.LagrangeBasis("GL", 4)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    mesh = F.CartesianMesh(2, (0, 0), (1, 1), (2, 2)).apply_periodicBCs(("1", "2"), ("3", "4"))
    gdisc = F.MultielementDisc(mesh, std, eq, F.HybridDivOperator(F.LxF(F.StdAverage(), 1.0), 1.0))
    Xg = mesh.node_coords(std)
    Qg = random_euler_state(Xg, 2, seed=1)
    with pytest.raises(NotImplementedError, match="Hybrid"):
        A.rhs_vjp(torch.zeros((4, gdisc.ndofs), dtype=torch.float64, device="cuda"), gdisc.to_device(Qg),
                  gdisc.to_device(Qg), gdisc)


GAUSS = [(1, 4, 8, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1.0),
         (1, 6, 5, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF, 1e-2),
         (2, 4, 3, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_SCALAR, 0.1),
         (2, 6, 2, O.TP_STDAVG, O.AVG_STD, O.DISS_LXF, 1.0),
         (3, 3, 3, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1e-3),
         (3, 5, 2, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF, 1.0),
         (3, 7, 2, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1.0)]


@pytest.mark.parametrize("case", GAUSS, ids=_ids)
def test_gauss_hybrid_rhs_fp64_matches_oracle(case):
    nd = case[0]
    disc, eq, orc = build(*case, nodes="GL")
    X, _ = orc.coords()
    Q = random_euler_state(X, nd, seed=12, amp=0.2)
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
    disc.check()
    got, ref = disc.to_host(dQ), orc.rhs(Q)
    err = np.max(np.abs(got - ref) / np.max(np.abs(ref), axis=0))
    print(f"Gauss {_ids(case)}: max rel err {err:.2e}, bit-identical {np.mean(got == ref):.4f}")
    assert err <= 1e-12


@pytest.mark.parametrize("case", [GAUSS[2], GAUSS[5]], ids=["2d", "3d"])
def test_gauss_hybrid_fp32_and_steps(case):
    nd = case[0]
    disc, eq, orc = build(*case, precision="fp32", nodes="GL")
    X, _ = orc.coords()
    Q = random_euler_state(X, nd, seed=12, amp=0.2).astype(np.float32).astype(np.float64)
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
    ref = orc.rhs(Q)
    err = np.max(np.abs(disc.to_host(dQ).astype(np.float64) - ref) / np.max(np.abs(ref), axis=0))
    print(f"FP32 Gauss hybrid {_ids(case)}: max rel err {err:.2e}")
    assert err <= 2e-5
    d64, _, _ = build(*case, nodes="GL")
    dt = 0.2 * orc.max_dt(Q, 1.0)
    ref5 = orc.rk_run(O.RK_ORK256, Q, dt, 5)
    Qd = d64.to_device(Q)
    F.timeintegrate(Qd, d64, eq, F.ORK256(), 5 * dt, dt=dt)
    err5 = np.max(np.abs(d64.to_host(Qd) - ref5) / np.max(np.abs(ref5), axis=0))
    print(f"Gauss hybrid ORK256 5 steps FP64: {err5:.2e}")
    assert err5 <= 1e-12


def test_gauss_hybrid_large_blend_is_the_gauss_split_form_on_gpu():
    nd, n, nel = 3, 4, 2
    case = (nd, n, nel, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1e300)
    disc, eq, orc = build(*case, nodes="GL")
    b = F.LagrangeBasis("GL", n)
    std = F.StdHex(b, F.DGSEMrec(b), 5)
    mesh = F.CartesianMesh(3, [0.0] * 3, [1.0, 1.5, 2.0], [nel] * 3)
    mesh.apply_periodicBCs(("1", "2"), ("3", "4"), ("5", "6"))
    split = F.MultielementDisc(mesh, std, eq, F.SplitDivOperator(F.ChandrasekharAverage(),
                                                                 F.MatrixDissipation(F.ChandrasekharAverage(), 1.0)))
    split.set_option(0, 0)  # the same corrected 3-D surface flux as `disc` (build)
    X, _ = orc.coords()
    Q = random_euler_state(X, nd, seed=21, amp=0.2)
    a, c = disc.new_state(), split.new_state()
    F.rhs_(a, disc.to_device(Q), F.EquationConfig(disc, eq))
    F.rhs_(c, split.to_device(Q), F.EquationConfig(split, eq))
    ha, hc = disc.to_host(a), split.to_host(c)
    err = np.max(np.abs(ha - hc)) / np.max(np.abs(hc))
    print(f"GPU Gauss hybrid(c=1e300) vs GPU Gauss split: {err:.2e}")
    assert err <= 1e-13
