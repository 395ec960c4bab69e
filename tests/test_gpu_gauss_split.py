"""Split form on Gauss nodes on the GPU (entropy-projected surface terms,
Equations/OpDivergence.jl:300-436) against the oracle.

FP64 plain kernels within 1e-12 of the oracle (pow, exp and log of the device library may
differ by an ulp); fused SSPRK33 steps within 1e-12; FP32 within 5e-5 of the FP64 oracle.
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA, random_euler_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TP = {O.TP_STDAVG: F.StdAverage, O.TP_CHANDRASEKHAR: F.ChandrasekharAverage,
      O.TP_KENNEDY_GRUBER: F.KennedyGruber}


def surface(savg, sdiss):
    avg = F.ChandrasekharAverage() if savg == O.AVG_CHANDRASEKHAR else F.StdAverage()
    return {O.DISS_NONE: lambda: avg, O.DISS_LXF: lambda: F.LxF(avg, 1.0),
            O.DISS_MATRIX: lambda: F.MatrixDissipation(avg, 1.0)}[sdiss]()


def build(nd, n, nel, tp, savg, sdiss, precision="fp64"):
    finish = [1.0, 1.5, 2.0][:nd]
    b = F.LagrangeBasis("GL", n)
    std = [None, F.StdSegment, F.StdQuad, F.StdHex][nd](b, F.DGSEMrec(b), nd + 2)
    mesh = F.CartesianMesh(nd, [0.0] * nd, finish, [nel] * nd)
    mesh.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
    eq = F.EulerEquation(nd, GAMMA)
    disc = F.MultielementDisc(mesh, std, eq, F.SplitDivOperator(TP[tp](), surface(savg, sdiss)),
                              precision=precision)
    orc = O.OracleDisc(nd, n, [nel] * nd, [0.0] * nd, finish, [1] * nd, O.build_ops(n, "GL"), O.EQ_EULER,
                       [GAMMA], O.VOL_SPLIT, tp, savg, sdiss, 1.0)
    orc.set_option(0, 0)  # corrected 3-D surface fluxes on both sides
    disc.set_option(0, 0)
    return disc, eq, orc


CASES = [(1, 4, 8, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_NONE),
         (2, 5, 3, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX),
         (2, 3, 4, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF),
         (2, 8, 2, O.TP_STDAVG, O.AVG_STD, O.DISS_LXF),
         (3, 4, 2, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF),
         (3, 6, 2, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX)]


def _ids(c):
    return f"{c[0]}d-n{c[1]}-tp{c[3]}-avg{c[4]}-diss{c[5]}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_gauss_split_rhs_fp64_matches_oracle(case):
    nd = case[0]
    disc, eq, orc = build(*case)
    X, _ = orc.coords()
    Q = random_euler_state(X, nd, seed=14, amp=0.2)
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
    disc.check()
    got, ref = disc.to_host(dQ), orc.rhs(Q)
    err = np.max(np.abs(got - ref) / np.max(np.abs(ref), axis=0))
    print(f"{_ids(case)}: max rel err {err:.2e}, bit-identical {np.mean(got == ref):.4f}")
    assert err <= 1e-12


@pytest.mark.parametrize("case", [CASES[1], CASES[4]], ids=["2d", "3d"])
def test_gauss_split_rhs_fp32_against_fp64_oracle(case):
    nd = case[0]
    disc, eq, orc = build(*case, precision="fp32")
    X, _ = orc.coords()
    Q = random_euler_state(X, nd, seed=14, amp=0.2).astype(np.float32).astype(np.float64)
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
    ref = orc.rhs(Q)
    err = np.max(np.abs(disc.to_host(dQ).astype(np.float64) - ref) / np.max(np.abs(ref), axis=0))
    print(f"FP32 Gauss split {_ids(case)}: max rel err {err:.2e}")
    assert err <= 5e-5


def test_gauss_split_ssprk33_steps_match_oracle():
    disc, eq, orc = build(*CASES[2])
    X, _ = orc.coords()
    Q = random_euler_state(X, 2, seed=15, amp=0.2)
    dt = 0.2 * orc.max_dt(Q, 1.0)
    ref = orc.rk_run(O.RK_SSPRK33, Q, dt, 5)
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, eq, F.SSPRK33, 5 * dt, dt=dt)
    err = np.max(np.abs(disc.to_host(Qd) - ref) / np.max(np.abs(ref), axis=0))
    print(f"Gauss split SSPRK33 5 steps: {err:.2e}")
    assert err <= 1e-12


def test_gauss_split_adjoint_is_refused():
    disc, eq, orc = build(*CASES[2])
    X, _ = orc.coords()
    Q = random_euler_state(X, 2, seed=1)
    from flouhip import adjoint as A
    Qbar = torch.zeros((4, orc.ndofs), dtype=torch.float64, device="cuda")
    with pytest.raises(NotImplementedError, match="Gauss"):
        A.rhs_vjp(Qbar, disc.to_device(Q), disc.to_device(Q), disc)
