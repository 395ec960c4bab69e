"""Curved (mapped) meshes on the GPU: per-node metric and J, per-face-node frames.

FP64 plain kernels against the oracle on the same per-node geometry (≤ 1e-13), FP32 within
2e-5, free-stream preservation, the integral with per-node Jω, the adjoint's dot-product
identity against the oracle's dual-number JVP, and the mixed-precision adjoint of the
curved 3-D configuration (FP32 forward and adjoint, FP64 checkpoints) against FP64.
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA, periodic_euler_state, warp
from flouhip import adjoint as A
from flouhip import mesh as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")



TP = {"ch": (F.ChandrasekharAverage, O.TP_CHANDRASEKHAR), "kg": (F.KennedyGruber, O.TP_KENNEDY_GRUBER),
      "std": (F.StdAverage, O.TP_STDAVG)}


def build_curved(nd, n, nel, vol="split", tp="ch", precision="fp64"):
    b = F.LagrangeBasis("GLL", n)
    std = [None, None, F.StdQuad, F.StdHex][nd](b, F.DGSEMrec(b), nd + 2)
    mesh = F.CartesianMesh(nd, (0.0,) * nd, (1.0,) * nd, (nel,) * nd)
    mesh.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
    mesh.apply_mapping(warp)
    eq = F.EulerEquation(nd, GAMMA)
    if vol == "strong":
        op = F.StrongDivOperator(F.LxF(F.StdAverage(), 1.0))
        volc, tpc = O.VOL_STRONG, 0
    else:
        op = F.SplitDivOperator(TP[tp][0](), F.LxF(F.StdAverage(), 1.0))
        volc, tpc = O.VOL_SPLIT, TP[tp][1]
    disc = F.MultielementDisc(mesh, std, eq, op, precision=precision)
    J, met, ev, fr, fj = M.curved_geometry(mesh, std)
    orc = O.OracleDisc(nd, n, [nel] * nd, [0.0] * nd, [1.0] * nd, [1] * nd, O.build_ops(n, "GLL"),
                       O.EQ_EULER, [GAMMA], volc, tpc, O.AVG_STD, O.DISS_LXF, 1.0)
    orc.set_curved(met, J, fr.reshape(-1), fj, ev)
    X = mesh.node_coords(std)
    return disc, orc, eq, X, J, std


def rel_err(a, b):
    scale = np.max(np.abs(b), axis=0)
    return np.max(np.abs(a - b) / scale)


CASES = [(2, 6, 4, "split", "ch"), (2, 5, 4, "split", "kg"), (2, 4, 4, "strong", "std"),
         (3, 5, 3, "split", "ch"), (3, 4, 3, "strong", "std"), (3, 7, 2, "split", "kg")]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_curved_rhs_fp64_matches_oracle(case):
    nd, n, nel, vol, tp = case
    disc, orc, eq, X, J, std = build_curved(nd, n, nel, vol, tp)
    Q = periodic_euler_state(X, nd, 1.0, seed=7, amp=0.2)
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
    disc.check()
    got, ref = disc.to_host(dQ), orc.rhs(Q)
    err = rel_err(got, ref)
    print(f"{case}: max rel err {err:.2e}, bit-identical {np.mean(got == ref):.4f}")
    assert err <= 1e-13


FP32_CURVED = [CASES[0], CASES[1], CASES[3], CASES[5], (3, 7, 2, "split", "ch"), (3, 6, 3, "split", "std"),
               (2, 8, 3, "split", "ch"), (3, 4, 3, "split", "kg")]


@pytest.mark.parametrize("case", FP32_CURVED, ids=lambda c: "-".join(map(str, c)))
def test_curved_rhs_fp32_against_fp64_oracle(case):
    """FP32 on the mapped mesh: the fast curved kernel and the plain FP32 kernel against the
    FP64 oracle."""
    nd, n, nel, vol, tp = case
    disc, orc, eq, X, J, std = build_curved(nd, n, nel, vol, tp, precision="fp32")
    assert disc.uses_fast_kernel()
    Q = periodic_euler_state(X, nd, 1.0, seed=7, amp=0.2)
    ref = orc.rhs(Q)
    Qd = disc.to_device(Q)
    # FP32 round-off of the volume term grows with |D♯| ~ N²: 2e-5 up to N = 6, 3e-5 at N = 7
    tol = 2e-5 if n <= 7 else 3e-5
    for plain in (0, 1):
        disc.set_option(1, plain)
        dQ = disc.new_state()
        F.rhs_(dQ, Qd, F.EquationConfig(disc, eq))
        disc.check()
        err = rel_err(disc.to_host(dQ).astype(np.float64), ref)
        print(f"FP32 curved {'plain' if plain else 'fast'} {case}: max rel err {err:.2e}")
        assert err <= tol


@pytest.mark.parametrize("nd,n,nel", [(3, 7, 3), (3, 6, 3), (2, 7, 5)], ids=["3d-n7", "3d-n6", "2d-n7"])
def test_curved_fast_kernel_stages_and_partial_blocks(nd, n, nel):
    """Every fused stage form of the fast curved kernel against the plain FP32 kernel on the
    same handle (3^3 elements: not a multiple of 2 elements per block at N=5)."""
    disc, orc, eq, X, J, std = build_curved(nd, n, nel, "split", "ch", precision="fp32")
    assert disc.uses_fast_kernel()
    Q = periodic_euler_state(X, nd, 1.0, seed=3, amp=0.2)
    dt = 0.1 * orc.max_dt(Q, 1.0)
    for solver in (F.SSPRK33, F.CarpenterKennedy2N54):
        outs = []
        for plain in (0, 1):
            disc.set_option(1, plain)
            Qd = disc.to_device(Q)
            F.timeintegrate(Qd, disc, eq, solver, 2 * dt, dt=dt)
            outs.append(disc.to_host(Qd).astype(np.float64))
        disc.set_option(1, 0)
        err = rel_err(outs[0], outs[1])
        print(f"curved fast vs plain FP32, {solver.__name__} x2 ({nd}-D n={n}): {err:.2e}")
        assert err < 1e-5


def test_curved_fast_kernel_wavy_slip_walls():
    """The fast curved kernel with slip, inflow and outflow boundaries on a mesh whose walls
    are wavy (rotated face frames), against the FP64 oracle; mass and energy through the
    slip walls vanish."""
    n, ne = 6, 4
    b = F.LagrangeBasis("GLL", n)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    res = {}
    for kind in ("slip", "inflow-outflow"):
        mesh = F.CartesianMesh(2, (0.0, 0.0), (1.0, 1.0), (ne, ne))
        mesh.apply_periodicBCs(("1", "2"))

        def wavy(X):
            Y = X.copy()
            Y[:, 1] = X[:, 1] + 0.08 * np.sin(2 * np.pi * X[:, 0])
            return Y

        mesh.apply_mapping(wavy)
        eq = F.EulerEquation(2, GAMMA)
        op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))
        qin = np.array([1.1, 0.2, 0.1, 2.6])
        if kind == "slip":
            bcs = {"3": F.EulerSlipBC(), "4": F.EulerSlipBC()}
        else:
            bcs = {"3": F.EulerInflowBC(qin), "4": F.EulerOutflowBC()}
        disc = F.MultielementDisc(mesh, std, eq, op, bcs, precision="fp32")
        assert disc.uses_fast_kernel()
        J, met, ev, fr, fj = M.curved_geometry(mesh, std)
        orc = O.OracleDisc(2, n, [ne, ne], [0.0, 0.0], [1.0, 1.0], [1, 0], O.build_ops(n, "GLL"), O.EQ_EULER,
                           [GAMMA], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF, 1.0)
        orc.set_curved(met, J, fr.reshape(-1), fj, ev)
        if kind == "slip":
            orc.set_bc(0, O.BC_SLIP)
            orc.set_bc(1, O.BC_SLIP)
        else:
            orc.set_bc(0, O.BC_INFLOW, qin)
            orc.set_bc(1, O.BC_OUTFLOW)
        X = mesh.node_coords(std)
        from cases import random_euler_state
        Q = random_euler_state(X, 2, seed=8, amp=0.2)
        ref = orc.rhs(Q)
        dQ = disc.new_state()
        F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
        disc.check()
        got = disc.to_host(dQ).astype(np.float64)
        err = rel_err(got, ref)
        print(f"fast curved FP32 with {kind} walls: max rel err vs FP64 oracle {err:.2e}")
        assert err <= 2e-5
        res[kind] = (got, J, std)
    got, J, std = res["slip"]
    jw = J * np.tile(std.ω, ne * ne)
    for v in (0, 3):
        tot = float(np.sum(jw * got[:, v]))
        assert abs(tot) <= 1e-5 * float(np.sum(jw * np.abs(got[:, v])))


def test_curved_fp32_ck2n54_steps_against_fp64_oracle():
    """3-D N=6 curved, FP32 fused CK2N54 steps (the register operand loaded before the
    staging barrier) against the FP64 oracle's steps."""
    disc, orc, eq, X, J, std = build_curved(3, 7, 2, precision="fp32")
    Q = periodic_euler_state(X, 3, 1.0, seed=5, amp=0.2)
    dt = 0.1 * orc.max_dt(Q, 1.0)
    ref = orc.rk_run(O.RK_CK2N54, Q, dt, 3)
    uT, _ = A.forward(disc, F.CarpenterKennedy2N54, disc.to_device(Q), dt, 3)
    got = disc.to_host(uT)
    err = rel_err(got.astype(np.float64), ref)
    print(f"FP32 curved CK2N54 x3: max rel err {err:.2e}")
    assert err <= 2e-5


def test_curved_free_stream_and_integral_on_gpu():
    disc, orc, eq, X, J, std = build_curved(3, 6, 3)
    Q = periodic_euler_state(X, 3, 1.0, seed=7, amp=0.2)
    Q0 = np.zeros_like(Q)
    Q0[:, 0], Q0[:, 1:4], Q0[:, 4] = 1.2, [0.3, -0.2, 0.1], 2.5
    dQ, dQ0 = disc.new_state(), disc.new_state()
    cfg = F.EquationConfig(disc, eq)
    F.rhs_(dQ, disc.to_device(Q), cfg)
    F.rhs_(dQ0, disc.to_device(Q0), cfg)
    fs = float(dQ0.abs().max() / dQ.abs().max())
    print(f"curved free stream on GPU: {fs:.2e}")
    assert fs <= 1e-12
    Qd = disc.to_device(Q)
    for v in range(5):
        got = F.integrate(Qd, disc, v)
        ref = orc.integrate(Q, v)
        assert abs(got - ref) <= 1e-14 * abs(ref)


@pytest.mark.parametrize("case", [CASES[0], CASES[3], CASES[4], CASES[5]], ids=["2d-ch", "3d-ch", "3d-strong", "3d-kg"])
def test_curved_vjp_dot_product_with_oracle_jvp(case):
    nd, n, nel, vol, tp = case
    disc, orc, eq, X, J, std = build_curved(nd, n, nel, vol, tp)
    Q = periodic_euler_state(X, nd, 1.0, seed=9, amp=0.2)
    rng = np.random.default_rng(4)
    V, W = rng.standard_normal(Q.shape), rng.standard_normal(Q.shape)
    _, JV = orc.rhs_jvp(Q, V)
    Qbar = torch.zeros((nd + 2, Q.shape[0]), dtype=torch.float64, device="cuda")
    A.rhs_vjp(Qbar, disc.to_device(W), disc.to_device(Q), disc)
    lhs, rhs = float(np.sum(W * JV)), float(np.sum(disc.to_host(Qbar) * V))
    err = abs(lhs - rhs) / float(np.sum(np.abs(W * JV)))
    print(f"curved VJP {case}: rel {err:.2e}")
    assert err <= 1e-12


def test_curved_mixed_precision_adjoint():
    """3-D curved mesh, N=6: FP32 forward and adjoint with FP64 checkpoints every 10 of 30
    SSPRK33 steps, against the FP64 adjoint of the same loss (mean-square density)."""
    res = {}
    for prec in ("fp64", "fp32"):
        disc, orc, eq, X, J, std = build_curved(3, 7, 3, precision=prec)
        Q = periodic_euler_state(X, 3, 1.0, seed=11, amp=0.1)
        dt = 0.1 * orc.max_dt(Q, 1.0)
        wq = torch.tensor(J * np.tile(std.ω, 27), device="cuda", dtype=disc.torch_dtype())

        def seed(u):
            g = torch.zeros_like(u)
            g[0] = 2.0 * wq * (u[0] - 1.0)
            return g

        res[prec] = disc.to_host(A.gradient(disc, F.SSPRK33, disc.to_device(Q), dt, 30, seed,
                                            checkpoint_every=10, checkpoint_dtype=torch.float64)).astype(np.float64)
    err = np.linalg.norm(res["fp32"] - res["fp64"]) / np.linalg.norm(res["fp64"])
    print(f"curved FP32 adjoint vs FP64 adjoint: relative L2 error {err:.2e}")
    assert err <= 1e-4
