"""The workloads named for this path, end to end through the C ABI, against the oracle.

C1  2-D linear advection (1, 0.5), 16x16 periodic quads on [0,1]^2, N=4 GLL, strong form,
    upwind LxF, FP64, Gaussian pulse + seeded U(-1e-3, 1e-3), SSPRK33 at CFL 0.3,
    100 steps: the full run on the GPU against the oracle.
C2  2-D Euler, N=7 GLL, split Kennedy-Gruber + Rusanov, FP64 and FP32, SSPRK33 (here on a
    32x32 sub-size of the 256x256 mesh, 10 steps): FP64 against the oracle, FP32 error vs
    FP64 reported.
C3  3-D Euler N=5 split Chandrasekhar + Rusanov, CarpenterKennedy2N54, FP32 vs FP64 (a 4^3
    sub-size, 4 steps; the 64^3 run is the bench line).
(C4 and C5 are covered by test_gpu_adjoint.py and test_gpu_curved.py.)
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA, c2_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_c1_advection_full_run():
    b = F.LagrangeBasis("GLL", 5)
    std = F.StdQuad(b, F.DGSEMrec(b), 1)
    mesh = F.CartesianMesh(2, (0, 0), (1, 1), (16, 16)).apply_periodicBCs(("1", "2"), ("3", "4"))
    eq = F.LinearAdvection(1.0, 0.5)
    disc = F.MultielementDisc(mesh, std, eq, F.StrongDivOperator(F.LxF(F.StdAverage(), 1.0)))
    orc = O.OracleDisc(2, 5, [16, 16], [0.0, 0.0], [1.0, 1.0], [1, 1], O.build_ops(5, "GLL"),
                       O.EQ_ADVECTION, [1.0, 0.5], O.VOL_STRONG, 0, O.AVG_STD, O.DISS_LXF, 1.0)
    X = mesh.node_coords(std)
    rng = np.random.default_rng(1)
    Q0 = (np.exp(-((X[:, 0] - 0.5) ** 2 + (X[:, 1] - 0.5) ** 2) / 0.01)
          + rng.uniform(-1e-3, 1e-3, X.shape[0]))[:, None]
    Q0 = np.asfortranarray(Q0)
    Qd = disc.to_device(Q0)
    dt = F.get_max_dt(Qd, disc, eq, 0.3)
    assert dt == orc.max_dt(Q0, 0.3)
    ref = orc.rk_run(O.RK_SSPRK33, Q0, dt, 100)
    F.timeintegrate(Qd, disc, eq, F.SSPRK33, 100 * dt, dt=dt)
    got = disc.to_host(Qd)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"C1: dt {dt:.6e}, 100 SSPRK33 steps, max rel err vs oracle {err:.2e}, "
          f"bit-identical {np.mean(got == ref):.4f}")
    assert err <= 1e-13


def test_c2_kennedy_gruber_fp64_and_fp32():
    nel, n, nsteps = 32, 8, 10
    b = F.LagrangeBasis("GLL", n)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    mesh = F.CartesianMesh(2, (0, 0), (2, 2), (nel, nel)).apply_periodicBCs(("1", "2"), ("3", "4"))
    eq = F.EulerEquation(2, GAMMA)
    op = F.SplitDivOperator(F.KennedyGruber(), F.LxF(F.StdAverage(), 1.0))
    orc = O.OracleDisc(2, n, [nel, nel], [0.0, 0.0], [2.0, 2.0], [1, 1], O.build_ops(n, "GLL"),
                       O.EQ_EULER, [GAMMA], O.VOL_SPLIT, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF, 1.0)
    Q0 = c2_state(mesh.node_coords(std))
    dt = orc.max_dt(Q0, 0.1)
    ref = orc.rk_run(O.RK_SSPRK33, Q0, dt, nsteps)
    out = {}
    for prec in ("fp64", "fp32"):
        disc = F.MultielementDisc(mesh, std, eq, op, precision=prec)
        Qd = disc.to_device(Q0)
        F.timeintegrate(Qd, disc, eq, F.SSPRK33, nsteps * dt, dt=dt)
        out[prec] = disc.to_host(Qd).astype(np.float64)
    e64 = np.max(np.abs(out["fp64"] - ref)) / np.max(np.abs(ref))
    e32 = np.linalg.norm(out["fp32"][:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0])
    print(f"C2 sub-size: FP64 max rel err {e64:.2e}; FP32 density rel L2 err vs FP64 oracle {e32:.2e}")
    assert e64 <= 1e-12
    assert e32 <= 1e-5


def test_c3_subsize_fp32_vs_fp64():
    nel, n, nsteps = 4, 6, 4
    b = F.LagrangeBasis("GLL", n)
    std = F.StdHex(b, F.DGSEMrec(b), 5)
    L = 2 * np.pi
    mesh = F.CartesianMesh(3, (0, 0, 0), (L, L, L), (nel,) * 3)
    mesh.apply_periodicBCs(("1", "2"), ("3", "4"), ("5", "6"))
    eq = F.EulerEquation(3, GAMMA)
    op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))
    orc = O.OracleDisc(3, n, [nel] * 3, [0.0] * 3, [L] * 3, [1, 1, 1], O.build_ops(n, "GLL"),
                       O.EQ_EULER, [GAMMA], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF, 1.0)
    X = mesh.node_coords(std)
    rng = np.random.default_rng(1234)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    M0 = 0.1
    prim = np.stack([np.ones_like(x), M0 * np.sin(x) * np.cos(y) * np.cos(z),
                     -M0 * np.cos(x) * np.sin(y) * np.cos(z), np.zeros_like(x),
                     1 / GAMMA + (M0 ** 2 / 16) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)], 1)
    prim += 1e-3 * rng.standard_normal(prim.shape)
    Q0 = np.zeros_like(prim)
    Q0[:, 0] = prim[:, 0]
    Q0[:, 1:4] = prim[:, :1] * prim[:, 1:4]
    Q0[:, 4] = prim[:, 4] / (GAMMA - 1) + prim[:, 0] * (prim[:, 1:4] ** 2).sum(1) / 2
    Q0 = np.asfortranarray(Q0)
    dt = orc.max_dt(Q0, 0.3)
    ref = orc.rk_run(O.RK_CK2N54, Q0, dt, nsteps)
    disc = F.MultielementDisc(mesh, std, eq, op, precision="fp32")
    assert disc.uses_fast_kernel()
    Qd = disc.to_device(Q0)
    F.timeintegrate(Qd, disc, eq, F.CarpenterKennedy2N54(), nsteps * dt, dt=dt)
    got = disc.to_host(Qd).astype(np.float64)
    e32 = np.linalg.norm(got[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0])
    print(f"C3 sub-size: FP32 (fast kernel) density rel L2 err vs FP64 oracle {e32:.2e}")
    assert e32 <= 1e-5
