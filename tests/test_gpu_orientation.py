"""Face orientations on the GPU (master2slave, StdSegment.jl:161-167, StdQuad.jl:161-181):
meshes whose elements number their nodes by rotations of the reference element
(tests/mesh_orient.py; every quad-face orientation 0..7 and both segment orientations
occur), against the oracle on the same mesh (pinned in tests/test_orientation.py):
the FP64 plain rhs with per-node (curved) and per-element (affine, non-diagonal metric)
geometry, FP32 (plain kernel: the fast kernels take orientation-0 meshes only), SSPRK33
steps, and the adjoint's dot-product identity against the oracle's dual-number JVP."""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA
from flouhip import adjoint as A
from test_orientation import build

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TPF = {O.TP_CHANDRASEKHAR: F.ChandrasekharAverage, O.TP_KENNEDY_GRUBER: F.KennedyGruber,
       O.TP_STDAVG: F.StdAverage}


def gpu_disc(nd, mesh, std, vol, tp, precision="fp64"):
    eq = F.EulerEquation(nd, GAMMA)
    if vol == "strong":
        op = F.StrongDivOperator(F.LxF(F.StdAverage(), 1.0))
    else:
        op = F.SplitDivOperator(TPF[tp](), F.LxF(F.StdAverage(), 1.0))
    return F.MultielementDisc(mesh, std, eq, op, precision=precision), eq


CASES = [(2, 4, 4, "split", O.TP_CHANDRASEKHAR, False), (2, 6, 3, "split", O.TP_KENNEDY_GRUBER, True),
         (2, 5, 3, "strong", 0, True), (3, 3, 3, "split", O.TP_CHANDRASEKHAR, False),
         (3, 4, 3, "split", O.TP_KENNEDY_GRUBER, True), (3, 6, 2, "split", O.TP_CHANDRASEKHAR, False),
         (3, 3, 2, "strong", 0, False)]


def _rel(a, b):
    return np.max(np.abs(a - b) / np.max(np.abs(b), axis=0))


# per-node geometry for every case; per-element geometry where the mesh is not warped
RHS_CASES = [(c, "curved") for c in CASES] + [(c, "affine") for c in CASES if not c[5]]


@pytest.mark.parametrize("case,geom", RHS_CASES,
                         ids=lambda c: "-".join(map(str, c)) if isinstance(c, tuple) else c)
def test_rhs_fp64_on_reoriented_mesh_matches_oracle(case, geom):
    nd, n, nel, vol, tp, curved = case
    base, mesh, ob, orr, Qb, Qr, std = build(nd, n, nel, vol, tp, curved=curved, affine=geom == "affine")
    disc, eq = gpu_disc(nd, mesh, std, vol, tp)
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Qr), F.EquationConfig(disc, eq))
    disc.check()
    got, ref = disc.to_host(dQ), orr.rhs(Qr)
    err = _rel(got, ref)
    print(f"{case} {geom}: orientations {sorted(set(mesh.face_orientation.tolist()))}, max rel err {err:.2e}")
    assert err <= 1e-13


@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[3], CASES[4]], ids=lambda c: "-".join(map(str, c)))
def test_rhs_fp32_on_reoriented_mesh(case):
    nd, n, nel, vol, tp, curved = case
    base, mesh, ob, orr, Qb, Qr, std = build(nd, n, nel, vol, tp, curved=curved)
    disc, eq = gpu_disc(nd, mesh, std, vol, tp, precision="fp32")
    assert not disc.uses_fast_kernel()
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Qr), F.EquationConfig(disc, eq))
    disc.check()
    err = _rel(disc.to_host(dQ).astype(np.float64), orr.rhs(Qr))
    print(f"{case} fp32: max rel err {err:.2e}")
    assert err <= 2e-5


def test_ssprk33_steps_on_reoriented_mesh():
    base, mesh, ob, orr, Qb, Qr, std = build(3, 4, 3, "split", O.TP_CHANDRASEKHAR, curved=True)
    disc, eq = gpu_disc(3, mesh, std, "split", O.TP_CHANDRASEKHAR)
    dt = 0.2 * orr.max_dt(Qr, 1.0)
    Q = disc.to_device(Qr)
    u1, u2 = disc.new_state(), disc.new_state()
    for _ in range(3):
        F.rk_stage(disc, F.SSPRK33.code, 0, dt, Q, u1)
        F.rk_stage(disc, F.SSPRK33.code, 1, dt, u1, u2, u_base=Q)
        F.rk_stage(disc, F.SSPRK33.code, 2, dt, u2, Q, u_base=Q)
    disc.check()
    ref = orr.rk_run(O.RK_SSPRK33, Qr, dt, 3)
    err = np.max(np.abs(disc.to_host(Q) - ref)) / np.max(np.abs(ref))
    print(f"3 SSPRK33 steps on a reoriented curved mesh: {err:.2e}")
    assert err <= 1e-13


VJP_CASES = [(CASES[0], "curved"), (CASES[3], "curved"), (CASES[4], "curved"), (CASES[5], "affine"),
             (CASES[0], "affine")]


@pytest.mark.parametrize("case,geom", VJP_CASES, ids=lambda c: "-".join(map(str, c)) if isinstance(c, tuple) else c)
def test_rhs_vjp_on_reoriented_mesh_matches_oracle_jvp(case, geom):
    nd, n, nel, vol, tp, curved = case
    base, mesh, ob, orr, Qb, Qr, std = build(nd, n, nel, vol, tp, curved=curved, affine=geom == "affine")
    disc, eq = gpu_disc(nd, mesh, std, vol, tp)
    rng = np.random.default_rng(4)
    V, W = rng.standard_normal(Qr.shape), rng.standard_normal(Qr.shape)
    _, JV = orr.rhs_jvp(Qr, V)
    lhs = float(np.sum(W * JV))
    Qbar = disc.to_device(np.zeros_like(Qr))
    A.rhs_vjp(Qbar, disc.to_device(W), disc.to_device(Qr), disc)
    rhs = float(np.sum(disc.to_host(Qbar) * V))
    err = abs(lhs - rhs) / float(np.sum(np.abs(W * JV)))
    print(f"{case} {geom}: <W,JV> = {lhs:.12e}  <J^T W,V> = {rhs:.12e}  rel {err:.2e}")
    assert err <= 1e-12
