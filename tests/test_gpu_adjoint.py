"""Adjoint of rhs! and of the RK steps on the GPU (flou_hip_rhs_vjp, flou_hip_rk_step_vjp).

Pins:
  * dot-product identity against the oracle's dual-number JVP: <W, J V> (oracle, CPU) equals
    <J^T W, V> (GPU) for random V, W -- exact up to round-off, so a wrong sign, a dropped
    term or a transposed operand fails it;
  * central finite differences of an L2 density misfit through 20 SSPRK33 steps along 16
    seeded unit directions (h = 1e-6), the adjoint gradient within 1e-6 relative;
  * the full gradient of that misfit on a 4x4, N=3 mesh, column by column from dual-number
    JVPs of the oracle;
  * FP32 adjoint against the FP64 adjoint.
Every Euler surface flux is covered: StdAverage with LxF or none by the hand-written face
VJP, the Chandrasekhar average, LxF on it, Scalar and Matrix dissipation by the dual-number
face VJP.
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from flouhip import adjoint as A
from cases import golden, periodic_euler_state
from test_gpu_parity import _state, build

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

VJP_CASES = [
    # nd, n, nel, vol, tp, surf, eq, nodes, bcs
    (1, 4, [5], "strong", "std", "lxf", "euler", "GLL", None),
    (1, 5, [4], "split", "ch", "lxf", "euler", "GLL", None),
    (1, 4, [6], "split", "kg", "avg", "euler", "GLL", "dirichlet"),
    (2, 3, [3, 3], "strong", "std", "lxf", "euler", "GLL", None),
    (2, 4, [3, 3], "split", "ch", "lxf", "euler", "GLL", None),
    (2, 6, [3, 2], "split", "ch", "lxf", "euler", "GLL", None),
    (2, 5, [3, 3], "split", "kg", "lxf", "euler", "GLL", None),
    (2, 4, [3, 3], "split", "std", "avg", "euler", "GLL", None),
    (2, 4, [3, 3], "split", "ch", "lxf", "euler", "GLL", "walls"),
    (2, 6, [3, 3], "split", "ch", "lxf", "euler", "GLL", "walls"),  # up-front loads, boundaries
    (3, 6, [2, 1, 2], "split", "ch", "lxf", "euler", "GLL", None),  # up-front loads, 3-D
    (3, 4, [2, 2, 2], "split", "ch", "lxf", "euler", "GLL", None),
    (3, 3, [2, 2, 2], "strong", "std", "lxf", "euler", "GL", None),
    (3, 5, [2, 1, 2], "split", "kg", "lxf", "euler", "GLL", None),
    (2, 5, [4, 4], "strong", "std", "lxf", "adv", "GLL", None),
    (2, 5, [4, 4], "split", "std", "lxf", "adv", "GLL", None),
    (2, 4, [4, 3], "strong", "std", "lxf", "adv", "GL", None),
    # surface fluxes whose face VJP is formed from dual-number evaluations of the flux
    (1, 4, [6], "split", "ch", "matrix", "euler", "GLL", None),
    (2, 4, [3, 3], "split", "ch", "scalar", "euler", "GLL", None),
    (2, 4, [3, 3], "split", "ch", "matrix", "euler", "GLL", "walls"),
    (2, 5, [3, 3], "split", "kg", "chavg", "euler", "GLL", "walls"),
    (3, 4, [2, 2, 2], "split", "ch", "matrix", "euler", "GLL", None),
    (3, 4, [2, 1, 2], "split", "ch", "lxf-ch", "euler", "GLL", None),
    (3, 3, [2, 2, 2], "strong", "std", "scalar-std", "euler", "GLL", None),
    (3, 4, [2, 2, 2], "split", "ch", "scalar", "euler", "GLL", None),
]


def _find(nd, n, vol, tp, bcs, eq="euler", surf="lxf"):
    """The VJP case with these values (picked by value, so inserting a case moves nothing)."""
    hits = [c for c in VJP_CASES if (c[0], c[1], c[3], c[4], c[5], c[6], c[8]) == (nd, n, vol, tp, surf, eq, bcs)]
    assert len(hits) == 1, hits
    return hits[0]


def _disc(case, precision="fp64"):
    nd, n, nel, vol, tp, surf, eq, nodes, bcs = case
    periodic = None
    if bcs == "walls":
        periodic = [0, 0]
    if bcs == "dirichlet":
        periodic = [0]
    mesh, std, equation, op, orc = build(nd, n, nel, vol, tp, surf, periodic=periodic, nodes=nodes, eq=eq)
    bc = {}
    if bcs == "walls":
        qin = np.array([1.1, 0.3, 0.0, 2.7])
        bc = {"1": F.EulerInflowBC(qin), "2": F.EulerOutflowBC(), "3": F.EulerSlipBC(), "4": F.EulerSlipBC()}
        orc.set_bc(0, O.BC_INFLOW, qin)
        orc.set_bc(1, O.BC_OUTFLOW)
        orc.set_bc(2, O.BC_SLIP)
        orc.set_bc(3, O.BC_SLIP)
    if bcs == "dirichlet":
        def qext(_, x, __, ___, eq_):
            return F.vars_prim2cons(np.array([1.0 + 0.1 * x[0], 0.2, 1.0]), eq_)
        bc = {"1": F.GenericBC(qext), "2": F.EulerOutflowBC()}
        _, fX = orc.coords()
        faces = np.array([qext(None, x, None, 0.0, equation) for x in fX.reshape(-1, nd)])
        orc.set_bc(0, O.BC_DIRICHLET, faces)
        orc.set_bc(1, O.BC_OUTFLOW)
    disc = F.MultielementDisc(mesh, std, equation, op, bc, precision=precision)
    return disc, orc, equation


def _dot_check(lhs, rhs, terms, tol):
    err = abs(lhs - rhs) / max(terms, 1e-300)
    return err, err <= tol


@pytest.mark.parametrize("case", VJP_CASES, ids=lambda c: "-".join(str(x) for x in c))
def test_rhs_vjp_dot_product_with_oracle_jvp(case):
    nd, eq = case[0], case[6]
    disc, orc, equation = _disc(case)
    Q = _state(disc, orc, nd, eq, seed=5, amp=0.2)
    rng = np.random.default_rng(11)
    V = rng.standard_normal(Q.shape)
    W = rng.standard_normal(Q.shape)
    _, JV = orc.rhs_jvp(Q, V)
    lhs = float(np.sum(W * JV))
    Qd, Wd = disc.to_device(Q), disc.to_device(W)
    Qbar0 = rng.standard_normal(Q.shape)
    Qbar = disc.to_device(Qbar0)
    A.rhs_vjp(Qbar, Wd, Qd, disc)
    JTW = disc.to_host(Qbar) - Qbar0  # rhs_vjp accumulates
    rhs = float(np.sum(JTW * V))
    err, ok = _dot_check(lhs, rhs, float(np.sum(np.abs(W * JV))), 1e-12)
    print(f"{case}: <W,JV> = {lhs:.12e}  <J^T W,V> = {rhs:.12e}  rel {err:.2e}")
    assert ok


def test_rhs_vjp_autograd_function():
    case = _find(2, 6, "split", "ch", None)
    disc, orc, equation = _disc(case)
    Q = _state(disc, orc, 2, "euler", seed=3)
    Qd = disc.to_device(Q).requires_grad_(True)
    dQ = A.Rhs.apply(Qd, disc)
    W = torch.randn_like(dQ)
    (dQ * W).sum().backward()
    Qbar = torch.zeros_like(Qd)
    A.rhs_vjp(Qbar, W.detach().contiguous(), Qd.detach(), disc)
    torch.cuda.synchronize()
    assert torch.equal(Qd.grad, Qbar)


SCHEMES = [(F.SSPRK33, O.RK_SSPRK33), (F.CarpenterKennedy2N54, O.RK_CK2N54), (F.ORK256, O.RK_ORK256)]


RK_VJP_CASES = {
    "2d-ch": _find(2, 4, "split", "ch", None),
    "3d-ch": _find(3, 4, "split", "ch", None),
    "3d-ch-n6": _find(3, 6, "split", "ch", None),    # FP64 3-D N=6: epilogue operands loaded early
    "2d-walls": _find(2, 4, "split", "ch", "walls"),
    "2d-walls-n6": _find(2, 6, "split", "ch", "walls"),
    "2d-matrix-walls": _find(2, 4, "split", "ch", "walls", surf="matrix"),  # dual-number face VJP
    "3d-scalar": _find(3, 4, "split", "ch", None, surf="scalar"),
}


@pytest.mark.parametrize("scheme", SCHEMES, ids=lambda s: s[0].__name__)
@pytest.mark.parametrize("case", list(RK_VJP_CASES.values()), ids=list(RK_VJP_CASES))
def test_rk_steps_vjp_dot_product_with_oracle_jvp(case, scheme):
    solver, ocode = scheme
    nd = case[0]
    disc, orc, equation = _disc(case)
    Q = _state(disc, orc, nd, "euler", seed=9, amp=0.2)
    # CFL 0.1: at 0.25 the 3-D N=5 state of the [2, 1, 2] mesh turns non-physical within
    # 5 SSPRK33/ORK256 steps (in the oracle as well)
    dt = 0.1 * orc.max_dt(Q, 1.0)
    nsteps = 5
    rng = np.random.default_rng(2)
    V = rng.standard_normal(Q.shape)
    W = rng.standard_normal(Q.shape)
    uref, du = orc.rk_run_jvp(ocode, Q, V, dt, nsteps)
    lhs = float(np.sum(W * du))
    Wd = disc.to_device(W)
    ubar, uT = A.gradient(disc, solver, disc.to_device(Q), dt, nsteps, lambda u: Wd,
                          checkpoint_every=2, return_state=True)
    rhs = float(np.sum(disc.to_host(ubar) * V))
    err, ok = _dot_check(lhs, rhs, float(np.sum(np.abs(W * du))), 1e-11)
    fwd = np.max(np.abs(disc.to_host(uT) - uref)) / np.max(np.abs(uref))
    print(f"{solver.__name__} {case[:7]}: dot-product rel {err:.2e}, forward rel {fwd:.2e}")
    assert fwd <= 1e-12
    assert ok


def _misfit_problem(nel, n, nsteps, precision="fp64"):
    """C4-shaped problem: 2-D Euler on [0,2]^2, split Chandrasekhar + LxF, SSPRK33, fixed
    dt; loss = ||ρ_A(T) - ρ_B(T)||²_{L2, Jω}."""
    mesh, std, equation, op, orc = build(2, n, [nel, nel], "split", "ch", "lxf", finish=[2.0, 2.0])
    disc = F.MultielementDisc(mesh, std, equation, op, precision=precision)
    X, _ = orc.coords()
    QA = periodic_euler_state(X, 2, 2.0, seed=101, amp=0.1)
    QB = periodic_euler_state(X, 2, 2.0, seed=202, amp=0.1)
    dt = 0.1 * orc.max_dt(QA, 1.0)
    w1 = np.asarray(std.basis.ω, dtype=np.float64)
    jac = (2.0 / nel / 2.0) ** 2  # Cartesian element: J = (Δx/2)(Δy/2)
    wnode = np.tile(np.kron(w1, w1) * jac, nel * nel)
    rhoB = orc.rk_run(O.RK_SSPRK33, QB, dt, nsteps)[:, 0]
    return disc, orc, QA, rhoB, wnode, dt


def _loss_host(u, rhoB, wnode):
    d = u[:, 0] - rhoB
    return float(np.sum(wnode * d * d))


def test_adjoint_gradient_matches_central_differences():
    nsteps = 20
    disc, orc, QA, rhoB, wnode, dt = _misfit_problem(8, 4, nsteps)
    wdev = torch.tensor(wnode, device="cuda")
    rdev = torch.tensor(rhoB, device="cuda")

    def seed(u):
        g = torch.zeros_like(u)
        g[0] = 2.0 * wdev * (u[0] - rdev)
        return g

    g = disc.to_host(A.gradient(disc, F.SSPRK33, disc.to_device(QA), dt, nsteps, seed, checkpoint_every=7))

    def L(Q):
        u, _ = A.forward(disc, F.SSPRK33, disc.to_device(Q), dt, nsteps)
        return _loss_host(disc.to_host(u), rhoB, wnode)

    rng = np.random.default_rng(16)
    h = 1e-6
    fds, ads = [], []
    for k in range(16):
        d = rng.standard_normal(QA.shape)
        d /= np.linalg.norm(d)
        fds.append((L(QA + h * d) - L(QA - h * d)) / (2 * h))
        ads.append(float(np.sum(g * d)))
    fds, ads = np.array(fds), np.array(ads)
    # relative to the RMS directional derivative: a direction nearly orthogonal to the
    # gradient has a tiny derivative that the FD round-off (~eps·L/h) cannot resolve
    scale = np.sqrt(np.mean(ads ** 2))
    rel = np.abs(fds - ads) / scale
    for k in range(16):
        print(f"direction {k}: FD {fds[k]:.10e} adjoint {ads[k]:.10e} "
              f"rel to RMS {rel[k]:.2e}, rel to itself {abs(fds[k] - ads[k]) / abs(fds[k]):.2e}")
    assert rel.max() <= 1e-6


def test_kept_segments_give_the_recomputed_gradient():
    """Reversing segments from the stage inputs the forward run kept gives bitwise the
    gradient of the recomputing sweep (17 steps, checkpoints every 5: a short last
    segment), for none, the last, two and all four segments kept and for keep_segments
    "auto" (sized from the free device memory), SSPRK33 and CK2N54."""
    nsteps = 17
    disc, orc, QA, rhoB, wnode, dt = _misfit_problem(8, 4, nsteps)
    wdev = torch.tensor(wnode, device="cuda")
    rdev = torch.tensor(rhoB, device="cuda")

    def seed(u):
        g = torch.zeros_like(u)
        g[0] = 2.0 * wdev * (u[0] - rdev)
        return g

    for solver in (F.SSPRK33, F.CarpenterKennedy2N54):
        res = {}
        for keep in (0, 1, 2, 4):
            uT, cps = A.forward(disc, solver, disc.to_device(QA), dt, nsteps, 5, keep_segments=keep)
            assert len(cps) == 4 and sorted(cps.stages) == list(range(4 - keep, 4))
            ubar = seed(uT)
            A.reverse(disc, solver, cps, dt, nsteps, ubar, 5)
            res[keep] = (disc.to_host(uT), disc.to_host(ubar))
        for keep in (1, 2, 4):
            assert np.array_equal(res[keep][0], res[0][0]), (solver, keep)
            assert np.array_equal(res[keep][1], res[0][1]), (solver, keep)
        g = disc.to_host(A.gradient(disc, solver, disc.to_device(QA), dt, nsteps, seed, checkpoint_every=5))
        assert np.array_equal(g, res[0][1])
        # "auto": as many segments as fit in half of the free device memory (here all four)
        assert A.auto_keep_segments(disc, solver, nsteps, 5) == 4
        uT, cps = A.forward(disc, solver, disc.to_device(QA), dt, nsteps, 5, keep_segments="auto")
        assert sorted(cps.stages) == [0, 1, 2, 3]
        g = disc.to_host(A.gradient(disc, solver, disc.to_device(QA), dt, nsteps, seed, checkpoint_every=5,
                                    keep_segments="auto"))
        assert np.array_equal(g, res[0][1])
        # a budget for one segment keeps one
        assert A.auto_keep_segments(disc, solver, nsteps, 5, fraction=1e-300) == 1


def test_adjoint_full_gradient_matches_dual_numbers():
    """Every component of the gradient on a 4x4, N=3 mesh from oracle dual-number JVPs."""
    nsteps = 5
    disc, orc, QA, rhoB, wnode, dt = _misfit_problem(4, 4, nsteps)
    u = orc.rk_run(O.RK_SSPRK33, QA, dt, nsteps)
    dLdu = np.zeros_like(u)
    dLdu[:, 0] = 2.0 * wnode * (u[:, 0] - rhoB)
    ref = np.zeros(QA.size)
    E = np.zeros(QA.shape, order="F")
    flat = E.reshape(-1, order="F")
    for i in range(QA.size):
        flat[:] = 0.0
        flat[i] = 1.0
        _, du = orc.rk_run_jvp(O.RK_SSPRK33, QA, E, dt, nsteps)
        ref[i] = np.sum(dLdu * du)
    ref = ref.reshape(QA.shape, order="F")
    seed_t = disc.to_device(dLdu)
    g = disc.to_host(A.gradient(disc, F.SSPRK33, disc.to_device(QA), dt, nsteps, lambda _: seed_t))
    err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
    print(f"full gradient vs dual numbers: max rel err {err:.2e}")
    assert err <= 1e-11


def test_c4_subsize_gradient_matches_dual_number_fixture():
    """C4 sub-size (8x8, N=3, 200 SSPRK33 steps, checkpoints every 20): every component of
    the gradient against the oracle's dual-number JVPs stored in
    tests/golden/c4_gradient_8x8_n3.npz (the ForwardDiff gradient of the reference path)."""
    g, meta = golden("c4_gradient_8x8_n3.npz")
    mesh, std, equation, op, orc = build(2, meta["n"], meta["nel"], "split", "ch", "lxf", finish=[2.0, 2.0])
    disc = F.MultielementDisc(mesh, std, equation, op)
    wdev = torch.tensor(g["wnode"], device="cuda")
    rdev = torch.tensor(g["rhoB"], device="cuda")

    def seed(u):
        s = torch.zeros_like(u)
        s[0] = 2.0 * wdev * (u[0] - rdev)
        return s

    grad = disc.to_host(A.gradient(disc, F.SSPRK33, disc.to_device(g["QA"]), float(g["dt"]), meta["steps"], seed,
                                   checkpoint_every=20))
    ref = g["grad"]
    err = np.max(np.abs(grad - ref)) / np.max(np.abs(ref))
    l2 = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    print(f"C4 8x8 N=3, 200 steps: gradient vs dual numbers max rel {err:.2e}, rel L2 {l2:.2e} "
          f"({ref.size} components)")
    assert err <= 1e-10


def test_fp32_adjoint_against_fp64_adjoint():
    nsteps = 10
    res = {}
    for prec in ("fp64", "fp32"):
        disc, orc, QA, rhoB, wnode, dt = _misfit_problem(8, 6, nsteps, precision=prec)
        wdev = torch.tensor(wnode, device="cuda", dtype=disc.torch_dtype())
        rdev = torch.tensor(rhoB, device="cuda", dtype=disc.torch_dtype())

        def seed(u):
            g = torch.zeros_like(u)
            g[0] = 2.0 * wdev * (u[0] - rdev)
            return g

        res[prec] = disc.to_host(A.gradient(disc, F.SSPRK33, disc.to_device(QA), dt, nsteps, seed,
                                            checkpoint_every=4)).astype(np.float64)
    err = np.linalg.norm(res["fp32"] - res["fp64"]) / np.linalg.norm(res["fp64"])
    print(f"FP32 adjoint vs FP64 adjoint: relative L2 error {err:.2e}")
    assert err <= 1e-4

