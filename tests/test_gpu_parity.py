"""Parity of the HIP path (through the C ABI) with the CPU oracle.

FP64 plain form: same arithmetic as the oracle term by term, so the bar is 1e-13 relative
to the field's max norm, and most cases must be bit-identical (reported per case).
FP32 chosen form: error against the FP64 oracle.
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA, isapprox, oracle_for, random_euler_state, reference_answer

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SURF = {
    "lxf": (lambda: F.LxF(F.StdAverage(), 1.0), O.AVG_STD, O.DISS_LXF),
    "avg": (lambda: F.StdAverage(), O.AVG_STD, O.DISS_NONE),
    "matrix": (lambda: F.MatrixDissipation(F.ChandrasekharAverage(), 1.0), O.AVG_CHANDRASEKHAR, O.DISS_MATRIX),
    "scalar": (lambda: F.ScalarDissipation(F.ChandrasekharAverage(), 1.0), O.AVG_CHANDRASEKHAR, O.DISS_SCALAR),
    "scalar-std": (lambda: F.ScalarDissipation(F.StdAverage(), 0.5), O.AVG_STD, O.DISS_SCALAR),
    "chavg": (lambda: F.ChandrasekharAverage(), O.AVG_CHANDRASEKHAR, O.DISS_NONE),
    "lxf-ch": (lambda: F.LxF(F.ChandrasekharAverage(), 0.7), O.AVG_CHANDRASEKHAR, O.DISS_LXF),
}
TP = {"std": (F.StdAverage, O.TP_STDAVG), "ch": (F.ChandrasekharAverage, O.TP_CHANDRASEKHAR),
      "kg": (F.KennedyGruber, O.TP_KENNEDY_GRUBER)}


def build(nd, n, nel, vol, tp="std", surf="lxf", periodic=None, nodes="GLL", eq="euler",
          precision="fp64", velocity=None, finish=None):
    periodic = [1] * nd if periodic is None else periodic
    start = [0.0] * nd
    finish = finish or [1.0, 1.5, 2.0][:nd]
    b = F.LagrangeBasis(nodes, n)
    nv = nd + 2 if eq == "euler" else 1
    std = [None, F.StdSegment, F.StdQuad, F.StdHex][nd](b, F.DGSEMrec(b), nv)
    mesh = F.CartesianMesh(nd, start, finish, nel)
    mesh.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd) if periodic[k]])
    sf, savg, sdiss = SURF[surf]
    if eq == "euler":
        equation = F.EulerEquation(nd, GAMMA)
        eqc, eqp = O.EQ_EULER, [GAMMA]
    else:
        velocity = velocity or [1.0, 0.5, -0.25][:nd]
        equation = F.LinearAdvection(*velocity)
        eqc, eqp = O.EQ_ADVECTION, velocity
    if vol == "strong":
        op = F.StrongDivOperator(sf())
        volc, tpc = O.VOL_STRONG, 0
    else:
        op = F.SplitDivOperator(TP[tp][0](), sf())
        volc, tpc = O.VOL_SPLIT, TP[tp][1]
    inten = getattr(sf(), "intensity", 1.0)
    return mesh, std, equation, op, oracle_for(nd, n, nel, start, finish, periodic, eqc, eqp, volc,
                                               tpc, savg, sdiss, inten, nodes=nodes)


def rel_err(a, b):
    scale = np.max(np.abs(b), axis=0)
    scale[scale == 0] = 1
    return np.max(np.abs(a - b) / scale)


RHS_CASES = [
    # nd, n, nel, vol, tp, surf, eq, nodes
    (1, 4, [7], "strong", "std", "lxf", "euler", "GLL"),
    (1, 6, [5], "split", "ch", "lxf", "euler", "GLL"),
    (1, 4, [6], "split", "ch", "matrix", "euler", "GLL"),
    (2, 3, [4, 3], "strong", "std", "lxf", "euler", "GLL"),
    (2, 8, [3, 2], "split", "kg", "lxf", "euler", "GLL"),
    (2, 5, [3, 4], "split", "ch", "avg", "euler", "GLL"),
    (2, 4, [3, 3], "split", "std", "lxf", "euler", "GLL"),
    (2, 4, [3, 3], "split", "ch", "matrix", "euler", "GLL"),
    (3, 2, [3, 2, 2], "strong", "std", "lxf", "euler", "GLL"),
    (3, 6, [2, 3, 2], "split", "ch", "lxf", "euler", "GLL"),
    (3, 5, [2, 2, 2], "strong", "std", "lxf", "euler", "GLL"),
    (3, 7, [2, 2, 1], "split", "kg", "lxf", "euler", "GLL"),
    (3, 8, [1, 2, 2], "split", "std", "avg", "euler", "GLL"),
    (3, 4, [2, 2, 2], "split", "ch", "matrix", "euler", "GLL"),
    (1, 5, [6], "split", "ch", "scalar", "euler", "GLL"),
    (2, 4, [3, 3], "split", "ch", "scalar", "euler", "GLL"),
    (3, 4, [2, 2, 2], "split", "ch", "scalar", "euler", "GLL"),
    (3, 5, [2, 2, 2], "strong", "std", "scalar-std", "euler", "GLL"),
    (2, 5, [4, 4], "strong", "std", "lxf", "adv", "GLL"),
    (2, 5, [4, 4], "split", "std", "lxf", "adv", "GLL"),
    (2, 5, [4, 3], "strong", "std", "lxf", "adv", "GL"),
    (3, 4, [2, 2, 2], "strong", "std", "lxf", "adv", "GL"),
    (3, 3, [2, 2, 2], "strong", "std", "lxf", "euler", "GL"),
]


def _state(disc, orc, nd, eq, seed, amp=0.1):
    X, _ = orc.coords()
    if eq == "euler":
        return random_euler_state(X, nd, seed=seed, amp=amp)
    rng = np.random.default_rng(seed)
    return np.asfortranarray(np.sin(X @ rng.uniform(1, 3, nd))[:, None] + 0.01 * rng.standard_normal((X.shape[0], 1)))


@pytest.mark.parametrize("case", RHS_CASES, ids=lambda c: "-".join(str(x) for x in c))
def test_rhs_fp64_matches_oracle(case):
    nd, n, nel, vol, tp, surf, eq, nodes = case
    mesh, std, equation, op, orc = build(nd, n, nel, vol, tp, surf, nodes=nodes, eq=eq)
    disc = F.MultielementDisc(mesh, std, equation, op, precision="fp64")
    Q = _state(disc, orc, nd, eq, seed=17, amp=0.3)
    for typo in ((0, 1) if surf in ("matrix", "scalar", "scalar-std") and nd == 3 else (0,)):
        # 3-D ChandrasekharAverage surface flux with and without the Euler.jl:216 typo
        orc.set_option(0, typo)
        disc.set_option(0, typo)
        ref = orc.rhs(Q)
        Qd = disc.to_device(Q)
        dQd = disc.new_state()
        F.rhs_(dQd, Qd, F.EquationConfig(disc, equation), 0.0)
        disc.check()
        got = disc.to_host(dQd)
        same = np.mean(got == ref)
        print(f"{case} typo={typo}: bit-identical fraction {same:.4f}, max rel err {rel_err(got, ref):.2e}")
        assert rel_err(got, ref) <= 1e-13
        # only the log branch of logarithmic_mean (libm vs device log, 1 ulp) may differ
        assert same >= (0.9 if (tp == "ch" or surf in ("matrix", "scalar", "scalar-std")) else 1.0)


def test_rhs_strong_form_is_bit_identical():
    """No log is evaluated on this path and sqrt/div are correctly rounded on both sides:
    the GPU must reproduce the oracle exactly."""
    mesh, std, equation, op, orc = build(3, 5, [3, 2, 2], "strong", surf="lxf")
    disc = F.MultielementDisc(mesh, std, equation, op)
    Q = _state(disc, orc, 3, "euler", seed=3)
    Qd = disc.to_device(Q)
    dQd = disc.new_state()
    F.rhs_(dQd, Qd, F.EquationConfig(disc, equation))
    assert np.array_equal(disc.to_host(dQd), orc.rhs(Q))


@pytest.mark.parametrize("solver,scheme", [(F.SSPRK33, O.RK_SSPRK33), (F.CarpenterKennedy2N54, O.RK_CK2N54),
                                           (F.ORK256, O.RK_ORK256)])
def test_fused_rk_steps_match_oracle(solver, scheme):
    mesh, std, equation, op, orc = build(2, 5, [4, 3], "split", "ch", "lxf")
    disc = F.MultielementDisc(mesh, std, equation, op)
    Q = _state(disc, orc, 2, "euler", seed=5)
    dt, steps = 2e-3, 4
    ref = orc.rk_run(scheme, Q, dt, steps)
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, equation, solver(), steps * dt, dt=dt)
    got = disc.to_host(Qd)
    print(f"{solver.__name__}: bit-identical {np.mean(got == ref):.4f}, rel err {rel_err(got - Q, ref - Q):.2e}")
    assert rel_err(got - Q, ref - Q) <= 1e-12


def test_rk_stage_api_and_aliasing():
    mesh, std, equation, op, orc = build(2, 4, [3, 3], "strong")
    disc = F.MultielementDisc(mesh, std, equation, op)
    Q = disc.to_device(_state(disc, orc, 2, "euler", seed=2))
    out, reg = disc.new_state(), disc.new_state()
    with pytest.raises(ValueError):
        F.rk_stage(disc, F.CarpenterKennedy2N54.code, 0, 1e-3, Q, Q, reg=reg)
    with pytest.raises(ValueError):
        F.rk_stage(disc, F.CarpenterKennedy2N54.code, 0, 1e-3, Q, out)  # no register
    F.rk_stage(disc, F.CarpenterKennedy2N54.code, 0, 1e-3, Q, out, reg=reg)
    # stage 0 of a 2N scheme: reg = dt*L(Q), out = Q + B1*reg
    dQ = disc.new_state()
    F.rhs_(dQ, Q, F.EquationConfig(disc, equation))
    A, B, _ = O.tableau_2n(O.RK_CK2N54)
    r = disc.to_host(reg)
    assert np.array_equal(r, 1e-3 * disc.to_host(dQ))
    assert np.array_equal(disc.to_host(out), disc.to_host(Q) + B[0] * r)

    # SSPRK33: u_base may be u_out only as the exact alias of stage 2; partial overlaps are
    # refused (one workgroup would read what another wrote)
    size = disc.ndofs * disc.nvars
    big = torch.zeros(2 * size, dtype=Q.dtype, device=Q.device)
    base = big[:size].view(disc.nvars, disc.ndofs)
    shifted = big[size // 2: size // 2 + size].view(disc.nvars, disc.ndofs)
    base.copy_(torch.from_numpy(np.ascontiguousarray(disc.to_host(Q).T)).to(Q.device))
    u1 = disc.new_state()
    F.rk_stage(disc, F.SSPRK33.code, 0, 1e-3, Q, u1)
    for stage, dst in ((1, shifted), (1, base), (2, shifted)):
        with pytest.raises(ValueError):
            F.rk_stage(disc, F.SSPRK33.code, stage, 1e-3, u1, dst, u_base=base)
    ref = disc.new_state()
    F.rk_stage(disc, F.SSPRK33.code, 2, 1e-3, u1, ref, u_base=base)
    F.rk_stage(disc, F.SSPRK33.code, 2, 1e-3, u1, base, u_base=base)  # in place
    assert np.array_equal(disc.to_host(base), disc.to_host(ref))

    # the adjoint step refuses a cotangent buffer that is one of the stage buffers
    from flouhip import adjoint as AD
    stages = [Q, disc.new_state(), disc.new_state()]
    unext = disc.new_state()
    AD.rk_step_record(disc, F.SSPRK33, 1e-3, stages, unext)
    with pytest.raises(ValueError):
        AD.rk_step_vjp(disc, F.SSPRK33, 1e-3, stages, stages[1])


def test_max_dt_and_integrate():
    mesh, std, equation, op, orc = build(3, 4, [3, 2, 2], "split", "ch")
    disc = F.MultielementDisc(mesh, std, equation, op)
    Q = _state(disc, orc, 3, "euler", seed=9)
    Qd = disc.to_device(Q)
    assert F.get_max_dt(Qd, disc, equation, 0.3) == orc.max_dt(Q, 0.3)
    for v in range(5):
        assert F.integrate(Qd, disc, v) == pytest.approx(orc.integrate(Q, v), rel=1e-14, abs=1e-15)


def test_timeintegrate_with_cfl_matches_oracle_dt():
    mesh, std, equation, op, orc = build(2, 4, [4, 4], "split", "ch")
    disc = F.MultielementDisc(mesh, std, equation, op)
    Q = _state(disc, orc, 2, "euler", seed=21)
    dt = orc.max_dt(Q, 0.2)
    ref = orc.rk_run(O.RK_SSPRK33, Q, dt, 1)
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, equation, F.SSPRK33(), dt, dt=1.0, cfl=0.2)
    assert rel_err(disc.to_host(Qd) - Q, ref - Q) <= 1e-12


def test_sod_tube_on_gpu_matches_reference_golden():
    """SodTube1D (test/tests.jl:90-134) through the product API -> runtests.jl:37-38."""
    g = GAMMA
    eq = F.EulerEquation(1, g)
    b = F.LagrangeBasis("GLL", 4)
    std = F.StdSegment(b, F.DGSEMrec(b), 3)
    mesh = F.CartesianMesh(1, [0.0], [1.0], [20])

    def Qext(_, x, __, ___, eq_):
        P = (1.0, 0.0, 100.0) if x[0] < 0.5 else (0.125, 0.0, 10.0)
        return F.vars_prim2cons(np.array(P), eq_)

    bcs = {"1": F.GenericBC(Qext), "2": F.GenericBC(Qext)}
    op = F.SplitDivOperator(F.MatrixDissipation(F.ChandrasekharAverage(), 1.0))
    disc = F.MultielementDisc(mesh, std, eq, op, bcs)
    X = mesh.node_coords(std)
    Q = np.array([Qext(None, x, None, 0.0, eq) for x in X])
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, eq, F.ORK256(williamson_condition=False), 0.018, dt=1e-4)
    u = disc.to_host(Qd)
    ref = reference_answer("sod_tube_1d")
    assert isapprox(u.min(), ref["min"], ref["rtol"]) and isapprox(u.max(), ref["max"], ref["rtol"])
    print("Sod GPU min/max rel dev:", abs(u.min() - ref["min"]) / abs(ref["min"]),
          abs(u.max() - ref["max"]) / abs(ref["max"]))


def test_advection_2d_full_period_gauss_nodes():
    """Advection2D (test/tests.jl:53-88) on the GPU: Gauss nodes use the projection path."""
    eq = F.LinearAdvection(3.0, 4.0)
    b = F.LagrangeBasis("GL", 5)
    std = F.StdQuad(b, F.DGSEMrec(b), 1)
    mesh = F.CartesianMesh(2, (0, 0), (1.5, 2), (20, 10)).apply_periodicBCs(("1", "2"), ("3", "4"))
    disc = F.MultielementDisc(mesh, std, eq, F.StrongDivOperator(F.LxF(F.StdAverage(), 1.0)))
    X = mesh.node_coords(std)
    Q0 = np.exp(-(X[:, 0] - 0.75) ** 2 / (2 * 0.2 ** 2) - (X[:, 1] - 1.0) ** 2 / (2 * 0.2 ** 2))[:, None]
    Qd = disc.to_device(Q0)
    F.timeintegrate(Qd, disc, eq, F.ORK256(), 0.5, dt=1e-3)
    u = disc.to_host(Qd)
    rtol = reference_answer("advection_2d")["rtol"]
    assert np.linalg.norm(u - Q0) <= rtol * max(np.linalg.norm(u), np.linalg.norm(Q0))


FP32_CASES = [
    # nd, n, tp, surf, periodic, amp
    (3, 6, "ch", "lxf", [1, 1, 1], 0.1),
    (3, 6, "ch", "lxf", [1, 1, 1], 0.5),  # large jumps: log branch of the log-mean
    (3, 5, "kg", "lxf", [1, 1, 1], 0.1),
    (3, 7, "std", "avg", [1, 1, 1], 0.1),
    (3, 4, "ch", "lxf", [1, 0, 1], 0.1),  # slip walls on y
    (3, 8, "ch", "lxf", [1, 1, 1], 0.1),
    (2, 8, "kg", "lxf", [1, 1], 0.1),     # the C2 operator
    (2, 8, "kg", "lxf", [1, 1], 0.5),
    (2, 7, "ch", "lxf", [1, 1], 0.1),
    (2, 7, "ch", "lxf", [1, 1], 0.5),
    (2, 6, "std", "avg", [1, 1], 0.1),
    (2, 5, "ch", "lxf", [1, 0], 0.1),     # slip walls on y
    (2, 4, "kg", "lxf", [1, 1], 0.1),
]


@pytest.mark.parametrize("case", FP32_CASES, ids=lambda c: "-".join(str(x) for x in c))
def test_fp32_rhs_error_against_fp64_oracle(case):
    """FP32 chosen form (fast kernel) and FP32 plain kernel against the FP64 oracle."""
    nd, n, tp, surf, periodic, amp = case
    nel = [2, 2, 2] if nd == 3 else [3, 4]
    mesh, std, equation, op, orc = build(nd, n, nel, "split", tp, surf, periodic=periodic,
                                         finish=[2 * np.pi] * nd)
    bcs = {}
    if not all(periodic):
        bcs = {"3": F.EulerSlipBC(), "4": F.EulerSlipBC()}
        orc.set_bc(0, O.BC_SLIP)
        orc.set_bc(1, O.BC_SLIP)
    disc = F.MultielementDisc(mesh, std, equation, op, bcs, precision="fp32")
    assert disc.uses_fast_kernel()
    Q = _state(disc, orc, nd, "euler", seed=4, amp=amp)
    ref = orc.rhs(Q)
    Qd = disc.to_device(Q)
    for plain in (0, 1):
        disc.set_option(1, plain)
        dQd = disc.new_state()
        F.rhs_(dQd, Qd, F.EquationConfig(disc, equation))
        disc.check()
        got = disc.to_host(dQd).astype(np.float64)
        err = rel_err(got, ref)
        print(f"FP32 {'plain' if plain else 'fast'} rhs {case}: max rel err vs FP64 oracle {err:.2e}")
        assert err < 2e-5


def test_fp32_fused_ck2n54_against_fp64_oracle():
    mesh, std, equation, op, orc = build(3, 6, [3, 2, 2], "split", "ch", finish=[2 * np.pi] * 3)
    disc = F.MultielementDisc(mesh, std, equation, op, precision="fp32")
    Q = _state(disc, orc, 3, "euler", seed=6)
    dt, steps = 1e-2, 5
    ref = orc.rk_run(O.RK_CK2N54, Q, dt, steps)
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, equation, F.CarpenterKennedy2N54(), steps * dt, dt=dt)
    got = disc.to_host(Qd).astype(np.float64)
    err = np.linalg.norm(got[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0])
    print(f"FP32 fused CK2N54 x{steps}: density rel L2 err vs FP64 oracle {err:.2e}")
    assert err < 1e-6
    # conservation of mass in FP32 (faces exactly antisymmetric): FP64 integrals
    m0 = orc.integrate(Q, 0)
    assert abs(F.integrate(Qd, disc, 0) - m0) < 1e-5 * abs(m0)


def test_domain_error_on_negative_pressure():
    mesh, std, equation, op, orc = build(2, 4, [3, 2], "strong")
    disc = F.MultielementDisc(mesh, std, equation, op)
    Q = _state(disc, orc, 2, "euler", seed=1)
    Q[0, 3] = -1.0
    with pytest.raises(ArithmeticError):
        orc.rhs(Q)
    Qd = disc.to_device(Q)
    dQd = disc.new_state()
    F.rhs_(dQd, Qd, F.EquationConfig(disc, equation))
    with pytest.raises(F.DomainError):
        disc.check()
    disc.check()  # the flag is cleared


def test_rhs_is_deterministic():
    mesh, std, equation, op, orc = build(3, 6, [2, 2, 2], "split", "ch")
    disc = F.MultielementDisc(mesh, std, equation, op)
    Qd = disc.to_device(_state(disc, orc, 3, "euler", seed=8))
    a, b = disc.new_state(), disc.new_state()
    cfg = F.EquationConfig(disc, equation)
    F.rhs_(a, Qd, cfg)
    F.rhs_(b, Qd, cfg)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("n,precision", [(4, "fp64"), (6, "fp64"), (6, "fp32")])
def test_non_periodic_boundaries_match_oracle(n, precision):
    """Inflow / outflow / slip boundaries on a 2-D box (Equations/Euler.jl:69-94).  n = 6:
    the split kernels that load each thread's first face node before staging."""
    mesh, std, equation, op, orc = build(2, n, [3, 3], "split", "ch", periodic=[0, 0])
    qin = np.array([1.1, 0.3, 0.0, 2.7])
    bcs = {"1": F.EulerInflowBC(qin), "2": F.EulerOutflowBC(), "3": F.EulerSlipBC(), "4": F.EulerSlipBC()}
    disc = F.MultielementDisc(mesh, std, equation, op, bcs, precision=precision)
    orc.set_bc(0, O.BC_INFLOW, qin)
    orc.set_bc(1, O.BC_OUTFLOW)
    orc.set_bc(2, O.BC_SLIP)
    orc.set_bc(3, O.BC_SLIP)
    Q = _state(disc, orc, 2, "euler", seed=12)
    Qd = disc.to_device(Q)
    dQd = disc.new_state()
    F.rhs_(dQd, Qd, F.EquationConfig(disc, equation))
    got, ref = disc.to_host(dQd).astype(np.float64), orc.rhs(Q)
    assert rel_err(got, ref) <= (1e-13 if precision == "fp64" else 2e-5)


@pytest.mark.parametrize("nd,axis", [(3, 0), (3, 2), (2, 0), (2, 1)], ids=["3d-x", "3d-z", "2d-x", "2d-y"])
@pytest.mark.parametrize("kind", ["inflow-outflow", "dirichlet", "slip-outflow"])
def test_fp32_fast_kernel_boundaries_match_oracle(kind, nd, axis):
    """Every boundary kind on the FP32 fast path (3-D N=5 and 2-D N=7, one axis
    non-periodic), against the FP64 oracle and against the FP32 plain kernel on the same
    handle.  3-D z: the lines of the block's last wave (z lines of its second element) end
    on the boundaries."""
    periodic = [1] * nd
    periodic[axis] = 0
    nel = [2] * nd if nd == 3 else [4, 5]
    nel[axis] = 3
    n = 6 if nd == 3 else 8
    mesh, std, equation, op, orc = build(nd, n, nel, "split", "ch", periodic=periodic,
                                         finish=[2 * np.pi] * nd)
    lo, hi = str(2 * axis + 1), str(2 * axis + 2)
    qin = np.array([1.1, 0.3, -0.05, 0.02, 2.7]) if nd == 3 else np.array([1.1, 0.3, -0.05, 2.7])
    X, FX = orc.coords()
    if kind == "inflow-outflow":
        bcs = {lo: F.EulerInflowBC(qin), hi: F.EulerOutflowBC()}
        orc.set_bc(0, O.BC_INFLOW, qin)
        orc.set_bc(1, O.BC_OUTFLOW)
    elif kind == "slip-outflow":
        bcs = {lo: F.EulerSlipBC(), hi: F.EulerOutflowBC()}
        orc.set_bc(0, O.BC_SLIP)
        orc.set_bc(1, O.BC_OUTFLOW)
    else:
        def qext(x):
            r = 1.0 + 0.1 * np.sin(x[1]) * np.cos(x[-1])
            mom = [0.2 * r, 0.05 * r, -0.03 * r][:nd]
            return np.array([r] + mom + [2.5 + 0.1 * np.cos(x[0] + x[-1])])

        bcs = {lo: F.GenericBC(lambda _q, x, _f, _t, _e: qext(x)),
               hi: F.GenericBC(lambda _q, x, _f, _t, _e: qext(x))}
        bc = np.zeros((orc.nf, orc.npf, nd + 2))
        for f in range(orc.nf):
            for k in range(orc.npf):
                bc[f, k] = qext(FX[f, k])
        orc.set_bc(0, O.BC_DIRICHLET, bc)
        orc.set_bc(1, O.BC_DIRICHLET, bc)
    disc = F.MultielementDisc(mesh, std, equation, op, bcs, precision="fp32")
    assert disc.uses_fast_kernel()
    Q = _state(disc, orc, nd, "euler", seed=21, amp=0.1)
    ref = orc.rhs(Q)
    Qd = disc.to_device(Q)
    out = {}
    for plain in (0, 1):
        disc.set_option(1, plain)
        dQd = disc.new_state()
        F.rhs_(dQd, Qd, F.EquationConfig(disc, equation))
        disc.check()
        out[plain] = disc.to_host(dQd).astype(np.float64)
    e_fast, e_plain = rel_err(out[0], ref), rel_err(out[1], ref)
    print(f"FP32 fast rhs with {kind} boundaries: {e_fast:.2e} (plain FP32 {e_plain:.2e})")
    assert e_fast < 2e-5 and e_plain < 2e-5


@pytest.mark.parametrize("n,nel", [(6, [3, 3, 3]), (5, [2, 2, 2]), (4, [3, 1, 2]), (6, [1, 1, 1]),
                                   (8, [3, 5]), (7, [2, 5]), (5, [5, 7]), (4, [3, 3]), (6, [1, 1])],
                         ids=["n6-27el", "n5-8el", "n4-6el", "n6-1el",
                              "2d-n8-15el", "2d-n7-10el", "2d-n5-35el", "2d-n4-9el", "2d-n6-1el"])
def test_fp32_fast_kernel_partial_blocks(n, nel):
    """Element counts that are not multiples of the elements per block (3-D: 2 at N=6, 3 at
    N=5, 4 at N=4; 2-D: 8 at N=8, 9 at N=7, 32 at N=5, 16 at N=4 and N=6), and a single
    element: the last block's missing elements must neither be read nor written, and the
    stage forms must agree with the plain FP32 kernel."""
    nd = len(nel)
    mesh, std, equation, op, orc = build(nd, n, nel, "split", "ch", finish=[2 * np.pi] * nd)
    disc = F.MultielementDisc(mesh, std, equation, op, precision="fp32")
    assert disc.uses_fast_kernel()
    Q = _state(disc, orc, nd, "euler", seed=31, amp=0.1)
    ref = orc.rhs(Q)
    Qd = disc.to_device(Q)
    dQd = disc.new_state()
    F.rhs_(dQd, Qd, F.EquationConfig(disc, equation))
    disc.check()
    err = rel_err(disc.to_host(dQd).astype(np.float64), ref)
    print(f"FP32 fast rhs n={n} nel={nel}: {err:.2e}")
    assert err < 2e-5
    # every stage form: fast against plain FP32 on the same handle
    dt = 0.1 * orc.max_dt(Q, 1.0)
    for solver in (F.SSPRK33, F.CarpenterKennedy2N54):
        outs = []
        for plain in (0, 1):
            disc.set_option(1, plain)
            Qd = disc.to_device(Q)
            F.timeintegrate(Qd, disc, equation, solver, 2 * dt, dt=dt)
            outs.append(disc.to_host(Qd).astype(np.float64))
        disc.set_option(1, 0)
        assert rel_err(outs[0], outs[1]) < 1e-5
