"""The FP32 fast adjoint kernel (flou.jl_amd/csrc/vjp_fast.hip) on curved and on
axis-aligned affine periodic meshes (the split form with the Chandrasekhar, Kennedy-Gruber
or StdAverage two-point flux and LxF; OpDivergence.jl:200-299, Interfaces.jl:111-136,
Equations/Euler.jl:385-536):

* the rhs VJP against the plain FP32 VJP (k_vjp, switched on with option 1) on the same
  inputs, over 2-D and 3-D, N = 4..8, element counts that leave the last block part-filled,
  and states whose density and β spreads put lines on the series and on the log branch of
  the log-mean;
* the dot-product identity <W, J V> = <J^T W, V> against the oracle's FP64 dual-number JVP
  (the oracle's rhs is pinned by tests/test_oracle.py and tests/test_curved.py);
* the reverse RK steps (SSPRK33 and CK2N54) against the FP64 adjoint at the same stage
  states, next to the plain FP32 path;
* boundaries (Interfaces.jl:25-49, Equations/Euler.jl:69-94): inflow, outflow, Dirichlet
  and slip walls on affine meshes, slip, inflow and outflow on a curved mesh with wavy
  walls -- against the plain FP32 VJP and the oracle's JVP.
The hand-derived reverse formulas were also checked against central differences of the
forward fluxes in float64 (series and log branches, the LxF face flux) before the kernel
ran.
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA, oracle_for, periodic_euler_state, random_euler_state
from flouhip import mesh as M
from flouhip import adjoint as A
from test_gpu_curved import TP, build_curved

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (nd, n, elements per direction): 3-D blocks take 4/3/2/1/1 elements at N = 4..8, 2-D
# blocks 16/32/16/9/8, so most of these leave the last block part-filled
CASES = [(3, 4, 3), (3, 5, 2), (3, 6, 3), (3, 7, 2), (3, 8, 2),
         (2, 4, 5), (2, 5, 6), (2, 6, 5), (2, 7, 4), (2, 8, 3)]


def _fp32_disc(nd, n, nel, tp="ch"):
    disc, orc, eq, X, J, std = build_curved(nd, n, nel, "split", tp, precision="fp32")
    assert disc.uses_fast_kernel() and disc.uses_fast_vjp()
    return disc, orc, X


def _affine_fp32_disc(nd, n, nel, lengths, tp="ch"):
    """Periodic Cartesian mesh on a box with unequal sides (Ja_dd differs per direction)."""
    b = F.LagrangeBasis("GLL", n)
    std = [None, None, F.StdQuad, F.StdHex][nd](b, F.DGSEMrec(b), nd + 2)
    mesh = F.CartesianMesh(nd, (0.0,) * nd, tuple(lengths), (nel,) * nd)
    mesh.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
    eq = F.EulerEquation(nd, GAMMA)
    op = F.SplitDivOperator(TP[tp][0](), F.LxF(F.StdAverage(), 1.0))
    disc = F.MultielementDisc(mesh, std, eq, op, precision="fp32")
    assert disc.uses_fast_kernel() and disc.uses_fast_vjp()
    orc = oracle_for(nd, n, [nel] * nd, [0.0] * nd, list(lengths), [1] * nd, O.EQ_EULER, [GAMMA], O.VOL_SPLIT,
                     TP[tp][1], O.AVG_STD, O.DISS_LXF, 1.0)
    return disc, orc, mesh.node_coords(std)


def _vjp(disc, Q, W, plain):
    disc.set_option(1, 1 if plain else 0)
    try:
        Qbar = disc.to_device(np.zeros_like(Q))
        A.rhs_vjp(Qbar, disc.to_device(W), disc.to_device(Q), disc)
        disc.check()
        return disc.to_host(Qbar).astype(np.float64)
    finally:
        disc.set_option(1, 0)


def _rel(a, b):
    return np.max(np.abs(a - b) / np.max(np.abs(b), axis=0))


def _rel2(a, b):
    """Relative L2 error: the measure for FP32 adjoints, where the LxF wave-speed maximum
    has a kink (λ = max of the two sides' |u·n| + c): at a face node where the two nearly
    tie, two FP32 evaluations may take different branches, and one node's cotangent then
    differs by the jump of the derivative."""
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("amp", [0.1, 0.6], ids=["series", "log"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "{}d-n{}-{}".format(*c))
def test_fast_vjp_matches_plain_fp32_vjp(case, amp):
    nd, n, nel = case
    disc, orc, X = _fp32_disc(nd, n, nel)
    Q = periodic_euler_state(X, nd, 1.0, seed=5, amp=amp)
    W = np.random.default_rng(6).standard_normal(Q.shape)
    got, ref = _vjp(disc, Q, W, plain=False), _vjp(disc, Q, W, plain=True)
    err, emax = _rel2(got, ref), _rel(got, ref)
    print(f"{case} amp {amp}: fast vs plain FP32 VJP, relative L2 {err:.2e}, max {emax:.2e}")
    assert err <= 2e-6


AFFINE = [(3, 6, 3, (1.0, 1.5, 0.75)), (3, 5, 3, (2.0, 1.0, 1.0)), (3, 7, 2, (1.0, 1.0, 1.0)),
          (2, 8, 4, (1.0, 2.0)), (2, 5, 6, (1.5, 1.0))]


@pytest.mark.parametrize("case", AFFINE, ids=lambda c: "{}d-n{}-{}".format(*c[:3]))
def test_fast_vjp_affine_matches_plain_fp32_vjp(case):
    nd, n, nel, lengths = case
    disc, orc, X = _affine_fp32_disc(nd, n, nel, lengths)
    Q = periodic_euler_state(X, nd, 1.0, seed=15, amp=0.2)
    W = np.random.default_rng(16).standard_normal(Q.shape)
    got, ref = _vjp(disc, Q, W, plain=False), _vjp(disc, Q, W, plain=True)
    err, emax = _rel2(got, ref), _rel(got, ref)
    print(f"affine {case}: fast vs plain FP32 VJP, relative L2 {err:.2e}, max {emax:.2e}")
    assert err <= 2e-6


@pytest.mark.parametrize("case", [AFFINE[0], AFFINE[3]], ids=["3d-n6", "2d-n8"])
def test_fast_vjp_affine_dot_product_with_oracle_jvp(case):
    nd, n, nel, lengths = case
    disc, orc, X = _affine_fp32_disc(nd, n, nel, lengths)
    Q = periodic_euler_state(X, nd, 1.0, seed=17, amp=0.2).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(18)
    V, W = rng.standard_normal(Q.shape), rng.standard_normal(Q.shape)
    _, JV = orc.rhs_jvp(np.asfortranarray(Q), V)
    JtW = _vjp(disc, Q, W, plain=False)
    lhs, rhs = float(np.sum(W * JV)), float(np.sum(JtW * V))
    err = abs(lhs - rhs) / float(np.sum(np.abs(W * JV)))
    print(f"affine {case}: <W,JV> = {lhs:.8e}  <J^T W,V> (fast FP32) = {rhs:.8e}  rel {err:.2e}")
    assert err <= 2e-5


@pytest.mark.parametrize("case", [(3, 7, 2), (3, 5, 2), (2, 6, 4)], ids=["3d-n7", "3d-n5", "2d-n6"])
def test_fast_vjp_dot_product_with_oracle_jvp(case):
    nd, n, nel = case
    disc, orc, X = _fp32_disc(nd, n, nel)
    Q = periodic_euler_state(X, nd, 1.0, seed=9, amp=0.2).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(4)
    V, W = rng.standard_normal(Q.shape), rng.standard_normal(Q.shape)
    _, JV = orc.rhs_jvp(np.asfortranarray(Q), V)
    JtW = _vjp(disc, Q, W, plain=False)
    lhs, rhs = float(np.sum(W * JV)), float(np.sum(JtW * V))
    err = abs(lhs - rhs) / float(np.sum(np.abs(W * JV)))
    print(f"{case}: <W,JV> = {lhs:.8e}  <J^T W,V> (fast FP32) = {rhs:.8e}  rel {err:.2e}")
    assert err <= 2e-5


@pytest.mark.parametrize("solver", [F.SSPRK33, F.CarpenterKennedy2N54], ids=["ssprk33", "ck2n54"])
def test_fast_vjp_reverse_rk_step_against_fp64(solver):
    """One reverse RK step at stages recorded in FP32: the fast and the plain FP32 adjoint
    against the FP64 adjoint at the same (FP32-rounded) stage states, relative L2 (see
    _rel2).  FP32 round-off alone gives about 1e-7 (the plain kernel's figure); the bound
    leaves room for a face node whose LxF wave-speed maximum takes the other branch (on the
    CK2N54 case one node of stage 2 has wave speeds 1.18949268 and 1.18949280, within an
    FP32 ulp: about 1e-6 relative L2)."""
    disc, orc, X = _fp32_disc(3, 6, 3)
    d64 = build_curved(3, 6, 3, "split", "ch", precision="fp64")[0]
    Q = periodic_euler_state(X, 3, 1.0, seed=12, amp=0.15)
    dt = 0.2 * orc.max_dt(Q, 1.0)
    S = 3 if solver is F.SSPRK33 else 5
    stages = [disc.to_device(Q)] + [disc.new_state() for _ in range(S - 1)]
    A.rk_step_record(disc, solver, dt, stages, disc.new_state())
    st64 = [d64.to_device(disc.to_host(x).astype(np.float64)) for x in stages]
    ybar = np.random.default_rng(2).standard_normal(Q.shape)
    ref = d64.to_host(A.rk_step_vjp(d64, solver, dt, st64, d64.to_device(ybar)))
    d64.check()
    err = {}
    for plain in (0, 1):
        disc.set_option(1, plain)
        ub = disc.to_device(ybar)
        A.rk_step_vjp(disc, solver, dt, stages, ub)
        disc.check()
        got = disc.to_host(ub).astype(np.float64)
        err[plain] = (_rel2(got, ref), _rel(got, ref))
    disc.set_option(1, 0)
    print(f"{'SSPRK33' if solver is F.SSPRK33 else 'CK2N54'} reverse step vs FP64, relative L2 (max): "
          f"fast {err[0][0]:.2e} ({err[0][1]:.2e}), plain FP32 {err[1][0]:.2e} ({err[1][1]:.2e})")
    assert err[0][0] <= 2e-6 and err[1][0] <= 2e-6


def _wall_case(kind, nd, axis, tp="ch"):
    """One axis non-periodic with the given boundary kind, as in the fast-rhs boundary test
    (tests/test_gpu_parity.py), FP32, with the FP64 oracle on the same boundaries."""
    from test_gpu_parity import build
    periodic = [1] * nd
    periodic[axis] = 0
    nel = [2] * nd if nd == 3 else [4, 5]
    nel[axis] = 3
    n = 6 if nd == 3 else 8
    mesh, std, equation, op, orc = build(nd, n, nel, "split", tp, periodic=periodic, finish=[2 * np.pi] * nd)
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
    assert disc.uses_fast_kernel() and disc.uses_fast_vjp()
    return disc, orc, random_euler_state(X, nd, seed=21, amp=0.1)


@pytest.mark.parametrize("nd,axis", [(3, 0), (3, 2), (2, 1)], ids=["3d-x", "3d-z", "2d-y"])
@pytest.mark.parametrize("kind", ["inflow-outflow", "dirichlet", "slip-outflow"])
def test_fast_vjp_boundaries(kind, nd, axis):
    disc, orc, Q = _wall_case(kind, nd, axis)
    Q = Q.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(22)
    V, W = rng.standard_normal(Q.shape), rng.standard_normal(Q.shape)
    got, ref = _vjp(disc, Q, W, plain=False), _vjp(disc, Q, W, plain=True)
    err = _rel2(got, ref)
    _, JV = orc.rhs_jvp(np.asfortranarray(Q), V)
    lhs, rhs = float(np.sum(W * JV)), float(np.sum(got * V))
    dot = abs(lhs - rhs) / float(np.sum(np.abs(W * JV)))
    print(f"{kind} {nd}-D axis {axis}: fast vs plain FP32 VJP relative L2 {err:.2e} (max {_rel(got, ref):.2e}); "
          f"dot product vs the oracle's JVP {dot:.2e}")
    assert err <= 2e-6 and dot <= 2e-5


@pytest.mark.parametrize("tp", ["ch", "kg", "std"])
@pytest.mark.parametrize("kind", ["slip", "inflow-outflow"])
def test_fast_vjp_curved_wavy_walls(kind, tp):
    """Curved 2-D mesh with wavy walls (rotated face frames at the slip reflection), each
    two-point flux."""
    n, ne = 6, 4
    b = F.LagrangeBasis("GLL", n)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    mesh = F.CartesianMesh(2, (0.0, 0.0), (1.0, 1.0), (ne, ne))
    mesh.apply_periodicBCs(("1", "2"))

    def wavy(X):
        Y = X.copy()
        Y[:, 1] = X[:, 1] + 0.08 * np.sin(2 * np.pi * X[:, 0])
        return Y

    mesh.apply_mapping(wavy)
    eq = F.EulerEquation(2, GAMMA)
    op = F.SplitDivOperator(TP[tp][0](), F.LxF(F.StdAverage(), 1.0))
    qin = np.array([1.1, 0.2, 0.1, 2.6])
    bcs = ({"3": F.EulerSlipBC(), "4": F.EulerSlipBC()} if kind == "slip"
           else {"3": F.EulerInflowBC(qin), "4": F.EulerOutflowBC()})
    disc = F.MultielementDisc(mesh, std, eq, op, bcs, precision="fp32")
    assert disc.uses_fast_kernel() and disc.uses_fast_vjp()
    J, met, ev, fr, fj = M.curved_geometry(mesh, std)
    orc = O.OracleDisc(2, n, [ne, ne], [0.0, 0.0], [1.0, 1.0], [1, 0], O.build_ops(n, "GLL"), O.EQ_EULER,
                       [GAMMA], O.VOL_SPLIT, TP[tp][1], O.AVG_STD, O.DISS_LXF, 1.0)
    orc.set_curved(met, J, fr.reshape(-1), fj, ev)
    if kind == "slip":
        orc.set_bc(0, O.BC_SLIP)
        orc.set_bc(1, O.BC_SLIP)
    else:
        orc.set_bc(0, O.BC_INFLOW, qin)
        orc.set_bc(1, O.BC_OUTFLOW)
    Q = random_euler_state(mesh.node_coords(std), 2, seed=8, amp=0.2).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(23)
    V, W = rng.standard_normal(Q.shape), rng.standard_normal(Q.shape)
    got, ref = _vjp(disc, Q, W, plain=False), _vjp(disc, Q, W, plain=True)
    err = _rel2(got, ref)
    _, JV = orc.rhs_jvp(np.asfortranarray(Q), V)
    lhs, rhs = float(np.sum(W * JV)), float(np.sum(got * V))
    dot = abs(lhs - rhs) / float(np.sum(np.abs(W * JV)))
    print(f"curved wavy {kind} walls, {tp}: fast vs plain FP32 VJP relative L2 {err:.2e}; dot product vs "
          f"the oracle's JVP {dot:.2e}")
    assert err <= 2e-6 and dot <= 2e-5


# The Kennedy-Gruber and StdAverage two-point fluxes (their reverse products checked
# against central differences of the forward fluxes in float64 before the kernel ran):
# curved and affine (where the StdAverage lines take the linear form), 2-D and 3-D, N = 4..8
OTHER = [(3, 4, 3, "kg"), (3, 5, 2, "std"), (3, 6, 2, "kg"), (3, 7, 2, "std"), (3, 8, 2, "kg"),
         (2, 4, 5, "std"), (2, 5, 6, "kg"), (2, 6, 5, "std"), (2, 7, 4, "kg"), (2, 8, 3, "std")]


@pytest.mark.parametrize("curved", [True, False], ids=["curved", "affine"])
@pytest.mark.parametrize("case", OTHER, ids=lambda c: "{}d-n{}-{}-{}".format(*c))
def test_fast_vjp_kg_std_matches_plain_fp32_vjp(case, curved):
    nd, n, nel, tp = case
    if curved:
        disc, orc, X = _fp32_disc(nd, n, nel, tp)
    else:
        disc, orc, X = _affine_fp32_disc(nd, n, nel, (1.0, 1.5, 0.75)[:nd], tp)
    Q = periodic_euler_state(X, nd, 1.0, seed=25, amp=0.4)
    W = np.random.default_rng(26).standard_normal(Q.shape)
    got, ref = _vjp(disc, Q, W, plain=False), _vjp(disc, Q, W, plain=True)
    err, emax = _rel2(got, ref), _rel(got, ref)
    print(f"{case} {'curved' if curved else 'affine'}: fast vs plain FP32 VJP, relative L2 {err:.2e}, "
          f"max {emax:.2e}")
    assert err <= 2e-6


@pytest.mark.parametrize("curved", [True, False], ids=["curved", "affine"])
@pytest.mark.parametrize("case", [(3, 6, 2, "kg"), (3, 5, 2, "std"), (2, 7, 4, "kg"), (2, 8, 3, "std")],
                         ids=lambda c: "{}d-n{}-{}-{}".format(*c))
def test_fast_vjp_kg_std_dot_product_with_oracle_jvp(case, curved):
    nd, n, nel, tp = case
    if curved:
        disc, orc, X = _fp32_disc(nd, n, nel, tp)
    else:
        disc, orc, X = _affine_fp32_disc(nd, n, nel, (1.0, 1.5, 0.75)[:nd], tp)
    Q = periodic_euler_state(X, nd, 1.0, seed=27, amp=0.2).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(28)
    V, W = rng.standard_normal(Q.shape), rng.standard_normal(Q.shape)
    _, JV = orc.rhs_jvp(np.asfortranarray(Q), V)
    JtW = _vjp(disc, Q, W, plain=False)
    lhs, rhs = float(np.sum(W * JV)), float(np.sum(JtW * V))
    err = abs(lhs - rhs) / float(np.sum(np.abs(W * JV)))
    print(f"{case} {'curved' if curved else 'affine'}: <W,JV> = {lhs:.8e}  <J^T W,V> (fast FP32) = {rhs:.8e}  "
          f"rel {err:.2e}")
    assert err <= 2e-5


@pytest.mark.parametrize("tp", ["kg", "std"])
@pytest.mark.parametrize("kind,nd,axis", [("inflow-outflow", 3, 0), ("dirichlet", 2, 1), ("slip-outflow", 3, 2)],
                         ids=["inflow-outflow-3d-x", "dirichlet-2d-y", "slip-outflow-3d-z"])
def test_fast_vjp_kg_std_boundaries(kind, nd, axis, tp):
    disc, orc, Q = _wall_case(kind, nd, axis, tp)
    Q = Q.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(29)
    V, W = rng.standard_normal(Q.shape), rng.standard_normal(Q.shape)
    got, ref = _vjp(disc, Q, W, plain=False), _vjp(disc, Q, W, plain=True)
    err = _rel2(got, ref)
    _, JV = orc.rhs_jvp(np.asfortranarray(Q), V)
    lhs, rhs = float(np.sum(W * JV)), float(np.sum(got * V))
    dot = abs(lhs - rhs) / float(np.sum(np.abs(W * JV)))
    print(f"{tp} {kind} {nd}-D axis {axis}: fast vs plain FP32 VJP relative L2 {err:.2e} "
          f"(max {_rel(got, ref):.2e}); dot product vs the oracle's JVP {dot:.2e}")
    assert err <= 2e-6 and dot <= 2e-5


@pytest.mark.parametrize("tp", ["kg", "std"])
def test_fast_vjp_kg_std_reverse_rk_step_against_fp64(tp):
    """One reverse SSPRK33 step at stages recorded in FP32, fast and plain FP32 adjoints
    against the FP64 adjoint at the same stage states (relative L2, see _rel2)."""
    disc, orc, X = _fp32_disc(2, 7, 4, tp)
    d64 = build_curved(2, 7, 4, "split", tp, precision="fp64")[0]
    Q = periodic_euler_state(X, 2, 1.0, seed=30, amp=0.15)
    dt = 0.2 * orc.max_dt(Q, 1.0)
    stages = [disc.to_device(Q)] + [disc.new_state() for _ in range(2)]
    A.rk_step_record(disc, F.SSPRK33, dt, stages, disc.new_state())
    st64 = [d64.to_device(disc.to_host(x).astype(np.float64)) for x in stages]
    ybar = np.random.default_rng(31).standard_normal(Q.shape)
    ref = d64.to_host(A.rk_step_vjp(d64, F.SSPRK33, dt, st64, d64.to_device(ybar)))
    d64.check()
    err = {}
    for plain in (0, 1):
        disc.set_option(1, plain)
        ub = disc.to_device(ybar)
        A.rk_step_vjp(disc, F.SSPRK33, dt, stages, ub)
        disc.check()
        err[plain] = _rel2(disc.to_host(ub).astype(np.float64), ref)
    disc.set_option(1, 0)
    print(f"{tp} SSPRK33 reverse step vs FP64, relative L2: fast {err[0]:.2e}, plain FP32 {err[1]:.2e}")
    assert err[0] <= 2e-6 and err[1] <= 2e-6
