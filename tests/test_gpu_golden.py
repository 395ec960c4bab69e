"""The HIP path against the committed golden fixtures (tests/golden/*.npz), through the
product API only: no oracle runs here, the expected values are the stored arrays.

FP64 (the plain form, the reference's operation order): within 1e-12 of the stored rhs and
states relative to each variable's largest value (device log/sqrt/pow may differ from the
host's by an ulp; most values are bit-identical, and the fraction is printed).  FP32: the
rhs within 2e-5 and a few RK steps within 5e-6 of the FP64 values, relative to each
variable's largest value (the FP32-rounded input alone moves the small momentum columns by
~1e-7 of their scale per step).
"""
import numpy as np
import pytest

import flouhip as F
from cases import GAMMA, golden, warp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STD = {1: F.StdSegment, 2: F.StdQuad, 3: F.StdHex}


def rel(a, b):
    return float(np.max(np.abs(a - b) / np.max(np.abs(b), axis=0)))


def disc_for(nd, n, nel, lo, hi, eq, op, nodes="GLL", precision="fp64", mapping=None):
    b = F.LagrangeBasis(nodes, n)
    std = STD[nd](b, F.DGSEMrec(b), eq.nvariables)
    mesh = F.CartesianMesh(nd, lo, hi, nel)
    mesh.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
    if mapping is not None:
        mesh.apply_mapping(mapping)
    return F.MultielementDisc(mesh, std, eq, op, precision=precision)


def rhs_of(disc, eq, Q):
    dQ = disc.new_state()
    F.rhs_(dQ, disc.to_device(Q), F.EquationConfig(disc, eq))
    disc.check()
    return disc.to_host(dQ).astype(np.float64)


def test_c1_advection_against_fixture():
    g, meta = golden("c1_advection_2d_n4.npz")
    eq = F.LinearAdvection(1.0, 0.5)
    disc = disc_for(2, 5, [16, 16], (0.0, 0.0), (1.0, 1.0), eq, F.StrongDivOperator(F.LxF(F.StdAverage(), 1.0)))
    Qd = disc.to_device(g["Q0"])
    dt = F.get_max_dt(Qd, disc, eq, meta["cfl"])
    assert dt == float(g["dt"])
    r = rhs_of(disc, eq, g["Q0"])
    assert rel(r, g["rhs0"]) <= 1e-13
    F.timeintegrate(Qd, disc, eq, F.SSPRK33(), meta["steps"] * dt, dt=dt)
    got = disc.to_host(Qd)
    print(f"C1 fixture: rhs bit-identical {np.mean(r == g['rhs0']):.4f}; 100 steps {rel(got, g['Q100']):.2e}, "
          f"bit-identical {np.mean(got == g['Q100']):.4f}")
    assert rel(got, g["Q100"]) <= 1e-13


OPS = {"ch_lxf": lambda: F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0)),
       "kg_lxf": lambda: F.SplitDivOperator(F.KennedyGruber(), F.LxF(F.StdAverage(), 1.0)),
       "strong_lxf": lambda: F.StrongDivOperator(F.LxF(F.StdAverage(), 1.0))}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_euler3d_n5_against_fixture(precision):
    g, meta = golden("euler3d_n5_2x2x2.npz")
    eq = F.EulerEquation(3, GAMMA)
    Q = g["Q"].astype(np.float32).astype(np.float64) if precision == "fp32" else g["Q"]
    for key in ("ch_lxf", "kg_lxf", "strong_lxf"):
        disc = disc_for(3, 6, [2, 2, 2], (0.0,) * 3, (1.0,) * 3, eq, OPS[key](), precision=precision)
        r = rhs_of(disc, eq, Q)
        err = rel(r, g["rhs_" + key])
        print(f"euler3d {precision} {key}: rhs {err:.2e}" + (f", bit-identical {np.mean(r == g['rhs_' + key]):.4f}"
                                                             if precision == "fp64" else
                                                             f", fast kernel {disc.uses_fast_kernel()}"))
        assert err <= (1e-12 if precision == "fp64" else 2e-5)
    disc = disc_for(3, 6, [2, 2, 2], (0.0,) * 3, (1.0,) * 3, eq, OPS["ch_lxf"](), precision=precision)
    Qd = disc.to_device(Q)
    dt = float(g["dt"])
    F.timeintegrate(Qd, disc, eq, F.CarpenterKennedy2N54(), dt, dt=dt)
    err = rel(disc.to_host(Qd).astype(np.float64), g["Q_ck2n54_1"])
    print(f"euler3d {precision}: one CK2N54 step {err:.2e}")
    assert err <= (1e-12 if precision == "fp64" else 5e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c2_kennedy_gruber_against_fixture(precision):
    g, meta = golden("c2_kennedy_gruber_n7_4x4.npz")
    eq = F.EulerEquation(2, GAMMA)
    disc = disc_for(2, 8, [4, 4], (0.0, 0.0), (2.0, 2.0), eq, OPS["kg_lxf"](), precision=precision)
    Q0 = g["Q0"].astype(np.float32).astype(np.float64) if precision == "fp32" else g["Q0"]
    r = rhs_of(disc, eq, Q0)
    err0 = rel(r, g["rhs0"])
    Qd = disc.to_device(Q0)
    dt = float(g["dt"])
    F.timeintegrate(Qd, disc, eq, F.SSPRK33(), meta["steps"] * dt, dt=dt)
    err3 = rel(disc.to_host(Qd).astype(np.float64), g["Q3"])
    print(f"C2 fixture {precision}: rhs {err0:.2e}, 3 SSPRK33 steps {err3:.2e}")
    if precision == "fp64":
        assert err0 <= 1e-12
    else:
        # the rhs of this smooth state is a small difference of O(1) fluxes through an N=7
        # derivative: FP32 rounding of the fluxes shows at ~1e-4 of the rhs.  Bounded by the
        # flux scale instead: |err| <= 5 ε32 ‖D♯‖∞ (2/Δx) max|F|, with max|F| <= 2 max|Q|
        dx = 2.0 / 4
        bound = (5 * np.finfo(np.float32).eps * np.max(np.sum(np.abs(disc.std.Dsharp), 1)) * (2 / dx)
                 * 2 * np.max(np.abs(Q0)))
        print(f"C2 fixture fp32: max abs rhs err {np.max(np.abs(r - g['rhs0'])):.2e} (bound {bound:.1e})")
        assert np.max(np.abs(r - g["rhs0"])) <= bound
    assert err3 <= (1e-12 if precision == "fp64" else 5e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c5_curved_against_fixture(precision):
    g, meta = golden("c5_curved_n6_2x2x2.npz")
    eq = F.EulerEquation(3, GAMMA)
    disc = disc_for(3, 7, [2, 2, 2], (0.0,) * 3, (1.0,) * 3, eq, OPS["ch_lxf"](), precision=precision, mapping=warp)
    Q = g["Q"].astype(np.float32).astype(np.float64) if precision == "fp32" else g["Q"]
    err = rel(rhs_of(disc, eq, Q), g["rhs"])
    print(f"C5 fixture {precision}: rhs {err:.2e}")
    assert err <= (1e-12 if precision == "fp64" else 5e-5)


FORMS = {"hybrid_gll6": ("GLL", 6, 1.0), "hybrid_gl4": ("GL", 4, 0.1), "split_gl5": ("GL", 5, None)}


@pytest.mark.parametrize("key", sorted(FORMS))
def test_subcell_forms_against_fixture(key):
    g, meta = golden("subcell_forms_2d.npz")
    nodes, n, blend = FORMS[key]
    assert meta["forms"][key]["nodes"] == nodes and meta["forms"][key]["n"] == n
    eq = F.EulerEquation(2, GAMMA)
    surf = F.MatrixDissipation(F.ChandrasekharAverage(), 1.0)
    op = (F.HybridDivOperator(F.ChandrasekharAverage(), surf, blend) if blend is not None
          else F.SplitDivOperator(F.ChandrasekharAverage(), surf))
    disc = disc_for(2, n, [3, 3], (0.0, 0.0), (1.0, 1.5), eq, op, nodes=nodes)
    r = rhs_of(disc, eq, g["Q_" + key])
    err = rel(r, g["rhs_" + key])
    print(f"{key} fixture: rhs {err:.2e}, bit-identical {np.mean(r == g['rhs_' + key]):.4f}")
    assert err <= 1e-12
