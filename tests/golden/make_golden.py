"""Write the golden fixtures of tests/golden/ with the CPU oracle.

    python tests/golden/make_golden.py

Each fixture is an .npz of FP64 arrays (inputs and the oracle's outputs) with a `meta`
entry (JSON text: configuration and provenance), loadable with allow_pickle=False.  The
oracle follows the reference's operation order, so these are the reference algorithm's
values on these inputs up to the unpinned round-off listed with the oracle
(Polynomials.fit operators, BLAS sums).  tests/test_golden.py regenerates every output
and requires the same bits; the GPU tests compare the HIP path with the stored arrays.

reference_answers.json holds the known answers of the reference's own test suite
(test/runtests.jl), which no code here produces.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
for p in (TESTS, os.path.join(ROOT, "flou.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle_lib as O  # noqa: E402
from cases import GAMMA, c2_state, random_euler_state, warp  # noqa: E402


def c1_advection():
    """C1: LinearAdvection(1, 0.5), 16x16 periodic on [0,1]^2, GLL 5, strong LxF, SSPRK33,
    dt = get_max_dt at CFL 0.3, 100 steps; Gaussian pulse + seeded U(-1e-3, 1e-3)."""
    d = O.OracleDisc(2, 5, [16, 16], [0.0, 0.0], [1.0, 1.0], [1, 1], O.build_ops(5, "GLL"), O.EQ_ADVECTION,
                     [1.0, 0.5], O.VOL_STRONG, 0, O.AVG_STD, O.DISS_LXF, 1.0)
    X, _ = d.coords()
    rng = np.random.default_rng(1)
    Q0 = np.asfortranarray((np.exp(-((X[:, 0] - 0.5) ** 2 + (X[:, 1] - 0.5) ** 2) / 0.01)
                            + rng.uniform(-1e-3, 1e-3, X.shape[0]))[:, None])
    dt = d.max_dt(Q0, 0.3)
    return dict(Q0=Q0, dt=np.array(dt), rhs0=d.rhs(Q0), Q100=d.rk_run(O.RK_SSPRK33, Q0, dt, 100)), dict(
        config="C1", nd=2, n=5, nel=[16, 16], domain=[[0, 0], [1, 1]], equation="LinearAdvection(1.0, 0.5)",
        operator="StrongDivOperator(LxF(StdAverage(), 1.0))", scheme="SSPRK33", cfl=0.3, steps=100)


EULER3D_OPS = {
    "ch_lxf": (O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF,
               "SplitDivOperator(ChandrasekharAverage(), LxF(StdAverage(), 1.0))"),
    "kg_lxf": (O.VOL_SPLIT, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF,
               "SplitDivOperator(KennedyGruber(), LxF(StdAverage(), 1.0))"),
    "strong_lxf": (O.VOL_STRONG, 0, O.AVG_STD, O.DISS_LXF, "StrongDivOperator(LxF(StdAverage(), 1.0))"),
}


def euler3d_n5():
    """3-D Euler N=5 (GLL 6), 2x2x2 periodic hexahedra on [0,1]^3: one seeded smooth state,
    the rhs of three operators, and one CarpenterKennedy2N54 step of the C3 operator."""
    arrays, ops = {}, {}
    Q = None
    for key, (vol, tp, savg, sdiss, text) in EULER3D_OPS.items():
        d = O.OracleDisc(3, 6, [2, 2, 2], [0.0] * 3, [1.0] * 3, [1, 1, 1], O.build_ops(6, "GLL"), O.EQ_EULER,
                         [GAMMA], vol, tp, savg, sdiss, 1.0)
        if Q is None:
            X, _ = d.coords()
            Q = random_euler_state(X, 3, seed=11, amp=0.2)
            arrays["Q"] = Q
        arrays["rhs_" + key] = d.rhs(Q)
        ops[key] = text
        if key == "ch_lxf":
            dt = 0.3 * d.max_dt(Q, 1.0)
            arrays["dt"] = np.array(dt)
            arrays["Q_ck2n54_1"] = d.rk_run(O.RK_CK2N54, Q, dt, 1)
    return arrays, dict(nd=3, n=6, nel=[2, 2, 2], domain=[[0, 0, 0], [1, 1, 1]], gamma=GAMMA, operators=ops,
                        state="random_euler_state(X, 3, seed=11, amp=0.2)",
                        step="one CarpenterKennedy2N54 step of ch_lxf, dt = 0.3 get_max_dt(Q, 1)")


def c2_kennedy_gruber():
    """C2 operator on a 4x4 sub-size: 2-D Euler N=7 (GLL 8) on [0,2]^2, split Kennedy-Gruber
    + LxF(StdAverage), SSPRK33 3 steps at CFL 0.1."""
    d = O.OracleDisc(2, 8, [4, 4], [0.0, 0.0], [2.0, 2.0], [1, 1], O.build_ops(8, "GLL"), O.EQ_EULER, [GAMMA],
                     O.VOL_SPLIT, O.TP_KENNEDY_GRUBER, O.AVG_STD, O.DISS_LXF, 1.0)
    X, _ = d.coords()
    Q0 = c2_state(X)
    dt = d.max_dt(Q0, 0.1)
    return dict(Q0=Q0, dt=np.array(dt), rhs0=d.rhs(Q0), Q3=d.rk_run(O.RK_SSPRK33, Q0, dt, 3)), dict(
        config="C2 sub-size", nd=2, n=8, nel=[4, 4], domain=[[0, 0], [2, 2]], gamma=GAMMA,
        operator="SplitDivOperator(KennedyGruber(), LxF(StdAverage(), 1.0))", scheme="SSPRK33", cfl=0.1,
        steps=3, state="c2_state(X, seed=2)")


def c5_curved():
    """C5 map x + 0.05 sin 2πx sin 2πy sin 2πz on [0,1]^3, 2x2x2 elements, GLL 7, split
    Chandrasekhar + LxF(StdAverage): per-node geometry from the mapped nodes."""
    import flouhip as F
    from flouhip import mesh as M
    b = F.LagrangeBasis("GLL", 7)
    std = F.StdHex(b, F.DGSEMrec(b), 5)
    mesh = F.CartesianMesh(3, (0.0,) * 3, (1.0,) * 3, (2, 2, 2))
    mesh.apply_periodicBCs(("1", "2"), ("3", "4"), ("5", "6"))
    mesh.apply_mapping(warp)
    J, met, ev, fr, fj = M.curved_geometry(mesh, std)
    d = O.OracleDisc(3, 7, [2, 2, 2], [0.0] * 3, [1.0] * 3, [1, 1, 1], O.build_ops(7, "GLL"), O.EQ_EULER, [GAMMA],
                     O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF, 1.0)
    d.set_curved(met, J, fr.reshape(-1), fj, ev)
    X = mesh.node_coords(std)
    Q = random_euler_state(X, 3, seed=13, amp=0.2)
    return dict(Q=Q, rhs=d.rhs(Q)), dict(config="C5 sub-size", nd=3, n=7, nel=[2, 2, 2], map="cases.warp",
                                         operator="SplitDivOperator(ChandrasekharAverage(), LxF(StdAverage(), 1.0))",
                                         state="random_euler_state(X, 3, seed=13, amp=0.2)", gamma=GAMMA)


def subcell_forms():
    """2-D Euler on [0,1]x[0,1.5], 3x3 periodic quads: HybridDivOperator on GLL 6 (c = 1) and
    on Gauss 4 (c = 0.1), and the split form on Gauss 5, each with Chandrasekhar +
    MatrixDissipation(Chandrasekhar, 1)."""
    out, meta = {}, {}
    for key, nodes, n, vol, blend in (("hybrid_gll6", "GLL", 6, O.VOL_HYBRID, 1.0),
                                      ("hybrid_gl4", "GL", 4, O.VOL_HYBRID, 0.1),
                                      ("split_gl5", "GL", 5, O.VOL_SPLIT, None)):
        d = O.OracleDisc(2, n, [3, 3], [0.0, 0.0], [1.0, 1.5], [1, 1], O.build_ops(n, nodes), O.EQ_EULER, [GAMMA],
                         vol, O.TP_CHANDRASEKHAR, O.AVG_CHANDRASEKHAR, O.DISS_MATRIX, 1.0)
        if blend is not None:
            d.set_blend(blend)
        X, _ = d.coords()
        Q = random_euler_state(X, 2, seed=17, amp=0.2)
        out["Q_" + key], out["rhs_" + key] = Q, d.rhs(Q)
        meta[key] = dict(nodes=nodes, n=n, blend=blend,
                         operator=("HybridDivOperator" if vol == O.VOL_HYBRID else "SplitDivOperator")
                         + "(ChandrasekharAverage(), MatrixDissipation(ChandrasekharAverage(), 1.0)"
                         + (f", {blend})" if blend is not None else ")"))
    return out, dict(nd=2, nel=[3, 3], domain=[[0, 0], [1, 1.5]], gamma=GAMMA, forms=meta,
                     state="random_euler_state(X, 2, seed=17, amp=0.2)")


C4_SUB = dict(nel=8, n=4, steps=200, seeds=(101, 202), cfl=0.1)


def c4_problem():
    """The C4 loss on its 8x8, N=3 sub-size: 2-D Euler on [0,2]^2 periodic, GLL 4, split
    Chandrasekhar + LxF(StdAverage, 1), SSPRK33 at dt = 0.1 get_max_dt(Q_A, 1) for 200
    steps; L(Q_A) = Σ Jω (ρ_A(T) − ρ_B(T))² with Q_A, Q_B seeded smooth periodic states."""
    from cases import periodic_euler_state
    c = C4_SUB
    nel, n = c["nel"], c["n"]

    def make():
        return O.OracleDisc(2, n, [nel, nel], [0.0, 0.0], [2.0, 2.0], [1, 1], O.build_ops(n, "GLL"), O.EQ_EULER,
                            [GAMMA], O.VOL_SPLIT, O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF, 1.0)
    d = make()
    X, _ = d.coords()
    QA = periodic_euler_state(X, 2, 2.0, seed=c["seeds"][0], amp=0.1)
    QB = periodic_euler_state(X, 2, 2.0, seed=c["seeds"][1], amp=0.1)
    dt = c["cfl"] * d.max_dt(QA, 1.0)
    w1 = O.build_ops(n, "GLL")["w"]
    wnode = np.tile(np.kron(w1, w1) * (2.0 / nel / 2.0) ** 2, nel * nel)
    rhoB = d.rk_run(O.RK_SSPRK33, QB, dt, c["steps"])[:, 0]
    uA = d.rk_run(O.RK_SSPRK33, QA, dt, c["steps"])
    dLdu = np.zeros_like(uA)
    dLdu[:, 0] = 2.0 * wnode * (uA[:, 0] - rhoB)
    return make, QA, dt, wnode, rhoB, dLdu


def c4_gradient_columns(make, QA, dt, dLdu, cols, workers=None):
    """∂L/∂Q_A for the flat (column-major) indices `cols`, one dual-number JVP of the
    200-step run per index (forward mode with one partial, as ForwardDiff with chunk 1);
    the indices are shared out over threads, each with its own oracle handle."""
    import concurrent.futures as cf
    cols = list(cols)
    workers = workers or min(16, os.cpu_count() or 1)
    chunks = [cols[i::workers] for i in range(workers)]

    def run(chunk):
        O.set_threads(1)
        d = make()
        E = np.zeros(QA.shape, order="F")
        flat = E.reshape(-1, order="F")
        out = {}
        for i in chunk:
            flat[:] = 0.0
            flat[i] = 1.0
            _, du = d.rk_run_jvp(O.RK_SSPRK33, QA, E, dt, C4_SUB["steps"])
            out[i] = float(np.sum(dLdu * du))
        return out

    res = {}
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for part in ex.map(run, [c for c in chunks if c]):
            res.update(part)
    return np.array([res[i] for i in cols])


def c4_gradient():
    """C4 sub-size: every component of the loss gradient from dual-number JVPs."""
    make, QA, dt, wnode, rhoB, dLdu = c4_problem()
    g = c4_gradient_columns(make, QA, dt, dLdu, range(QA.size)).reshape(QA.shape, order="F")
    return dict(QA=QA, dt=np.array(dt), wnode=wnode, rhoB=rhoB, grad=g), dict(
        config="C4 sub-size", nd=2, n=C4_SUB["n"], nel=[C4_SUB["nel"]] * 2, domain=[[0, 0], [2, 2]], gamma=GAMMA,
        operator="SplitDivOperator(ChandrasekharAverage(), LxF(StdAverage(), 1.0))", scheme="SSPRK33",
        steps=C4_SUB["steps"], cfl=C4_SUB["cfl"],
        states=f"periodic_euler_state(X, 2, 2.0, seed={C4_SUB['seeds']}, amp=0.1)",
        loss="sum(wnode * (rho_A(T) - rhoB)^2)", gradient="oracle dual-number JVP per component")


FIXTURES = {"c1_advection_2d_n4.npz": c1_advection, "euler3d_n5_2x2x2.npz": euler3d_n5,
            "c2_kennedy_gruber_n7_4x4.npz": c2_kennedy_gruber, "c5_curved_n6_2x2x2.npz": c5_curved,
            "subcell_forms_2d.npz": subcell_forms, "c4_gradient_8x8_n3.npz": c4_gradient}


def generate(name):
    arrays, meta = FIXTURES[name]()
    meta = dict(meta, generator="tests/golden/make_golden.py", oracle="oracle/flou_oracle.{hpp,cpp}")
    return {k: np.asarray(v, dtype=np.float64) for k, v in arrays.items()}, meta


def main(names=None):
    for name in names or FIXTURES:
        arrays, meta = generate(name)
        path = os.path.join(HERE, name)
        np.savez_compressed(path, meta=np.array(json.dumps(meta, sort_keys=True)), **arrays)
        print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    main(sys.argv[1:])
