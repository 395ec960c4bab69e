"""VTKHDF output of device states on the GPU (src/FlouSpatial/IO.jl:16-126,
src/FlouBiz/FlouBiz.jl:16-115, src/FlouTime/FlouTime.jl:59-90).

* flou_hip_equispaced (pointdata2VTKHDF) against the oracle's dense node2eq product on the
  same input: FP64 within 1e-14 of max|Q| (the kernel sum-factorises, so the roundings
  differ), FP32 within 5e-6 of max|Q| against the FP64 oracle on the FP32-rounded input;
  1-D to 3-D, GLL and Gauss nodes, m = n and m != n, states and gradient-like blocks,
  device and host inputs;
* the save callback through timeintegrate: the files it writes, read back with h5dump,
  hold the mesh of the host arrays and, for every saved iteration, Time and the projected
  solution of the oracle's own RK run (1e-12 in FP64);
* argument errors.
"""
import os

import numpy as np
import pytest

import h5read
import oracle_lib as O
from cases import GAMMA, oracle_for, random_euler_state

import flouhip as F
from flouhip import io as FIO

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STD = {1: F.StdSegment, 2: F.StdQuad, 3: F.StdHex}


def build(nd, n, nel, nodes="GLL", m=None, precision="fp64"):
    b = F.LagrangeBasis(nodes, n)
    std = STD[nd](b, F.DGSEMrec(b), nd + 2, nequispaced=m)
    finish = [1.0, 1.5, 2.0][:nd]
    mesh = F.CartesianMesh(nd, [0.0] * nd, finish, [nel] * nd)
    mesh.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
    eq = F.EulerEquation(nd, GAMMA)
    op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))
    return F.MultielementDisc(mesh, std, eq, op, precision=precision), mesh, std, eq


CASES = [(1, 5, 6, "GLL", None), (1, 4, 5, "GL", 7), (2, 8, 3, "GLL", None), (2, 4, 3, "GL", 2),
         (3, 6, 2, "GLL", None), (3, 7, 2, "GLL", None), (3, 4, 2, "GL", 9), (3, 6, 3, "GLL", 3)]


def _ids(c):
    return f"{c[0]}d-n{c[1]}-{c[3]}-m{c[4] or c[1]}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_equispaced_kernel_matches_oracle(case, precision):
    nd, n, nel, nodes, m = case
    disc, mesh, std, eq = build(nd, n, nel, nodes, m, precision)
    X = mesh.node_coords(std)
    Q = random_euler_state(X, nd, seed=3, amp=0.3, smooth=False)
    if precision == "fp32":
        Q = Q.astype(np.float32).astype(np.float64)
    ref = O.equispaced(nd, n, std.node2eq1, mesh.nelements, Q)
    tol = (1e-14 if precision == "fp64" else 5e-6) * np.max(np.abs(Q))
    got = np.stack(FIO.pointdata2VTKHDF(disc.to_device(Q), disc), 1)
    assert got.dtype == disc.dtype and got.shape == ref.shape
    err = np.max(np.abs(got - ref))
    print(f"{_ids(case)} {precision}: max abs err {err:.2e} (tol {tol:.1e})")
    assert err <= tol
    # host input (uploaded first) and a gradient-like block (nd, nvars, ndofs) on the device
    got_h = np.stack(FIO.pointdata2VTKHDF(Q.astype(disc.dtype), disc), 1)
    assert np.array_equal(got_h, got)
    blk = np.random.default_rng(4).standard_normal((nd, nd + 2, disc.ndofs))
    if precision == "fp32":
        blk = blk.astype(np.float32).astype(np.float64)
    cols = FIO.pointdata2VTKHDF(torch.from_numpy(blk.astype(disc.dtype)).cuda(), disc)
    assert len(cols) == nd * (nd + 2)
    refb = O.equispaced(nd, n, std.node2eq1, mesh.nelements, blk.reshape(-1, disc.ndofs).T)
    assert np.max(np.abs(np.stack(cols, 1) - refb)) <= (1e-14 if precision == "fp64" else 5e-6) * np.max(np.abs(blk))


needs_h5 = pytest.mark.skipif(not (h5read.available() and os.path.exists(FIO.VTKHDF_LIB_PATH)),
                              reason="h5dump or libflouvtkhdf.so not present")


@needs_h5
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_save_callback_files(tmp_path, precision):
    nd, n, nel = 2, 5, 3
    disc, mesh, std, eq = build(nd, n, nel, precision=precision)
    orc = oracle_for(nd, n, [nel] * nd, [0.0] * nd, [1.0, 1.5], [1] * nd, O.EQ_EULER, [GAMMA], O.VOL_SPLIT,
                     O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF, 1.0)
    X = mesh.node_coords(std)
    Q0 = random_euler_state(X, nd, seed=8, amp=0.2)
    if precision == "fp32":
        Q0 = Q0.astype(np.float32).astype(np.float64)
    dt = 0.2 * orc.max_dt(Q0, 1.0)
    base = str(tmp_path / "sol")
    cb = F.get_save_callback(base, iter=[2, 3])
    Qd = disc.to_device(Q0)
    F.timeintegrate(Qd, disc, eq, F.SSPRK33(), 4 * dt, dt=dt, callback=cb)
    want = [f"{base}_{it:010d}.hdf" for it in (0, 2, 3)]
    assert cb.files == want and all(os.path.exists(f) for f in want)
    assert not os.path.exists(f"{base}_{1:010d}.hdf")

    pts, conn, offs, types, reg = FIO.mesh_arrays(mesh, std, disc.dtype)
    real = "H5T_IEEE_F64LE" if precision == "fp64" else "H5T_IEEE_F32LE"
    tol = 1e-12 if precision == "fp64" else 2e-5
    for f, it in zip(want, (0, 2, 3)):
        assert h5read.dataset_header(f, "/VTKHDF/Points") == (real, (pts.shape[0], 3))
        assert np.array_equal(h5read.dataset(f, "/VTKHDF/Points"), pts)
        assert np.array_equal(h5read.dataset(f, "/VTKHDF/Connectivity"), conn)
        assert np.array_equal(h5read.dataset(f, "/VTKHDF/Offsets"), offs)
        assert np.array_equal(h5read.dataset(f, "/VTKHDF/Types"), types)
        t = h5read.dataset(f, "/VTKHDF/FieldData/Time")
        assert t.shape == (1,) and abs(t[0] - it * dt) <= 1e-15 * max(1.0, it * dt)
        Qit = orc.rk_run(O.RK_SSPRK33, Q0, dt, it) if it else Q0
        ref = O.equispaced(nd, n, std.node2eq1, mesh.nelements, Qit)
        for iv, name in enumerate(eq.variablenames()):
            got = h5read.dataset(f, "/VTKHDF/PointData/" + name)
            assert h5read.dataset_header(f, "/VTKHDF/PointData/" + name) == (real, (pts.shape[0],))
            err = np.max(np.abs(got - ref[:, iv])) / np.max(np.abs(ref[:, iv]))
            assert err <= tol, (f, name, err)
    print(f"save callback {precision}: files {[os.path.basename(x) for x in want]}")


def test_equispaced_argument_errors():
    disc, mesh, std, eq = build(3, 8, 1)
    Qd = disc.new_state()
    Qd.fill_(1.0)
    A = np.ascontiguousarray(std.node2eq1)
    out = torch.empty((5, 8 ** 3), dtype=torch.float64, device="cuda")
    import ctypes
    from flouhip import _capi as C
    dp = A.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert C.lib.flou_hip_equispaced(disc.h, 0, dp, ctypes.c_void_p(Qd.data_ptr()), 5,
                                     ctypes.c_void_p(out.data_ptr())) == C.EINVAL
    big = np.zeros((17, 8))
    assert C.lib.flou_hip_equispaced(disc.h, 17, big.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ctypes.c_void_p(Qd.data_ptr()), 5, ctypes.c_void_p(out.data_ptr())) == C.EUNSUPPORTED
    nan = A.copy()
    nan[0, 0] = np.nan
    assert C.lib.flou_hip_equispaced(disc.h, 8, nan.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ctypes.c_void_p(Qd.data_ptr()), 5, ctypes.c_void_p(out.data_ptr())) == C.EINVAL
    # the valid call still works after the rejected ones: a constant stays constant
    cols = FIO.pointdata2VTKHDF(Qd, disc)
    assert all(np.max(np.abs(c - 1.0)) <= 1e-14 for c in cols)
