"""flou_hip_rk_step_vjp_noalloc: the adjoint RK step with its work states in the library's
static per-GPU scratch buffer instead of allocated memory.

Called through the C ABI and compared with flou_hip_rk_step_vjp (same kernels, so the
results must be identical bit for bit) on two calls in a row, the second on a smaller
mesh, for the three RK schemes in FP64 and FP32; a state too large for the buffer is
refused with EINVAL.
"""
import ctypes

import numpy as np
import pytest

import flouhip as F
from flouhip import _capi as C
from flouhip import adjoint as A
from cases import periodic_euler_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _disc(nel, n, precision):
    b = F.LagrangeBasis("GLL", n)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    mesh = F.CartesianMesh(2, (0.0, 0.0), (2.0, 2.0), (nel, nel)).apply_periodicBCs(("1", "2"), ("3", "4"))
    eq = F.EulerEquation(2, 1.4)
    op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))
    return F.MultielementDisc(mesh, std, eq, op, precision=precision), eq, std, mesh


def _noalloc(disc, code, dt, stages, ubar):
    arr = (ctypes.c_void_p * len(stages))()
    for i, s in enumerate(stages):
        arr[i] = s.data_ptr()
    disc._sync_stream()
    return C.lib.flou_hip_rk_step_vjp_noalloc(disc.h, code, float(dt), 0.0, arr, ctypes.c_void_p(ubar.data_ptr()))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_step_vjp_equals_allocating_entry_point(precision):
    # the second mesh is smaller: the scratch buffer keeps the first call's contents
    for nel, n in ((24, 6), (5, 4)):
        disc, eq, std, mesh = _disc(nel, n, precision)
        X = mesh.node_coords(std)
        Q = disc.to_device(periodic_euler_state(X, 2, 2.0, seed=nel, amp=0.1))
        dt = 0.1 * F.get_max_dt(Q, disc, eq, 1.0)
        rng = np.random.default_rng(n)
        for solver in (F.SSPRK33, F.CarpenterKennedy2N54, F.ORK256):
            S = 3 if solver is F.SSPRK33 else 5
            stages = [Q.clone()] + [disc.new_state() for _ in range(S - 1)]
            A.rk_step_record(disc, solver, dt, stages, disc.new_state())
            ubar0 = disc.to_device(rng.standard_normal((disc.ndofs, 4)))
            got = ubar0.clone()
            rc = _noalloc(disc, solver.code, dt, stages, got)
            assert rc == C.OK, C.lib.flou_hip_last_error(disc.h)
            ref = A.rk_step_vjp(disc, solver, dt, stages, ubar0.clone())
            torch.cuda.synchronize()
            same = torch.equal(got, ref)
            print(f"{precision} {nel}x{nel} N={n - 1} {solver.__name__}: noalloc == allocating: {same}")
            assert same
            assert float(got.abs().max()) > 0.0


def test_noalloc_refuses_states_beyond_the_scratch_buffer():
    # 150^2 quads at N=5: 810,000 nodes, 25.9 MB per FP64 state, three exceed 64 MiB
    disc, eq, std, mesh = _disc(150, 6, "fp64")
    S = 3
    stages = [disc.new_state() for _ in range(S)]
    ubar = disc.new_state()
    rc = _noalloc(disc, F.SSPRK33.code, 1e-3, stages, ubar)
    assert rc == C.EINVAL
    assert b"scratch" in C.lib.flou_hip_last_error(disc.h)
