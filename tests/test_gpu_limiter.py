"""Zhang & Shu stage limiter and the kinetic-energy / entropy monitors on the GPU, through
the C ABI, against the oracle (Equations/Euler.jl:541-654, FlouTime.jl:107-148).

FP64: the limiter is bit-identical to the oracle (same fold order, no contraction), limited
RK runs agree to 1e-13, monitors to 1e-13 (a different summation tree).  FP32: 1e-5
against the FP64 oracle.
"""
import ctypes

import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA, random_euler_state
from flouhip import _capi as C
from flouhip import adjoint as A
from test_gpu_curved import build_curved
from test_gpu_parity import build

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def make(nd, n, nel, precision="fp64", tp="ch"):
    mesh, std, eq, op, orc = build(nd, n, [nel] * nd, "split", tp=tp)
    disc = F.MultielementDisc(mesh, std, eq, op, precision=precision)
    X, _ = orc.coords()
    return disc, eq, orc, X


def negative_state(X, nd, ndofs, which, seed=21):
    Q = random_euler_state(X, nd, seed=seed, amp=0.5, smooth=False)
    rng = np.random.default_rng(seed + 1)
    bad = rng.choice(ndofs, ndofs // 10, replace=False)
    if which == "rho":
        Q[bad, 0] *= -0.2
    else:
        Q[bad, nd + 1] = np.sum(Q[bad, 1:nd + 1] ** 2, 1) / (2 * Q[bad, 0]) - 0.05
    return np.asfortranarray(Q)


LCASES = [(1, 6, 16), (2, 5, 6), (3, 4, 4), (3, 6, 3)]


@pytest.mark.parametrize("which", ["rho", "p"])
@pytest.mark.parametrize("case", LCASES, ids=lambda c: "-".join(map(str, c)))
def test_limiter_once_bit_identical_fp64(case, which):
    nd, n, nel = case
    disc, eq, orc, X = make(nd, n, nel)
    Q = negative_state(X, nd, orc.ndofs, which)
    ref = orc.limit(Q, 1e-6)
    assert not np.array_equal(ref, Q)
    Qd = disc.to_device(Q)
    F.get_limiter(disc, eq, "zhang_shu", 1e-6)(Qd, disc, eq)
    got = disc.to_host(Qd)
    print(f"{case} {which}: bit-identical {np.mean(got == ref):.4f}, "
          f"limited nodes {np.mean(ref != Q):.3f}")
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("case", [LCASES[1], LCASES[2]], ids=["2d", "3d"])
def test_limiter_once_fp32(case):
    nd, n, nel = case
    disc, eq, orc, X = make(nd, n, nel, precision="fp32")
    Q = negative_state(X, nd, orc.ndofs, "rho")
    Q32 = Q.astype(np.float32).astype(np.float64)
    ref = orc.limit(Q32, 1e-6)
    Qd = disc.to_device(Q32)
    F.get_limiter(disc, eq, "zhang_shu", 1e-6)(Qd, disc, eq)
    got = disc.to_host(Qd).astype(np.float64)
    err = np.max(np.abs(got - ref) / np.max(np.abs(ref), axis=0))
    print(f"FP32 limiter {case}: max rel err {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("solver,scheme", [(F.SSPRK33, O.RK_SSPRK33), (F.CarpenterKennedy2N54, O.RK_CK2N54),
                                           (F.ORK256, O.RK_ORK256)], ids=["ssprk33", "ck2n54", "ork256"])
def test_limited_timeintegrate_matches_oracle(solver, scheme):
    """Stage limiter after every stage with minval above the smallest nodal density and
    pressure, so that it acts in every stage."""
    disc, eq, orc, X = make(2, 4, 6)
    Q = random_euler_state(X, 2, seed=8, amp=0.3)
    minval, nsteps = 0.85, 5
    dt = 0.2 * orc.max_dt(Q, 1.0)
    orc.set_limiter(True, minval)
    ref = orc.rk_run(scheme, Q, dt, nsteps)
    orc.set_limiter(False)
    free = orc.rk_run(scheme, Q, dt, nsteps)
    assert np.abs(ref - free).max() > 1e-3
    Qd = disc.to_device(Q)
    lim = F.get_limiter_callback(disc, eq, "zhang_shu", minval)
    F.timeintegrate(Qd, disc, eq, solver(stage_limiter=lim), nsteps * dt, dt=dt)
    got = disc.to_host(Qd)
    err = np.max(np.abs(got - ref) / np.max(np.abs(ref), axis=0))
    print(f"limited {solver.__name__}: max rel err {err:.2e}, bit-identical {np.mean(got == ref):.4f}")
    assert err <= 1e-13
    # the handle forgets the limiter after the call
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, eq, solver, nsteps * dt, dt=dt)
    assert np.max(np.abs(disc.to_host(Qd) - free)) <= 1e-13 * np.abs(free).max()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
@pytest.mark.parametrize("nd", [2, 3])
def test_monitors_match_oracle(nd, precision, tol):
    disc, eq, orc, X = make(nd, 5, 4, precision=precision)
    Q = random_euler_state(X, nd, seed=5, amp=0.3)
    if precision == "fp32":
        Q = Q.astype(np.float32).astype(np.float64)
    Qd = disc.to_device(Q)
    for name, kind in (("kinetic_energy", O.MON_KINETIC_ENERGY), ("energy", O.MON_KINETIC_ENERGY),
                       ("entropy", O.MON_ENTROPY)):
        got = F.get_monitor(disc, eq, name)(Qd, disc, eq)
        ref = orc.monitor(Q, kind)
        print(f"{nd}-D {precision} {name}: {got:.16e} vs {ref:.16e}")
        assert abs(got - ref) <= tol * abs(ref)


def test_monitors_on_curved_mesh():
    disc, orc, eq, X, J, std = build_curved(3, 5, 3)
    Q = random_euler_state(X, 3, seed=6, amp=0.2)
    Qd = disc.to_device(Q)
    for name, kind in (("kinetic_energy", O.MON_KINETIC_ENERGY), ("entropy", O.MON_ENTROPY)):
        got, ref = F.get_monitor(disc, eq, name)(Qd, disc, eq), orc.monitor(Q, kind)
        assert abs(got - ref) <= 1e-13 * abs(ref)


def test_limiter_on_curved_mesh_bit_identical():
    disc, orc, eq, X, J, std = build_curved(2, 6, 4)
    Q = negative_state(X, 2, orc.ndofs, "rho", seed=31)
    ref = orc.limit(Q, 1e-6)
    Qd = disc.to_device(Q)
    F.get_limiter(disc, eq, "zhang_shu", 1e-6)(Qd, disc, eq)
    assert np.array_equal(disc.to_host(Qd), ref)


def test_monitor_callback_records_every_step():
    disc, eq, orc, X = make(2, 4, 5)
    Q = random_euler_state(X, 2, seed=9, amp=0.2)
    dt = 0.3 * orc.max_dt(Q, 1.0)
    cb, out = F.get_monitor_callback(float, float, disc, eq, "entropy")
    cb2, out2 = F.get_monitor_callback(float, float, disc, eq, "kinetic_energy", iter=[2, 4])
    Qd = disc.to_device(Q)
    F.timeintegrate(Qd, disc, eq, F.CarpenterKennedy2N54(), 4 * dt, dt=dt,
                    callback=F.make_callback_list(cb, cb2))
    assert out.iter == [1, 2, 3, 4] and out2.iter == [2, 4]
    u = Q
    for k in range(4):
        u = orc.rk_run(O.RK_CK2N54, u, dt, 1)
        assert abs(out.value[k] - orc.monitor(u, O.MON_ENTROPY)) <= 1e-12 * abs(out.value[k])
        assert abs(out.time[k] - (k + 1) * dt) <= 1e-14
        if k in (1, 3):
            i = out2.iter.index(k + 1)
            assert abs(out2.value[i] - orc.monitor(u, O.MON_KINETIC_ENERGY)) <= 1e-12 * out2.value[i]


def test_entropy_monitor_domain_error():
    disc, eq, orc, X = make(2, 3, 2)
    Q = np.asfortranarray(np.tile([1.0, 0.0, 0.0, -1.0], (orc.ndofs, 1)))
    with pytest.raises(F.DomainError):
        F.get_monitor(disc, eq, "entropy")(disc.to_device(Q), disc, eq)
    # the flag is cleared: a valid state afterwards is fine
    Q[:, 3] = 2.5
    F.get_monitor(disc, eq, "entropy")(disc.to_device(Q), disc, eq)


def test_limiter_errors():
    disc, eq, orc, X = make(2, 3, 2)
    with pytest.raises(ValueError):
        F.get_limiter(disc, eq, "zhang_shu")
    with pytest.raises(ValueError):
        F.get_monitor(disc, eq, "vorticity")
    assert F.list_limiters(disc, eq) == ("zhang_shu",)
    assert F.list_monitors(disc, eq) == ("kinetic_energy", "entropy")
    # C ABI: non-finite minval, unknown kind, and the adjoint refuses a limited handle
    L = C.lib
    assert L.flou_hip_set_limiter(disc.h, C.LIMITER_ZHANG_SHU, float("nan")) == C.EINVAL
    assert L.flou_hip_set_limiter(disc.h, 7, 1e-6) == C.EINVAL
    assert L.flou_hip_set_limiter(disc.h, C.LIMITER_ZHANG_SHU, 1e-6) == 0
    Q = random_euler_state(X, 2, seed=1)
    st = [disc.to_device(Q) for _ in range(3)]
    ptrs = (ctypes.c_void_p * 3)(*[disc._dptr(s) for s in st])
    nxt = disc.new_state()
    assert L.flou_hip_rk_step_record(disc.h, C.SSPRK33, 1e-3, 0.0, ptrs, disc._dptr(nxt)) == C.EUNSUPPORTED
    assert L.flou_hip_set_limiter(disc.h, C.LIMITER_NONE, 0.0) == 0
    assert L.flou_hip_rk_step_record(disc.h, C.SSPRK33, 1e-3, 0.0, ptrs, disc._dptr(nxt)) == 0
    with pytest.raises(NotImplementedError):
        A.forward(disc, F.SSPRK33(stage_limiter=F.get_limiter_callback(disc, eq, "zhang_shu", 1e-6)),
                  disc.to_device(Q), 1e-3, 2)
    # linear advection has no limiter or monitors
    mesh, std, eqa, op, _ = build(2, 3, [2, 2], "strong", eq="advection")
    da = F.MultielementDisc(mesh, std, eqa, op)
    assert L.flou_hip_set_limiter(da.h, C.LIMITER_ZHANG_SHU, 1e-6) == C.EUNSUPPORTED
    out = ctypes.c_double()
    assert L.flou_hip_monitor(da.h, da._dptr(da.new_state()), 0, ctypes.byref(out)) == C.EUNSUPPORTED
