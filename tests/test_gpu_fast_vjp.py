"""The FP32 fast adjoint kernel (flou.jl_amd/csrc/vjp_fast.hip) on curved periodic meshes
(the split form with the Chandrasekhar flux and LxF; OpDivergence.jl:200-299,
Interfaces.jl:111-136, Equations/Euler.jl:474-536):

* the rhs VJP against the plain FP32 VJP (k_vjp, switched on with option 1) on the same
  inputs, over 2-D and 3-D, N = 4..8, element counts that leave the last block part-filled,
  and states whose density and β spreads put lines on the series and on the log branch of
  the log-mean;
* the dot-product identity <W, J V> = <J^T W, V> against the oracle's FP64 dual-number JVP
  (the oracle's rhs is pinned by tests/test_oracle.py and tests/test_curved.py);
* the reverse RK steps (SSPRK33 and CK2N54) against the plain FP32 path.
The hand-derived reverse formulas were also checked against central differences of the
forward fluxes in float64 (series and log branches, the LxF face flux) before the kernel
ran.
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import periodic_euler_state
from flouhip import adjoint as A
from test_gpu_curved import build_curved

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (nd, n, elements per direction): 3-D blocks take 4/3/2/1/1 elements at N = 4..8, 2-D
# blocks 16/32/16/9/8, so most of these leave the last block part-filled
CASES = [(3, 4, 3), (3, 5, 2), (3, 6, 3), (3, 7, 2), (3, 8, 2),
         (2, 4, 5), (2, 5, 6), (2, 6, 5), (2, 7, 4), (2, 8, 3)]


def _fp32_disc(nd, n, nel):
    disc, orc, eq, X, J, std = build_curved(nd, n, nel, "split", "ch", precision="fp32")
    assert disc.uses_fast_kernel() and disc.uses_fast_vjp()
    return disc, orc, X


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


@pytest.mark.parametrize("amp", [0.1, 0.6], ids=["series", "log"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "{}d-n{}-{}".format(*c))
def test_fast_vjp_matches_plain_fp32_vjp(case, amp):
    nd, n, nel = case
    disc, orc, X = _fp32_disc(nd, n, nel)
    Q = periodic_euler_state(X, nd, 1.0, seed=5, amp=amp)
    W = np.random.default_rng(6).standard_normal(Q.shape)
    got, ref = _vjp(disc, Q, W, plain=False), _vjp(disc, Q, W, plain=True)
    err = _rel(got, ref)
    print(f"{case} amp {amp}: fast vs plain FP32 VJP, max rel err {err:.2e}")
    assert err <= 3e-5


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
def test_fast_vjp_reverse_rk_step_matches_plain(solver):
    disc, orc, X = _fp32_disc(3, 6, 3)
    Q = periodic_euler_state(X, 3, 1.0, seed=12, amp=0.15)
    dt = 0.2 * orc.max_dt(Q, 1.0)
    S = 3 if solver is F.SSPRK33 else 5
    stages = [disc.to_device(Q)] + [disc.new_state() for _ in range(S - 1)]
    u_next = disc.new_state()
    A.rk_step_record(disc, solver, dt, stages, u_next)
    ybar = np.random.default_rng(2).standard_normal(Q.shape)
    out = {}
    for plain in (0, 1):
        disc.set_option(1, plain)
        ub = disc.to_device(ybar)
        A.rk_step_vjp(disc, solver, dt, stages, ub)
        disc.check()
        out[plain] = disc.to_host(ub).astype(np.float64)
    disc.set_option(1, 0)
    err = _rel(out[0], out[1])
    print(f"{solver.__name__ if hasattr(solver, '__name__') else solver}: fast vs plain reverse step {err:.2e}")
    assert err <= 3e-5
