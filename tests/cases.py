"""Shared builders for the tests: the same discretisation on the oracle and on the GPU."""
import json
import os

import numpy as np

import oracle_lib as O

GAMMA = 1.4


def random_euler_state(X, nd, seed, amp=0.1, smooth=True):
    """Seeded positive state (ndofs, nd+2): ρ ≈ 1 ± amp, |u| ≲ amp, p ≈ 1 ± amp."""
    rng = np.random.default_rng(seed)
    nn = X.shape[0]
    prim = np.zeros((nn, nd + 2))
    if smooth:
        k = rng.uniform(1, 3, size=(nd + 2, nd))
        ph = rng.uniform(0, 2 * np.pi, size=(nd + 2,))
        for v in range(nd + 2):
            prim[:, v] = np.sin(X @ k[v] + ph[v])
        prim += 0.1 * rng.standard_normal(prim.shape) * amp
    else:
        prim = rng.uniform(-1, 1, size=(nn, nd + 2))
    prim[:, 0] = 1.0 + amp * prim[:, 0]
    prim[:, 1:nd + 1] *= amp
    prim[:, nd + 1] = 1.0 + amp * prim[:, nd + 1]
    Q = np.zeros_like(prim)
    Q[:, 0] = prim[:, 0]
    u2 = prim[:, 1] * prim[:, 1]
    for k in range(1, nd):
        u2 = u2 + prim[:, 1 + k] * prim[:, 1 + k]
    for k in range(nd):
        Q[:, 1 + k] = prim[:, 0] * prim[:, 1 + k]
    Q[:, nd + 1] = prim[:, nd + 1] / (GAMMA - 1) + prim[:, 0] * u2 / 2
    return np.asfortranarray(Q)


def periodic_euler_state(X, nd, length, seed, amp=0.1, nmodes=4):
    """Seeded smooth state that is periodic on [0, length]^nd: each primitive is a sum of
    nmodes Fourier modes with integer wavenumbers 1..2 per axis and random phases."""
    rng = np.random.default_rng(seed)
    nn = X.shape[0]
    prim = np.zeros((nn, nd + 2))
    for v in range(nd + 2):
        for _ in range(nmodes):
            k = rng.integers(1, 3, size=nd) * (2 * np.pi / length)
            prim[:, v] += np.sin(X @ k + rng.uniform(0, 2 * np.pi)) / nmodes
    prim[:, 0] = 1.0 + amp * prim[:, 0]
    prim[:, 1:nd + 1] *= 3 * amp
    prim[:, nd + 1] = 1.0 + amp * prim[:, nd + 1]
    Q = np.zeros_like(prim)
    Q[:, 0] = prim[:, 0]
    u2 = np.sum(prim[:, 1:nd + 1] ** 2, axis=1)
    Q[:, 1:nd + 1] = prim[:, :1] * prim[:, 1:nd + 1]
    Q[:, nd + 1] = prim[:, nd + 1] / (GAMMA - 1) + prim[:, 0] * u2 / 2
    return np.asfortranarray(Q)


def oracle_for(nd, n, nel, start, finish, periodic, eq, eqparams, volop, tpflux, savg, sdiss,
               intensity=1.0, nodes="GLL"):
    ops = O.build_ops(n, nodes)
    return O.OracleDisc(nd, n, nel, start, finish, periodic, ops, eq, eqparams, volop, tpflux,
                        savg, sdiss, intensity)


def c2_state(X, seed=2):
    """The C2 initial state on [0,2]^2: ρ = 1 + 0.2 sin(π(x+y)) + 8 seeded Fourier modes of
    amplitude 0.01, u, v = 0.5 + 0.05 seeded smooth perturbations, p = 1."""
    rng = np.random.default_rng(seed)
    x, y = X[:, 0], X[:, 1]
    rho = 1 + 0.2 * np.sin(np.pi * (x + y))
    for _ in range(8):
        kx, ky = rng.integers(1, 4, 2)
        rho = rho + 0.01 / 8 * np.sin(np.pi * (kx * x + ky * y) + rng.uniform(0, 2 * np.pi))
    u = 0.5 + 0.05 * np.sin(np.pi * x + rng.uniform(0, 2 * np.pi)) * np.cos(np.pi * y)
    v = 0.5 + 0.05 * np.cos(np.pi * x) * np.sin(np.pi * y + rng.uniform(0, 2 * np.pi))
    p = np.ones_like(x)
    Q = np.stack([rho, rho * u, rho * v, p / (GAMMA - 1) + rho * (u * u + v * v) / 2], 1)
    return np.asfortranarray(Q)


WARP_AMP = 0.05


def warp(X, amp=WARP_AMP):
    """The C5 map x -> x + amp sin(2πx) sin(2πy) sin(2πz) (each component) on [0,1]^d."""
    s = np.ones((X.shape[0], 1))
    for k in range(X.shape[1]):
        s = s * np.sin(2 * np.pi * X[:, [k]])
    return X + amp * s


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reference_answer(name):
    """A known answer of the reference's test suite (tests/golden/reference_answers.json)."""
    with open(os.path.join(GOLDEN, "reference_answers.json")) as f:
        return json.load(f)[name]


def golden(name):
    """(arrays, meta) of a committed fixture of tests/golden (written by make_golden.py)."""
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files if k != "meta"}
        meta = json.loads(str(z["meta"]))
    return arrays, meta


def isapprox(a, b, rtol):
    """Julia's isapprox for scalars: |a - b| <= rtol * max(|a|, |b|)."""
    return abs(a - b) <= rtol * max(abs(a), abs(b))
