"""The committed golden fixtures (tests/golden/*.npz) against the oracle, and their own
consistency.

Each fixture holds inputs and the oracle's outputs for one configuration (written by
tests/golden/make_golden.py).  Regenerating them must give the same bits: a change to the
oracle that moves any result shows up here, before it can move the GPU parity tests that
read the stored arrays.  The known answers of the reference's test suite are pinned by
test_oracle.py and test_hybrid.py from tests/golden/reference_answers.json.
"""
import json
import os
import sys

import numpy as np
import pytest

from cases import GOLDEN, golden, reference_answer

sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402  (tests/golden/make_golden.py)


# fixtures whose full regeneration takes minutes: their inputs are regenerated in full and
# a seeded sample of their outputs
SAMPLED = {"c4_gradient_8x8_n3.npz"}


@pytest.mark.parametrize("name", sorted(set(make_golden.FIXTURES) - SAMPLED))
def test_fixture_regenerates_bit_for_bit(name):
    stored, meta = golden(name)
    arrays, meta2 = make_golden.generate(name)
    assert meta == meta2
    assert set(stored) == set(arrays)
    for k in arrays:
        assert stored[k].shape == arrays[k].shape, k
        assert np.array_equal(stored[k], arrays[k]), k


def test_c4_gradient_fixture_regenerates_on_a_sample():
    stored, meta = golden("c4_gradient_8x8_n3.npz")
    assert meta["steps"] == 200 and meta["nel"] == [8, 8] and meta["n"] == 4
    make, QA, dt, wnode, rhoB, dLdu = make_golden.c4_problem()
    for k, v in (("QA", QA), ("dt", np.array(dt)), ("wnode", wnode), ("rhoB", rhoB)):
        assert np.array_equal(stored[k], v), k
    g = stored["grad"].reshape(-1, order="F")
    cols = np.random.default_rng(4).choice(g.size, 24, replace=False)
    got = make_golden.c4_gradient_columns(make, QA, dt, dLdu, cols)
    assert np.array_equal(got, g[cols])
    # the gradient is a real signal in every variable, not round-off
    G = stored["grad"]
    assert all(np.abs(G[:, v]).max() > 1e-6 for v in range(4))


def test_fixture_contents_are_physical():
    arrays, meta = golden("euler3d_n5_2x2x2.npz")
    Q = arrays["Q"]
    assert Q.shape == (8 * 216, 5) and np.all(Q[:, 0] > 0)
    # the three operators differ, and one CK2N54 step moves the state by O(dt·rhs)
    r = [arrays["rhs_" + k] for k in ("ch_lxf", "kg_lxf", "strong_lxf")]
    assert all(np.max(np.abs(a - b)) > 1e-6 for a, b in ((r[0], r[1]), (r[0], r[2])))
    dt = float(arrays["dt"])
    step = np.max(np.abs(arrays["Q_ck2n54_1"] - Q))
    assert 0.1 * dt * np.max(np.abs(r[0])) < step < 10 * dt * np.max(np.abs(r[0]))
    c1, _ = golden("c1_advection_2d_n4.npz")
    assert c1["Q0"].shape == (6400, 1) and 3.3e-3 < float(c1["dt"]) < 3.4e-3


def test_reference_answers_file():
    with open(os.path.join(GOLDEN, "reference_answers.json")) as f:
        ans = json.load(f)
    assert set(ans) >= {"advection_1d", "advection_2d", "sod_tube_1d", "shockwave_2d"}
    assert reference_answer("sod_tube_1d")["max"] == 254.90504152149907
    assert reference_answer("shockwave_2d")["max"] == 12.977466260673845
