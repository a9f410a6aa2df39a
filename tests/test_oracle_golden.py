"""The oracle against the committed golden vectors.

The vectors in tests/golden/ were written by tests/golden/make_golden.py from
the reference's own hot-path components (compiled from its sources at one
OpenMP thread with -O2 -ffp-contract=off).  The oracle restates the same
arithmetic in the same order, so the comparison is bitwise.
"""
import json
import os

import numpy as np
import pytest

from tests.cases import oracle_sim, oracle_state

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLDEN, "cases.json")) as fh:
    CASES = json.load(fh)

KEYS = ("rho", "ux", "uy", "fx", "fy", "x", "y", "vx", "vy", "fx_node", "fy_node", "C")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixture_index_complete():
    names = {f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz")}
    assert names == set(CASES)
    assert len(CASES) >= 8


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden_bitwise(name):
    spec = CASES[name]
    sim = oracle_sim(spec["case"])
    sim.step(spec["steps"])
    got = oracle_state(sim)
    ref = load(name)
    for k in KEYS:
        if k not in ref:
            continue
        a, b = np.asarray(got[k]).ravel(), ref[k].ravel()
        assert a.shape == b.shape, k
        assert np.array_equal(a, b), (k, float(np.max(np.abs(a - b))))
    assert int(got["capped"]) == int(ref["capped"])
    # the fixtures are non-trivial: flow developed, capsules moved
    assert np.max(np.abs(ref["ux"])) > 1e-6
    if ref["x"].size:
        assert np.max(np.abs(ref["vx"])) > 0.0
