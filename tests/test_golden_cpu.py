"""The committed golden fixtures (tests/golden/fdtd3d_golden.npz) are the
oracle's float32 output on the cases of tests/golden/cases.py: re-derived
here bit for bit, so a change to the oracle cannot go unnoticed."""

import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)

import cases as K  # noqa: E402
import make_fixtures as M  # noqa: E402


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "fdtd3d_golden.npz"))


@pytest.mark.parametrize("name", list(K.CASES))
def test_oracle_reproduces_golden(golden, name):
    rec = M.record(name, M.run_case(name))
    for key, val in rec.items():
        np.testing.assert_array_equal(val, golden[key], err_msg=key)
    # the fixtures hold real signal
    assert max(float(golden[f"{name}/{n}/norm"].max()) for n in M.NAMES) > 0
