"""The committed golden fixtures (tests/golden/step_*.npz, written by make_golden.py) pin the
oracle against drift: rebuilding a fixture's inputs (masks through the restated masker,
prescribed values, initial state) and rerunning the oracle must reproduce every stored array bit
for bit."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as mg  # noqa: E402


@pytest.mark.parametrize("name", sorted(mg.GOLDEN))
def test_oracle_reproduces_golden(name):
    fx = dict(np.load(mg.path(name), allow_pickle=False))
    fresh = mg.build(name)
    for key in ("bc_mask", "missing", "f_init", "aux"):
        np.testing.assert_array_equal(fresh[key], fx[key], err_msg=f"{name}: {key}")
    assert str(fresh["bcs"]) == str(fx["bcs"])
    np.testing.assert_array_equal(mg.run_oracle(fx), fx["f_final"], err_msg=f"{name}: f_final")
