"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py from the oracle) are reproduced by the oracle
as it stands: a change to the oracle that moves any stored output fails
here before it can move the bars of the GPU parity tests."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures as MF  # noqa: E402
from oracle import reference_step as R  # noqa: E402


@pytest.mark.parametrize("name", sorted(MF.FIXTURES))
def test_oracle_reproduces_golden_fixture(name):
    d = np.load(os.path.join(HERE, "golden", name))
    mk, run = MF.FIXTURES[name]
    inputs = mk()
    for k, v in inputs.items():
        assert np.array_equal(np.asarray(v), d[k]), (name, "input", k)
    out = run(inputs)
    assert set(out) | set(inputs) == set(d.files)
    for k, v in out.items():
        e = R.rel_l2(v, d[k])
        assert e <= 1e-13, (name, k, e)
