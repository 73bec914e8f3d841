"""The oracle reproduces its committed golden fixture bitwise (regression
pin; the fixture and its generator live in tests/golden/)."""
import os

import numpy as np

from tests.golden import make_golden

PATH = os.path.join(os.path.dirname(__file__), "golden", "dam_break_12.npz")


def test_oracle_reproduces_golden(oracle):
    ref = np.load(PATH)
    new = make_golden.generate()
    for k in ("x0", "v0", "rowstart", "col", "val", "rhs", "pressure", "dof", "x3", "v3"):
        assert np.array_equal(new[k], ref[k]), k
