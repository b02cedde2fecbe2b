"""The oracle reproduces its frozen regression vectors (tests/golden/oracle_small.npz)."""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402


def test_oracle_reproduces_golden_vectors():
    ref = np.load(make_golden.OUT, allow_pickle=False)
    got = make_golden.compute()
    for k in ref.files:
        a, b = np.asarray(got[k]), ref[k]
        if a.dtype.kind in "iu":
            assert np.array_equal(a, b), k
        else:
            assert np.allclose(a, b, rtol=1e-12, atol=1e-12), k
