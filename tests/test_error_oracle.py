"""Pins of the mesh error functional oracle (oracle/error_oracle.py), CPU."""
import numpy as np

from oracle import error_oracle as E


def test_terms_nonnegative_and_refinement():
    L, kappa, alpha = 12.0, 1.0, 0.8
    prev = None
    for M in (8, 16):
        terms, _ = E.pm_error_terms(kappa, alpha, 6, 2, M, L)
        assert terms.min() >= -1e-12 * terms.max()
        q = E.pm_error_sum(kappa, alpha, 6, 2, M, L)
        if prev is not None:
            assert q < prev / 10
        prev = q


def test_single_alias_vanishes():
    terms, _ = E.pm_error_terms(1.0, 0.8, 6, 0, 8, 10.0)
    S1_scale = 1.0
    assert np.max(np.abs(terms)) <= 1e-12 * S1_scale
