"""The oracle against the committed brute-force fixtures (tests/golden/,
written by tests/golden/make_golden.py from the closed forms: direct SO(3)
synthesis and analysis with Wigner's explicit d and the defining-sum weights)."""
import glob
import os

import numpy as np
import pytest

import oracle_bind as O

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "brute_B*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_fixtures_present():
    assert len(FIXTURES) >= 5


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_matches_brute_force_fixture(path):
    f = load(path)
    B = int(f["B"])
    assert rel(O.idsoft(B, f["coeffs"]), f["grid_from_coeffs"]) < 1e-14
    assert rel(O.dsoft(B, f["grid"]), f["coeffs_from_grid"]) < 1e-14
    assert np.abs(O.quadrature_weights(B) - f["weights"]).max() < 1e-15
