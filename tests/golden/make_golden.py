"""Writes the brute-force golden fixtures of the DSOFT / iDSOFT path.

PFSOFTlib ships no golden vectors (its inputs are time-seeded; its only check
is the round trip of examples/example.cpp:45-74), so these come from the
mathematics, independently of the oracle and of the device code:

  inverse   f(alpha_c, beta_k, gamma_r) = 1/(2 pi) sum_{l,M,M'} c^l_{M M'}
            sqrt((2l+1)/2) d^l_{M' M}(beta_k) e^{i (M alpha_c + M' gamma_r)}
  forward   c^l_{M M'} = pi/(2B^2) sum_k w_B(k) sqrt((2l+1)/2) d^l_{M' M}(beta_k)
            sum_{c,r} f e^{-i (M alpha_c + M' gamma_r)}

with Wigner's explicit sum for d (40-digit mpmath) and the Driscoll-Healy
weights from their defining sum (tests/closed_forms.py), in the reference's
grid layout [k][c][r] and packed coefficient order.  Inputs: the seeded
coefficients of pfsoftlib_amd.layout (SplitMix64, uniform[-1,1]/(l+1)) and a
seeded grid that is not band-limited.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import closed_forms as CF  # noqa: E402
from pfsoftlib_amd.layout import coeff_count, coeff_index, seeded_coeffs  # noqa: E402

BANDWIDTHS = (1, 2, 3, 4, 6, 8)


def make(B):
    N = 2 * B
    c = seeded_coeffs(B, 2024 + B)
    g_inv = CF.brute_inverse(B, lambda l, M, Mp: c[coeff_index(l, M, Mp)])
    rng = np.random.default_rng(7000 + B)
    g_in = rng.uniform(-1, 1, N ** 3) + 1j * rng.uniform(-1, 1, N ** 3)
    c_fwd = CF.brute_forward(B, g_in, coeff_index, coeff_count(B))
    w = CF.weights_closed(B)
    return dict(B=np.array(B), coeffs=c, grid_from_coeffs=g_inv, grid=g_in, coeffs_from_grid=c_fwd, weights=w)


if __name__ == "__main__":
    for B in BANDWIDTHS:
        d = make(B)
        path = os.path.join(HERE, "brute_B%d.npz" % B)
        np.savez_compressed(path, **d)
        print("wrote", path, {k: v.shape for k, v in d.items()})
