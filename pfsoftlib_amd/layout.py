"""Packed coefficient layout and the seeded synthetic-coefficient generator.

Layout: degree-l block at offset(l) = (4l^3 - l)/3, element (l, M, M') at
offset(l) + iM + (2l+1)*iM' with iX = X if X >= 0 else 2l+1+X -- the
reference's per-degree column-major (2l+1)^2 matrices laid end to end
(src/dsoft_fourier_coefficients.cpp:48-112).

Generator: SplitMix64 from a seed, u = (x >> 11) * 2^-53, value =
(2u - 1) * (l+1)^(-power), filled in the order of the reference's rand()
(include/PFSOFTlib_headers/random.hpp:105-116): l up, M = -l..l, M' = -l..l,
re then im.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def coeff_offset(l):
    return (4 * l * l * l - l) // 3


def coeff_count(B):
    return coeff_offset(B)


def coeff_index(l, M, Mp):
    n = 2 * l + 1
    iM = M if M >= 0 else n + M
    iMp = Mp if Mp >= 0 else n + Mp
    return coeff_offset(l) + iM + n * iMp


def _splitmix(states):
    z = states
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def seeded_coeffs(B, seed, power=1.0):
    """Packed complex128 coefficients for bandwidth B."""
    n = coeff_count(B)
    t = np.arange(1, 2 * n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        states = np.uint64(seed) + t * _GOLDEN
        u = (_splitmix(states) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    vals = 2.0 * u - 1.0
    out = np.empty(n, dtype=np.complex128)
    pos = 0
    for l in range(B):
        m = 2 * l + 1
        cnt = m * m
        v = vals[2 * pos: 2 * (pos + cnt)]
        scale = float((l + 1) ** (-power)) if power != 1.0 else 1.0 / (l + 1)
        blk = (v[0::2] * scale) + 1j * (v[1::2] * scale)
        # fill order: M outer, M' inner -> packed column-major (iM fastest)
        Mo = np.repeat(np.arange(-l, l + 1), m)
        Mpo = np.tile(np.arange(-l, l + 1), m)
        iM = np.where(Mo >= 0, Mo, m + Mo)
        iMp = np.where(Mpo >= 0, Mpo, m + Mpo)
        out[coeff_offset(l) + iM + m * iMp] = blk
        pos += cnt
    return out
