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


def _fill_order(B, power):
    """Packed destination index and scale of every fill-order position."""
    n = coeff_count(B)
    dest = np.empty(n, dtype=np.int64)
    scale = np.empty(n, dtype=np.float64)
    pos = 0
    for l in range(B):
        m = 2 * l + 1
        cnt = m * m
        Mo = np.repeat(np.arange(-l, l + 1), m)
        Mpo = np.tile(np.arange(-l, l + 1), m)
        iM = np.where(Mo >= 0, Mo, m + Mo)
        iMp = np.where(Mpo >= 0, Mpo, m + Mpo)
        dest[pos:pos + cnt] = coeff_offset(l) + iM + m * iMp
        scale[pos:pos + cnt] = float((l + 1) ** (-power)) if power != 1.0 else 1.0 / (l + 1)
        pos += cnt
    return dest, scale


def _i64(u):
    """uint64 constant as the int64 with the same bits."""
    u = int(u)
    return u - (1 << 64) if u >= (1 << 63) else u


def seeded_coeffs_torch(B, seeds, device, power=1.0, dtype=None):
    """The generator of seeded_coeffs, vectorised over signals in torch.

    Returns a (len(seeds), coeff_count(B)) complex128 tensor (or `dtype`) whose
    row i equals seeded_coeffs(B, seeds[i], power) bit for bit.  SplitMix64 runs
    in int64 with wrapping multiplication; logical shifts are arithmetic shifts
    masked to the low bits."""
    import torch
    n = coeff_count(B)
    dest, scale = _fill_order(B, power)
    dest_t = torch.from_numpy(dest).to(device)
    scale_t = torch.from_numpy(scale).to(device)
    t = torch.arange(1, 2 * n + 1, dtype=torch.int64, device=device)
    golden, m1, m2 = _i64(_GOLDEN), _i64(_M1), _i64(_M2)

    def lsr(z, k):
        return (z >> k) & ((1 << (64 - k)) - 1)

    out = torch.empty((len(seeds), n), dtype=dtype or torch.complex128, device=device)
    for i, seed in enumerate(seeds):
        z = t * golden + _i64(seed % (1 << 64))
        z = (z ^ lsr(z, 30)) * m1
        z = (z ^ lsr(z, 27)) * m2
        z = z ^ lsr(z, 31)
        u = lsr(z, 11).to(torch.float64) * 2.0 ** -53
        v = (2.0 * u - 1.0).view(n, 2)
        row = torch.empty(n, dtype=torch.complex128, device=device)
        row[dest_t] = torch.complex(v[:, 0] * scale_t, v[:, 1] * scale_t)
        out[i] = row
    return out
