"""Pins the CPU oracle against closed forms, brute force and invariants.

The reference (PFSOFTlib) ships no golden vectors or known-answer tests: its
only check is the IDSOFT->DSOFT round trip at 1e-11 absolute per component
(examples/example.cpp:45-74), which test_roundtrip_reference_check restates.
"""
import numpy as np
import pytest

import closed_forms as CF
import oracle_bind as O


@pytest.mark.parametrize("B", [1, 2, 4, 8, 32, 128])
def test_weights_sum_and_symmetry(B):
    w = O.quadrature_weights(B)
    assert abs(w.sum() - 2.0) < 1e-14
    assert np.array_equal(w, w[::-1])


@pytest.mark.parametrize("B", [2, 5, 16])
def test_weights_match_defining_sum(B):
    assert np.abs(O.quadrature_weights(B) - CF.weights_closed(B)).max() < 1e-15


@pytest.mark.parametrize("B", [4, 16])
def test_weights_integrate_legendre_exactly(B):
    # Chebyshev-node quadrature is exact for sum_k w_k P_n(cos beta_k) = 2 delta_n0, n < 2B
    w = O.quadrature_weights(B)
    x = np.cos(np.pi * (2 * np.arange(2 * B) + 1) / (4 * B))
    for n in range(2 * B):
        P = np.polynomial.legendre.Legendre.basis(n)(x)
        assert abs(np.dot(w, P) - (2.0 if n == 0 else 0.0)) < 1e-13


@pytest.mark.parametrize("M,Mp", [(0, 0), (1, 0), (3, 3), (5, 2), (7, 1), (6, 5)])
def test_wigner_table_matches_explicit_formula(M, Mp):
    B = 8
    D = O.wigner_d(B, M, Mp)
    ref = np.array([CF.dtilde_table(B, M + i, M, Mp) for i in range(B - M)])
    # the reference builds the table with base sign -1 (fn_dwt.hpp:124-125)
    assert np.abs(D + ref).max() < 1e-13


def test_weighted_table_is_table_times_weights():
    B = 16
    w = O.quadrature_weights(B)
    for M, Mp in [(0, 0), (9, 4), (15, 15)]:
        D = O.wigner_d(B, M, Mp)
        Dw = O.wigner_d(B, M, Mp, weighted=True)
        assert np.abs(Dw - D * w[None, :]).max() < 1e-13


@pytest.mark.parametrize("B,M,Mp", [(32, 0, 0), (32, 7, 3), (128, 0, 0), (128, 40, 17), (128, 100, 99)])
def test_wigner_discrete_orthonormality(B, M, Mp):
    D = O.wigner_d(B, M, Mp)
    Dw = O.wigner_d(B, M, Mp, weighted=True)
    G = D @ Dw.T
    assert np.abs(G - np.eye(G.shape[0])).max() < 1e-12


@pytest.mark.parametrize("shape", [(3, 8, 8), (2, 16, 16), (2, 12, 12), (1, 6, 10)])
def test_fft2_layers_matches_numpy(shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    fwd = O.fft2_layers(a, -1)
    assert np.abs(fwd - np.fft.fft2(a, axes=(1, 2))).max() < 1e-12
    bwd = O.fft2_layers(a, +1)
    assert np.abs(bwd - np.fft.ifft2(a, axes=(1, 2)) * shape[1] * shape[2]).max() < 1e-12


def test_fft2_layers_brute_force_dft():
    rng = np.random.default_rng(2)
    n0, n1 = 4, 6
    a = rng.standard_normal((1, n0, n1)) + 1j * rng.standard_normal((1, n0, n1))
    out = O.fft2_layers(a, -1)[0]
    for k0 in range(n0):
        for k1 in range(n1):
            s = sum(a[0, i0, i1] * np.exp(-2j * np.pi * (k0 * i0 / n0 + k1 * i1 / n1))
                    for i0 in range(n0) for i1 in range(n1))
            assert abs(out[k0, k1] - s) < 1e-12


@pytest.mark.parametrize("B", [1, 2, 3, 4])
def test_inverse_matches_brute_force_synthesis(B):
    c = O.seeded_coeffs(B, 99)
    g = O.idsoft(B, c)
    gb = CF.brute_inverse(B, lambda l, M, Mp: c[O.coeff_index(l, M, Mp)])
    assert np.linalg.norm(g - gb) / np.linalg.norm(gb) < 1e-14


@pytest.mark.parametrize("B", [2, 3, 4])
def test_forward_matches_brute_force_analysis(B):
    # arbitrary (not band-limited) grid
    rng = np.random.default_rng(B)
    N = 2 * B
    g = rng.uniform(-1, 1, N ** 3) + 1j * rng.uniform(-1, 1, N ** 3)
    c = O.dsoft(B, g)
    cb = CF.brute_forward(B, g, O.coeff_index, O.coeff_count(B))
    assert np.linalg.norm(c - cb) / np.linalg.norm(cb) < 1e-14


def test_unit_coefficient_known_answer():
    # a single coefficient synthesises the closed-form grid
    B, N = 6, 12
    for (l, M, Mp) in [(0, 0, 0), (3, -2, 1), (5, 5, -4), (4, -3, -3)]:
        c = np.zeros(O.coeff_count(B), complex)
        c[O.coeff_index(l, M, Mp)] = 1.0
        g = O.idsoft(B, c).reshape(N, N, N)
        ang = 2 * np.pi * np.arange(N) / N
        dv = CF.dtilde_table(B, l, M, Mp)
        ref = dv[:, None, None] * np.exp(1j * M * ang)[None, :, None] * np.exp(1j * Mp * ang)[None, None, :] / (2 * np.pi)
        assert np.abs(g - ref).max() < 1e-14
        back = O.dsoft(B, g)
        assert np.abs(back - c).max() < 1e-13


@pytest.mark.parametrize("B", [2, 8, 19, 20, 32])
def test_roundtrip_reference_check(B):
    # examples/example.cpp:45-74: every |re|, |im| difference <= 1e-11
    c = O.seeded_coeffs(B, 7)
    r = O.dsoft(B, O.idsoft(B, c))
    assert np.abs(r.real - c.real).max() <= 1e-11
    assert np.abs(r.imag - c.imag).max() <= 1e-11
    assert np.linalg.norm(r - c) / np.linalg.norm(c) < 1e-13


def test_serial_equals_threaded():
    B = 24
    c = O.seeded_coeffs(B, 3)
    g1 = O.idsoft(B, c, threads=1)
    g4 = O.idsoft(B, c, threads=4)
    assert np.array_equal(g1, g4)
    assert np.array_equal(O.dsoft(B, g1, threads=1), O.dsoft(B, g1, threads=4))


def test_inverse_reads_stale_nyquist_entries():
    # the inverse DWT never writes row B or column B; the inverse FFT reads what
    # the caller's grid held there (src/fn_idsoft.cpp:289, grid3D_cx.hpp:400)
    B, N = 4, 8
    c = O.seeded_coeffs(B, 5)
    fresh = O.idsoft(B, c)
    stale = np.zeros((N, N, N), complex)
    stale[:, B, :] = 1.0 + 2.0j
    stale[:, :, B] = -0.5j
    got = O.idsoft(B, c, grid=stale.reshape(-1))
    assert not np.allclose(got, fresh)
    # linearity: the difference is the inverse FFT of the stale entries alone
    nyq = np.zeros((N, N, N), complex)
    nyq[:, B, :] = stale[:, B, :]
    nyq[:, :, B] = stale[:, :, B]
    diff = np.fft.ifft2(nyq, axes=(1, 2)).reshape(-1)
    assert np.abs((got - fresh) - diff).max() < 1e-14


def test_seeded_generator_fill_order_and_range():
    B = 5
    c = O.seeded_coeffs(B, 42)
    assert c.size == B * (4 * B * B - 1) // 3
    for l in range(B):
        blk = [c[O.coeff_index(l, M, Mp)] for M in range(-l, l + 1) for Mp in range(-l, l + 1)]
        blk = np.array(blk)
        assert np.all(np.abs(blk.real) <= 1.0 / (l + 1)) and np.all(np.abs(blk.imag) <= 1.0 / (l + 1))
    # first draw belongs to (0,0,0).re: SplitMix64(seed=42) first output
    z = (42 + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    z ^= z >> 31
    assert c[0].real == 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0
