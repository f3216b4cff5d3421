"""Independent closed forms used to pin the oracle (tests only).

Nothing here follows the reference's code: the Wigner d-function comes from
Wigner's explicit sum, the quadrature weights from their defining sum, the
transforms from direct (brute-force) summation over the SO(3) sample grid.
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40


def wigner_small_d(j, m1, m2, beta):
    """d^j_{m1 m2}(beta) by Wigner's explicit formula (high precision)."""
    s = mp.mpf(0)
    pre = mp.sqrt(mp.factorial(j + m1) * mp.factorial(j - m1) * mp.factorial(j + m2) * mp.factorial(j - m2))
    c, sn = mp.cos(beta / 2), mp.sin(beta / 2)
    for k in range(0, 2 * j + 1):
        a, b, d = j + m2 - k, m1 - m2 + k, j - m1 - k
        if a < 0 or b < 0 or d < 0:
            continue
        s += (-1) ** (m1 - m2 + k) * c ** (2 * j + m2 - m1 - 2 * k) * sn ** (m1 - m2 + 2 * k) / (
            mp.factorial(a) * mp.factorial(k) * mp.factorial(b) * mp.factorial(d))
    return pre * s


def betas(B):
    return [mp.pi * (2 * k + 1) / (4 * B) for k in range(2 * B)]


def dtilde_table(B, l, M, Mp):
    """L2-normalised sqrt((2l+1)/2) d^l_{M' M}(beta_k) on the 2B Chebyshev nodes:
    the function the transform pairs coefficient (l, M, M') with."""
    f = mp.sqrt((2 * l + 1) / mp.mpf(2))
    return np.array([float(f * wigner_small_d(l, Mp, M, b)) for b in betas(B)])


def weights_closed(B):
    """Driscoll-Healy weights w_B(j) from their defining sum, in mpmath."""
    out = []
    for j in range(2 * B):
        s = mp.mpf(0)
        for k in range(B):
            s += mp.sin((2 * j + 1) * (2 * k + 1) * mp.pi / (4 * B)) / (2 * k + 1)
        out.append(float(mp.mpf(2) / B * mp.sin(mp.pi * (2 * j + 1) / (4 * B)) * s))
    return np.array(out)


def brute_inverse(B, coeff_at):
    """f(alpha_c, beta_k, gamma_r) = 1/(2 pi) sum c^l_{M M'} dtilde^l_{M M'}(beta_k)
    e^{i (M alpha_c + M' gamma_r)}, returned in grid layout [k][c][r] flattened."""
    N = 2 * B
    ang = 2 * np.pi * np.arange(N) / N
    f = np.zeros((N, N, N), complex)
    for l in range(B):
        for M in range(-l, l + 1):
            for Mp in range(-l, l + 1):
                dv = dtilde_table(B, l, M, Mp)
                f += coeff_at(l, M, Mp) * dv[:, None, None] * np.exp(1j * M * ang)[None, :, None] * \
                    np.exp(1j * Mp * ang)[None, None, :]
    return (f / (2 * np.pi)).reshape(-1)


def brute_forward(B, grid, coeff_index, ncoeff):
    """c^l_{M M'} = pi/(2B^2) sum_k w_k dtilde(beta_k) sum_{c,r} f e^{-i(M alpha_c + M' gamma_r)}."""
    N = 2 * B
    f = np.asarray(grid).reshape(N, N, N)
    ang = 2 * np.pi * np.arange(N) / N
    w = weights_closed(B)
    out = np.zeros(ncoeff, complex)
    for l in range(B):
        for M in range(-l, l + 1):
            ea = np.exp(-1j * M * ang)
            for Mp in range(-l, l + 1):
                eg = np.exp(-1j * Mp * ang)
                proj = np.einsum("kcr,c,r->k", f, ea, eg)
                dv = dtilde_table(B, l, M, Mp)
                out[coeff_index(l, M, Mp)] = np.pi / (2 * B * B) * np.sum(w * dv * proj)
    return out
