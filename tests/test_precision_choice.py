"""Why the batched DWT contracts in FP32 on MFMA and not in split-BF16x3.

Emulates, on the CPU, the contraction C = D_w x over the 2B = 128 beta nodes
for Wigner tables of the oracle (weighted, as the forward transform uses
them) and 16 right-hand sides, accumulating in FP32 the way the MFMA does:
  fp32     products of FP32 operands;
  bf16x3   a = a_hi + a_lo (bf16 each), a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi;
  bf16x6   a three-way split with the six products of order <= 2^-16.
The path's error bound is 1e-6 relative L2: FP32 and bf16x6 meet it, bf16x3
(2^-16-level truncation of every product) does not."""
import numpy as np

import oracle_bind as O


def bf16(x):
    """Round-to-nearest-even FP32 -> bfloat16, returned as FP32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def contract(D, x, mode):
    R, N = D.shape
    D32, x32 = D.astype(np.float32), x.astype(np.float32)
    if mode != "fp32":
        dh, xh = bf16(D32), bf16(x32)
        dl, xl = bf16(D32 - dh), bf16(x32 - xh)
        dm, xm = bf16(D32 - dh - dl), bf16(x32 - xh - xl)
    acc = np.zeros((R, x.shape[1]), np.float32)
    for k in range(N):
        if mode == "fp32":
            p = np.outer(D32[:, k], x32[k])
        else:
            p = np.outer(dh[:, k], xh[k]) + np.outer(dh[:, k], xl[k]) + np.outer(dl[:, k], xh[k])
            if mode == "bf16x6":
                p = p + np.outer(dl[:, k], xl[k]) + np.outer(dh[:, k], xm[k]) + np.outer(dm[:, k], xh[k])
        acc += p.astype(np.float32)
    return acc


def test_fp32_meets_the_bound_and_bf16x3_does_not():
    rng = np.random.default_rng(1)
    B = 64
    worst = {"fp32": 0.0, "bf16x3": 0.0, "bf16x6": 0.0}
    for M, Mp in [(0, 0), (3, 1), (20, 0), (40, 33), (60, 2)]:
        D = O.wigner_d(B, M, Mp, weighted=True)
        x = rng.standard_normal((2 * B, 16))
        ref = D @ x
        for mode in worst:
            e = np.linalg.norm(contract(D, x, mode) - ref) / np.linalg.norm(ref)
            worst[mode] = max(worst[mode], e)
    assert worst["fp32"] < 5e-7
    assert worst["bf16x6"] < 5e-7
    assert worst["bf16x3"] > 2e-6  # ~4.6e-6: outside the 1e-6 bound
