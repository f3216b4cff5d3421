"""The allocation-free forward transform (pfs_dsoft_f64_dev_noalloc) through
the C ABI: the same results as pfs_dsoft_f64_dev bit for bit, on two calls in
a row with the second at a smaller bandwidth (the shared scratch array keeps
the first call's contents), within 1e-11 of the oracle, and its limits."""
import ctypes

import numpy as np
import pytest

import oracle_bind as O

pytestmark = pytest.mark.gpu


def test_noalloc_forward_matches_allocating_entry_point():
    import torch
    import pfsoftlib_amd as P
    assert torch.cuda.is_available()
    L = P.lib()
    ctx = P.Context(0)
    s = torch.cuda.current_stream()
    for B, seed in ((64, 71), (16, 72)):
        N = 2 * B
        c = O.seeded_coeffs(B, seed)
        g = O.idsoft(B, c)
        gd = torch.from_numpy(g).cuda()
        want = torch.empty(O.coeff_count(B), dtype=torch.complex128, device="cuda")
        got = torch.full_like(want, float("nan"))
        ctx.warm(B)
        ctx.dsoft_dev(B, gd, want, stream=s)
        rc = L.pfs_dsoft_f64_dev_noalloc(ctx.handle, B, ctypes.c_void_p(gd.data_ptr()),
                                         ctypes.c_void_p(got.data_ptr()), ctypes.c_void_p(s.cuda_stream))
        assert rc == 0, L.pfs_last_error().decode()
        torch.cuda.synchronize()
        assert torch.equal(got, want), B
        e = np.linalg.norm(got.cpu().numpy() - O.dsoft(B, g)) / np.linalg.norm(c)
        print("B=%d: no-alloc forward equals the allocating one; rel L2 vs oracle %.3e" % (B, e), flush=True)
        assert e <= 1e-11
    # limits: beyond the scratch, and a bandwidth whose plan is not prepared
    big = torch.empty(1, dtype=torch.complex128, device="cuda")
    assert L.pfs_dsoft_f64_dev_noalloc(ctx.handle, 128, ctypes.c_void_p(big.data_ptr()),
                                       ctypes.c_void_p(big.data_ptr()), None) == 5
    assert "64 MiB" in L.pfs_last_error().decode()
    assert L.pfs_dsoft_f64_dev_noalloc(ctx.handle, 24, ctypes.c_void_p(big.data_ptr()),
                                       ctypes.c_void_p(big.data_ptr()), None) == 1
    assert "pfs_plan_warm" in L.pfs_last_error().decode()
    ctx.close()
