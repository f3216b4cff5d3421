"""GPU tests at the workloads' real sizes, against the FP64 oracle.

- Plain FP64 at B = 256 (512^3 complex128 grid, 2.1 GB): DSOFT and IDSOFT
  against the oracle on one seed, bar 1e-11 relative L2 (src/fn_dsoft.cpp:76,
  src/fn_idsoft.cpp:72), plus the reference's round-trip check
  (examples/example.cpp:45-54).
- The 512-signal batch at B = 64 (complex64, 8.6 GB of grids): every signal's
  round trip within 2e-6, and a sample of signals against the FP64 oracle
  within 1e-6 per direction.
- The serving bandwidth set {16, 32, 64, 128} with a 64 GB budget: B = 64 and
  B = 128 forward and inverse requests through the server, each equal bit for
  bit alone and in a mix with smaller requests, and within 1e-6 of the oracle.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


@pytest.fixture(scope="module")
def ctx(torch):
    import pfsoftlib_amd as P
    c = P.Context(0)
    yield c
    c.close()


def test_b256_dsoft_and_idsoft_match_oracle(ctx, torch):
    B, N = 256, 512
    c = O.seeded_coeffs(B, 256256, 1.5)
    say("B=256: oracle IDSOFT ...")
    g_ref = O.idsoft(B, c)
    say("B=256: device IDSOFT")
    cd = torch.from_numpy(c).cuda()
    gd = torch.zeros(N ** 3, dtype=torch.complex128, device="cuda")
    ctx.idsoft_dev(B, cd, gd, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    e_inv = rel(gd.cpu().numpy(), g_ref)
    say("B=256: IDSOFT rel L2 %.3e" % e_inv)
    assert e_inv <= 1e-11
    # forward on the oracle's grid (the same input on both sides)
    say("B=256: oracle DSOFT ...")
    r_ref = O.dsoft(B, g_ref)
    gd.copy_(torch.from_numpy(g_ref))
    rd = torch.empty_like(cd)
    ctx.dsoft_dev(B, gd, rd, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    r = rd.cpu().numpy()
    e_fwd = rel(r, r_ref)
    say("B=256: DSOFT rel L2 %.3e, round trip max abs %.3e" % (e_fwd, np.abs(r - c).max()))
    assert e_fwd <= 1e-11
    # the reference's own criterion: every |re|, |im| difference <= 1e-11
    assert np.abs(r.real - c.real).max() <= 1e-11 and np.abs(r.imag - c.imag).max() <= 1e-11


def test_full_512_signal_batch_at_b64(ctx, torch):
    from pfsoftlib_amd.layout import seeded_coeffs_torch
    B, N, nsig = 64, 128, 512
    n = O.coeff_count(B)
    seeds = [4000 + i for i in range(nsig)]
    say("512 x B=64: generating coefficients on the device")
    c = seeded_coeffs_torch(B, seeds, "cuda", 1.0, torch.complex64).reshape(-1)
    g = torch.empty(nsig * N ** 3, dtype=torch.complex64, device="cuda")
    s = torch.cuda.current_stream()
    ctx.idsoft_batch_f32(B, nsig, c, g, stream=s)
    back = torch.empty_like(c)
    ctx.dsoft_batch_f32(B, nsig, g, back, stream=s)
    torch.cuda.synchronize()
    cs, bs = c.reshape(nsig, n), back.reshape(nsig, n)
    rt = (torch.linalg.vector_norm((bs - cs).to(torch.complex128), dim=1) /
          torch.linalg.vector_norm(cs.to(torch.complex128), dim=1)).cpu().numpy()
    say("512 x B=64: round trip rel L2 max %.3e median %.3e" % (rt.max(), np.median(rt)))
    assert np.all(np.isfinite(rt)) and rt.max() <= 2e-6
    # a sample of signals against the FP64 oracle, each direction
    gs = g.reshape(nsig, N ** 3)
    for i in (0, 1, 255, 511):
        ci = O.seeded_coeffs(B, seeds[i])
        e_inv = rel(gs[i].cpu().numpy().astype(np.complex128), O.idsoft(B, ci))
        gi = gs[i].cpu().numpy().astype(np.complex128)
        e_fwd = rel(bs[i].cpu().numpy().astype(np.complex128), O.dsoft(B, gi))
        say("  signal %d: inverse %.3e forward %.3e" % (i, e_inv, e_fwd))
        assert e_inv <= 1e-6 and e_fwd <= 1e-6


def test_serving_bandwidth_set_b64_b128_alone_and_in_mix(ctx, torch):
    import pfsoftlib_amd as P
    bws = [16, 32, 64, 128]
    plan = [(P.PFS_REQ_FORWARD, 128), (P.PFS_REQ_INVERSE, 128), (P.PFS_REQ_FORWARD, 64),
            (P.PFS_REQ_INVERSE, 64), (P.PFS_REQ_INVERSE, 128), (P.PFS_REQ_FORWARD, 64)]
    plan += [(k, b) for b in (16, 32) for k in (P.PFS_REQ_FORWARD, P.PFS_REQ_INVERSE) for _ in range(3)]
    reqs = []
    say("serving set: %d requests, oracle references ..." % len(plan))
    for j, (kind, B) in enumerate(plan):
        N = 2 * B
        c = O.seeded_coeffs(B, 9100 + j).astype(np.complex64)
        if kind == P.PFS_REQ_INVERSE:
            inp = torch.from_numpy(c).cuda()
            out = torch.full((N ** 3,), float("nan"), dtype=torch.complex64, device="cuda")
            ref = O.idsoft(B, c.astype(np.complex128))
        else:
            gin = O.idsoft(B, c.astype(np.complex128)).astype(np.complex64)
            inp = torch.from_numpy(gin).cuda()
            out = torch.full((O.coeff_count(B),), float("nan"), dtype=torch.complex64, device="cuda")
            ref = O.dsoft(B, gin.astype(np.complex128))
        reqs.append((kind, B, inp, out, ref))
    torch.cuda.synchronize()
    # the serving configuration: four slots, batches of up to 64, a 64 GB budget
    # (shapes are not autotuned here: every shape gives the same bits)
    ctx.tune("server_autotune", 0)
    try:
        srv = P.Server(ctx, bws, slots=4, max_batch=64, hbm_budget=64 << 30)
    finally:
        ctx.tune("server_autotune", 1)
    say("server: capacity per batch", {B: srv.capacity(B) for B in bws}, "device bytes",
        srv.stats()["device_bytes"])
    assert srv.stats()["device_bytes"] <= 64 << 30
    order = np.random.default_rng(5).permutation(len(reqs))
    tickets = {int(i): srv.submit(reqs[i][0], reqs[i][1], reqs[i][2], reqs[i][3]) for i in order}
    for i in order:
        srv.wait(tickets[int(i)], timeout=120.0)
    mixed = [r[3].clone() for r in reqs]
    for (k, B, inp, out, _) in reqs:
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        srv.wait(srv.submit(k, B, inp, out), timeout=120.0)
    st = srv.stats()
    srv.close()
    assert st["tracked"] == 0
    for (k, B, _, out, ref), m in zip(reqs, mixed):
        assert torch.equal(out, m), (k, B)
        e = rel(m.cpu().numpy().astype(np.complex128), ref)
        if B >= 64:
            say("  %s B=%d rel L2 %.3e" % ("forward" if k == P.PFS_REQ_FORWARD else "inverse", B, e))
        assert e <= 1e-6, (k, B)


def test_server_wait_only_client_keeps_no_tickets(ctx, torch):
    import pfsoftlib_amd as P
    B, N = 8, 16
    srv = P.Server(ctx, [B], slots=2, max_batch=4)
    c = torch.from_numpy(O.seeded_coeffs(B, 3).astype(np.complex64)).cuda()
    g = torch.empty(N ** 3, dtype=torch.complex64, device="cuda")
    for _ in range(50):
        srv.wait(srv.submit(P.PFS_REQ_INVERSE, B, c, g), timeout=60.0)
    st = srv.stats()
    assert st["completed"] == 50 and st["tracked"] == 0
    assert srv.collect() == []
    srv.close()


def test_cpp_dropin_api_program():
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_api")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")


@pytest.mark.parametrize("inverse", [False, True])
def test_reference_seam_symbols_through_ctypes(inverse):
    import ctypes
    import pfsoftlib_amd as P
    L = P.lib()
    f = L.uzl_fftw_layer_wise_IDFT2_grid3D if inverse else L.uzl_fftw_layer_wise_DFT2_grid3D
    rng = np.random.default_rng(9 + inverse)
    lays, n = 5, 32
    a = rng.standard_normal((lays, n, n)) + 1j * rng.standard_normal((lays, n, n))
    ref = O.fft2_layers(a, +1 if inverse else -1)
    f(n, n, lays, a.view(np.float64).reshape(-1).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1)
    assert rel(a, ref) < 1e-14


@pytest.mark.parametrize("scale", [0.25, 0.5 - 0.75j])
def test_grid3d_layer_wise_members_with_scale(scale):
    import pfsoftlib_amd as P
    n = 16
    g = P.grid3D(n)
    rng = np.random.default_rng(3)
    g.mem[:] = rng.standard_normal(n ** 3) + 1j * rng.standard_normal(n ** 3)
    x = g.mem.copy()
    g.layer_wise_DFT2(scale)
    ref = O.fft2_layers(x.reshape(n, n, n), -1).reshape(-1) * scale
    assert rel(g.mem, ref) < 1e-14
    g.layer_wise_IDFT2(1.0 / (n * n * scale))
    assert rel(g.mem, x) < 1e-14
