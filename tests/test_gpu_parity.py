"""GPU parity: the HIP path against the CPU oracle, through the C ABI.

Bar (FP64 plain path): relative L2 <= 1e-11 against the oracle on the same
seeded inputs, and the reference's own round-trip check (every |re|, |im|
difference <= 1e-11, examples/example.cpp:45-74).  At the largest size the
oracle is too slow for a test, so B=256 is checked through size-independent
properties (round trip, linearity).
"""
import numpy as np
import pytest

import oracle_bind as O

pytestmark = pytest.mark.gpu

TOL = 1e-11


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


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


def test_context_and_mfma_layout_selftest(ctx, torch):
    # Context creation runs the FP64 MFMA lane-layout self-test on the device
    assert ctx.handle
    p = torch.cuda.get_device_properties(0)
    print("device:", p.name, "CUs:", p.multi_processor_count, "LDS/block:",
          getattr(p, "shared_memory_per_block", None), "L2:", getattr(p, "L2_cache_size", None))


@pytest.mark.parametrize("B", [1, 2, 3, 8, 64, 128, 256])
def test_quadrature_weights_on_device(ctx, B):
    w = ctx.quadrature_weights(B)
    assert np.abs(w - O.quadrature_weights(B)).max() < 1e-15


@pytest.mark.parametrize("B,M,Mp", [(8, 0, 0), (8, 5, 2), (64, 0, 0), (64, 31, 7), (64, 63, 63),
                                    (256, 0, 0), (256, 100, 37), (256, 200, 0)])
@pytest.mark.parametrize("weighted", [False, True])
def test_wigner_table_on_device(ctx, B, M, Mp, weighted):
    d = ctx.wigner_d_matrix(B, M, Mp, weighted)
    ref = O.wigner_d(B, M, Mp, weighted)
    assert np.linalg.norm(d - ref) / np.linalg.norm(ref) < 1e-12


@pytest.mark.parametrize("bw", [2, 16, 128, 256])
def test_dwt_accuracy_roundtrip_on_device_tables(ctx, bw):
    # benchmark/benchmark_dwt_accuracy.cpp's round trip f -> d_w (d^T f) on the
    # device tables (with its :60 table mix-up corrected): identity to rounding
    rng = np.random.default_rng(bw)
    for M, Mp in [(0, 0), (bw // 2, 0), (bw // 2, bw // 2)]:
        dw = ctx.wigner_d_matrix(bw, M, Mp, True)
        dt = ctx.wigner_d_matrix(bw, M, Mp, False)
        f = rng.uniform(-1, 1, dw.shape[0]) + 1j * rng.uniform(-1, 1, dw.shape[0])
        assert rel(dw @ (dt.T @ f), f) < 1e-12


@pytest.mark.parametrize("n,lays", [(2, 2), (4, 3), (6, 4), (8, 8), (12, 5), (16, 16), (20, 3), (64, 8),
                                    (256, 4), (512, 2)])
@pytest.mark.parametrize("inverse", [False, True])
def test_layer_wise_dft2_seam(n, lays, inverse):
    import pfsoftlib_amd as P
    rng = np.random.default_rng(n * 7 + lays)
    a = rng.standard_normal((lays, n, n)) + 1j * rng.standard_normal((lays, n, n))
    got = P.layer_wise_DFT2_grid3D(a.copy(), inverse=inverse)
    ref = O.fft2_layers(a, +1 if inverse else -1)
    assert rel(got, ref) < 1e-14


@pytest.mark.parametrize("B", [4, 16, 128])
def test_fft_stage_pencil_layout(ctx, torch, B):
    N = 2 * B
    rng = np.random.default_rng(B)
    g = rng.standard_normal(N ** 3) + 1j * rng.standard_normal(N ** 3)
    gd = torch.from_numpy(g).cuda()
    pen = torch.empty_like(gd)
    s = torch.cuda.current_stream()
    ctx.stage_fft_fwd(B, gd, pen, stream=s)
    torch.cuda.synchronize()
    F = O.fft2_layers(g.reshape(N, N, N), -1)           # [k][q][p]
    ref = np.transpose(F, (2, 1, 0)).reshape(-1)         # [p][q][k]
    assert rel(pen.cpu().numpy(), ref) < 1e-14
    back = torch.empty_like(gd)
    ctx.stage_fft_inv(B, pen, back, stream=s)
    torch.cuda.synchronize()
    assert rel(back.cpu().numpy(), g) < 1e-13


def test_fft_stage_512_lines_on_slices(ctx, torch):
    """2B = 512 (B = 256): both FFT stages -- row-in/tile-out and tile-in/row-out
    512-point lines -- against numpy on three beta-slices of a random grid."""
    B, N = 256, 512
    gen = torch.Generator(device="cuda").manual_seed(512)
    s = torch.cuda.current_stream()
    g = torch.randn(N ** 3, dtype=torch.complex128, device="cuda", generator=gen)
    pen = torch.empty_like(g)
    ctx.stage_fft_fwd(B, g, pen, stream=s)
    back = torch.empty_like(g)
    ctx.stage_fft_inv(B, pen, back, stream=s)
    torch.cuda.synchronize()
    g3, p3, b3 = g.view(N, N, N), pen.view(N, N, N), back.view(N, N, N)
    for k in (0, 257, 511):
        F = np.fft.fft2(g3[k].cpu().numpy())          # [q][p] of slice k
        assert rel(p3[:, :, k].cpu().numpy(), F.T) < 1e-14
        P = p3[:, :, k].cpu().numpy().T
        assert rel(b3[k].cpu().numpy(), np.fft.ifft2(P)) < 1e-14
    del g, pen, back


@pytest.mark.parametrize("B", [1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 19, 20, 32, 33, 64])
def test_idsoft_matches_oracle(ctx, B):
    c = O.seeded_coeffs(B, 1000 + B)
    g = ctx.idsoft(B, c)
    assert rel(g, O.idsoft(B, c)) < TOL


@pytest.mark.parametrize("B", [1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 19, 20, 32, 33, 64])
def test_dsoft_matches_oracle_on_arbitrary_grid(ctx, B):
    N = 2 * B
    rng = np.random.default_rng(B)
    g = rng.uniform(-1, 1, N ** 3) + 1j * rng.uniform(-1, 1, N ** 3)
    assert rel(ctx.dsoft(B, g), O.dsoft(B, g)) < TOL


@pytest.fixture(scope="module")
def b128(ctx):
    B = 128
    c = O.seeded_coeffs(B, 2026)
    g_ref = O.idsoft(B, c)
    return B, c, g_ref, O.dsoft(B, g_ref)


def test_idsoft_b128_matches_oracle(ctx, b128):
    B, c, g_ref, _ = b128
    assert rel(ctx.idsoft(B, c), g_ref) < TOL


def test_dsoft_b128_matches_oracle(ctx, b128):
    B, c, g_ref, c_ref = b128
    got = ctx.dsoft(B, g_ref)
    assert rel(got, c_ref) < TOL
    # the reference's own pass criterion on the round trip (example.cpp:45-54)
    assert np.abs(got.real - c.real).max() <= 1e-11 and np.abs(got.imag - c.imag).max() <= 1e-11


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
def test_dwt_work_shapes_match_oracle(ctx, b128, variant):
    # every work shape of the FP64 DWT kernels (chains per lane x waves per
    # pair; 4 / 5 the node-blocked forward at three / two waves per SIMD, 6 the
    # chain-parallel forward's default shape), at B = 128 and at grids that
    # leave part of the nodes a shape covers empty (2B = 64, 80, 128)
    B, c, g_ref, c_ref = b128
    ctx.tune("dwt_variant", variant)
    try:
        assert rel(ctx.idsoft(B, c), g_ref) < TOL
        assert rel(ctx.dsoft(B, g_ref), c_ref) < TOL
        for Bs, seed in ((64, 77), (40, 78), (32, 79)):
            cs = O.seeded_coeffs(Bs, seed)
            gs = O.idsoft(Bs, cs)
            assert rel(ctx.idsoft(Bs, cs), gs) < TOL
            assert rel(ctx.dsoft(Bs, gs), O.dsoft(Bs, gs)) < TOL
    finally:
        ctx.tune("dwt_variant", 0)


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 5, 6, 7, 13, 15])
def test_inverse_dwt_kernels_match_oracle(ctx, torch, b128, kind):
    # the MFMA inverse DWT (0) and the vector-ALU one in its three work shapes
    # (1: 2 chains x 2 waves over the nodes, 2: 4 chains x 2 waves over the
    # right-hand sides, 3: 4 x 1 per pair; 5-7 the same with the
    # row broadcasts folded into the FMAs; 13 and 15 the staged kernel in
    # shapes 1 and 3; grids of 2B <= 128
    # use the shape that covers their nodes), at B = 128, at grids that leave
    # part of a shape's nodes empty (2B = 40, 80, 200), at B = 1..4, and through
    # the device route with stale Nyquist entries (copied by the same launch)
    B, c, g_ref, _ = b128
    default = ctx.tune_get("dwt_inv_kernel")
    ctx.tune("dwt_inv_kernel", kind)
    try:
        assert rel(ctx.idsoft(B, c), g_ref) < TOL
        for Bs in (1, 2, 3, 4, 20, 40, 100):
            cs = O.seeded_coeffs(Bs, 300 + Bs)
            assert rel(ctx.idsoft(Bs, cs), O.idsoft(Bs, cs)) < TOL, Bs
        Bs = 64
        N = 2 * Bs
        cs = O.seeded_coeffs(Bs, 6)
        stale = np.random.default_rng(7).standard_normal(N ** 3) + 0j
        gd = torch.from_numpy(stale.copy()).cuda()
        ctx.idsoft_dev(Bs, torch.from_numpy(cs).cuda(), gd, stream=torch.cuda.current_stream())
        torch.cuda.synchronize()
        assert rel(gd.cpu().numpy(), O.idsoft(Bs, cs, grid=stale)) < TOL
    finally:
        ctx.tune("dwt_inv_kernel", default)


@pytest.mark.parametrize("pair", [(5, 13), (7, 15)])
def test_staged_inverse_matches_valu(ctx, torch, pair):
    # the staged inverse runs the vector-ALU kernel's operations in the same
    # order, so the grids agree exactly (zeros up to their sign), at B = 128,
    # sizes that leave chains or chunk rows empty, and through the device
    # route with stale Nyquist entries; at 2B = 512 (four chains x two waves
    # for both) it runs the rescaled recurrence (d_r = P_r e_r) and agrees to
    # rounding
    base, staged = pair
    default = ctx.tune_get("dwt_inv_kernel")
    try:
        for B in (1, 3, 17, 40, 64, 100, 128, 256):
            if B == 256 and base != 7:
                continue
            cs = O.seeded_coeffs(B, 500 + B)
            out = []
            for kind in (base, staged):
                ctx.tune("dwt_inv_kernel", kind)
                out.append(ctx.idsoft(B, cs))
            if B > 128:
                assert rel(out[0], out[1]) < 1e-13, B
            else:
                assert np.array_equal(out[0], out[1]), B
        Bs = 64
        N = 2 * Bs
        cs = torch.from_numpy(O.seeded_coeffs(Bs, 8)).cuda()
        stale = np.random.default_rng(9).standard_normal(N ** 3) + 0j
        grids = []
        for kind in (base, staged):
            ctx.tune("dwt_inv_kernel", kind)
            gd = torch.from_numpy(stale.copy()).cuda()
            ctx.idsoft_dev(Bs, cs, gd, stream=torch.cuda.current_stream())
            torch.cuda.synchronize()
            grids.append(gd.cpu().numpy())
        assert np.array_equal(grids[0], grids[1])
    finally:
        ctx.tune("dwt_inv_kernel", default)


def test_dwt_pair_orders_bit_identical(ctx, b128):
    # the FP64 DWT kernels' pair launch order (per-XCD M' ranges, or runs of 8
    # consecutive M' per window of 64) only moves pairs between workgroups: the
    # results are the same bit for bit, at B = 128 and at pair counts that leave
    # the last window partial (B = 5: 15 pairs, B = 20: 210 pairs) or the XCD
    # queues uneven (B = 100: 5050 pairs)
    B, c, g_ref, c_ref = b128
    default = ctx.tune_get("dwt_pair_order")
    cases = [(B, c, g_ref)] + [(Bs, O.seeded_coeffs(Bs, 90 + Bs), None) for Bs in (5, 20, 100)]
    try:
        for Bs, cs, gs in cases:
            if gs is None:
                gs = O.idsoft(Bs, cs)
            out = {}
            for order in (0, 3):  # both directions in one order, then in the other
                ctx.tune("dwt_pair_order", order)
                out[order] = (ctx.idsoft(Bs, cs), ctx.dsoft(Bs, gs))
            assert np.array_equal(out[0][0], out[3][0]) and np.array_equal(out[0][1], out[3][1]), Bs
            assert rel(out[0][0], gs) < TOL
            assert rel(out[0][1], cs) < 1e-11
    finally:
        ctx.tune("dwt_pair_order", default)


@pytest.mark.parametrize("B", [4, 32])
def test_idsoft_reads_stale_nyquist_like_reference(ctx, B):
    N = 2 * B
    c = O.seeded_coeffs(B, 5)
    rng = np.random.default_rng(3)
    stale = rng.standard_normal(N ** 3) + 1j * rng.standard_normal(N ** 3)
    got = ctx.idsoft(B, c, grid=stale)
    ref = O.idsoft(B, c, grid=stale)
    assert rel(got, ref) < TOL


@pytest.mark.parametrize("B,variant", [(4, 0), (32, 0), (64, 1), (128, 0), (128, 1)])
def test_idsoft_dev_reads_stale_nyquist_like_reference(ctx, torch, B, variant):
    # device buffers: the inverse DWT launch itself copies the stale entries
    # (one to four waves per workgroup across these shapes)
    N = 2 * B
    c = O.seeded_coeffs(B, 6)
    rng = np.random.default_rng(4)
    stale = rng.standard_normal(N ** 3) + 1j * rng.standard_normal(N ** 3)
    gd = torch.from_numpy(stale.copy()).cuda()
    cd = torch.from_numpy(c).cuda()
    ctx.tune("dwt_variant", variant)
    try:
        ctx.idsoft_dev(B, cd, gd, stream=torch.cuda.current_stream())
        torch.cuda.synchronize()
    finally:
        ctx.tune("dwt_variant", 0)
    assert rel(gd.cpu().numpy(), O.idsoft(B, c, grid=stale)) < TOL


def test_dsoft_leaves_input_grid_untouched(ctx, torch):
    B = 16
    N = 2 * B
    g = O.idsoft(B, O.seeded_coeffs(B, 8))
    gd = torch.from_numpy(g.copy()).cuda()
    cd = torch.empty(O.coeff_count(B), dtype=torch.complex128, device="cuda")
    ctx.dsoft_dev(B, gd, cd, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    assert np.array_equal(gd.cpu().numpy(), g)


def test_device_api_on_torch_stream_matches_host_api(ctx, torch):
    B = 32
    c = O.seeded_coeffs(B, 11)
    cd = torch.from_numpy(c).cuda()
    gd = torch.zeros((2 * B) ** 3, dtype=torch.complex128, device="cuda")
    s = torch.cuda.current_stream()
    ctx.idsoft_dev(B, cd, gd, stream=s)
    back = torch.empty_like(cd)
    ctx.dsoft_dev(B, gd, back, stream=s)
    torch.cuda.synchronize()
    assert np.array_equal(gd.cpu().numpy(), ctx.idsoft(B, c))
    assert np.array_equal(back.cpu().numpy(), ctx.dsoft(B, gd.cpu().numpy()))
    assert rel(back.cpu().numpy(), c) < TOL


def test_api_mirror_roundtrip_like_reference_example(ctx):
    # examples/example.cpp:13-85 through the mirrored interface
    import pfsoftlib_amd as P
    B = 24
    sample = P.grid3D(2 * B)
    coef = P.DSOFTFourierCoefficients(B)
    rec = P.DSOFTFourierCoefficients(B)
    coef.mem[:] = O.seeded_coeffs(B, 99)
    P.IDSOFT(coef, sample, ctx=ctx)
    P.DSOFT(sample, rec, ctx=ctx)
    for l in range(B):
        for M in range(-l, l + 1):
            for Mp in range(-l, l + 1):
                d = coef(l, M, Mp) - rec(l, M, Mp)
                assert abs(d.real) <= 1e-11 and abs(d.imag) <= 1e-11


def test_b256_roundtrip_and_linearity(ctx, torch):
    # every device call runs on torch's current stream, so torch's own ops
    # (zero_, clone, arithmetic) are ordered with the transforms
    B = 256
    n = O.coeff_count(B)
    import pfsoftlib_amd as P
    s = torch.cuda.current_stream()
    c1 = torch.from_numpy(P.seeded_coeffs(B, 7, 1.5)).cuda()
    c2 = torch.from_numpy(P.seeded_coeffs(B, 8, 1.5)).cuda()
    N = 2 * B
    g = torch.zeros(N ** 3, dtype=torch.complex128, device="cuda")
    r = torch.empty(n, dtype=torch.complex128, device="cuda")
    ctx.idsoft_dev(B, c1, g, stream=s)
    ctx.dsoft_dev(B, g, r, stream=s)
    torch.cuda.synchronize()
    assert (torch.linalg.vector_norm(r - c1) / torch.linalg.vector_norm(c1)).item() < TOL
    # linearity: IDSOFT(c1 + 2 c2) = IDSOFT(c1) + 2 IDSOFT(c2)
    g1 = g.clone()
    g.zero_()
    ctx.idsoft_dev(B, c2, g, stream=s)
    g2 = g.clone()
    g.zero_()
    c3 = c1 + 2 * c2
    ctx.idsoft_dev(B, c3, g, stream=s)
    lin = g1 + 2 * g2
    torch.cuda.synchronize()
    assert (torch.linalg.vector_norm(g - lin) / torch.linalg.vector_norm(lin)).item() < 1e-13


def test_b256_inverse_kernels_agree(ctx, torch):
    # 2B = 512: the vector-ALU inverse DWT with separate broadcast moves
    # (dwt_inv_kernel 3) against the staged default (15), four chains x two
    # waves per pair for both, and its grid back through DSOFT (the
    # node-blocked forward, eight 64-node blocks per pair); both directions
    # are pinned to the oracle at B = 256 in test_gpu_configs
    B, N = 256, 512
    import pfsoftlib_amd as P
    s = torch.cuda.current_stream()
    c = torch.from_numpy(P.seeded_coeffs(B, 9, 1.5)).cuda()
    out = {}
    default = ctx.tune_get("dwt_inv_kernel")
    try:
        for kind in (3, 15):
            ctx.tune("dwt_inv_kernel", kind)
            g = torch.zeros(N ** 3, dtype=torch.complex128, device="cuda")
            ctx.idsoft_dev(B, c, g, stream=s)
            out[kind] = g
    finally:
        ctx.tune("dwt_inv_kernel", default)
    nrm = torch.linalg.vector_norm
    assert (nrm(out[15] - out[3]) / nrm(out[3])).item() < 1e-13
    r = torch.empty(O.coeff_count(B), dtype=torch.complex128, device="cuda")
    ctx.dsoft_dev(B, out[3], r, stream=s)
    torch.cuda.synchronize()
    assert (nrm(r - c) / nrm(c)).item() < TOL
    del out, r


def test_b256_stale_nyquist_on_both_routes(ctx, torch):
    # 2B = 512 (separate gather kernel, then the inverse DWT): with zero
    # coefficients the inverse is the inverse FFT of the caller's entries at
    # row B or column B alone, every other entry of the given grid ignored --
    # slice k of the output is ifft2 of slice k masked to that row and column
    B, N = 256, 512
    gen = torch.Generator(device="cuda").manual_seed(256)
    stale = torch.randn(N ** 3, dtype=torch.complex128, device="cuda", generator=gen)
    zero = torch.zeros(O.coeff_count(B), dtype=torch.complex128, device="cuda")
    s = torch.cuda.current_stream()
    dev = stale.clone()
    ctx.idsoft_dev(B, zero, dev, stream=s)
    torch.cuda.synchronize()
    host = ctx.idsoft(B, np.zeros(O.coeff_count(B), complex), grid=stale.cpu().numpy())
    st3, d3, h3 = stale.view(N, N, N), dev.view(N, N, N), host.reshape(N, N, N)
    for k in (0, 255, 511):
        a = st3[k].cpu().numpy()                      # [c][r]
        z = np.zeros_like(a)
        z[B, :] = a[B, :]
        z[:, B] = a[:, B]
        ref = np.fft.ifft2(z)
        assert rel(d3[k].cpu().numpy(), ref) < 1e-13
        assert rel(h3[k], ref) < 1e-13
    del stale, dev, host


def test_cpp_dropin_example_like_reference(ctx):
    # tests/cpp/example_roundtrip.cpp mirrors examples/example.cpp through include/pfsoft_amd.hpp
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "example_roundtrip")
    r = subprocess.run([exe, "32"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "Correct result:  Yes" in r.stdout
    assert "[PFSOFTLib warning]" in r.stderr and "not even" in r.stderr


@pytest.mark.parametrize("B", [64, 128])
def test_fft_chunking_is_bitwise_invariant(ctx, torch, B):
    # each line transform is the same whatever the slice chunking
    c = O.seeded_coeffs(B, 31)
    outs = []
    for chunk in [8, 16, 2 * B, 0]:
        ctx.tune("fft_chunk_slices", chunk)
        g = ctx.idsoft(B, c)
        outs.append((g, ctx.dsoft(B, g)))
    ctx.tune("fft_chunk_slices", 0)
    for g, r in outs[1:]:
        assert np.array_equal(g, outs[0][0]) and np.array_equal(r, outs[0][1])


def test_timing_select_times_only_the_named_kernels(ctx, torch):
    B = 16
    N = 2 * B
    c = torch.from_numpy(O.seeded_coeffs(B, 9)).cuda()
    g = torch.zeros(N ** 3, dtype=torch.complex128, device="cuda")
    back = torch.empty_like(c)
    s = torch.cuda.current_stream()
    try:
        ctx.timing(True, only=["dwt_inv", "fft_fwd_cols"])
        for _ in range(3):
            ctx.idsoft_dev(B, c, g, stream=s)
            ctx.dsoft_dev(B, g, back, stream=s)
        ctx.timing(False)
        t = ctx.timing_collect()
        assert sorted(t) == ["dwt_inv", "fft_fwd_cols"] and all(v[0] == 3 and v[1] > 0 for v in t.values())
        ctx.timing(True)  # no list: every kernel again
        ctx.idsoft_dev(B, c, g, stream=s)
        ctx.dsoft_dev(B, g, back, stream=s)
        ctx.timing(False)
        assert {"dwt_inv", "dwt_fwd", "fft_inv_cols", "fft_inv_rows", "fft_fwd_rows", "fft_fwd_cols"} <= set(ctx.timing_collect())
    finally:
        ctx.timing(False)
        ctx.timing_collect()
