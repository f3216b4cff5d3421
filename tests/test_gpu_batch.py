"""GPU parity of the batched mixed-precision path against the FP64 oracle.

Bar: relative L2 <= 1e-6 per signal (FP32 FFT and contraction, FP64-computed
Wigner tables, weights and twiddles)."""
import numpy as np
import pytest

import oracle_bind as O

pytestmark = pytest.mark.gpu

TOL = 1e-6


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


@pytest.mark.parametrize("B,nsig,chunk", [(8, 5, 64), (16, 7, 3), (32, 9, 4), (64, 5, 2), (128, 2, 64)])
def test_batch_inverse_then_forward(ctx, torch, B, nsig, chunk):
    N = 2 * B
    n = O.coeff_count(B)
    s = torch.cuda.current_stream()
    ctx.tune("batch_chunk_signals", chunk)
    cs = [O.seeded_coeffs(B, 500 + i) for i in range(nsig)]
    cdev = torch.from_numpy(np.concatenate(cs).astype(np.complex64)).cuda()
    gdev = torch.empty(nsig * N ** 3, dtype=torch.complex64, device="cuda")
    ctx.idsoft_batch_f32(B, nsig, cdev, gdev, stream=s)
    back = torch.empty_like(cdev)
    ctx.dsoft_batch_f32(B, nsig, gdev, back, stream=s)
    torch.cuda.synchronize()
    ctx.tune("batch_chunk_signals", 64)
    g = gdev.cpu().numpy().reshape(nsig, -1)
    r = back.cpu().numpy().reshape(nsig, -1)
    for i in range(nsig):
        gref = O.idsoft(B, cs[i])
        assert rel(g[i].astype(np.complex128), gref) < TOL, ("inverse", i)
        # forward on the same complex64 grid, against the FP64 oracle
        cref = O.dsoft(B, g[i].astype(np.complex128))
        assert rel(r[i].astype(np.complex128), cref) < TOL, ("forward", i)
        assert rel(r[i].astype(np.complex128), cs[i]) < 2 * TOL, ("roundtrip", i)


def test_batch_signal_is_same_alone_and_in_batch(ctx, torch):
    # a signal's result does not depend on its neighbours in the batch
    B, nsig = 32, 6
    N = 2 * B
    s = torch.cuda.current_stream()
    cs = np.concatenate([O.seeded_coeffs(B, 900 + i) for i in range(nsig)]).astype(np.complex64)
    cdev = torch.from_numpy(cs).cuda()
    g_all = torch.empty(nsig * N ** 3, dtype=torch.complex64, device="cuda")
    ctx.idsoft_batch_f32(B, nsig, cdev, g_all, stream=s)
    n = O.coeff_count(B)
    g_one = torch.empty(N ** 3, dtype=torch.complex64, device="cuda")
    ctx.idsoft_batch_f32(B, 1, cdev[3 * n:4 * n], g_one, stream=s)
    torch.cuda.synchronize()
    assert torch.equal(g_all[3 * N ** 3:4 * N ** 3], g_one)


def test_batch_rejects_unsupported_bandwidth(ctx, torch):
    import pfsoftlib_amd as P
    x = torch.zeros(8, dtype=torch.complex64, device="cuda")
    with pytest.raises(P.PfsError):
        ctx.dsoft_batch_f32(12, 1, x, x)


@pytest.mark.parametrize("B,nsig,chunk", [(16, 3, 1), (32, 11, 5), (64, 3, 64)])
def test_batch_interleaved_and_direct_coefficients_agree(ctx, torch, B, nsig, chunk):
    # the contraction on signal-interleaved coefficients (with transposes) and on
    # the caller's per-signal layout sum in the same order: results are identical
    N = 2 * B
    s = torch.cuda.current_stream()
    cs = np.concatenate([O.seeded_coeffs(B, 40 + i) for i in range(nsig)]).astype(np.complex64)
    cdev = torch.from_numpy(cs).cuda()
    out = {}
    ctx.tune("batch_chunk_signals", chunk)
    for mode in (0, 1):
        ctx.tune("batch_interleave", mode)
        g = torch.empty(nsig * N ** 3, dtype=torch.complex64, device="cuda")
        ctx.idsoft_batch_f32(B, nsig, cdev, g, stream=s)
        c = torch.empty_like(cdev)
        ctx.dsoft_batch_f32(B, nsig, g, c, stream=s)
        out[mode] = (g, c)
    torch.cuda.synchronize()
    ctx.tune("batch_interleave", 1)
    ctx.tune("batch_chunk_signals", 64)
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])
    gref = O.idsoft(B, cs[:O.coeff_count(B)].astype(np.complex128))
    assert rel(out[1][0][:N ** 3].cpu().numpy().astype(np.complex128), gref) < TOL


@pytest.mark.parametrize("B,nsig", [(16, 5), (64, 3), (128, 2)])
def test_every_contraction_shape_gives_identical_results(ctx, torch, B, nsig):
    # waves per workgroup x signals per wave only move work between waves and
    # registers: every valid shape gives the same bits
    N = 2 * B
    s = torch.cuda.current_stream()
    cs = np.concatenate([O.seeded_coeffs(B, 60 + i) for i in range(nsig)]).astype(np.complex64)
    cdev = torch.from_numpy(cs).cuda()
    outs = {}
    for w in (1, 2, 4):
        for sp in (1, 2, 4):
            # (a shape a direction cannot take -- four signals per wave at 2B = 256,
            # or for the inverse at 2B = 128 -- falls back to the default there)
            ctx.tune("batch_force_shape", 10 * w + sp)
            g = torch.empty(nsig * N ** 3, dtype=torch.complex64, device="cuda")
            ctx.idsoft_batch_f32(B, nsig, cdev, g, stream=s)
            c = torch.empty_like(cdev)
            ctx.dsoft_batch_f32(B, nsig, g, c, stream=s)
            outs[(w, sp)] = (g, c)
    ctx.tune("batch_force_shape", 0)
    torch.cuda.synchronize()
    ref = outs[(4, 2)]
    for k, (g, c) in outs.items():
        assert torch.equal(g, ref[0]) and torch.equal(c, ref[1]), k


def test_autotuner_picks_valid_shapes(ctx, torch):
    ctx.autotune_batch(32, 16)
    for fwd in (True, False):
        for n in (1, 2, 3, 4, 8, 16, 40):
            w, sp = ctx.batch_shape(32, fwd, n)
            assert w in (1, 2, 4) and sp in (1, 2, 4)


@pytest.mark.parametrize("B,nsig", [(64, 5), (128, 3)])
def test_chunked_batch_fft_is_bitwise_invariant(ctx, torch, B, nsig):
    # the slice FFTs in chunks of signals x slices do the same arithmetic per line
    N = 2 * B
    s = torch.cuda.current_stream()
    cs = np.concatenate([O.seeded_coeffs(B, 80 + i) for i in range(nsig)]).astype(np.complex64)
    cdev = torch.from_numpy(cs).cuda()
    outs = []
    ctx.tune("batch_fft_onepass", 0)  # the chunk knobs steer the two-pass kernels (the default)
    for S, K in [(0, 0), (2, 32), (3, 64), (1, 0), (0, 64)]:
        ctx.tune("batch_fft_signals", S)
        ctx.tune("batch_fft_slices", K)
        g = torch.empty(nsig * N ** 3, dtype=torch.complex64, device="cuda")
        ctx.idsoft_batch_f32(B, nsig, cdev, g, stream=s)
        c = torch.empty_like(cdev)
        ctx.dsoft_batch_f32(B, nsig, g, c, stream=s)
        outs.append((g, c))
    ctx.tune("batch_fft_signals", 0)
    ctx.tune("batch_fft_slices", 0)
    torch.cuda.synchronize()
    for g, c in outs[1:]:
        assert torch.equal(g, outs[0][0]) and torch.equal(c, outs[0][1])


@pytest.mark.parametrize("B,nsig,offset", [(32, 3, 0), (64, 5, 0), (128, 2, 0), (64, 3, 1)])
def test_wide_fft_passes_match_narrow(ctx, torch, B, nsig, offset):
    """The complex64 FFT passes with 16-byte accesses (two values per thread,
    tuning key batch_fft_wide = 1: every pass; the default 2 takes them for
    the forward passes of 2B <= 128) run every line's operations in the same
    order as the 8-byte kernel: grids and coefficients equal bit for bit in both
    directions at 2B = 64, 128, 256.  offset = 1 starts the grids one complex64
    value past a 16-byte boundary, which the wide kernel moves 8 bytes at a
    time."""
    N = 2 * B
    s = torch.cuda.current_stream()
    cs = np.concatenate([O.seeded_coeffs(B, 400 + i) for i in range(nsig)]).astype(np.complex64)
    cdev = torch.from_numpy(cs).cuda()
    rng = np.random.default_rng(11 + nsig)
    gin = (rng.uniform(-1, 1, nsig * N ** 3) + 1j * rng.uniform(-1, 1, nsig * N ** 3)).astype(np.complex64)
    gbig = torch.zeros(nsig * N ** 3 + 1, dtype=torch.complex64, device="cuda")
    gbig[offset:offset + nsig * N ** 3] = torch.from_numpy(gin).cuda()
    gdev = gbig[offset:offset + nsig * N ** 3]
    outs = {}
    try:
        for wide in (0, 1):
            ctx.tune("batch_fft_wide", wide)
            obig = torch.full((nsig * N ** 3 + 1,), float("nan"), dtype=torch.complex64, device="cuda")
            g = obig[offset:offset + nsig * N ** 3]
            ctx.idsoft_batch_f32(B, nsig, cdev, g, stream=s)
            c = torch.full_like(cdev, float("nan"))
            ctx.dsoft_batch_f32(B, nsig, gdev, c, stream=s)
            outs[wide] = (g.clone(), c)
    finally:
        ctx.tune("batch_fft_wide", 2)
    torch.cuda.synchronize()
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])
    assert not torch.isnan(outs[1][0]).any() and not torch.isnan(outs[1][1]).any()


@pytest.mark.parametrize("B,nsig", [(32, 3), (32, 8), (64, 5), (64, 17)])
def test_onepass_slice_fft_matches_two_pass(ctx, torch, B, nsig):
    """The one-pass slice FFT (both line passes of a slice's 2D FFT inside one
    workgroup) runs every line with the two-pass kernels' operations in the
    same order, so grids and coefficients equal the two-pass results bit for
    bit in both directions.  Signal counts whose k-blocks do not fill the
    last group of eight XCD slots (B = 32 x 3, B = 64 x 17) exercise the
    workgroups that find no slice."""
    N = 2 * B
    s = torch.cuda.current_stream()
    cs = np.concatenate([O.seeded_coeffs(B, 300 + i) for i in range(nsig)]).astype(np.complex64)
    cdev = torch.from_numpy(cs).cuda()
    # a seeded non-band-limited grid batch for the forward direction
    rng = np.random.default_rng(7 + nsig)
    gin = (rng.uniform(-1, 1, nsig * N ** 3) + 1j * rng.uniform(-1, 1, nsig * N ** 3)).astype(np.complex64)
    gdev = torch.from_numpy(gin).cuda()
    outs = {}
    for one in (0, 1):
        ctx.tune("batch_fft_onepass", one)
        g = torch.full((nsig * N ** 3,), float("nan"), dtype=torch.complex64, device="cuda")
        ctx.idsoft_batch_f32(B, nsig, cdev, g, stream=s)
        c = torch.full_like(cdev, float("nan"))
        ctx.dsoft_batch_f32(B, nsig, gdev, c, stream=s)
        outs[one] = (g, c)
    ctx.tune("batch_fft_onepass", 0)
    torch.cuda.synchronize()
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])
    assert not torch.isnan(outs[1][0]).any() and not torch.isnan(outs[1][1]).any()


@pytest.mark.parametrize("prec,bound", [(6, 1e-6), (3, 2e-5)])
@pytest.mark.parametrize("B,nsig", [(16, 5), (64, 6), (128, 2)])
def test_split_bf16_contraction_against_fp64_oracle(ctx, torch, B, nsig, prec, bound):
    """The split-bf16 contractions (batch_precision 3 and 6: three or six bf16
    MFMAs per FP32 step group, FP32 accumulation) against the FP64 oracle per
    signal and direction.  bf16 x6 keeps the 1e-6 bar of the batched path;
    bf16 x3 drops terms of order 2^-16 and is held to 2e-5."""
    N = 2 * B
    s = torch.cuda.current_stream()
    cs = [O.seeded_coeffs(B, 900 + i) for i in range(nsig)]
    cdev = torch.from_numpy(np.concatenate(cs).astype(np.complex64)).cuda()
    g = torch.empty(nsig * N ** 3, dtype=torch.complex64, device="cuda")
    back = torch.empty_like(cdev)
    ctx.tune("batch_precision", prec)
    try:
        ctx.idsoft_batch_f32(B, nsig, cdev, g, stream=s)
        ctx.dsoft_batch_f32(B, nsig, g, back, stream=s)
        torch.cuda.synchronize()
    finally:
        ctx.tune("batch_precision", 0)
    gh = g.cpu().numpy().reshape(nsig, -1)
    rh = back.cpu().numpy().reshape(nsig, -1)
    worst = 0.0
    for i in range(nsig):
        e_inv = rel(gh[i].astype(np.complex128), O.idsoft(B, cs[i]))
        e_fwd = rel(rh[i].astype(np.complex128), O.dsoft(B, gh[i].astype(np.complex128)))
        worst = max(worst, e_inv, e_fwd)
        assert e_inv < bound and e_fwd < bound, (prec, i, e_inv, e_fwd)
    print("bf16 x%d B=%d: worst rel L2 vs FP64 oracle %.3e" % (prec, B, worst))


def test_split_bf16_signal_is_same_alone_and_in_batch(ctx, torch):
    B, N, nsig = 32, 64, 7
    n = O.coeff_count(B)
    cs = np.concatenate([O.seeded_coeffs(B, 70 + i) for i in range(nsig)]).astype(np.complex64)
    cdev = torch.from_numpy(cs).cuda()
    g_all = torch.empty(nsig * N ** 3, dtype=torch.complex64, device="cuda")
    g_one = torch.empty(N ** 3, dtype=torch.complex64, device="cuda")
    ctx.tune("batch_precision", 6)
    try:
        ctx.idsoft_batch_f32(B, nsig, cdev, g_all)
        ctx.idsoft_batch_f32(B, 1, cdev[4 * n:5 * n], g_one)
        torch.cuda.synchronize()
    finally:
        ctx.tune("batch_precision", 0)
    assert torch.equal(g_all[4 * N ** 3:5 * N ** 3], g_one)


def test_batch_precision_key_rejects_other_values(ctx, torch):
    import pfsoftlib_amd as P
    with pytest.raises(P.PfsError):
        ctx.tune("batch_precision", 4)
    assert ctx.tune_get("batch_precision") == 0


MIX = [64, 16, 128, 32, 16, 64, 32, 16, 128, 32, 16]  # the serving bandwidths, shuffled


def _mix_inputs(torch, seed0):
    cs = [O.seeded_coeffs(B, seed0 + i) for i, B in enumerate(MIX)]
    cd = [torch.from_numpy(c.astype(np.complex64)).cuda() for c in cs]
    gd = [torch.full(((2 * B) ** 3,), float("nan"), dtype=torch.complex64, device="cuda") for B in MIX]
    return cs, cd, gd


def test_mixed_bandwidth_batch_against_oracle_and_single_bandwidth_path(ctx, torch):
    """pfs_{id,d}soft_batch_multi_f32_dev: every signal of a {16, 32, 64, 128}
    mix within 1e-6 of the FP64 oracle per direction, and bit-equal to the
    single-bandwidth batched entry points on the same inputs."""
    cs, cd, gd = _mix_inputs(torch, 5100)
    s = torch.cuda.current_stream()
    ctx.idsoft_batch_multi(MIX, cd, gd, stream=s)
    back = [torch.full_like(c, float("nan")) for c in cd]
    ctx.dsoft_batch_multi(MIX, gd, back, stream=s)
    torch.cuda.synchronize()
    for i, B in enumerate(MIX):
        g = gd[i].cpu().numpy().astype(np.complex128)
        e_inv = rel(g, O.idsoft(B, cs[i]))
        e_fwd = rel(back[i].cpu().numpy().astype(np.complex128), O.dsoft(B, g))
        assert e_inv < TOL and e_fwd < TOL, (i, B, e_inv, e_fwd)
    # the single-bandwidth path, one call per bandwidth, gives the same bits
    for B in sorted(set(MIX)):
        idx = [i for i, b in enumerate(MIX) if b == B]
        n, N3 = O.coeff_count(B), (2 * B) ** 3
        c_cat = torch.cat([cd[i] for i in idx])
        g_cat = torch.empty(len(idx) * N3, dtype=torch.complex64, device="cuda")
        ctx.idsoft_batch_f32(B, len(idx), c_cat, g_cat, stream=s)
        r_cat = torch.empty_like(c_cat)
        ctx.dsoft_batch_f32(B, len(idx), torch.cat([gd[i] for i in idx]), r_cat, stream=s)
        torch.cuda.synchronize()
        for j, i in enumerate(idx):
            assert torch.equal(g_cat[j * N3:(j + 1) * N3], gd[i]), ("inverse", B, i)
            assert torch.equal(r_cat[j * n:(j + 1) * n], back[i]), ("forward", B, i)


def test_mixed_bandwidth_graph_replays_match_direct_launches(ctx, torch):
    """The captured pipeline of a composition, replayed on new inputs with the
    program's own work and synchronisations in between, equals the directly
    launched sequence on every replay."""
    s = torch.cuda.current_stream()
    ref = None
    for rnd in range(4):
        cs, cd, gd = _mix_inputs(torch, 6200 + rnd)
        scratch = torch.randn(1 << 20, device="cuda")  # the program's own work in between
        scratch.mul_(2.0).sum().item()
        ctx.tune("batch_graphs", 1)
        ctx.idsoft_batch_multi(MIX, cd, gd, stream=s)
        torch.cuda.synchronize()
        direct = [torch.full_like(g, float("nan")) for g in gd]
        ctx.tune("batch_graphs", 0)
        ctx.idsoft_batch_multi(MIX, cd, direct, stream=s)
        torch.cuda.synchronize()
        ctx.tune("batch_graphs", 1)
        for g, d in zip(gd, direct):
            assert torch.equal(g, d), rnd


@pytest.mark.parametrize("key,value", [("batch_fft_onepass", 1), ("batch_fft_wide", 1), ("batch_fft_wide", 0)])
def test_mixed_bandwidth_batch_under_fft_keys(ctx, torch, key, value):
    """The mixed-bandwidth entry points (per-signal address tables, graph
    capture per composition) with the FFT kernels a tuning key selects: the
    one-pass slice kernel, the 16-byte kernel in every pass, the 8-byte
    kernel only.  Every grid and coefficient equals the default's bit for bit."""
    s = torch.cuda.current_stream()
    cs, cd, gd0 = _mix_inputs(torch, 7300)
    default = ctx.tune_get(key)
    outs = []
    try:
        for v in (default, value):
            ctx.tune(key, v)
            gd = [torch.full_like(g, float("nan")) for g in gd0]
            ctx.idsoft_batch_multi(MIX, cd, gd, stream=s)
            back = [torch.full_like(c, float("nan")) for c in cd]
            ctx.dsoft_batch_multi(MIX, gd, back, stream=s)
            outs.append((gd, back))
    finally:
        ctx.tune(key, default)
    torch.cuda.synchronize()
    for i in range(len(MIX)):
        assert torch.equal(outs[0][0][i], outs[1][0][i]), ("inverse", i)
        assert torch.equal(outs[0][1][i], outs[1][1][i]), ("forward", i)
        assert not torch.isnan(outs[1][1][i]).any()
