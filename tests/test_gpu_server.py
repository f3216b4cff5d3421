"""GPU tests of the request server: continuous batching of mixed-bandwidth
forward / inverse requests.  Every result equals the one the same request
gives alone (bit for bit: a signal's arithmetic does not depend on its batch
neighbours) and is within 1e-6 relative L2 of the FP64 oracle."""
import threading

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


def make_requests(torch, seed, n, bandwidths):
    """n requests: (kind, B, input tensor, output tensor, FP64 reference)."""
    import pfsoftlib_amd as P
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        B = int(rng.choice(bandwidths))
        kind = int(rng.integers(0, 2))
        N = 2 * B
        c = O.seeded_coeffs(B, 7000 + 31 * seed + i).astype(np.complex64)
        if kind == P.PFS_REQ_INVERSE:
            inp = torch.from_numpy(c).cuda()
            out = torch.full((N ** 3,), float("nan"), dtype=torch.complex64, device="cuda")
            ref = O.idsoft(B, c.astype(np.complex128))
        else:
            g = O.idsoft(B, c.astype(np.complex128)).astype(np.complex64)
            inp = torch.from_numpy(g).cuda()
            out = torch.full((O.coeff_count(B),), float("nan"), dtype=torch.complex64, device="cuda")
            ref = O.dsoft(B, g.astype(np.complex128))
        reqs.append((kind, B, inp, out, ref))
    torch.cuda.synchronize()
    return reqs


def test_mixed_stream_matches_alone_and_oracle(ctx, torch):
    import pfsoftlib_amd as P
    bws = [8, 16, 32]
    reqs = make_requests(torch, 1, 40, bws)
    srv = P.Server(ctx, bws, slots=3, max_batch=8)
    tickets = [srv.submit(k, B, i, o) for (k, B, i, o, _) in reqs]
    lats = [srv.wait(t, timeout=60.0) for t in tickets]
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["pending"] == 0
    assert st["batches"] < len(reqs)  # requests were batched together
    assert all(l > 0 for l in lats)
    mixed = [o.clone() for (_, _, _, o, _) in reqs]
    # the same requests one at a time
    for (k, B, i, o, _) in reqs:
        o.fill_(float("nan"))
        torch.cuda.synchronize()
        srv.wait(srv.submit(k, B, i, o), timeout=60.0)
    srv.close()
    for (k, B, i, o, ref), m in zip(reqs, mixed):
        assert torch.equal(o, m), (k, B)
        assert rel(m.cpu().numpy().astype(np.complex128), ref) < TOL, (k, B)


def test_collect_and_concurrent_submitters(ctx, torch):
    import pfsoftlib_amd as P
    bws = [8, 16]
    reqs = make_requests(torch, 2, 24, bws)
    srv = P.Server(ctx, bws, slots=2, max_batch=4)
    got = {}
    lock = threading.Lock()

    def client(part):
        for (k, B, i, o, ref) in part:
            t = srv.submit(k, B, i, o)
            with lock:
                got[t] = (o, ref)

    th = [threading.Thread(target=client, args=(reqs[j::2],)) for j in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.drain(timeout=60.0)
    done = srv.collect()
    assert sorted(t for t, _, _ in done) == sorted(got)
    assert all(s == 0 for _, _, s in done)
    for t, lat, _ in done:
        o, ref = got[t]
        assert lat > 0
        assert rel(o.cpu().numpy().astype(np.complex128), ref) < TOL
    assert srv.stats()["max_inflight_batches"] <= 2
    srv.close()


def test_budget_limits_batch_size(ctx, torch):
    import pfsoftlib_amd as P
    B = 32
    N = 2 * B
    grid_bytes = N ** 3 * 8
    # room for the tables plus two signals per slot (two work grids per signal)
    probe = P.Server(ctx, [B], slots=1, max_batch=1)
    tables = probe.stats()["device_bytes"] - 2 * grid_bytes
    probe.close()
    budget = tables + 2 * 2 * grid_bytes * 2  # two slots of two signals
    srv = P.Server(ctx, [B], slots=2, max_batch=16, hbm_budget=budget)
    assert srv.capacity(B) == 2
    assert srv.stats()["device_bytes"] <= budget
    reqs = make_requests(torch, 3, 7, [B])
    for r in reqs:
        srv.submit(r[0], B, r[2], r[3])
    srv.drain(timeout=60.0)
    st = srv.stats()
    assert st["completed"] == 7 and st["signals"] == 7 and st["batches"] >= 4
    for (k, _, _, o, ref) in reqs:
        assert rel(o.cpu().numpy().astype(np.complex128), ref) < TOL
    srv.close()
    with pytest.raises(P.PfsError):
        P.Server(ctx, [B], slots=2, max_batch=4, hbm_budget=tables // 2)


def test_rejects_unserved_bandwidth_and_bad_kind(ctx, torch):
    import pfsoftlib_amd as P
    srv = P.Server(ctx, [8], slots=1, max_batch=2)
    x = torch.zeros(16 ** 3, dtype=torch.complex64, device="cuda")
    with pytest.raises(P.PfsError):
        srv.submit(P.PFS_REQ_FORWARD, 16, x, x)
    with pytest.raises(P.PfsError):
        srv.submit(7, 8, x, x)
    srv.close()
    with pytest.raises(P.PfsError):
        P.Server(ctx, [12], slots=1, max_batch=2)


def test_graph_replays_match_eager_batches(ctx, torch):
    # batches replayed from captured graphs, with other work and synchronisations
    # between the rounds, equal the eager batched entry points bit for bit
    import pfsoftlib_amd as P
    assert ctx is not None
    for B in (16, 32):
        N, n = 2 * B, O.coeff_count(B)
        k = 3
        c = torch.from_numpy(np.concatenate([O.seeded_coeffs(B, 400 + i) for i in range(k)]).astype(np.complex64)).cuda()
        g_eager = torch.empty(k * N ** 3, dtype=torch.complex64, device="cuda")
        c_eager = torch.empty_like(c)
        ctx.idsoft_batch_f32(B, k, c, g_eager)
        ctx.dsoft_batch_f32(B, k, g_eager, c_eager)
        torch.cuda.synchronize()
        for graphs in (1, 0):
            ctx.tune("server_graphs", graphs)
            srv = P.Server(ctx, [B], slots=1, max_batch=k)
            for rnd in range(3):
                gs = [torch.full((N ** 3,), float("nan"), dtype=torch.complex64, device="cuda") for _ in range(k)]
                cs = [torch.full((n,), float("nan"), dtype=torch.complex64, device="cuda") for _ in range(k)]
                inv = [srv.submit(P.PFS_REQ_INVERSE, B, c[i * n:(i + 1) * n], gs[i]) for i in range(k)]
                for t in inv:
                    srv.wait(t, timeout=60.0)
                fwd = [srv.submit(P.PFS_REQ_FORWARD, B, g_eager[i * N ** 3:(i + 1) * N ** 3], cs[i]) for i in range(k)]
                for t in fwd:
                    srv.wait(t, timeout=60.0)
                for i in range(k):
                    assert torch.equal(gs[i], g_eager[i * N ** 3:(i + 1) * N ** 3]), (B, graphs, rnd, i)
                    assert torch.equal(cs[i], c_eager[i * n:(i + 1) * n]), (B, graphs, rnd, i)
                # eager work of the program between replays
                ctx.idsoft_batch_f32(B, 1, c[:n], g_eager[:N ** 3])
                torch.cuda.synchronize()
            st = srv.stats()
            assert (st["graphs"] > 0) == (graphs == 1)
            srv.close()
        ctx.tune("server_graphs", 1)
