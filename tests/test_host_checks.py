"""Host-side argument checks and ticket bookkeeping of the Python binding
(no GPU: every call here is rejected, or answered by a stand-in library,
before any device work)."""
import ctypes

import numpy as np
import pytest
import torch

import pfsoftlib_amd as P


class _NoDevContext(P.Context):
    def close(self):  # holds no library context
        self._h = None


class _NoDevServer(P.Server):
    def close(self):  # holds no library server
        self._h = None


def fake_context(device=0):
    c = object.__new__(_NoDevContext)
    c._h = ctypes.c_void_p(1)
    c.device = device
    return c


def fake_server():
    srv = object.__new__(_NoDevServer)
    srv.ctx = fake_context()
    srv._h = ctypes.c_void_p(1)
    srv._keep = {}
    return srv


class FakeCudaTensor:
    """Stands in for a CUDA tensor: the attributes the checks read."""

    def __init__(self, n, dtype=torch.complex128, index=0, contiguous=True):
        self._n, self.dtype, self._c = n, dtype, contiguous
        self.device = torch.device("cuda", index)

    def numel(self):
        return self._n

    def is_contiguous(self):
        return self._c

    def data_ptr(self):
        return 0x1000  # only reached for a buffer that passes; its call is rejected by another argument


def test_fp64_entry_points_reject_wrong_dtype_before_any_call():
    ctx = fake_context()
    B = 4
    g = torch.zeros((2 * B) ** 3, dtype=torch.complex64)
    c = torch.zeros(P.coeff_count(B), dtype=torch.complex128)
    with pytest.raises(P.PfsArgumentError, match="complex128"):
        ctx.dsoft_dev(B, g, c)
    with pytest.raises(ValueError):
        ctx.idsoft_dev(B, c.to(torch.complex64), g.to(torch.complex128))


def test_rejects_host_tensors_and_other_devices():
    ctx = fake_context(device=1)
    B = 4
    g = torch.zeros((2 * B) ** 3, dtype=torch.complex128)  # CPU tensor
    c = torch.zeros(P.coeff_count(B), dtype=torch.complex128)
    with pytest.raises(P.PfsArgumentError, match="cuda:1"):
        ctx.dsoft_dev(B, g, c)
    with pytest.raises(P.PfsArgumentError, match="cuda:1"):
        ctx.stage_fft_fwd(B, FakeCudaTensor((2 * B) ** 3, index=0), FakeCudaTensor((2 * B) ** 3, index=0))


def test_rejects_undersized_and_strided_buffers():
    ctx = fake_context()
    B = 8
    N3, nc = (2 * B) ** 3, P.coeff_count(B)
    with pytest.raises(P.PfsArgumentError, match="at least %d" % N3):
        ctx.dsoft_dev(B, FakeCudaTensor(N3 - 1), FakeCudaTensor(nc))
    with pytest.raises(P.PfsArgumentError, match="at least %d" % nc):
        ctx.idsoft_dev(B, FakeCudaTensor(nc - 1), FakeCudaTensor(N3))
    with pytest.raises(P.PfsArgumentError, match="contiguous"):
        ctx.dsoft_dev(B, FakeCudaTensor(N3, contiguous=False), FakeCudaTensor(nc))
    # batched: nsig signals of complex64
    with pytest.raises(P.PfsArgumentError, match="at least %d" % (3 * N3)):
        ctx.idsoft_batch_f32(B, 3, FakeCudaTensor(3 * nc, torch.complex64), FakeCudaTensor(2 * N3, torch.complex64))
    with pytest.raises(P.PfsArgumentError, match="complex64"):
        ctx.dsoft_batch_f32(B, 1, FakeCudaTensor(N3), FakeCudaTensor(nc))
    with pytest.raises(P.PfsArgumentError):
        ctx.layer_wise_dft2_dev(16, 4, FakeCudaTensor(16 * 16 * 3), -1)


def test_server_submit_checks_request_buffers():
    srv = fake_server()
    B = 16
    with pytest.raises(P.PfsArgumentError, match="at least %d" % (2 * B) ** 3):
        srv.submit(P.PFS_REQ_FORWARD, B, FakeCudaTensor(B ** 3, torch.complex64),
                   FakeCudaTensor(P.coeff_count(B), torch.complex64))
    with pytest.raises(P.PfsArgumentError, match="complex64"):
        srv.submit(P.PFS_REQ_INVERSE, B, FakeCudaTensor(P.coeff_count(B)), FakeCudaTensor((2 * B) ** 3))
    assert srv._keep == {}


class StubLib:
    """A library stand-in whose pfs_server_wait returns a scripted status."""

    def __init__(self, statuses):
        self.statuses = list(statuses)

    def pfs_server_wait(self, h, ticket, timeout, lat):
        return self.statuses.pop(0)

    def pfs_last_error(self):
        return b"scripted"


def test_server_wait_keeps_buffers_of_a_timed_out_request(monkeypatch):
    stub = StubLib([P.PFS_ERR_TIMEOUT, 0, 3])
    monkeypatch.setattr(P, "_lib", stub)
    srv = fake_server()
    bufs = (np.zeros(4), np.zeros(4))
    srv._keep = {5: bufs, 6: bufs}
    with pytest.raises(P.PfsError) as e:
        srv.wait(5, timeout=0.0)
    assert e.value.code == P.PFS_ERR_TIMEOUT
    assert 5 in srv._keep                # still queued or running on the device
    srv.wait(5)                          # finished: released
    assert 5 not in srv._keep
    with pytest.raises(P.PfsError) as e:  # failed on the device: terminal, released
        srv.wait(6)
    assert e.value.code == 3 and srv._keep == {}
