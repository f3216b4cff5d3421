"""The device transforms against the committed brute-force fixtures
(tests/golden/: direct SO(3) synthesis and analysis from the closed forms,
independent of the oracle).  Bars: 1e-11 relative L2 for the FP64 path,
1e-6 for the complex64 batched path."""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "brute_B*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def ctx():
    import pfsoftlib_amd as P
    c = P.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_fp64_path_matches_brute_force_fixture(ctx, path):
    f = load(path)
    B = int(f["B"])
    assert rel(ctx.idsoft(B, f["coeffs"]), f["grid_from_coeffs"]) < 1e-11
    assert rel(ctx.dsoft(B, f["grid"]), f["coeffs_from_grid"]) < 1e-11
    assert np.abs(ctx.quadrature_weights(B) - f["weights"]).max() < 1e-15


@pytest.mark.parametrize("path", [p for p in FIXTURES if p.endswith("brute_B8.npz")], ids=os.path.basename)
def test_batched_path_matches_brute_force_fixture(ctx, path):
    import torch
    f = load(path)
    B, N = int(f["B"]), 2 * int(f["B"])
    s = torch.cuda.current_stream()
    c = torch.from_numpy(f["coeffs"].astype(np.complex64)).cuda()
    g = torch.empty(N ** 3, dtype=torch.complex64, device="cuda")
    ctx.idsoft_batch_f32(B, 1, c, g, stream=s)
    gi = torch.from_numpy(f["grid"].astype(np.complex64)).cuda()
    r = torch.empty_like(c)
    ctx.dsoft_batch_f32(B, 1, gi, r, stream=s)
    torch.cuda.synchronize()
    assert rel(g.cpu().numpy().astype(np.complex128), f["grid_from_coeffs"]) < 1e-6
    assert rel(r.cpu().numpy().astype(np.complex128), f["coeffs_from_grid"]) < 1e-6
