"""Process-level behaviour of the Python package on a GPU."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_LIBRARY_FIRST = """
import sys
sys.path.insert(0, %r)
import numpy as np
import pfsoftlib_amd as P
out = P.layer_wise_DFT2_grid3D(np.ones((2, 4, 4), complex))
assert abs(out[0, 0, 0] - 16) < 1e-12
import torch
assert torch.cuda.is_available(), "torch lost the GPU after the library initialised HIP"
maps = open("/proc/self/maps").read()
runtimes = sorted(set(l.split()[-1] for l in maps.splitlines() if "libamdhip64" in l))
assert len(runtimes) == 1, runtimes
print("ok", runtimes[0])
"""


def test_library_used_before_torch_keeps_one_hip_runtime():
    # a fresh interpreter, so that nothing has loaded torch before the library
    r = subprocess.run([sys.executable, "-c", _LIBRARY_FIRST % ROOT], capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")


@pytest.mark.parametrize("width", [8, 16])
@pytest.mark.parametrize("nontemporal", [False, True])
def test_copy_kernel_used_for_counter_calibration(width, nontemporal):
    # pfs_copy_dev: the 8- and 16-byte copy kernels whose FETCH_SIZE / WRITE_SIZE
    # calibrate the byte counters must move exactly the bytes asked for
    import torch
    import pfsoftlib_amd as P
    ctx = P.Context(0)
    try:
        n = (1 << 20) + 48  # bytes, a multiple of 16 that is not a multiple of the grid stride
        src = torch.randint(0, 255, (n + 64,), dtype=torch.uint8, device="cuda")
        dst = torch.zeros(n + 64, dtype=torch.uint8, device="cuda")
        ctx.copy_dev(dst, src, n, nontemporal=nontemporal, width=width)
        torch.cuda.synchronize()
        assert torch.equal(dst[:n], src[:n])
        assert not dst[n:].any()  # nothing past the end
        with pytest.raises(ValueError):
            ctx.copy_dev(dst, src, n, width=4)
    finally:
        ctx.close()
