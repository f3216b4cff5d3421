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
