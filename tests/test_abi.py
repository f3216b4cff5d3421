"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pfsoftlib_amd", "libpfsoft_amd.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b((?:pfs|uzl)_\w+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["pfs_dsoft_f64", "pfs_idsoft_f64", "pfs_dsoft_f64_dev", "pfs_idsoft_f64_dev",
              "pfs_layer_wise_DFT2_grid3D", "pfs_layer_wise_IDFT2_grid3D", "pfs_ctx_create", "pfs_last_error",
              "uzl_fftw_layer_wise_DFT2_grid3D", "uzl_fftw_layer_wise_IDFT2_grid3D", "pfs_layer_wise_dft2"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_pure_host_entry_points():
    import pfsoftlib_amd as P
    L = P.lib()
    assert L.pfs_version().decode().startswith("pfsoft_amd")
    for B in [1, 2, 7, 64, 128]:
        assert L.pfs_coeff_count(B) == B * (4 * B * B - 1) // 3 == P.coeff_count(B)
    for (l, M, Mp) in [(0, 0, 0), (3, -2, 1), (5, 5, -5), (7, -7, 7)]:
        assert L.pfs_coeff_index(l, M, Mp) == P.coeff_index(l, M, Mp)
    assert L.pfs_coeff_index(2, 3, 0) == -1


def test_packed_layout_is_per_degree_column_major():
    import pfsoftlib_amd as P
    # degree blocks tile [0, count) exactly; inside a block iM is fastest
    B = 6
    seen = np.zeros(P.coeff_count(B), int)
    for l in range(B):
        for M in range(-l, l + 1):
            for Mp in range(-l, l + 1):
                seen[P.coeff_index(l, M, Mp)] += 1
    assert np.all(seen == 1)
    assert P.coeff_index(3, 1, 0) == P.coeff_index(3, 0, 0) + 1
    assert P.coeff_index(3, 0, 1) == P.coeff_index(3, 0, 0) + 7
    assert P.coeff_index(3, -1, 0) == P.coeff_index(3, 0, 0) + 6


def test_generator_matches_oracle_generator():
    import oracle_bind as O
    import pfsoftlib_amd as P
    for B, pw in [(9, 1.0), (24, 1.5)]:
        assert np.array_equal(P.seeded_coeffs(B, 314, pw), O.seeded_coeffs(B, 314, pw))


def test_torch_generator_is_bit_identical():
    import torch
    import oracle_bind as O
    from pfsoftlib_amd.layout import seeded_coeffs_torch
    seeds = [0, 11, 2 ** 63 + 5]
    for B, pw in [(1, 1.0), (7, 1.0), (20, 1.5)]:
        got = seeded_coeffs_torch(B, seeds, "cpu", pw).numpy()
        for i, s in enumerate(seeds):
            assert np.array_equal(got[i].view(np.uint64), O.seeded_coeffs(B, s, pw).view(np.uint64))


def test_api_mirror_shape_checks_warn_and_return(capsys):
    import pfsoftlib_amd as P
    fc = P.DSOFTFourierCoefficients(4)
    before = fc.mem.copy()
    P.DSOFT(P.grid3D(8, 8, 6), fc)      # not cubic
    P.DSOFT(P.grid3D(7), fc)            # odd
    P.DSOFT(P.grid3D(10), fc)           # bandwidth mismatch
    P.IDSOFT(fc, P.grid3D(10))
    err = capsys.readouterr().err
    assert err.count("[PFSOFTLib warning]") == 4
    assert "dimensions should be equal" in err and "not even" in err and "does not match" in err
    assert np.array_equal(fc.mem, before)
    with pytest.raises(IndexError):
        fc(2, 3, 0)
