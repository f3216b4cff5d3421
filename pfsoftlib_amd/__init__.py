"""MI355X DSOFT / iDSOFT (discrete SO(3) Fourier transform) -- Python host side.

The compute lives in libpfsoft_amd.so (hand-written HIP for gfx950) behind the
C ABI of include/pfsoft_amd.h; this module binds it with ctypes and mirrors
the reference's C++ interface for the path:

    DSOFT(sample, fc, threads)       pfsoft::FourierTransforms::DSOFT
                                     (include/PFSOFTlib_headers/fn_fourier_transforms.hpp:75)
    IDSOFT(fc, synthesis, threads)   pfsoft::FourierTransforms::IDSOFT (:78)
    DSOFTFourierCoefficients(B)      include/PFSOFTlib_headers/dsoft_fourier_coefficients.hpp:50
    grid3D(n)                        grid3D<complex<double>> (grid3D_cx.hpp:47)

There is no CPU fallback: loading fails loudly when the library is missing.
"""
import ctypes
import os
import sys

import numpy as np

from .layout import coeff_count, coeff_index, coeff_offset, seeded_coeffs, seeded_coeffs_torch  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libpfsoft_amd.so")
_lib = None

PFS_OK = 0
PFS_F64 = 0
PFS_F32 = 1


PFS_REQ_FORWARD = 0
PFS_REQ_INVERSE = 1


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("signals", ctypes.c_uint64),
                ("pending", ctypes.c_uint64), ("max_inflight_batches", ctypes.c_int),
                ("device_bytes", ctypes.c_size_t), ("slot_bytes", ctypes.c_size_t), ("graphs", ctypes.c_uint64),
                ("tracked", ctypes.c_uint64)]


PFS_ERR_ARG = 1
PFS_ERR_TIMEOUT = 7


class PfsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pfsoft_amd error %d: %s" % (code, msg))
        self.code = code


class PfsArgumentError(PfsError, ValueError):
    """A buffer argument rejected on the host before any device work (wrong
    dtype, layout, device or size)."""

    def __init__(self, msg):
        PfsError.__init__(self, PFS_ERR_ARG, msg)


def lib():
    """The loaded C-ABI library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise ImportError("pfsoftlib_amd: %s is not built; run __graft_entry__.build()" % _LIBPATH)
        # One HIP runtime per process: PyTorch bundles its own libamdhip64
        # (SONAME libamdhip64.so.7, needed by torch as plain libamdhip64.so).
        # Loaded after this library, torch would map a second runtime next to
        # the system one and find no GPU; loaded first, it is the runtime this
        # library binds to.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(_LIBPATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
        sig = {
            "pfs_last_error": (ctypes.c_char_p, []),
            "pfs_version": (ctypes.c_char_p, []),
            "pfs_ctx_create": (ip, [ip, ctypes.c_size_t, ctypes.POINTER(vp)]),
            "pfs_ctx_destroy": (ip, [vp]),
            "pfs_plan_warm": (ip, [vp, ip]),
            "pfs_coeff_count": (ctypes.c_long, [ip]),
            "pfs_coeff_index": (ctypes.c_long, [ip, ip, ip]),
            "pfs_dsoft_f64": (ip, [vp, ip, dp, dp]),
            "pfs_idsoft_f64": (ip, [vp, ip, dp, dp]),
            "pfs_dsoft_f64_dev": (ip, [vp, ip, vp, vp, vp]),
            "pfs_idsoft_f64_dev": (ip, [vp, ip, vp, vp, vp]),
            "pfs_dsoft_f64_dev_noalloc": (ip, [vp, ip, vp, vp, vp]),
            "pfs_layer_wise_DFT2_grid3D": (None, [ip, ip, ip, dp, ip]),
            "pfs_layer_wise_IDFT2_grid3D": (None, [ip, ip, ip, dp, ip]),
            "pfs_layer_wise_dft2_dev": (ip, [vp, ip, ip, vp, ip, ctypes.c_double, vp]),
            "pfs_layer_wise_dft2": (ip, [ip, ip, ip, dp, ip, ctypes.c_double, ctypes.c_double, ip]),
            "uzl_fftw_layer_wise_DFT2_grid3D": (None, [ip, ip, ip, dp, ip]),
            "uzl_fftw_layer_wise_IDFT2_grid3D": (None, [ip, ip, ip, dp, ip]),
            "pfs_quadrature_weights": (ip, [vp, ip, dp]),
            "pfs_wigner_d_matrix": (ip, [vp, ip, ip, ip, ip, dp]),
            "pfs_stage_fft_fwd_dev": (ip, [vp, ip, vp, vp, vp]),
            "pfs_stage_dwt_fwd_dev": (ip, [vp, ip, vp, vp, vp]),
            "pfs_stage_dwt_inv_dev": (ip, [vp, ip, vp, vp, vp]),
            "pfs_stage_fft_inv_dev": (ip, [vp, ip, vp, vp, vp]),
            "pfs_dsoft_batch_f32_dev": (ip, [vp, ip, ip, vp, vp, vp]),
            "pfs_idsoft_batch_f32_dev": (ip, [vp, ip, ip, vp, vp, vp]),
            "pfs_tune_set": (ip, [vp, ctypes.c_char_p, ctypes.c_long]),
            "pfs_tune_get": (ip, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]),
            "pfs_timing_enable": (ip, [vp, ip]),
            "pfs_timing_collect": (ip, [vp, ctypes.c_char_p, ctypes.c_size_t]),
            "pfs_timing_select": (ip, [vp, ctypes.c_char_p]),
            "pfs_device_alloc": (ip, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
            "pfs_device_free": (ip, [vp, vp]),
            "pfs_memcpy_h2d": (ip, [vp, vp, vp, ctypes.c_size_t]),
            "pfs_memcpy_d2h": (ip, [vp, vp, vp, ctypes.c_size_t]),
            "pfs_synchronize": (ip, [vp, vp]),
            "pfs_copy_dev": (ip, [vp, vp, vp, ctypes.c_size_t, ip, vp]),
            "pfs_dsoft_batch_multi_f32_dev": (ip, [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(vp),
                                                  ctypes.POINTER(vp), vp]),
            "pfs_idsoft_batch_multi_f32_dev": (ip, [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(vp),
                                                   ctypes.POINTER(vp), vp]),
            "pfs_batch_autotune": (ip, [vp, ip, ip]),
            "pfs_batch_shape": (ip, [vp, ip, ip, ip, ctypes.POINTER(ip), ctypes.POINTER(ip)]),
            "pfs_server_create": (ip, [vp, ctypes.POINTER(ip), ip, ip, ip, ctypes.c_size_t, ctypes.POINTER(vp)]),
            "pfs_server_destroy": (ip, [vp]),
            "pfs_server_submit": (ip, [vp, ip, ip, vp, vp, ctypes.POINTER(ctypes.c_uint64)]),
            "pfs_server_wait": (ip, [vp, ctypes.c_uint64, ctypes.c_double, dp]),
            "pfs_server_collect": (ip, [vp, ctypes.POINTER(ctypes.c_uint64), dp, ctypes.POINTER(ip), ip,
                                        ctypes.POINTER(ip)]),
            "pfs_server_drain": (ip, [vp, ctypes.c_double]),
            "pfs_server_stats": (ip, [vp, ctypes.POINTER(ServerStats)]),
            "pfs_server_capacity": (ip, [vp, ip, ctypes.POINTER(ip)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != PFS_OK:
        raise PfsError(rc, lib().pfs_last_error().decode())


def _dptr(a):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _as_f64(a, n_complex):
    a = np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
    if a.size != n_complex:
        raise ValueError("expected %d complex values, got %d" % (n_complex, a.size))
    return a


def _stream_handle(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)  # torch.cuda.Stream


def _devptr(t):
    return t if isinstance(t, int) else int(t.data_ptr())


def _devbuf(t, what, dtype, n_complex, device):
    """Device address of a buffer argument.  A torch tensor is checked before
    its address goes to the library: complex dtype (complex128 for the FP64
    path, complex64 for the mixed-precision one), contiguous, on this
    context's device, and at least n_complex elements -- an undersized or
    mistyped buffer would otherwise make the kernels read and write out of
    its bounds.  A raw integer address is passed as given: its owner vouches
    for it."""
    if isinstance(t, int):
        return t
    import torch
    want = torch.complex128 if dtype == "c128" else torch.complex64
    if t.dtype != want:
        raise PfsArgumentError("%s: expected a %s tensor, got %s" % (what, want, t.dtype))
    if not t.is_contiguous():
        raise PfsArgumentError("%s: tensor must be contiguous" % what)
    if t.device.type != "cuda" or t.device.index != device:
        raise PfsArgumentError("%s: tensor must live on cuda:%d, got %s" % (what, device, t.device))
    if t.numel() < n_complex:
        raise PfsArgumentError("%s: needs at least %d elements, got %d" % (what, n_complex, t.numel()))
    return int(t.data_ptr())


class Context:
    """One device, its stream, the resident per-bandwidth tables and work buffers."""

    def __init__(self, device=0, hbm_budget=0):
        h = ctypes.c_void_p()
        _check(lib().pfs_ctx_create(device, hbm_budget, ctypes.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            lib().pfs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def warm(self, B):
        _check(lib().pfs_plan_warm(self._h, B))

    def tune(self, key, value):
        _check(lib().pfs_tune_set(self._h, key.encode(), int(value)))

    def tune_get(self, key):
        v = ctypes.c_long()
        _check(lib().pfs_tune_get(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def timing(self, on, only=None):
        """Per-kernel timing on or off; only: kernel names to time (None: all)."""
        _check(lib().pfs_timing_select(self._h, ",".join(only).encode() if only else None))
        _check(lib().pfs_timing_enable(self._h, 1 if on else 0))

    def timing_collect(self):
        """{kernel: (launches, total_ms)} for the launches recorded since the last call."""
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().pfs_timing_collect(self._h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().strip().splitlines():
            name, cnt, tot = line.split(",")
            out[name] = (int(cnt), float(tot))
        return out

    # ---- host-buffer transforms (numpy complex128) ----------------------
    def dsoft(self, B, grid):
        """grid: (2B)^3 complex128 in reference layout -> packed coefficients."""
        g = _as_f64(grid, (2 * B) ** 3)
        out = np.empty(coeff_count(B), dtype=np.complex128)
        _check(lib().pfs_dsoft_f64(self._h, B, _dptr(g.view(np.float64)), _dptr(out.view(np.float64))))
        return out

    def idsoft(self, B, coeffs, grid=None):
        """Packed coefficients -> (2B)^3 grid.  grid, if given, supplies the
        entries with row or column index B (as the reference's in-place call)."""
        c = _as_f64(coeffs, coeff_count(B))
        N = 2 * B
        if grid is None:
            g = np.zeros(N ** 3, dtype=np.complex128)
        else:
            g = np.array(grid, dtype=np.complex128, copy=True).reshape(-1)
            if g.size != N ** 3:
                raise ValueError("grid must hold (2B)^3 values")
        _check(lib().pfs_idsoft_f64(self._h, B, _dptr(c.view(np.float64)), _dptr(g.view(np.float64))))
        return g

    # ---- device-buffer transforms (torch tensors or raw pointers) -------
    # Tensor arguments are checked on the host (_devbuf) before any device work.
    def _grid(self, t, what, B, dtype="c128", nsig=1):
        return _devbuf(t, what, dtype, nsig * (2 * B) ** 3 if B > 0 else 0, self.device)

    def _coef(self, t, what, B, dtype="c128", nsig=1):
        return _devbuf(t, what, dtype, nsig * coeff_count(B) if B > 0 else 0, self.device)

    def dsoft_dev(self, B, grid_dev, coeffs_dev, stream=None):
        _check(lib().pfs_dsoft_f64_dev(self._h, B, self._grid(grid_dev, "grid", B),
                                       self._coef(coeffs_dev, "coeffs", B), _stream_handle(stream)))

    def idsoft_dev(self, B, coeffs_dev, grid_dev, stream=None):
        _check(lib().pfs_idsoft_f64_dev(self._h, B, self._coef(coeffs_dev, "coeffs", B),
                                        self._grid(grid_dev, "grid", B), _stream_handle(stream)))

    def dsoft_dev_noalloc(self, B, grid_dev, coeffs_dev, stream=None):
        """dsoft_dev without any device allocation (work buffers in the
        library's 64 MiB per-GPU scratch array): B <= 64, plan warmed first
        (warm(B)); concurrent calls on one GPU race."""
        _check(lib().pfs_dsoft_f64_dev_noalloc(self._h, B, self._grid(grid_dev, "grid", B),
                                               self._coef(coeffs_dev, "coeffs", B), _stream_handle(stream)))

    def dsoft_batch_f32(self, B, nsig, grids_dev, coeffs_dev, stream=None):
        """nsig complex64 grids -> nsig packed complex64 coefficient sets (mixed precision)."""
        n = max(int(nsig), 0)
        _check(lib().pfs_dsoft_batch_f32_dev(self._h, B, nsig, self._grid(grids_dev, "grids", B, "c64", n),
                                             self._coef(coeffs_dev, "coeffs", B, "c64", n), _stream_handle(stream)))

    def idsoft_batch_f32(self, B, nsig, coeffs_dev, grids_dev, stream=None):
        n = max(int(nsig), 0)
        _check(lib().pfs_idsoft_batch_f32_dev(self._h, B, nsig, self._coef(coeffs_dev, "coeffs", B, "c64", n),
                                              self._grid(grids_dev, "grids", B, "c64", n), _stream_handle(stream)))

    def _multi(self, fwd, bandwidths, ins, outs, stream):
        n = len(bandwidths)
        if len(ins) != n or len(outs) != n:
            raise PfsArgumentError("one input and one output per signal")
        bw = (ctypes.c_int * n)(*[int(b) for b in bandwidths])
        pin, pout = (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)()
        for i, (b, x, y) in enumerate(zip(bandwidths, ins, outs)):
            ng, nc = (2 * b) ** 3, coeff_count(b)
            pin[i] = _devbuf(x, "in[%d]" % i, "c64", ng if fwd else nc, self.device)
            pout[i] = _devbuf(y, "out[%d]" % i, "c64", nc if fwd else ng, self.device)
        f = lib().pfs_dsoft_batch_multi_f32_dev if fwd else lib().pfs_idsoft_batch_multi_f32_dev
        _check(f(self._h, n, bw, pin, pout, _stream_handle(stream)))

    def dsoft_batch_multi(self, bandwidths, grids, coeffs, stream=None):
        """Forward transforms of complex64 signals of mixed bandwidth: signal i
        (bandwidth bandwidths[i]) from grids[i] into coeffs[i], one call."""
        self._multi(True, bandwidths, grids, coeffs, stream)

    def idsoft_batch_multi(self, bandwidths, coeffs, grids, stream=None):
        """Inverse transforms of mixed bandwidth: coeffs[i] -> grids[i]."""
        self._multi(False, bandwidths, coeffs, grids, stream)

    def autotune_batch(self, B, max_signals):
        """Times the batched contraction shapes for 1, 2, 4, ... max_signals signals."""
        _check(lib().pfs_batch_autotune(self._h, B, max_signals))

    def batch_shape(self, B, forward, nsig):
        """(waves per workgroup, signals per wave) a batch of nsig signals uses."""
        w, s = ctypes.c_int(), ctypes.c_int()
        _check(lib().pfs_batch_shape(self._h, B, 1 if forward else 0, nsig, ctypes.byref(w), ctypes.byref(s)))
        return w.value, s.value

    def stage_fft_fwd(self, B, grid_dev, pencils_dev, stream=None):
        _check(lib().pfs_stage_fft_fwd_dev(self._h, B, self._grid(grid_dev, "grid", B),
                                           self._grid(pencils_dev, "pencils", B), _stream_handle(stream)))

    def stage_dwt_fwd(self, B, pencils_dev, coeffs_dev, stream=None):
        _check(lib().pfs_stage_dwt_fwd_dev(self._h, B, self._grid(pencils_dev, "pencils", B),
                                           self._coef(coeffs_dev, "coeffs", B), _stream_handle(stream)))

    def stage_dwt_inv(self, B, coeffs_dev, pencils_dev, stream=None):
        _check(lib().pfs_stage_dwt_inv_dev(self._h, B, self._coef(coeffs_dev, "coeffs", B),
                                           self._grid(pencils_dev, "pencils", B), _stream_handle(stream)))

    def stage_fft_inv(self, B, pencils_dev, grid_dev, stream=None):
        _check(lib().pfs_stage_fft_inv_dev(self._h, B, self._grid(pencils_dev, "pencils", B),
                                           self._grid(grid_dev, "grid", B), _stream_handle(stream)))

    def layer_wise_dft2_dev(self, n, lays, arr_dev, sign, scale=1.0, stream=None):
        ptr = _devbuf(arr_dev, "arr", "c128", max(int(n), 0) ** 2 * max(int(lays), 0), self.device)
        _check(lib().pfs_layer_wise_dft2_dev(self._h, n, lays, ptr, sign, scale, _stream_handle(stream)))

    def copy_dev(self, dst, src, nbytes, nontemporal=False, stream=None, width=16):
        """dst <- src (device addresses or tensors), 16- or 8-byte accesses, optionally non-temporal."""
        if width not in (8, 16):
            raise ValueError("copy_dev: width must be 8 or 16 bytes")
        for t, what in ((dst, "dst"), (src, "src")):
            if not isinstance(t, int) and (t.device.type != "cuda" or not t.is_contiguous()
                                           or t.numel() * t.element_size() < nbytes):
                raise ValueError("copy_dev: %s must be a contiguous device tensor of at least %d bytes" % (what, nbytes))
        flags = (1 if nontemporal else 0) | (2 if width == 8 else 0)
        _check(lib().pfs_copy_dev(self._h, _devptr(dst), _devptr(src), nbytes, flags, _stream_handle(stream)))

    def synchronize(self, stream=None):
        _check(lib().pfs_synchronize(self._h, _stream_handle(stream)))

    # ---- public Wigner helpers (fn_dwt.hpp:64,107,239) ------------------
    def quadrature_weights(self, B):
        out = np.empty(2 * B, dtype=np.float64)
        _check(lib().pfs_quadrature_weights(self._h, B, _dptr(out)))
        return out

    def wigner_d_matrix(self, B, M, Mp, weighted=False):
        R = B - max(abs(M), abs(Mp))
        out = np.empty(R * 2 * B, dtype=np.float64)
        _check(lib().pfs_wigner_d_matrix(self._h, B, M, Mp, 1 if weighted else 0, _dptr(out)))
        return out.reshape(2 * B, R).T.copy()


def layer_wise_DFT2_grid3D(arr, inverse=False):
    """In-place layer-wise 2D DFT of a complex128 array shaped (lays, cols, rows)
    (src/fftw_wrapper.cpp:47-172).  Returns the array."""
    a = np.ascontiguousarray(arr, dtype=np.complex128)
    lays, cols, rows = a.shape
    f = lib().pfs_layer_wise_IDFT2_grid3D if inverse else lib().pfs_layer_wise_DFT2_grid3D
    f(cols, rows, lays, _dptr(a.view(np.float64).reshape(-1)), 1)
    err = lib().pfs_last_error().decode()
    if err:
        raise PfsError(-1, err)
    return a


# --------------------------------------------------------------------------
# Mirror of the reference's API types and functions.

_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _warn(msg):
    # same text shape as pfsoft_cond_w_ret (cmake_aux/compiler_config.hpp.in:112-120)
    f = sys._getframe(2)
    sys.stderr.write("** [PFSOFTLib warning] %s:%d:%s(): %s **\n" % (
        os.path.basename(f.f_code.co_filename), f.f_lineno, f.f_code.co_name, msg))


class DSOFTFourierCoefficients:
    """Coefficients f^l_{M,M'} for 0 <= l < bandwidth, |M|,|M'| <= l, packed by
    degree (dsoft_fourier_coefficients.cpp:48-112).  fc(l, M, Mp) is the
    element; negative orders count from the end of the degree block.  With no
    argument the container is empty (bandwidth 0, :31-35)."""

    def __init__(self, bandwidth=0):
        self.bandwidth = int(bandwidth)
        self.mem = np.zeros(coeff_count(self.bandwidth), dtype=np.complex128)

    def _idx(self, l, M, Mp):
        if M > l or Mp > l or M < -l or Mp < -l or l < 0 or l >= self.bandwidth:
            # the reference prints and calls exit(EXIT_FAILURE) (dsoft_fourier_coefficients.cpp:82)
            raise IndexError("illegal parameter for DSOFTFourierCoefficients. Condition |M|,|Mp| <= l violated.")
        return coeff_index(l, M, Mp)

    def __call__(self, l, M, Mp):
        return self.mem[self._idx(l, M, Mp)]

    def set(self, l, M, Mp, value):
        self.mem[self._idx(l, M, Mp)] = value


class grid3D:
    """rows x cols x lays complex128 grid, element (row, col, lay) at
    rows*cols*lay + cols*col + row (grid3D_cx.hpp:293-303).  New grids hold
    `initial` (0 by default; a real initial value has imaginary part 0,
    grid3D_cx.hpp:61-70)."""

    def __init__(self, rows=0, cols=None, lays=None, initial=0):
        self.rows = int(rows)
        self.cols = int(cols if cols is not None else rows)
        self.lays = int(lays if lays is not None else rows)
        self.mem = np.full(self.rows * self.cols * self.lays, complex(initial), dtype=np.complex128)

    def layer_wise_DFT2(self, scale=1.0, threads=1):
        """Unnormalised 2D DFT (sign -1) of every slice, in place, times
        scale (grid3D_cx.hpp:309-366); on the device."""
        self._dft2(-1, scale, threads)

    def layer_wise_IDFT2(self, scale=1.0, threads=1):
        """Unnormalised inverse 2D DFT (sign +1), in place, times scale
        (grid3D_cx.hpp:370-427)."""
        self._dft2(+1, scale, threads)

    def _dft2(self, sign, scale, threads):
        if self.mem.size == 0:
            return
        sc = complex(scale)
        _check(lib().pfs_layer_wise_dft2(self.cols, self.rows, self.lays, _dptr(self.mem.view(np.float64)), sign,
                                         sc.real, sc.imag, threads))

    def __call__(self, row, col, lay):
        return self.mem[self.rows * self.cols * lay + self.cols * col + row]

    def set(self, row, col, lay, value):
        self.mem[self.rows * self.cols * lay + self.cols * col + row] = value


def DSOFT(sample, fc, threads=None, ctx=None):
    """Forward transform (src/fn_dsoft.cpp:76).  sample is not modified."""
    if sample.rows != sample.cols or sample.rows != sample.lays:
        return _warn("all DSOFT sample grid dimensions should be equal.")
    if sample.rows & 1:
        return _warn("DSOFT sample grid dimensions are not even.")
    B = sample.cols // 2
    if B != fc.bandwidth:
        return _warn("DSOFT Fourier coefficients container bandwidth does not match to sample grid bandwidth.")
    fc.mem[:] = (ctx or default_context()).dsoft(B, sample.mem)


def IDSOFT(fc, synthesis, threads=None, ctx=None):
    """Inverse transform (src/fn_idsoft.cpp:72), written into synthesis."""
    if synthesis.rows != synthesis.cols or synthesis.rows != synthesis.lays:
        return _warn("all IDSOFT synthesis grid dimensions should be equal.")
    if synthesis.rows & 1:
        return _warn("IDSOFT synthesis grid dimensions are not even.")
    B = synthesis.cols // 2
    if B != fc.bandwidth:
        return _warn("IDSOFT Fourier coefficients container bandwidth does not match to synthesis grid bandwidth.")
    synthesis.mem[:] = (ctx or default_context()).idsoft(B, fc.mem, synthesis.mem)


class Server:
    """Request server on a context (pfs_server_* in the C ABI): continuous
    batching of forward / inverse complex64 requests of mixed bandwidth on
    caller-owned device buffers (torch tensors or device addresses)."""

    def __init__(self, ctx, bandwidths, slots=4, max_batch=64, hbm_budget=0):
        self.ctx = ctx
        arr = (ctypes.c_int * len(bandwidths))(*bandwidths)
        h = ctypes.c_void_p()
        _check(lib().pfs_server_create(ctx._h, arr, len(bandwidths), slots, max_batch, hbm_budget, ctypes.byref(h)))
        self._h = h
        self._keep = {}  # ticket -> buffers, kept alive until the request is consumed

    def submit(self, kind, B, inp, out):
        # complex64 buffers: forward grid -> coefficients, inverse the reverse
        if kind in (PFS_REQ_FORWARD, PFS_REQ_INVERSE) and B > 0:
            ng, nc = (2 * B) ** 3, coeff_count(B)
            n_in, n_out = (ng, nc) if kind == PFS_REQ_FORWARD else (nc, ng)
            pin = _devbuf(inp, "in", "c64", n_in, self.ctx.device)
            pout = _devbuf(out, "out", "c64", n_out, self.ctx.device)
        else:  # the library rejects the kind or bandwidth
            pin, pout = _devptr(inp), _devptr(out)
        t = ctypes.c_uint64()
        _check(lib().pfs_server_submit(self._h, kind, B, pin, pout, ctypes.byref(t)))
        self._keep[t.value] = (inp, out)
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """Blocks until the request is done; returns its latency in seconds.
        On a timeout (PfsError with code PFS_ERR_TIMEOUT) the request is still
        queued or running: its buffers stay referenced and the ticket stays
        valid for another wait or for collect()."""
        lat = ctypes.c_double()
        rc = lib().pfs_server_wait(self._h, ticket, timeout, ctypes.byref(lat))
        if rc != PFS_ERR_TIMEOUT:
            # terminal: done or failed on the device -- the buffers are free
            self._keep.pop(ticket, None)
        _check(rc)
        return lat.value

    def collect(self, cap=4096):
        """[(ticket, latency_s, status)] of the requests finished since the last call."""
        tk = (ctypes.c_uint64 * cap)()
        lat = (ctypes.c_double * cap)()
        st = (ctypes.c_int * cap)()
        n = ctypes.c_int()
        _check(lib().pfs_server_collect(self._h, tk, lat, st, cap, ctypes.byref(n)))
        out = [(tk[i], lat[i], st[i]) for i in range(n.value)]
        for t, _, _ in out:
            self._keep.pop(t, None)
        return out

    def drain(self, timeout=-1.0):
        _check(lib().pfs_server_drain(self._h, timeout))

    def stats(self):
        s = ServerStats()
        _check(lib().pfs_server_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def capacity(self, B):
        n = ctypes.c_int()
        _check(lib().pfs_server_capacity(self._h, B, ctypes.byref(n)))
        return n.value

    def close(self):
        if self._h:
            _check(lib().pfs_server_destroy(self._h))
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Context", "PfsError", "PfsArgumentError", "PFS_ERR_TIMEOUT", "DSOFT", "IDSOFT", "DSOFTFourierCoefficients", "grid3D",
           "coeff_count", "coeff_index", "coeff_offset", "seeded_coeffs", "layer_wise_DFT2_grid3D", "lib",
           "Server", "PFS_REQ_FORWARD", "PFS_REQ_INVERSE"]
