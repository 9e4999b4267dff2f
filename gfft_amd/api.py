"""Python binding of the C ABI in include/gfftx.h (ctypes over the in-tree libgfftx.so).

Mirrors GFFT's plan-and-call surface: a ``Plan`` is what ``CreateTransformObject`` returns
(src/gfftgen.h:222-231) and ``fft(data)`` / ``fft(src, dst)`` are the in-place and
out-of-place calls (src/gfftpolicy.h:163-172, :203-212).  Arguments may be torch CUDA
tensors (device pointers; the call is enqueued on the current torch stream) or numpy
arrays (host pointers; the call is synchronous, as GFFT's is).

There is no fallback: if libgfftx.so is missing this module raises on import of ``lib()``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GFFTX_LIBRARY", os.path.join(HERE, "libgfftx.so"))

# IDs as in GFFT (src/gfftpolicy.h:79,95,111,127 / :226,245,263,283 / :144,181)
DOUBLE, FLOAT, COMPLEX_DOUBLE, COMPLEX_FLOAT = 0, 1, 2, 3
DFT, IDFT, RDFT, IRDFT = 0, 1, 2, 3
IN_PLACE, OUT_OF_PLACE = 0, 1

FLAG_R2C_BINS = 0x1
FLAG_SPLIT_MFMA = 0x2
FLAG_NO_AUTOTUNE = 0x4
FLAG_GRAPH = 0x8
FLAG_PROFILE = 0x10

SUCCESS, ERR_INVALID_ARG, ERR_UNSUPPORTED, ERR_ALLOC, ERR_DEVICE, ERR_ALIAS, ERR_LAUNCH = range(7)

EXPORTED = [
    "gfftx_plan_create", "gfftx_plan_create_batched", "gfftx_exec_inplace", "gfftx_exec_oop",
    "gfftx_plan_destroy", "gfftx_plan_describe", "gfftx_plan_profile", "gfftx_ragged_create", "gfftx_ragged_exec",
    "gfftx_ragged_destroy", "gfftx_ragged_describe", "gfftx_ragged_profile", "gfftx_digit_reverse",
    "gfftx_stream_synchronize", "gfftx_strerror", "gfftx_last_error", "gfftx_exec_batch",
    "gfftx_exec_inplace_noalloc",
]

_lib = None


class GfftxError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        L = lib()
        msg = L.gfftx_strerror(code).decode()
        detail = L.gfftx_last_error().decode()
        super().__init__(f"{where}: {msg} [{code}] {detail}")


class UnsupportedTransform(GfftxError):
    """GFFT raises TransformFactoryError::Exception for a transform that was not generated
    (src/gfftgen.h:133-145); this is its counterpart."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (python -m gfft_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, ci, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.gfftx_plan_create.argtypes = [pp, u64, ci, ci, ci, ci, u32]
        L.gfftx_plan_create_batched.argtypes = [pp, u64, u64, ci, ci, ci, ci, u32]
        L.gfftx_exec_inplace.argtypes = [vp, vp, vp]
        L.gfftx_exec_oop.argtypes = [vp, vp, vp, vp]
        L.gfftx_plan_destroy.argtypes = [vp]
        L.gfftx_plan_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.gfftx_plan_profile.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64),
                                         ctypes.POINTER(ctypes.c_char_p), ci, ctypes.POINTER(ci)]
        L.gfftx_ragged_create.argtypes = [pp, ctypes.POINTER(u64), u64, ci, ci, ci, u32]
        L.gfftx_ragged_exec.argtypes = [vp, vp, vp]
        L.gfftx_ragged_destroy.argtypes = [vp]
        L.gfftx_ragged_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.gfftx_ragged_profile.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64)]
        L.gfftx_digit_reverse.argtypes = [vp, vp, u64, u32, ci, vp]
        L.gfftx_stream_synchronize.argtypes = [vp]
        L.gfftx_exec_batch.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_size_t, vp]
        L.gfftx_exec_inplace_noalloc.argtypes = [vp, vp, vp]
        L.gfftx_strerror.argtypes = [ci]
        L.gfftx_strerror.restype = ctypes.c_char_p
        L.gfftx_last_error.restype = ctypes.c_char_p
        for name in EXPORTED:
            if name not in ("gfftx_strerror", "gfftx_last_error"):
                getattr(L, name).restype = ci
        _lib = L
    return _lib


def _check(rc: int, where: str):
    if rc != SUCCESS:
        if rc == ERR_UNSUPPORTED:
            raise UnsupportedTransform(rc, where)
        raise GfftxError(rc, where)


def _ptr(x):
    """Device pointer of a torch tensor, host pointer of a numpy array, or an int."""
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return x.data_ptr()
    raise TypeError(f"unsupported buffer type {type(x)}")


def _scalar_bytes(x):
    """Bytes of one real scalar of a float/complex torch tensor or numpy array, else None."""
    if isinstance(x, np.ndarray):
        if x.dtype.kind == "f":
            return x.dtype.itemsize
        if x.dtype.kind == "c":
            return x.dtype.itemsize // 2
        return None
    dt = getattr(x, "dtype", None)
    if dt is None or not (dt.is_floating_point or dt.is_complex):
        return None
    es = x.element_size()
    return es // 2 if dt.is_complex else es


def _check_buffer(x, need: int, fp64: bool, device: int, what: str):
    """Rejects a buffer the plan would read or write past, or whose element type or device
    does not match the plan (raw integer pointers are passed through unchecked)."""
    if isinstance(x, int):
        return
    sb = _scalar_bytes(x)
    want = 8 if fp64 else 4
    if sb is None:
        raise ValueError(f"{what}: float or complex data expected, got {getattr(x, 'dtype', type(x))}")
    if sb != want:
        raise ValueError(f"{what}: the plan computes in {'FP64' if fp64 else 'FP32'} "
                         f"({want}-byte scalars), the buffer holds {sb}-byte scalars")
    nbytes = x.nbytes if isinstance(x, np.ndarray) else x.numel() * x.element_size()
    if nbytes < need:
        raise ValueError(f"{what}: buffer of {nbytes} bytes, the plan needs {need}")
    if hasattr(x, "is_cuda") and x.is_cuda and x.device.index != device:
        raise ValueError(f"{what}: tensor on cuda:{x.device.index}, plan on device {device}")


def _stream(x, stream):
    if stream is not None:
        return stream if isinstance(stream, int) else stream.cuda_stream
    if hasattr(x, "is_cuda") and x.is_cuda:
        import torch

        return torch.cuda.current_stream(x.device).cuda_stream
    return None


def plan_bytes(n: int, vtype: int, ttype: int, flags: int = 0, batch: int = 1):
    """(source, destination, in-place) bytes a plan reads or writes, the library's own rule:
    dist complex elements per batch item (n, or n+1 for the bins layout of a real transform);
    the real side of an out-of-place real transform holds n per item, its last item ending
    at n complex elements."""
    real = ttype in (RDFT, IRDFT)
    eb = 16 if vtype in (DOUBLE, COMPLEX_DOUBLE) else 8
    dist = n + (1 if real and (flags & FLAG_R2C_BINS) else 0)
    full = dist * batch * eb
    real_side = (dist * (batch - 1) + n) * eb
    return (real_side if ttype == RDFT else full, real_side if ttype == IRDFT else full, full)


class Plan:
    """One planned transform: GFFT's Transform<N, VType, Type, Dim, HIP<dev>, Place>."""

    def __init__(self, n: int, vtype: int = COMPLEX_DOUBLE, ttype: int = DFT,
                 place: int = OUT_OF_PLACE, device: int = 0, flags: int = 0, batch: int = 1):
        self.n, self.vtype, self.ttype, self.place = n, vtype, ttype, place
        self.batch, self.flags, self.device = batch, flags, device
        h = ctypes.c_void_p()
        if batch == 1:
            rc = lib().gfftx_plan_create(ctypes.byref(h), n, vtype, ttype, place, device, flags)
        else:
            rc = lib().gfftx_plan_create_batched(ctypes.byref(h), n, batch, vtype, ttype, place,
                                                 device, flags)
        _check(rc, f"plan_create(n={n}, vtype={vtype}, ttype={ttype}, place={place}, batch={batch})")
        self._h = h
        self.fp64 = vtype in (DOUBLE, COMPLEX_DOUBLE)
        self.in_bytes, self.out_bytes, self.inplace_bytes = plan_bytes(n, vtype, ttype, flags, batch)

    def fft(self, data, dst=None, stream=None):
        """fft(data) in place, or fft(src, dst) out of place."""
        if dst is None:
            _check_buffer(data, self.inplace_bytes, self.fp64, self.device, "fft(data)")
            rc = lib().gfftx_exec_inplace(self._h, _ptr(data), _stream(data, stream))
            _check(rc, "exec_inplace")
        else:
            _check_buffer(data, self.in_bytes, self.fp64, self.device, "fft(src, dst): src")
            _check_buffer(dst, self.out_bytes, self.fp64, self.device, "fft(src, dst): dst")
            rc = lib().gfftx_exec_oop(self._h, _ptr(data), _ptr(dst), _stream(data, stream))
            _check(rc, "exec_oop")
        return data if dst is None else dst

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(8192)
        _check(lib().gfftx_plan_describe(self._h, buf, len(buf)), "describe")
        return buf.value.decode()

    def profile(self):
        """[(kernel name, total ms, launches)] since the last call (needs FLAG_PROFILE)."""
        mx = 16
        tot = (ctypes.c_double * mx)()
        cnt = (ctypes.c_uint64 * mx)()
        names = (ctypes.c_char_p * mx)()
        ns = ctypes.c_int()
        _check(lib().gfftx_plan_profile(self._h, tot, cnt, names, mx, ctypes.byref(ns)), "profile")
        return [(names[i].decode(), tot[i], int(cnt[i])) for i in range(min(ns.value, mx))]

    def close(self):
        if getattr(self, "_h", None):
            lib().gfftx_plan_destroy(self._h)
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


class RaggedPlan:
    """Many transforms of individual lengths, back to back in one buffer, one launch."""

    def __init__(self, lengths, vtype: int = COMPLEX_FLOAT, ttype: int = DFT, device: int = 0,
                 flags: int = 0):
        arr = (ctypes.c_uint64 * len(lengths))(*[int(v) for v in lengths])
        h = ctypes.c_void_p()
        _check(lib().gfftx_ragged_create(ctypes.byref(h), arr, len(lengths), vtype, ttype, device,
                                         flags), "ragged_create")
        self._h = h
        self.lengths = list(int(v) for v in lengths)
        self.device = device
        self.fp64 = vtype in (DOUBLE, COMPLEX_DOUBLE)
        self.nbytes = sum(self.lengths) * (16 if self.fp64 else 8)

    def fft(self, data, stream=None):
        _check_buffer(data, self.nbytes, self.fp64, self.device, "ragged fft(data)")
        _check(lib().gfftx_ragged_exec(self._h, _ptr(data), _stream(data, stream)), "ragged_exec")
        return data

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().gfftx_ragged_describe(self._h, buf, len(buf)), "ragged_describe")
        return buf.value.decode()

    def profile(self):
        """(total ms, launches) since the last call (needs FLAG_PROFILE)."""
        tot = ctypes.c_double()
        cnt = ctypes.c_uint64()
        _check(lib().gfftx_ragged_profile(self._h, ctypes.byref(tot), ctypes.byref(cnt)), "ragged_profile")
        return tot.value, int(cnt.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().gfftx_ragged_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def exec_batch(plans, buffers, stream=None):
    """gfftx_exec_batch: plans[i] transforms buffers[i] in place, in order, in one call."""
    if len(plans) != len(buffers):
        raise ValueError("one buffer per plan")
    for p, b in zip(plans, buffers):
        _check_buffer(b, p.inplace_bytes, p.fp64, p.device, "exec_batch")
    n = len(plans)
    hs = (ctypes.c_void_p * n)(*[p._h.value if isinstance(p._h, ctypes.c_void_p) else p._h for p in plans])
    ds = (ctypes.c_void_p * n)(*[_ptr(b) for b in buffers])
    _check(lib().gfftx_exec_batch(hs, ds, n, _stream(buffers[0], stream) if n else None), "exec_batch")
    return buffers


def digit_reverse(src, dst, radix: int, digits: int, vtype: int, stream=None):
    _check(lib().gfftx_digit_reverse(_ptr(src), _ptr(dst), radix, digits, vtype,
                                     _stream(src, stream)), "digit_reverse")
    return dst
