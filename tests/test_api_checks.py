"""Buffer checks of the Python binding: a call whose buffer is too short, of the wrong
element type or on another device is refused before any pointer reaches the library.
CPU only (the checks run on numpy arrays and CPU tensors; no plan is created)."""
import numpy as np
import pytest
import torch

from gfft_amd import api


def test_plan_bytes_complex_and_batched():
    # 1024 complex FP64 in place: 16 KiB; batch 3 FP32: 3 * 1024 * 8
    assert api.plan_bytes(1024, api.COMPLEX_DOUBLE, api.DFT) == (16384, 16384, 16384)
    assert api.plan_bytes(1024, api.COMPLEX_FLOAT, api.IDFT, batch=3) == (24576,) * 3


def test_plan_bytes_real_sides():
    # RDFT n=8 (16 reals) FP32 in NR packing: both sides 64 bytes
    assert api.plan_bytes(8, api.FLOAT, api.RDFT) == (64, 64, 64)
    # bins layout: the complex side holds n+1 elements
    assert api.plan_bytes(8, api.FLOAT, api.RDFT, api.FLAG_R2C_BINS) == (64, 72, 72)
    assert api.plan_bytes(8, api.DOUBLE, api.IRDFT, api.FLAG_R2C_BINS) == (144, 128, 144)


def test_short_buffer_is_refused():
    with pytest.raises(ValueError, match="needs"):
        api._check_buffer(np.zeros(1023, np.complex128), 16384, True, 0, "x")
    with pytest.raises(ValueError, match="needs"):
        api._check_buffer(torch.zeros(100, dtype=torch.complex64), 1024, False, 0, "x")
    api._check_buffer(np.zeros(1024, np.complex128), 16384, True, 0, "x")
    api._check_buffer(np.zeros(2048, np.float64), 16384, True, 0, "x")  # interleaved reals


def test_wrong_precision_is_refused():
    with pytest.raises(ValueError, match="FP64"):
        api._check_buffer(np.zeros(4096, np.complex64), 16384, True, 0, "x")
    with pytest.raises(ValueError, match="FP32"):
        api._check_buffer(torch.zeros(4096, dtype=torch.float64), 8192, False, 0, "x")
    with pytest.raises(ValueError, match="float or complex"):
        api._check_buffer(np.zeros(4096, np.int64), 8192, True, 0, "x")


def test_raw_pointers_pass_unchecked():
    api._check_buffer(12345, 1 << 40, True, 0, "x")
