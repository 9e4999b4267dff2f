"""The C ABI library loads and exports every symbol include/gfftx.h declares; argument and
device errors come back as status codes.  CPU only (no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT
from gfft_amd import api

HEADER = os.path.join(ROOT, "include", "gfftx.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(gfftx_[a-z_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert set(names) == set(api.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = api.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gfftx_\w+)", out))
    assert set(declared_functions()) <= exported


def test_strerror_texts():
    L = api.lib()
    assert L.gfftx_strerror(0) == b"success"
    assert b"not compiled" in L.gfftx_strerror(api.ERR_UNSUPPORTED)
    assert L.gfftx_strerror(999) == b"unknown status"


def test_argument_errors_without_device():
    L = api.lib()
    h = ctypes.c_void_p()
    # NULL plan pointer and bad IDs are rejected before any device work
    assert L.gfftx_plan_create(None, 1024, 0, 0, 0, 0, 0) == api.ERR_INVALID_ARG
    assert L.gfftx_plan_create(ctypes.byref(h), 0, 0, 0, 0, 0, 0) == api.ERR_INVALID_ARG
    assert L.gfftx_plan_create(ctypes.byref(h), 1024, 9, 0, 0, 0, 0) == api.ERR_INVALID_ARG
    assert L.gfftx_plan_create(ctypes.byref(h), 1024, 0, 7, 0, 0, 0) == api.ERR_INVALID_ARG
    assert L.gfftx_plan_create(ctypes.byref(h), 1024, 0, 0, 5, 0, 0) == api.ERR_INVALID_ARG
    assert L.gfftx_plan_create(ctypes.byref(h), 1024, 0, 0, 0, 0, 0x100) == api.ERR_INVALID_ARG
    assert L.gfftx_exec_inplace(None, None, None) == api.ERR_INVALID_ARG
    assert L.gfftx_plan_destroy(None) == api.SUCCESS
    assert L.gfftx_digit_reverse(None, None, 2, 3, 0, None) == api.ERR_INVALID_ARG


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is visible")
def test_no_device_is_a_clean_error():
    L = api.lib()
    h = ctypes.c_void_p()
    rc = L.gfftx_plan_create(ctypes.byref(h), 1024, 0, 0, 0, 0, 0)
    assert rc == api.ERR_DEVICE
    assert b"device" in L.gfftx_last_error()
    with pytest.raises(api.GfftxError):
        api.Plan(1024)


def test_cpp_policy_header_compiles():
    """The C++ policy layer (GFFT surface + HIP<Dev>) compiles as plain host C++."""
    src = os.path.join(ROOT, "tests", "cpp", "gfft_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_header_is_plain_c(tmp_path):
    """include/gfftx.h is a C header (the boundary any FFI binds): a C99 translation unit that
    calls every entry point compiles and links against the library with gcc."""
    src = tmp_path / "use_gfftx.c"
    src.write_text(r"""
#include <stdio.h>
#include "gfftx.h"
int main(void)
{
    gfftx_plan p = 0;
    gfftx_ragged r = 0;
    uint64_t lens[2] = {256, 512};
    double ms[4];
    uint64_t cnt[4];
    const char* names[4];
    int nslots = 0;
    char buf[64];
    int rc = gfftx_plan_create(&p, 1024, GFFTX_COMPLEX_FLOAT, GFFTX_DFT, GFFTX_IN_PLACE, 0, 0);
    if (rc == GFFTX_SUCCESS) {
        (void)gfftx_plan_describe(p, buf, sizeof buf);
        (void)gfftx_plan_profile(p, ms, cnt, names, 4, &nslots);
        (void)gfftx_exec_inplace(p, 0, 0);
        (void)gfftx_exec_oop(p, 0, 0, 0);
        (void)gfftx_plan_destroy(p);
    }
    rc = gfftx_plan_create_batched(&p, 1024, 2, GFFTX_COMPLEX_DOUBLE, GFFTX_IDFT, GFFTX_OUT_OF_PLACE, 0, 0);
    if (rc == GFFTX_SUCCESS) (void)gfftx_plan_destroy(p);
    rc = gfftx_ragged_create(&r, lens, 2, GFFTX_COMPLEX_FLOAT, GFFTX_DFT, 0, 0);
    if (rc == GFFTX_SUCCESS) {
        (void)gfftx_ragged_exec(r, 0, 0);
        (void)gfftx_ragged_describe(r, buf, sizeof buf);
        (void)gfftx_ragged_profile(r, ms, cnt);
        (void)gfftx_ragged_destroy(r);
    }
    (void)gfftx_digit_reverse(0, 0, 2, 4, GFFTX_COMPLEX_FLOAT, 0);
    (void)gfftx_stream_synchronize(0);
    printf("%s | %s\n", gfftx_strerror(rc), gfftx_last_error());
    return 0;
}
""")
    exe = tmp_path / "use_gfftx"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", os.path.dirname(api.LIB_PATH), "-lgfftx", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_policy_header_compiles_cleanly():
    """include/gfftx/gfft.hpp (GFFT's names and the HIP<Dev> policy) compiles without
    warnings under -Wall -Wextra, as the C++ check program that drives it."""
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "gfft_check.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_exec_batch_argument_errors():
    L = api.lib()
    vp = ctypes.c_void_p
    # an empty batch is a no-op; NULL arrays and NULL entries are rejected before any work
    assert L.gfftx_exec_batch(None, None, 0, None) == api.SUCCESS
    assert L.gfftx_exec_batch(None, None, 2, None) == api.ERR_INVALID_ARG
    plans = (vp * 2)(None, None)
    data = (vp * 2)(None, None)
    assert L.gfftx_exec_batch(plans, data, 2, None) == api.ERR_INVALID_ARG
    assert b"plans[0]" in L.gfftx_last_error()
