"""C-ABI boundary checks that need no GPU: the library loads, exports every function that
include/mibwt.h declares, and rejects bad parameters before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    src = open(os.path.join(ROOT, "include", "mibwt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.findall(r"^\s*(?:int64_t|void|int)\s+\**(\w+)\s*\(", src, flags=re.M)


def test_header_declares_libcubwt_abi():
    names = header_functions()
    for f in ("libcubwt_allocate_device_storage", "libcubwt_free_device_storage", "libcubwt_bwt"):
        assert f in names


def test_library_exports_every_declared_function(mibwt):
    lib = mibwt.lib()
    for name in header_functions():
        assert hasattr(lib, name), name
    assert sorted(header_functions()) == sorted(mibwt.EXPORTED)


def test_error_codes_match_libcubwt():
    src = open(os.path.join(ROOT, "include", "mibwt.h")).read()
    codes = dict(re.findall(r"#define (LIBCUBWT_\w+)\s+(-?\d+)", src))
    # values of libcubwt.cuh:32-38
    assert codes == {"LIBCUBWT_NO_ERROR": "0", "LIBCUBWT_BAD_PARAMETER": "-1",
                     "LIBCUBWT_NOT_ENOUGH_MEMORY": "-2", "LIBCUBWT_GPU_ERROR": "-7",
                     "LIBCUBWT_GPU_NOT_SUPPORTED": "-8", "LIBCUBWT_GPU_NOT_ENOUGH_MEMORY": "-9",
                     "LIBCUBWT_VERSION_MAJOR": "1", "LIBCUBWT_VERSION_MINOR": "6",
                     "LIBCUBWT_VERSION_PATCH": "3"}


def test_bad_parameters_rejected_without_device(mibwt):
    lib = mibwt.lib()
    T = np.zeros(32, dtype=np.uint8)
    L = np.zeros(32, dtype=np.uint8)
    u8 = T.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    l8 = L.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    assert lib.libcubwt_allocate_device_storage(None, 100) == mibwt.BAD_PARAMETER
    h = ctypes.c_void_p()
    assert lib.libcubwt_allocate_device_storage(ctypes.byref(h), -1) == mibwt.BAD_PARAMETER
    assert lib.libcubwt_allocate_device_storage(ctypes.byref(h), 1 << 31) == mibwt.BAD_PARAMETER
    assert lib.libcubwt_free_device_storage(None) == mibwt.BAD_PARAMETER
    assert lib.libcubwt_bwt(None, u8, l8, 32) == mibwt.BAD_PARAMETER
    assert lib.mibwt_pfp_bwt(None, u8, l8, 32, 5, 16) == mibwt.BAD_PARAMETER
    assert lib.mibwt_lyndon_bwt(None, u8, l8, 32) == mibwt.BAD_PARAMETER
    assert lib.mibwt_pfp_bwt_device(None, None, None, 32, 5, 16, None) == mibwt.BAD_PARAMETER
    assert lib.mibwt_last_phase_stats(None, None) == mibwt.BAD_PARAMETER
    assert lib.mibwt_set_kernel_timing(None, None) == mibwt.BAD_PARAMETER
    assert lib.mibwt_time_kernel(None, b"scan", 1, None, None) == mibwt.BAD_PARAMETER
    W, P = ctypes.c_int32(), ctypes.c_int32()
    assert lib.mibwt_select_wp(None, 32, ctypes.byref(W), ctypes.byref(P)) == mibwt.BAD_PARAMETER
    assert lib.mibwt_select_wp(u8, 0, ctypes.byref(W), ctypes.byref(P)) == mibwt.BAD_PARAMETER


def test_reference_format_conversion(mibwt):
    import oracle
    for T in (b"mississippi", b"zzzzzzzz", b"banana", b"abcabcabcx"):
        O, idx = oracle.bwt(T)
        L = oracle.ref_to_cubwt(O, idx)
        assert (mibwt.reference_to_cubwt(O, idx) == L).all()
        assert (mibwt.cubwt_to_reference(L, idx) == O).all()
