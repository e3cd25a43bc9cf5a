"""The C ABI library loads and exports every entry point its header declares (CPU, no compute)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpsparse_amd.h")
LIB = os.path.join(ROOT, "mpsparse_amd", "libmpsparse_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msp_[a-z0-9_]+)\s*\(", text)))


class Opts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxiter", ctypes.c_int32), ("check_every", ctypes.c_int32),
                ("criterion", ctypes.c_int32), ("use_x0", ctypes.c_int32), ("precond", ctypes.c_int32),
                ("values", ctypes.c_int32), ("verify", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libmpsparse_amd.so is not built (python -m mpsparse_amd.build)")
    return ctypes.CDLL(LIB)


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_defaults(lib):
    assert lib.msp_abi_version() == 1
    o = Opts()
    lib.msp_default_opts(ctypes.byref(o))
    assert o.tol == 1e-6 and o.maxiter == 10000 and o.check_every == 16
    assert o.criterion == 0 and o.use_x0 == 0 and o.verify == 1


def test_no_device_is_an_error_not_a_fallback(lib):
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    assert lib.msp_device_count() == 0
    h = ctypes.c_void_p()
    keys = (ctypes.c_uint64 * 1)(0)
    vals = (ctypes.c_float * 1)(1.0)
    lib.msp_last_error.restype = ctypes.c_char_p
    st = lib.msp_csr_create_coo(ctypes.byref(h), keys, vals, ctypes.c_uint64(1), 1, 1, None)
    assert st == 4  # MSP_ERR_NO_DEVICE
    assert b"no HIP device" in lib.msp_last_error()
    assert not h.value


def test_null_handle_rejected(lib):
    lib.msp_last_error.restype = ctypes.c_char_p
    assert lib.msp_csr_shape(None, None, None, None) == 1
    assert lib.msp_csr_destroy(None) == 0


def test_extension_imports_without_gpu():
    from mpsparse_amd import coo_csr, spmv

    assert hasattr(spmv.csr_tensor, "cgd") and hasattr(spmv.csr_tensor, "bicgstab")
    assert hasattr(spmv.csr_tensor, "mv") and callable(coo_csr.coo_csr)
