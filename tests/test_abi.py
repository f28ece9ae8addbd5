"""The C ABI: both libraries export every entry point include/pimc_gpu.h declares."""

import ctypes
import os
import re

import pytest

from pimcplusplus_amd import abi, native
from tests.helpers import ROOT, oracle_cdll

HEADER = os.path.join(ROOT, "include", "pimc_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pimc_\w+)\s*\(", text, flags=re.M)))


def test_header_and_bindings_agree():
    names = declared_symbols()
    assert len(names) >= 20
    assert set(names) == set(abi.SIGNATURES)


def test_gpu_library_loads_and_exports_every_symbol():
    if not os.path.exists(native.LIB_PATH):
        from pimcplusplus_amd import build
        build.build()
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    abi.bind(lib)
    assert lib.pimc_abi_version() == abi.PIMC_ABI_VERSION


def test_oracle_exports_every_symbol():
    lib = oracle_cdll()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_gpu_context_without_a_device_fails_loudly():
    lib = abi.bind(ctypes.CDLL(native.LIB_PATH))
    ctx = ctypes.c_void_p()
    rc = lib.pimc_ctx_create(0, ctypes.byref(ctx))
    if rc == abi.OK:          # a GPU is present: nothing to check here
        lib.pimc_ctx_destroy(ctx)
        pytest.skip("a GPU is visible")
    assert rc == abi.ERR_DEVICE
    assert b"device" in lib.pimc_last_error()
