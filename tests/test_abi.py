"""CPU checks of the C-ABI library: it loads, exports every declared symbol, and refuses to
run without a device instead of falling back to the CPU."""
import ctypes
import os

import pytest

import ssmhip
from ssmhip import _lib


def test_library_is_built():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    L = _lib.load()
    for n in _lib.header_symbols():
        assert getattr(L, n).restype is not None or n in (), n


def test_version():
    assert _lib.load().ssm_version() >= 10000


def test_no_cpu_fallback_without_device():
    # Without a HIP device a context cannot be created: the library has no CPU path.
    try:
        c = ssmhip.Context(0)
    except ssmhip.SSMError as e:
        assert e.code == _lib.SSM_ERR_NO_DEVICE
        return
    c.close()
    pytest.skip("a device is present")
