"""The C-ABI library loads and exports every symbol include/qfvm.h declares (CPU; no compute)."""

import ctypes
import os
import re

import pytest

from pyquafu_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qfvm.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qfvm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    syms = header_symbols()
    for s in ["qfvm_create", "qfvm_apply", "qfvm_run", "qfvm_expect_paulis", "qfvm_sample", "qfvm_store"]:
        assert s in syms


def test_library_exports_every_header_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libqfvm_hip.so not built")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding covers exactly the header's entry points
    assert sorted(_native.SIGNATURES) == header_symbols()


def test_create_reports_missing_device_cleanly():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libqfvm_hip.so not built")
    if _native.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_native.EngineError) as e:
        _native.DeviceState(4)
    assert e.value.code == _native.ERR_DEVICE
    assert "no HIP device" in str(e.value)


def test_product_refuses_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(ImportError):
        _native.lib()
