"""The C-ABI library loads and exports every entry point include/kupsmi.h declares
(CPU only: no compute call is made)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "kupsmi.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(kupsmi_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "kupsmi_md_lj" in syms and "kupsmi_neighborlist" in syms


def test_library_exports_declared_symbols():
    lib_path = ROOT / "kups_amd" / "libkupsmi.so"
    if not lib_path.exists():
        pytest.fail("kups_amd/libkupsmi.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(lib_path))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_python_binding_covers_header():
    from kups_amd import _lib

    assert set(declared_symbols()) <= set(_lib.exported_symbols())
    assert _lib.load().kupsmi_version() >= 100


def test_no_gpu_raises_instead_of_falling_back():
    import torch

    from kups_amd import _lib

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_lib.KupsmiError):
        _lib.ctx()
