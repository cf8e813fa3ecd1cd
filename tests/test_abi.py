"""The C ABI: the library loads on the CPU and exports every symbol ``include/cdft_hip.h`` declares
(no compute call is made here), and the ctypes table matches the header."""

from __future__ import annotations

import ctypes
import re

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "cdft_hip.h"


def declared_symbols() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(cdft_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "cdft_hdiv_apply_f64" in names and "cdft_last_error" in names


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), f"{name} declared in the header but not exported"


def test_ctypes_table_covers_header(pkg):
    bound = set(pkg._lib.exported_symbols())
    assert set(declared_symbols()) <= bound


def test_abi_version_and_error_string(pkg):
    lib = pkg.load_library()
    assert lib.cdft_abi_version() == 1
    assert isinstance(lib.cdft_last_error(), bytes)


def test_null_arguments_are_rejected_before_any_launch(pkg):
    """Validation runs on the host: a null pointer returns CDFT_E_ARG without touching a device."""
    with pytest.raises(pkg._lib.HipLibraryError, match="null pointer"):
        pkg._lib.call("cdft_hdiv_apply_f64", 0, 0, 0, 0, 0, 0, 0, 1, 4, 4, 4, 0, 8, 0)
