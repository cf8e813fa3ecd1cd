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


def test_header_declares_every_bound_symbol(pkg):
    """The header is the whole ABI: nothing the host binding calls is missing from it."""
    assert set(pkg._lib.exported_symbols()) <= set(declared_symbols())


def test_launch_shape_is_host_only_and_tiles_odd_grids(pkg):
    """The stencil launch plan (a host computation) covers 65 = 5 x 13 with 13-wide tiles."""
    shape = pkg.ops.hdiv_launch_shape((65, 65, 65), 5, 8)
    assert (shape["tile_y"], shape["tile_z"]) == (13, 13)
    assert shape["threads"] % 64 == 0 and shape["blocks"] == 25 * 5 * shape["segments"]
    assert shape["segment"] >= 8


def test_abi_version_and_error_string(pkg):
    lib = pkg.load_library()
    assert lib.cdft_abi_version() == 1
    assert isinstance(lib.cdft_last_error(), bytes)


def test_null_arguments_are_rejected_before_any_launch(pkg):
    """Validation runs on the host: a null pointer returns CDFT_E_ARG without touching a device."""
    with pytest.raises(pkg._lib.HipLibraryError, match="null pointer"):
        pkg._lib.call("cdft_hdiv_apply_f64", 0, 0, 0, 0, 0, 0, 0, 0, 1, 4, 4, 4, 0, 8, 0)


def test_vector_launch_shape_and_settings_are_host_only(pkg):
    """The vector kernel's launch (a host computation): 65 columns take 22-wide float64 tiles (three
    tiles, one idle column) and 24-wide float32 tiles; a per-shape setting is used and removed; widths
    the kernel is not built for are rejected."""
    import torch

    ops = pkg.ops
    shape = ops.hdiv_vec_shape((65, 65, 65), 5, torch.float64)
    assert shape["tile_z"] == 22 and shape["tile_y"] == 17 and shape["threads"] % 64 == 0
    assert shape["blocks"] == 4 * 3 * 5 * shape["segments"]
    assert ops.hdiv_vec_shape((65, 65, 65), 5, torch.float32)["tile_z"] == 24
    ops.hdiv_vec_set((65, 65, 65), 5, torch.float64, 11, 24, 6)
    try:
        s = ops.hdiv_vec_shape((65, 65, 65), 5, torch.float64)
        assert (s["tile_y"], s["tile_z"], s["segment"]) == (11, 24, 6)
        assert ops.hdiv_vec_shape((65, 65, 65), 4, torch.float64)["tile_z"] == 22  # other k: untouched
    finally:
        ops.hdiv_vec_set((65, 65, 65), 5, torch.float64, 0, 0, 0)
    assert ops.hdiv_vec_shape((65, 65, 65), 5, torch.float64)["tile_z"] == 22
    with pytest.raises(pkg._lib.HipLibraryError):
        ops.hdiv_vec_set((65, 65, 65), 5, torch.float64, 11, 13, 0)
    with pytest.raises(pkg._lib.HipLibraryError):
        ops.hdiv_vec_set((65, 65, 65), 5, torch.float32, 11, 22, 0)
    with pytest.raises(pkg._lib.HipLibraryError):
        ops.hdiv_set_kernel(2)


def test_batch_plan_picks_the_vector_kernel_for_order_8(pkg):
    """A batch plan is host-only (it writes descriptors and reads no device memory, so placeholder
    addresses do): order 8 runs on the vector kernel (launch tile < 0) at the width with the fewest
    idle columns over the batch, a column-only width or the column-kernel switch selects the column
    kernel, other orders stay on it, and a vector-only width never reaches a column launch."""
    import ctypes

    from cdft_amd import _lib

    ops = pkg.ops
    lib = _lib.load()
    shapes_list = [(61, 57, 57), (71, 57, 57), (81, 57, 57)]

    def plan(order=8):
        n = len(shapes_list)
        shapes = (ctypes.c_int * (3 * n))(*[v for s in shapes_list for v in s])
        coefs = (ctypes.c_double * (order * n))(*([0.5] * (order * n)))
        ptrs = (ctypes.c_void_p * (11 * n))(*([0x100000] * (11 * n)))
        nbytes = int(lib.cdft_batch_desc_bytes(n))
        host = (ctypes.c_uint8 * nbytes)()
        launch = (ctypes.c_int * 3)()
        _lib.call("cdft_batch_plan_f64", n, ctypes.addressof(shapes), 5, order, ctypes.addressof(coefs),
                  ctypes.addressof(ptrs), ctypes.addressof(host), nbytes, ctypes.addressof(launch))
        return list(launch)

    blocks, threads, tile = plan()
    assert tile == -20 and threads % 64 == 0 and blocks > 0  # 57 = 3 x 20 - 3: the fewest idle columns
    assert plan(order=6)[2] in (13, 16)
    try:
        ops.hdiv_set_tile(0, 13)
        assert plan()[2] == 13
        ops.hdiv_set_tile(0, 28)
        assert plan()[2] == -28
        assert ops.hdiv_launch_shape((61, 57, 57), 5)["tile_z"] in (13, 16)  # the column kernel's own choice
        ops.hdiv_set_tile(0, 0)
        ops.hdiv_set_kernel(1)
        assert plan()[2] in (13, 16)
    finally:
        ops.hdiv_set_kernel(0)
        ops.hdiv_set_tile(0, 0)
    with pytest.raises(pkg._lib.HipLibraryError):
        ops.hdiv_set_tile(0, 21)
