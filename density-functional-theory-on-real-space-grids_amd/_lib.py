"""ctypes binding of ``libcdft_hip.so`` (the C ABI declared in ``include/cdft_hip.h``).

The library is built in-tree by ``_build.py``. Loading fails loudly when it is missing: there is no
CPU fallback anywhere in this package.
"""

from __future__ import annotations

import ctypes
import pathlib
import threading

_LIB_PATH = pathlib.Path(__file__).resolve().parent / "libcdft_hip.so"
_lock = threading.Lock()
_lib: ctypes.CDLL | None = None

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
L = ctypes.c_long

#: Entry point -> argument types. Pointers are raw device (or host, where the header says so)
#: addresses passed as integers; every function returns an int status.
SIGNATURES: dict[str, list] = {
    "cdft_abi_version": [],
    "cdft_hdiv_apply_f64": [P, P, P, P, P, P, P, I, I, I, I, P, I, P],
    "cdft_hdiv_apply_f32": [P, P, P, P, P, P, P, I, I, I, I, P, I, P],
    "cdft_cheb_step_f64": [P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, I, P],
    "cdft_cheb_step_f32": [P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, I, P],
    "cdft_cheb_filter_f64": [P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, I, P],
    "cdft_cheb_filter_f32": [P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, I, P],
}


class HipLibraryError(RuntimeError):
    """A non-zero status from the HIP library, with its recorded message."""


def library_path() -> pathlib.Path:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load the library once; raise if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not _LIB_PATH.exists():
            raise ImportError(
                f"{_LIB_PATH.name} is not built; run __graft_entry__.build() (hipcc, gfx950). "
                "This package has no CPU fallback."
            )
        lib = ctypes.CDLL(str(_LIB_PATH))
        lib.cdft_last_error.restype = ctypes.c_char_p
        lib.cdft_last_error.argtypes = []
        for name, argtypes in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = argtypes
        _lib = lib
        return lib


def exported_symbols() -> list[str]:
    return ["cdft_last_error", *SIGNATURES]


def call(name: str, *args) -> None:
    """Call an entry point and raise :class:`HipLibraryError` on a non-zero status."""
    lib = load()
    status = getattr(lib, name)(*args)
    if status != 0:
        message = lib.cdft_last_error().decode(errors="replace")
        raise HipLibraryError(f"{name} returned {status}: {message}")
