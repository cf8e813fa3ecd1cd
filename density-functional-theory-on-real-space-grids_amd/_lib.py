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
    "cdft_hdiv_pack_f64": [P, P, P, P, I, I, I, I, P, P],
    "cdft_hdiv_pack_f32": [P, P, P, P, I, I, I, I, P, P],
    "cdft_hdiv_apply_f64": [P, P, P, P, P, P, P, P, I, I, I, I, P, I, P],
    "cdft_hdiv_apply_f32": [P, P, P, P, P, P, P, P, I, I, I, I, P, I, P],
    "cdft_cheb_step_f64": [P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, I, P],
    "cdft_cheb_step_f32": [P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, I, P],
    "cdft_cheb_filter_f64": [P, P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, I, P],
    "cdft_cheb_filter_f32": [P, P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, I, P],
    "cdft_hdiv_apply_adjoint_f64": [P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P],
    "cdft_hdiv_apply_adjoint_f32": [P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P],
    "cdft_batch_plan_f64": [I, P, I, I, P, P, P, L, P],
    "cdft_cheb_filter_batched_f64": [P, I, P, I, I, P],
    "cdft_hdiv_apply_batched_f64": [P, I, P, I, P],
    "cdft_batch_plan_f32": [I, P, I, I, P, P, P, L, P],
    "cdft_cheb_filter_batched_f32": [P, I, P, I, I, P],
    "cdft_hdiv_apply_batched_f32": [P, I, P, I, P],
    "cdft_hdiv_set_segment": [I],
    "cdft_hdiv_set_tile": [I, I],
    "cdft_hdiv_launch_shape": [I, I, I, I, I, P],
    "cdft_hdiv_set_kernel": [I],
    "cdft_hdiv_vec_set": [I, I, I, I, I, I, I, I],
    "cdft_hdiv_vec_shape": [I, I, I, I, I, P],
    "cdft_wgram_f64": [P, I, P, I, P, L, D, I, P, L, P, P],
    "cdft_wgram_noalloc_f64": [P, I, P, I, P, L, D, I, P, P],
    "cdft_rotate_f64": [P, I, I, I, P, P, L, P],
    "cdft_lincomb_f64": [P, I, P, P, P, P, P, L, P],
    "cdft_lincomb_dots_f64": [P, I, P, P, P, P, P, L, P, I, P, I, P, L, P, P],
    "cdft_resnorm_f64": [P, P, P, P, I, L, P, L, P, P],
    "cdft_chol_inv_f64": [P, I, P, P, P],
    "cdft_eigh_f64": [P, I, P, P, P],
    "cdft_orth_error_f64": [P, I, P, P],
    "cdft_cheb_table_f64": [P, I, P, I, P, P],
    "cdft_rotate_gram_f64": [P, I, P, P, P, L, P, L, P, P],
    "cdft_hdiv_resnorm_f64": [P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, L, P, P],
    "cdft_hdiv_resnorm_tail_f64": [P, P, P, P, P, P, P, P, P, I, I, I, I, P, I, P, L, P, I, P, I, P, P, P, P],
    "cdft_local_apply_f64": [P, P, P, I, I, I, I, P, I, I, P],
    "cdft_local_filter_f64": [P, P, P, P, P, I, I, I, I, P, I, I, P, I, P],
    "cdft_local_resnorm_f64": [P, P, P, P, I, I, I, I, P, I, I, P, L, P, I, P, I, P, P, P, P],
    "cdft_cusp_grid_fields_f64": [I, P, P, P, D, P, D, P, P, D, P, P, P, P, P],
    "cdft_face_weights_f64": [I, P, P, P, D, P, I, P, P, P, P, P],
    "cdft_node_fields_f64": [P, P, P, L, I, P, P, P, D, P, P, P, P, P, P],
    "cdft_far_factor_f64": [P, D, P, I, P, P, P],
    "cdft_node_stencils_f64": [P, L, P, D, P, P, P, P, P],
    "cdft_interpolate_f64": [P, I, P, P, L, P, D, P, P, P],
    "cdft_interpolate_chunks_f64": [P, I, P, I, P, P, L, P, D, P, P, P],
    "cdft_lift_f64": [P, I, L, P, P, P, D, P, P, I, P, P, P, P, P, L, P, P, L, P, P],
    "cdft_gradient_f64": [P, I, P, D, P, P, I, P, P],
    "cdft_far_fine_f64": [I, P, P, P, P, P, P, D, P, P],
    "cdft_interp_transpose_axis_f64": [P, P, I, I, P, P, P, P, P],
    "cdft_far_blend_f64": [P, D, P, D, D, P, P, P, P],
    "cdft_mt19937_states": [ctypes.c_uint64, I, P, P, P],
    "cdft_node_table": [I, P, P, P, P, P, D, P, I, P, P, P, P, P, L, P, P, P, P, P, P, P],
    "cdft_hydrogenic_fields_f64": [P, L, P, D, P, I, P, P, P, P, P],
    "cdft_lda_vwn_f64": [P, L, P, P, P],
    "cdft_pbe_f64": [P, P, L, P, P, P, P],
    "cdft_gga_points_f64": [P, P, P, P, P, P, L, P, P, P, P, P],
    "cdft_divergence_f64": [P, P, P, I, P, P],
    "cdft_gga_combine_f64": [P, P, P, P, D, L, P, P],
    "cdft_ewise_f64": [I, P, P, P, D, P, L, P],
    "cdft_density_f64": [P, I, L, P, P, P, P, P],
    "cdft_atom_combine_f64": [P, P, P, I, L, D, P, P],
    "cdft_small_solve_f64": [P, P, P, I, P, P],
    "cdft_pulay_coefficients_f64": [P, I, P, I, D, P, P],
    "cdft_reduce_f64": [I, P, P, P, L, D, I, P, L, P, I, P],
    "cdft_reduce_dots_f64": [I, P, P, P, L, P, P, P, P, L, P, P],
    "cdft_poisson_create": [I, I, I, I, I, I, P],
    "cdft_poisson_destroy": [P],
    "cdft_poisson_kernel_f64": [P, D, D, P, P],
    "cdft_poisson_solve_f64": [P, P, P, P, P, P, P, P, P],
    "cdft_poisson_set_x_path": [I],
    "cdft_tpa_create": [I, I, I, P],
    "cdft_tpa_destroy": [P],
    "cdft_tpa_filter_f64": [P, D, P, P, P],
    "cdft_tpa_apply_f64": [P, P, I, P, P, P, P, P, P],
}

#: Entry points that return a value rather than a status.
VALUED = {
    "cdft_reduction_workspace_doubles": (ctypes.c_long, [L]),
    "cdft_reduce_workspace_doubles": (ctypes.c_long, [L]),
    "cdft_batch_desc_bytes": (ctypes.c_long, [I]),
    "cdft_hdiv_pack_len": (ctypes.c_long, [I, I, I, I]),
    "cdft_local_partials_len": (ctypes.c_long, [I, I, I, I]),
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
        for name, (restype, argtypes) in VALUED.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
        return lib


def exported_symbols() -> list[str]:
    return ["cdft_last_error", *SIGNATURES, *VALUED]


def call(name: str, *args) -> None:
    """Call an entry point and raise :class:`HipLibraryError` on a non-zero status."""
    lib = load()
    status = getattr(lib, name)(*args)
    if status != 0:
        message = lib.cdft_last_error().decode(errors="replace")
        raise HipLibraryError(f"{name} returned {status}: {message}")
