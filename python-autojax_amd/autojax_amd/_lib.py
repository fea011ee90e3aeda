"""ctypes binding of libautojax_amd.so (declared in include/autojax_amd.h).

Loading fails loudly: there is no CPU fallback anywhere in the product.
Status codes map to exceptions the way the reference's callers see errors:
bad arguments -> ValueError, not positive definite -> numpy.linalg.LinAlgError,
HIP failures and short workspaces -> RuntimeError.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libautojax_amd.so"

AJ_OK, AJ_EINVAL, AJ_ENOTPD, AJ_EHIP, AJ_ESPACE = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "int8x3": 3, "int8x4": 4, "int8x5": 5, "int8crt": 6, "int8crt11": 7}

_P = C.c_void_p
_I = C.c_int64
_S = C.c_size_t
_D = C.c_double
_INT = C.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "aj_last_error": (C.c_char_p, []),
    "aj_abi_version": (_INT, []),
    "aj_timing_enable": (_INT, [_INT]),
    "aj_timing_read": (_INT, [_INT, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "aj_mfma_peak_probe": (_INT, [_INT, _I, _I, _P, _P]),
    "aj_w_tilde_set_variant": (_INT, [_INT]),
    "aj_w_tilde_variant_count": (_INT, []),
    "aj_w_tilde_set_group_bytes": (_INT, [_S]),
    "aj_crt_constants": (_INT, [_INT, _P, _P, _P, _P]),
    "aj_crt_sk_schedule": (_INT, [_INT, _INT, _INT, _INT, _P, _I, _P, _P, _I, _P]),
    "aj_w_tilde_crt_configure": (_INT, [_INT, _I]),
    "aj_w_tilde_workspace_bytes": (_S, [_I, _I, _INT]),
    "aj_w_tilde_curvature_interferometer_from": (_INT, [_P, _P, _P, _I, _I, _INT, _P, _P, _S, _P]),
    "aj_w_compact_workspace_bytes": (_S, [_I, _I]),
    "aj_w_compact_curvature_interferometer_from": (_INT, [_P, _P, _I, _I, _D, _P, _P, _S, _P]),
    "aj_w_tilde_via_compact_from": (_INT, [_P, _I, _P, _I, _P, _P]),
    "aj_w_tilde_via_preload_from": (_INT, [_P, _I, _I, _P, _I, _P, _P]),
    "aj_w_tilde_preload_workspace_bytes": (_S, [_I, _I, _I]),
    "aj_w_tilde_curvature_preload_interferometer_from": (_INT, [_P, _P, _I, _I, _I, _P, _I, _I, _P, _P, _S, _P]),
    "aj_visibility_sum_workspace_bytes": (_S, [_I, _I]),
    "aj_w_tilde_data_interferometer_from": (_INT, [_P, _P, _P, _P, _I, _I, _P, _P, _S, _P]),
    "aj_dirty_image_from": (_INT, [_P, _P, _P, _P, _I, _I, _P, _P, _S, _P]),
    "aj_w_tilde_with_data_workspace_bytes": (_S, [_I, _I, _INT]),
    "aj_w_tilde_with_data_interferometer_from": (_INT, [_P, _P, _P, _I, _I, _INT, _P, _P, _P, _P, _P, _P, _S, _P]),
    "aj_mapping_matrix_from": (_INT, [_P, _P, _I, _I, _I, _P, _P]),
    "aj_data_vector_from": (_INT, [_P, _P, _I, _I, _P, _P]),
    "aj_curvature_dense_workspace_bytes": (_S, [_I, _I]),
    "aj_curvature_matrix_via_w_tilde_from": (_INT, [_P, _P, _I, _I, _P, _P, _S, _P]),
    "aj_sparse_transpose_matmul_workspace_bytes": (_S, [_I, _I, _I]),
    "aj_sparse_mapping_matrix_transpose_matmul": (_INT, [_P, _P, _P, _I, _I, _I, _I, _P, _P, _S, _P]),
    "aj_curvature_sparse_workspace_bytes": (_S, [_I, _I, _I]),
    "aj_curvature_matrix_via_w_wilde_sparse_mapping_matrix_from": (_INT, [_P, _P, _P, _I, _I, _I, _P, _P, _S, _P]),
    "aj_curvature_matrix_via_w_compact_sparse_mapping_matrix_from": (
        _INT, [_P, _I, _P, _P, _P, _I, _I, _I, _P, _P, _S, _P]),
    "aj_curvature_matrix_via_w_tilde_curvature_preload_interferometer_from": (
        _INT, [_P, _I, _I, _P, _P, _P, _I, _I, _I, _P, _P, _S, _P]),
    "aj_constant_regularization_matrix_from": (_INT, [_D, _P, _P, _I, _I, _P, _P]),
    "aj_solve_workspace_bytes": (_S, [_I]),
    "aj_reconstruction_positive_negative_from": (_INT, [_P, _P, _I, _P, _P, _P, _P, _S, _P]),
    "aj_reconstruction_positive_negative_from_noalloc": (_INT, [_P, _P, _I, _P, _P, _P, _P]),
    "aj_noise_normalization_complex_from": (_INT, [_P, _I, _P, _P]),
    "aj_chi_squared_noise_normalization": (_INT, [_P, _P, _I, _P, _P]),
    "aj_log_likelihood_workspace_bytes": (_S, [_I, _I, _I, _I, _I]),
    "aj_log_likelihood_function_via_w_tilde_from": (
        _INT, [_P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _S, _P]),
    "aj_log_likelihood_function_via_w_compact_from": (
        _INT, [_P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _S, _P]),
    "aj_log_evidence_batched_workspace_bytes": (_S, [_I, _I, _I, _P]),
    "aj_log_evidence_configure": (_INT, [_INT]),
    "aj_log_evidence_flow_configure": (_INT, [_INT, _INT]),
    "aj_log_evidence_flow_trace": (_INT, [_P, _I]),
    "aj_log_evidence_batched_compact": (
        _INT, [_P, _I, _P, _I, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _S, _P]),
    "aj_log_evidence_batched": (
        _INT, [_P, _I, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _S, _P]),
    "aj_log_evidence_backward_workspace_bytes": (_S, [_I, _I, _P]),
    "aj_log_evidence_batched_backward": (
        _INT, [_P, _I, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _S, _P, _P, _P, _S, _P]),
    "aj_barycentric_weights": (_INT, [_P, _P, _P, _I, _P, _P]),
    "aj_barycentric_vjp_workspace_bytes": (_S, [_I, _I]),
    "aj_barycentric_vjp": (_INT, [_P, _P, _P, _I, _I, _P, _P, _P, _S, _P]),
    "aj_delaunay_mesh_batched": (_INT, [_P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P]),
    "aj_delaunay_locate_batched": (_INT, [_P, _P, _I, _P, _P, _P, _I, _I, _P, _P, _P]),
}

_lib = None


def load() -> C.CDLL:
    """Load the library (once).  Raises RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m autojax_amd.build` (hipcc, gfx950); "
            "there is no CPU fallback"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().aj_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    if rc == AJ_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if rc == AJ_EINVAL:
        raise ValueError(msg)
    if rc == AJ_ENOTPD:
        raise np.linalg.LinAlgError(msg)
    raise RuntimeError(f"libautojax_amd status {rc}: {msg}")


def call(name: str, *args):
    """Call an entry point that returns a status and raise on failure."""
    check(getattr(load(), name)(*args), name)


TIMERS = {"wt_syrk": 0, "wt_features": 1, "omega": 2, "fgather": 3, "chol": 4, "solve": 5, "wt_combine": 6,
          "batch": 7}


def timing_enable(on: bool = True) -> None:
    check(load().aj_timing_enable(1 if on else 0), "aj_timing_enable")


def timing_read(name: str):
    """(total milliseconds, launch count) of a timer since the last timing_enable (synchronises)."""
    ms = C.c_double()
    n = C.c_int64()
    check(load().aj_timing_read(TIMERS[name], C.byref(ms), C.byref(n)), "aj_timing_read")
    return ms.value, n.value


def query(name: str, *args) -> int:
    return int(getattr(load(), name)(*args))
