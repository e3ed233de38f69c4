"""ctypes binding of libtsn.so (C ABI declared in include/tsn.h).

The library is loaded from this package directory.  There is no fallback:
if the shared library is missing or fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIBRARY_PATH = Path(__file__).resolve().parent / "libtsn.so"

TSN_OK, TSN_ERR_RUNTIME, TSN_ERR_INCOMPATIBLE, TSN_ERR_CG, TSN_ERR_ARGUMENT = 0, 1, 2, 3, 4
PREC_SPLIT16, PREC_DENSE16, PREC_FP32, PREC_HI16, PREC_SPLITBF16 = 0, 1, 2, 3, 4
VALUES_FP16, VALUES_BF16 = 0, 1
OUT_COIL_MAJOR, OUT_ROW_MAJOR, OUT_SPLIT_IL = 0, 1, 2
DC_MATERIALIZED, DC_FUSED = 1, 2


class PackHeader(C.Structure):
    _pack_ = 1
    _fields_ = [("magic", C.c_uint64), ("version", C.c_uint32), ("n_channels", C.c_uint32),
                ("output_size", C.c_uint32), ("n_cols", C.c_uint32), ("active_rows", C.c_uint32),
                ("nnz", C.c_uint32), ("groups", C.c_uint32), ("tiles", C.c_uint32),
                ("residual_nnz", C.c_uint32), ("m_tile", C.c_uint32), ("k_tile", C.c_uint32)]


class PlanConfig(C.Structure):
    _fields_ = [("image0", C.c_uint32), ("image1", C.c_uint32), ("grid0", C.c_uint32),
                ("grid1", C.c_uint32), ("coils", C.c_uint32), ("precision", C.c_int)]


class BatchProblem(C.Structure):
    _fields_ = [("forward", C.c_void_p), ("adjoint", C.c_void_p), ("smaps", C.c_void_p), ("coils", C.c_uint32)]


class TsnError(RuntimeError):
    """Raised for a non-zero status; ``code`` keeps the reference exit-code meaning."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[tsn status {code}] {message}")
        self.code = code


P = C.c_void_p
U32 = C.c_uint32
PP = C.POINTER(C.c_void_p)

# name: (restype, argtypes); restype None means void
SIGNATURES = {
    "tsn_last_error": (C.c_char_p, []),
    "tsn_version": (C.c_char_p, []),
    "tsn_host_pack_read": (C.c_int, [C.c_char_p, C.c_char_p, PP]),
    "tsn_host_pack_write": (C.c_int, [P, C.c_char_p, C.c_char_p]),
    "tsn_host_pack_create": (C.c_int, [C.POINTER(PackHeader), P, P, P, P, P, P, P, PP]),
    "tsn_host_pack_header": (C.c_int, [P, C.POINTER(PackHeader)]),
    "tsn_host_pack_arrays": (C.c_int, [P, PP, PP, PP, PP, PP, PP, PP]),
    "tsn_host_pack_free": (None, [P]),
    "tsn_kb_build": (C.c_int, [P, U32, U32, U32, U32, U32, U32, C.c_int, PP]),
    "tsn_kb_arrays": (C.c_int, [P, C.POINTER(U32), C.POINTER(U32), PP, PP, PP]),
    "tsn_kb_apodization": (C.c_int, [U32, U32, U32, U32, U32, P]),
    "tsn_kb_pack": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, C.c_int, PP]),
    "tsn_kb_pack_values": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, PP]),
    "tsn_kb_free": (None, [P]),
    "tsn_direction_upload": (C.c_int, [P, PP]),
    "tsn_direction_load": (C.c_int, [C.c_char_p, C.c_char_p, PP]),
    "tsn_direction_upload_csr": (C.c_int, [P, C.c_int, PP]),
    "tsn_direction_header": (C.c_int, [P, C.POINTER(PackHeader)]),
    "tsn_direction_bytes": (C.c_size_t, [P]),
    "tsn_direction_expand_dense": (C.c_int, [P]),
    "tsn_direction_schedule": (C.c_int, [P, C.c_uint32]),
    "tsn_direction_unit_tiles": (C.c_uint32, [P]),
    "tsn_direction_set_coil_waves": (C.c_int, [P, C.c_uint32]),
    "tsn_direction_coil_waves": (C.c_uint32, [P]),
    "tsn_direction_set_pipe_depth": (C.c_int, [P, C.c_uint32]),
    "tsn_direction_pipe_depth": (C.c_uint32, [P]),
    "tsn_batch_autotune": (C.c_int, [P, P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_float)]),
    "tsn_direction_autotune": (C.c_int, [P, C.c_uint32, P, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "tsn_direction_free": (None, [P]),
    "tsn_interp": (C.c_int, [P, P, P, P, U32, C.c_int, C.c_int, C.c_float, C.c_int, P]),
    "tsn_interp_noalloc": (C.c_int, [P, P, P, P, U32, C.c_int, C.c_int, C.c_float, C.c_int, P]),
    "tsn_interp_dense": (C.c_int, [P, P, P, P, U32, C.c_int, C.c_int, C.c_float, C.c_int, P]),
    "tsn_interp_f32": (C.c_int, [P, P, P, U32, C.c_int, C.c_int, C.c_float, C.c_int, P]),
    "tsn_split_panel": (C.c_int, [P, U32, U32, C.c_float, P, P, P]),
    "tsn_split_panel_il": (C.c_int, [P, U32, U32, C.c_float, C.c_int, P, P]),
    "tsn_interp_il": (C.c_int, [P, P, P, U32, C.c_int, C.c_int, C.c_float, C.c_int, P]),
    "tsn_interp_dc": (C.c_int, [P, P, P, C.c_int, P, U32, C.c_float, C.c_float, C.c_int, P, P, P]),
    "tsn_plan_create": (C.c_int, [C.POINTER(PlanConfig), P, P, P, P, PP]),
    "tsn_plan_free": (None, [P]),
    "tsn_plan_bytes": (C.c_size_t, [P]),
    "tsn_forward": (C.c_int, [P, P, P, P]),
    "tsn_adjoint": (C.c_int, [P, P, P, P]),
    "tsn_normal": (C.c_int, [P, P, P, C.c_float, C.c_float, P]),
    "tsn_data_consistency": (C.c_int, [P, P, P, P, C.c_float, P, P]),
    "tsn_cg": (C.c_int, [P, P, U32, C.c_float, C.c_float, P, P, P, P]),
    "tsn_cg_graph_capture": (C.c_int, [P, P, U32, C.c_float, C.c_float, P, P, P, P]),
    "tsn_cg_graph_launch": (C.c_int, [P, P]),
    "tsn_batch_create": (C.c_int, [C.POINTER(PlanConfig), C.POINTER(BatchProblem), U32, P, PP]),
    "tsn_batch_free": (None, [P]),
    "tsn_batch_bytes": (C.c_size_t, [P]),
    "tsn_batch_cg": (C.c_int, [P, C.POINTER(C.c_void_p), U32, C.POINTER(C.c_float), C.c_float,
                               C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), P, P]),
    "tsn_batch_cg_graph_capture": (C.c_int, [P, C.POINTER(C.c_void_p), U32, C.POINTER(C.c_float), C.c_float,
                                             C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), P, P]),
    "tsn_batch_cg_graph_launch": (C.c_int, [P, P]),
    "jkf_initialize_custom_fft_twiddles": (C.c_int, []),
    "jkf_launch_custom_fft_f30_forward": (None, [P, P, P, P, P]),
    "jkf_launch_custom_fft_f30_forward_split2": (None, [P, P, P, P, P, P]),
    "jkf_launch_custom_fft_f30_forward_split2_debug": (None, [P, P, P, P, P, P, P]),
    "jkf_launch_custom_fft_f30_forward_sense": (None, [P, P, P, P, P, P]),
    "jkf_launch_custom_fft_f30_forward_sense_split2": (None, [P, P, P, P, P, P, P]),
    "jkf_launch_custom_ifft_f31_adjoint": (None, [P, P, P, P, P, P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libtsn.so once; raise if it is absent (no CPU fallback exists)."""

    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: build it with `python -m trajsparsenufft_amd.build` "
                "(the HIP path has no fallback)")
        handle = C.CDLL(os.fspath(LIBRARY_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != TSN_OK:
        message = lib().tsn_last_error().decode(errors="replace")
        raise TsnError(status, message)


def exported_symbols() -> list[str]:
    return list(SIGNATURES)
