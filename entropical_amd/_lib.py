"""ctypes binding of libentropical_kde.so (include/entropical_kde.h).

The library is loaded from this package directory only; there is no CPU
fallback.  If it is missing, every call raises ``EntropicalError``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libentropical_kde.so"

ENT_OK = 0
ENT_ERR_ARG = 1
ENT_ERR_NO_DEVICE = 2
ENT_ERR_HIP = 3
ENT_ERR_OOM = 4
ENT_ERR_NONFINITE = 5
ENT_ERR_UNSUPPORTED = 6

METHOD_CELL = 0
METHOD_BRUTE = 1
HEST_THUMB = 0
HEST_AMISE = 1

_STATUS_NAMES = {
    ENT_ERR_ARG: "ENT_ERR_ARG",
    ENT_ERR_NO_DEVICE: "ENT_ERR_NO_DEVICE",
    ENT_ERR_HIP: "ENT_ERR_HIP",
    ENT_ERR_OOM: "ENT_ERR_OOM",
    ENT_ERR_NONFINITE: "ENT_ERR_NONFINITE",
    ENT_ERR_UNSUPPORTED: "ENT_ERR_UNSUPPORTED",
}


class EntropicalError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class ent_job(C.Structure):
    _fields_ = [
        ("n_dim", C.c_uint32),
        ("n", C.c_uint32),
        ("n_rows", C.c_uint32),
        ("reserved", C.c_uint32),
        ("off", C.c_uint64 * 3),
        ("h", C.c_float * 3),
        ("reserved2", C.c_float),
        ("out_off", C.c_uint64),
    ]


P = C.c_void_p
U32P = C.POINTER(C.c_uint32)
F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol of include/entropical_kde.h
SIGNATURES = {
    "ent_version": (C.c_int, []),
    "ent_last_error": (C.c_char_p, []),
    "ent_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ent_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "ent_ctx_destroy": (C.c_int, [P]),
    "ent_ctx_set_tuning": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ent_ctx_last_stats": (C.c_int, [P, F64P, F64P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "ent_ctx_last_pairs": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "ent_ctx_last_sort": (C.c_int, [P, C.POINTER(C.c_uint32)]),
    "ent_ctx_last_moments": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "ent_ctx_last_tiles": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "ent_ctx_last_graph_replays": (C.c_int, [P, C.POINTER(C.c_uint32)]),
    "ent_column_stats_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P]),
    "ent_ctx_set_graph": (C.c_int, [P, C.c_int]),
    "ent_dev_alloc": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
    "ent_dev_free": (C.c_int, [P, P]),
    "ent_h2d": (C.c_int, [P, P, P, C.c_size_t]),
    "ent_d2h": (C.c_int, [P, P, P, C.c_size_t]),
    "ent_ctx_sync": (C.c_int, [P]),
    "ent_coord_prep": (C.c_int, [P, C.c_size_t, P, P, P, C.c_uint32, C.c_int, P, C.POINTER(C.c_size_t)]),
    "ent_combined_densities": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P, P, C.c_uint32, C.c_int, P]),
    "ent_combined_densities_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P, P, C.c_uint32, C.c_int, P]),
    "ent_combined_densities_dev_noalloc": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P, P, C.c_uint32, C.c_int, P]),
    "ent_batch_densities_dev": (C.c_int, [P, P, C.POINTER(ent_job), C.c_size_t, C.c_int, P]),
    "ent_epa_convolution": (C.c_int, [P, P, C.c_size_t, C.c_float, F32P]),
    "ent_bin": (C.c_int, [P, P, C.c_size_t, C.c_float, P, U32P, F32P, F32P]),
    "ent_cell_sort": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P, P, C.c_uint32, P, P, P, P, P]),
    "ent_mi_matrix_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, C.c_int, C.c_int, P]),
    "ent_te_matrix_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P, C.c_size_t, C.c_int, C.c_int, P]),
    "ent_batch_entropy_dev": (C.c_int, [P, P, C.POINTER(ent_job), C.c_size_t, C.c_int, P]),
    "ent_ctx_last_driver_ms": (C.c_int, [P, F64P]),
    "ent_batch_loo_dev": (C.c_int, [P, P, C.POINTER(ent_job), C.c_size_t, C.c_double, C.c_int, P]),
    "ent_hestimate_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, C.c_int, P]),
    "ent_read_xtc": (C.c_int, [C.c_char_p, P, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)]),
    "ent_autoent_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, C.c_uint32, C.c_int, C.c_int, P]),
    "ent_kldiv_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, C.c_int, C.c_int, P]),
    "ent_negs_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, C.c_int, C.c_int, P]),
    "ent_mi_autotune_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P]),
    "ent_te_autotune_dev": (C.c_int, [P, P, C.c_size_t, C.c_size_t, P, P, C.c_size_t, P]),
    "ent_autotune_dev": (C.c_int, [P, P, C.POINTER(ent_job), C.c_size_t, P, F64P]),
    "ent_ctx_get_tuning": (C.c_int, [P, P]),
    "ent_mlp_create": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(P)]),
    "ent_mlp_destroy": (C.c_int, [P]),
    "ent_mlp_num_params": (C.c_int, [P, C.POINTER(C.c_size_t)]),
    "ent_mlp_get_params": (C.c_int, [P, P, P]),
    "ent_mlp_set_params": (C.c_int, [P, P, P]),
    "ent_mlp_predict": (C.c_int, [P, P, P, C.c_size_t, P]),
    "ent_mlp_gradient": (C.c_int, [P, P, P, P, C.c_size_t, P, F32P]),
    "ent_mlp_train": (C.c_int, [P, P, P, P, C.c_size_t, P, C.c_int, C.c_int, C.c_float, P]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("ENT_LIB", LIB_PATH))
    if not path.exists():
        raise EntropicalError(
            ENT_ERR_UNSUPPORTED,
            f"HIP library not built: {path} is missing (run python -m entropical_amd.build)",
        )
    lib = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:  # an older library build; the symbol test covers the shipped one
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != ENT_OK:
        msg = load().ent_last_error()
        raise EntropicalError(status, msg.decode() if msg else "")
