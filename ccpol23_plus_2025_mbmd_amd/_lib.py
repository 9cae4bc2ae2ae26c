"""ctypes binding of libccpol3b.so (include/ccpol3b.h).

The library is the only compute path: when it is missing this module raises,
it never falls back to another implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libccpol3b.so"

FP64 = 0
MIXED = 1

OK, EINVAL, EHIP, ENOMEM, ECAPACITY, ENODEV = 0, -1, -2, -3, -4, -5


class Params(C.Structure):
    _fields_ = [
        ("beta_sr", C.POINTER(C.c_double)), ("gamma_sr", C.POINTER(C.c_double)), ("r0_sr", C.POINTER(C.c_double)),
        ("beta_lr", C.POINTER(C.c_double)), ("gamma_lr", C.POINTER(C.c_double)), ("r0_lr", C.POINTER(C.c_double)),
        ("ld_pair", C.c_int),
        ("cs_sr", C.POINTER(C.c_double)), ("cs_lr", C.POINTER(C.c_double)),
        ("nlin", C.c_int),
        ("alpha3", C.c_double), ("delta33", C.c_double), ("rcut3", C.c_double),
        ("s2_term", C.c_int), ("s3_term", C.c_int),
        ("precision", C.c_int), ("device", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n_fs3", C.c_long), ("n_fs3_sr", C.c_long), ("n_fs3_switch", C.c_long), ("n_fs3_lr", C.c_long),
        ("n_fs2_items", C.c_long), ("n_fs2_pairs", C.c_long),
        ("ms_total", C.c_double), ("ms_lists", C.c_double), ("ms_fs2", C.c_double), ("ms_fs3", C.c_double),
        ("ms_reduce", C.c_double), ("ms_h2d", C.c_double), ("ms_d2h", C.c_double),
        ("ms_fs3_eval", C.c_double), ("n_cells", C.c_long),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("fs3_lds_coef", C.c_int), ("fs2_parts", C.c_int), ("fs2_pair_parts", C.c_int),
                ("neighbour_cells", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CcpolError(RuntimeError):
    pass


_lib = None


def load():
    """Load libccpol3b.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise CcpolError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                         "(make -C ccpol23_plus_2025_mbmd_amd/csrc)")
    # PyTorch-ROCm ships its own HIP runtime under the same soname; loading it
    # first makes the library bind to that runtime, so one process never holds
    # two HIP runtimes (torch would otherwise see no devices).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.ccpol3b_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
    L.ccpol3b_create.restype = C.c_int
    L.ccpol3b_destroy.argtypes = [vp]
    L.ccpol3b_destroy.restype = None
    L.ccpol3b_set_precision.argtypes = [vp, C.c_int]
    L.ccpol3b_set_precision.restype = C.c_int
    L.ccpol3b_threebody_term.argtypes = [vp, C.c_int, C.c_int] + [dp] * 6 + [C.c_double] + [dp] * 18
    L.ccpol3b_threebody_term.restype = C.c_int
    L.ccpol3b_eval_device.argtypes = [vp, C.c_int, C.c_int] + [vp] * 6 + [C.c_double] + [vp] * 6 + [vp]
    L.ccpol3b_eval_device.restype = C.c_int
    L.ccpol3b_eval_device_noalloc.argtypes = [vp, C.c_int, C.c_int] + [vp] * 6 + [C.c_double] + [vp] * 6 + [vp]
    L.ccpol3b_eval_device_noalloc.restype = C.c_int
    L.ccpol3b_eval_batch_device.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp, C.c_int] + [vp] * 12 + [vp]
    L.ccpol3b_eval_batch_device.restype = C.c_int
    L.ccpol3b_eval_batch_grad_device.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp, C.c_int] + [vp] * 13 + [vp]
    L.ccpol3b_eval_batch_grad_device.restype = C.c_int
    L.ccpol3b_set_coefficients.argtypes = [vp, dp, dp]
    L.ccpol3b_set_coefficients.restype = C.c_int
    L.ccpol3b_gram_device.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp]
    L.ccpol3b_gram_device.restype = C.c_int
    L.ccpol3b_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.ccpol3b_get_tuning.restype = C.c_int
    L.ccpol3b_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.ccpol3b_set_tuning.restype = C.c_int
    L.ccpol3b_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp, C.c_int] + [vp] * 6 + \
        [C.c_int, vp, C.POINTER(Tuning), dp]
    L.ccpol3b_autotune.restype = C.c_int
    L.ccpol3b_graph_create.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp, C.c_int] + [vp] * 13 + \
        [C.c_double, vp, C.POINTER(C.c_void_p)]
    L.ccpol3b_graph_create.restype = C.c_int
    L.ccpol3b_graph_launch.argtypes = [vp, vp]
    L.ccpol3b_graph_launch.restype = C.c_int
    L.ccpol3b_graph_status.argtypes = [vp]
    L.ccpol3b_graph_status.restype = C.c_int
    L.ccpol3b_graph_destroy.argtypes = [vp]
    L.ccpol3b_graph_destroy.restype = None
    L.ccpol3b_md_create.argtypes = [vp, C.c_int, C.c_double, C.c_int, dp, C.c_double, dp, C.POINTER(C.c_void_p)]
    L.ccpol3b_md_create.restype = C.c_int
    L.ccpol3b_md_set_state.argtypes = [vp, dp, dp, dp, dp]
    L.ccpol3b_md_set_state.restype = C.c_int
    L.ccpol3b_md_get_state.argtypes = [vp, dp, dp, dp, dp]
    L.ccpol3b_md_get_state.restype = C.c_int
    L.ccpol3b_md_run.argtypes = [vp, C.c_int, C.c_double, dp]
    L.ccpol3b_md_run.restype = C.c_int
    L.ccpol3b_md_destroy.argtypes = [vp]
    L.ccpol3b_md_destroy.restype = None
    L.ccpol3b_last_stats.argtypes = [vp, C.POINTER(Stats)]
    L.ccpol3b_last_stats.restype = C.c_int
    L.ccpol3b_get_fs3_list.argtypes = [vp, C.POINTER(C.c_int), C.c_long]
    L.ccpol3b_get_fs3_list.restype = C.c_long
    L.ccpol3b_get_fs2_list.argtypes = [vp, C.POINTER(C.c_int), C.c_long]
    L.ccpol3b_get_fs2_list.restype = C.c_long
    L.ccpol3b_read_data3b.argtypes = [C.c_char_p, dp, dp, dp, C.c_int, dp, C.POINTER(C.c_int)]
    L.ccpol3b_read_data3b.restype = C.c_int
    L.ccpol3b_orbit_index.argtypes = [C.c_int] * 6
    L.ccpol3b_orbit_index.restype = C.c_int
    L.ccpol3b_status_string.argtypes = [C.c_int]
    L.ccpol3b_status_string.restype = C.c_char_p
    L.ccpol3b_last_error.argtypes = [vp]
    L.ccpol3b_last_error.restype = C.c_char_p
    _lib = L
    return L


EXPORTED = [
    "ccpol3b_create", "ccpol3b_destroy", "ccpol3b_set_precision", "ccpol3b_threebody_term",
    "ccpol3b_eval_device", "ccpol3b_eval_batch_device", "ccpol3b_last_stats", "ccpol3b_get_fs3_list",
    "ccpol3b_get_fs2_list", "ccpol3b_read_data3b", "ccpol3b_orbit_index", "ccpol3b_status_string",
    "ccpol3b_last_error", "ccpol3b_eval_batch_grad_device", "ccpol3b_set_coefficients", "ccpol3b_gram_device",
    "ccpol3b_get_tuning", "ccpol3b_set_tuning", "ccpol3b_autotune", "ccpol3b_graph_create", "ccpol3b_graph_launch",
    "ccpol3b_eval_device_noalloc",
    "ccpol3b_graph_status", "ccpol3b_graph_destroy", "ccpol3b_md_create", "ccpol3b_md_set_state",
    "ccpol3b_md_get_state", "ccpol3b_md_run", "ccpol3b_md_destroy",
]


def check(rc: int, ctx=None, what: str = ""):
    if rc != OK:
        L = load()
        msg = L.ccpol3b_status_string(rc).decode()
        if ctx:
            detail = L.ccpol3b_last_error(ctx).decode()
            if detail:
                msg += f" ({detail})"
        raise CcpolError(f"{what}: {msg}" if what else msg)
