"""ctypes wrapper around liboracle3b.so, the CPU FP64 oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path
(ccpol23_plus_2025_mbmd_amd) never imports this module.

The C code restates source/threebody_potential.f90; see ccpol3b_oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle3b.so"


class _Params(C.Structure):
    _fields_ = [
        ("beta_sr", C.c_double * 289), ("gamma_sr", C.c_double * 289), ("r0_sr", C.c_double * 289),
        ("beta_lr", C.c_double * 289), ("gamma_lr", C.c_double * 289), ("r0_lr", C.c_double * 289),
        ("cs_sr", C.c_double * 364), ("cs_lr", C.c_double * 364),
        ("alpha3", C.c_double), ("delta33", C.c_double), ("rcut3", C.c_double),
        ("s2_term", C.c_int), ("s3_term", C.c_int),
    ]


class _Opts(C.Structure):
    _fields_ = [("rebuild_table_per_triplet", C.c_int), ("nthreads", C.c_int),
                ("fs3_first", C.c_long), ("fs3_last", C.c_long),
                ("fs2_first", C.c_long), ("fs2_last", C.c_long), ("optimised", C.c_int)]


class _Stats(C.Structure):
    _fields_ = [("n_fs3", C.c_long), ("n_fs3_sr", C.c_long), ("n_fs3_sw", C.c_long),
                ("n_fs3_lr", C.c_long), ("n_fs2", C.c_long),
                ("t_fs3_list", C.c_double), ("t_fs3_eval", C.c_double),
                ("t_fs2_list", C.c_double), ("t_fs2_eval", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        dp = C.POINTER(C.c_double)
        L.orc_threebody_term.argtypes = [C.POINTER(_Params), C.c_int, C.c_int, dp, dp, dp, dp, dp, dp,
                                         C.c_double, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                         C.POINTER(_Opts), C.POINTER(_Stats)]
        L.orc_threebody_term.restype = None
        L.orc_build_ind_lin.argtypes = [C.POINTER(C.c_int)]
        L.orc_build_ind_lin.restype = C.c_int
        for fn in (L.orc_fs3_list, L.orc_fs2_list):
            fn.argtypes = [C.POINTER(_Params), C.c_int, dp, dp, dp, C.c_double, C.POINTER(C.c_int), C.c_long]
            fn.restype = C.c_long
        L.orc_fs3_rc2_bohr2.restype = C.c_double
        L.orc_fs2_rc2_bohr2.restype = C.c_double
        L.orc_fs2_rc2_bohr2.argtypes = [C.c_double]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def to_cparams(p) -> _Params:
    """ThreeBodyParams -> C struct."""
    cp = _Params()
    for name, tab in (("sr", p.sr), ("lr", p.lr)):
        for f in ("beta", "gamma", "r0"):
            arr = np.ascontiguousarray(getattr(tab, f), dtype=np.float64).reshape(-1)
            C.memmove(getattr(cp, f"{f}_{name}"), arr.ctypes.data, 289 * 8)
        cs = np.ascontiguousarray(tab.cs, dtype=np.float64)
        C.memmove(getattr(cp, f"cs_{name}"), cs.ctypes.data, 364 * 8)
    cp.alpha3 = p.alpha3
    cp.delta33 = p.delta33
    cp.rcut3 = p.rcut3
    cp.s2_term = int(bool(p.s2_term))
    cp.s3_term = int(bool(p.s3_term))
    return cp


def ind_lin_2() -> np.ndarray:
    """The 39304-entry orbit table (values 1..364) and the orbit count."""
    t = np.zeros(39304, dtype=np.int32)
    n = lib().orc_build_ind_lin(t.ctypes.data_as(C.POINTER(C.c_int)))
    return t, n


def threebody_term(p, box_obj=None, *, x=None, y=None, z=None, xs=None, ys=None, zs=None,
                   box=None, ld=None, nthreads=1, rebuild=False,
                   fs3_window=(0, -1), fs2_window=(0, -1), optimised=False):
    """Run the oracle on one box.  Returns a dict with qfs2, qfs3 (kcal/mol),
    site forces fxs/fys/fzs [nm, ld] (+= onto zeros), COM forces fxfs2.. [nm],
    vir (+= onto 0) and the list statistics.  optimised=True is the optimised
    CPU baseline: the same lists from neighbour lists, and FS3 by the factorised
    contraction instead of the reference's loop nest."""
    if box_obj is not None:
        x, y, z = box_obj.x, box_obj.y, box_obj.z
        xs, ys, zs = box_obj.xs, box_obj.ys, box_obj.zs
        box, ld = box_obj.box, box_obj.ld
    nm = len(x)
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    xs, ys, zs = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (xs, ys, zs))
    out = {k: np.zeros(nm * ld) for k in ("fxs", "fys", "fzs")}
    for k in ("fxfs2", "fyfs2", "fzfs2", "fxfs3", "fyfs3", "fzfs3"):
        out[k] = np.zeros(nm)
    q2, q3, vir = C.c_double(0.0), C.c_double(0.0), C.c_double(0.0)
    o = _Opts(int(rebuild), int(nthreads), fs3_window[0], fs3_window[1], fs2_window[0], fs2_window[1],
              int(optimised))
    st = _Stats()
    cp = to_cparams(p)
    lib().orc_threebody_term(C.byref(cp), nm, ld, _dp(x), _dp(y), _dp(z), _dp(xs), _dp(ys), _dp(zs),
                             float(box), C.byref(q2), C.byref(q3),
                             _dp(out["fxs"]), _dp(out["fys"]), _dp(out["fzs"]),
                             _dp(out["fxfs2"]), _dp(out["fyfs2"]), _dp(out["fzfs2"]),
                             _dp(out["fxfs3"]), _dp(out["fyfs3"]), _dp(out["fzfs3"]),
                             C.byref(vir), C.byref(o), C.byref(st))
    for k in ("fxs", "fys", "fzs"):
        out[k] = out[k].reshape(nm, ld)
    out.update(qfs2=q2.value, qfs3=q3.value, vir=vir.value, n_fs3=st.n_fs3, n_fs3_sr=st.n_fs3_sr,
               n_fs3_sw=st.n_fs3_sw, n_fs3_lr=st.n_fs3_lr, n_fs2=st.n_fs2,
               t_fs3_list=st.t_fs3_list, t_fs3_eval=st.t_fs3_eval,
               t_fs2_list=st.t_fs2_list, t_fs2_eval=st.t_fs2_eval)
    return out


def fs3_list(p, x, y, z, box) -> np.ndarray:
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    cp = to_cparams(p)
    n = lib().orc_fs3_list(C.byref(cp), len(x), _dp(x), _dp(y), _dp(z), float(box), None, 0)
    out = np.zeros((max(n, 1), 3), dtype=np.int32)
    lib().orc_fs3_list(C.byref(cp), len(x), _dp(x), _dp(y), _dp(z), float(box),
                       out.ctypes.data_as(C.POINTER(C.c_int)), n)
    return out[:n]


def fs2_list(p, x, y, z, box) -> np.ndarray:
    """Ordered FS2 items as rows (k, i, j) in the reference's loop order."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    cp = to_cparams(p)
    n = lib().orc_fs2_list(C.byref(cp), len(x), _dp(x), _dp(y), _dp(z), float(box), None, 0)
    out = np.zeros((max(n, 1), 3), dtype=np.int32)
    lib().orc_fs2_list(C.byref(cp), len(x), _dp(x), _dp(y), _dp(z), float(box),
                       out.ctypes.data_as(C.POINTER(C.c_int)), n)
    return out[:n]


def rc2_fs3() -> float:
    return lib().orc_fs3_rc2_bohr2()


def rc2_fs2(rcut3: float) -> float:
    return lib().orc_fs2_rc2_bohr2(rcut3)
