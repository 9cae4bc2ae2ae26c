"""Host-side mirror of the reference's ``threebody_term`` (threebody_potential.f90:17-99).

``ThreeBody`` owns one library context (device copies of the SR/LR tables).
``ThreeBody.threebody_term`` has the reference's argument meaning and side
effects on numpy arrays (site forces +=, COM forces =, torques = 0, virial +=,
energies returned), ``ThreeBody.eval_device`` runs on device-resident torch
tensors for the gradient path and the benchmarks.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .params import ThreeBodyParams

PRECISIONS = {"fp64": _lib.FP64, "mixed": _lib.MIXED}
NLIN = 364  # linear coefficients per region (cs3_s, cs3_l)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fortran_pair_table(t: np.ndarray) -> np.ndarray:
    """[is, js] -> Fortran order with ld_pair = 17: element (is,js) at is + 17*js."""
    return np.ascontiguousarray(np.asarray(t, dtype=np.float64).T).reshape(-1)


class ThreeBody:
    """One device context for a parameter set (replaces the module globals the
    reference routine reads, common_arrays.f90:201-218)."""

    def __init__(self, params: ThreeBodyParams, precision: str = "fp64", device: int = 0):
        L = _lib.load()
        self.params = params
        keep = {}
        for name, tab in (("sr", params.sr), ("lr", params.lr)):
            keep[f"beta_{name}"] = _fortran_pair_table(tab.beta)
            keep[f"gamma_{name}"] = _fortran_pair_table(tab.gamma)
            keep[f"r0_{name}"] = _fortran_pair_table(tab.r0)
            keep[f"cs_{name}"] = np.ascontiguousarray(tab.cs, dtype=np.float64)
        self._keep = keep
        p = _lib.Params()
        for k, v in keep.items():
            setattr(p, k, _dp(v))
        p.ld_pair = 17
        p.nlin = len(params.sr.cs)
        p.alpha3, p.delta33, p.rcut3 = params.alpha3, params.delta33, params.rcut3
        p.s2_term, p.s3_term = int(bool(params.s2_term)), int(bool(params.s3_term))
        p.precision = PRECISIONS[precision]
        p.device = device
        ctx = C.c_void_p()
        _lib.check(L.ccpol3b_create(C.byref(p), C.byref(ctx)), None, "ccpol3b_create")
        self._ctx = ctx
        self.precision = precision
        self._coef_ref, self._coef_version = None, None  # the last tensor loaded by ThreeBodyEnergy

    def close(self):
        if getattr(self, "_ctx", None):
            _lib.load().ccpol3b_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision: str):
        _lib.check(_lib.load().ccpol3b_set_precision(self._ctx, PRECISIONS[precision]), self._ctx)
        self.precision = precision

    # ------------------------------------------------------------------ drop-in
    def threebody_term(self, x, y, z, xs, ys, zs, box, fxs, fys, fzs, fxfs2, fyfs2, fzfs2,
                       fxfs3, fyfs3, fzfs3, vir=0.0, txfs2=None, tyfs2=None, tzfs2=None,
                       txfs3=None, tyfs3=None, tzfs3=None, ld=None):
        """Same inputs/outputs as the reference routine on numpy arrays.
        Site arrays are Fortran-ordered (``xs[is + ld*i]``) or 2-D [nm, ld].
        Returns (qfs2, qfs3, vir) -- vir is the input value plus the increment."""
        L = _lib.load()
        nm = len(x)
        if ld is None:
            ld = np.asarray(xs).shape[-1] if np.asarray(xs).ndim == 2 else len(np.asarray(xs).reshape(-1)) // max(nm, 1)
        arrs_in = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (x, y, z, xs, ys, zs)]
        outs = [fxs, fys, fzs, fxfs2, fyfs2, fzfs2, fxfs3, fyfs3, fzfs3]
        for o in outs:
            if not (isinstance(o, np.ndarray) and o.dtype == np.float64 and o.flags.c_contiguous):
                raise TypeError("output arrays must be C-contiguous float64 numpy arrays (updated in place)")
        torq = [txfs2, tyfs2, tzfs2, txfs3, tyfs3, tzfs3]
        q2, q3, v = C.c_double(0.0), C.c_double(0.0), C.c_double(float(vir))
        tp = [(_dp(t) if t is not None else None) for t in torq]
        rc = L.ccpol3b_threebody_term(self._ctx, nm, ld, *[_dp(a) for a in arrs_in], float(box),
                                      C.byref(q2), C.byref(q3), *[_dp(o) for o in outs], *tp, C.byref(v))
        _lib.check(rc, self._ctx, "ccpol3b_threebody_term")
        return q2.value, q3.value, v.value

    def evaluate(self, box_obj):
        """Convenience: evaluate a geometry.WaterBox; returns a dict like the oracle's."""
        nm, ld = box_obj.nm, box_obj.ld
        out = {k: np.zeros(nm * ld) for k in ("fxs", "fys", "fzs")}
        for k in ("fxfs2", "fyfs2", "fzfs2", "fxfs3", "fyfs3", "fzfs3"):
            out[k] = np.zeros(nm)
        q2, q3, vir = self.threebody_term(box_obj.x, box_obj.y, box_obj.z, box_obj.xs, box_obj.ys, box_obj.zs,
                                          box_obj.box, out["fxs"], out["fys"], out["fzs"], out["fxfs2"],
                                          out["fyfs2"], out["fzfs2"], out["fxfs3"], out["fyfs3"], out["fzfs3"],
                                          0.0, ld=ld)
        for k in ("fxs", "fys", "fzs"):
            out[k] = out[k].reshape(nm, ld)
        out.update(qfs2=q2, qfs3=q3, vir=vir)
        out.update(self.stats())
        return out

    # ------------------------------------------------------------------ device path
    def eval_device(self, com, site, box, fsite=None, fcom2=None, fcom3=None, energies=None, stream=None):
        """Device-resident evaluation on torch tensors (float64, cuda):
        com [3, nm] (A), site [3, nm, ld] (A).  Site forces fsite [3, nm, ld] are
        accumulated, fcom2/fcom3 [3, nm] overwritten, energies [3] =
        (qfs2, qfs3, vir increment).  Returns (fsite, fcom2, fcom3, energies)."""
        import torch
        nm, ld = com.shape[1], site.shape[2]
        assert com.is_cuda and com.dtype == torch.float64 and com.is_contiguous() and com.shape[0] == 3
        assert site.is_cuda and site.dtype == torch.float64 and site.is_contiguous() and site.shape[:2] == (3, nm)
        dev = com.device
        if fsite is None:
            fsite = torch.zeros((3, nm, ld), dtype=torch.float64, device=dev)
        if fcom2 is None:
            fcom2 = torch.zeros((3, nm), dtype=torch.float64, device=dev)
        if fcom3 is None:
            fcom3 = torch.zeros((3, nm), dtype=torch.float64, device=dev)
        if energies is None:
            energies = torch.zeros(3, dtype=torch.float64, device=dev)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        p = lambda t, *i: C.c_void_p(t[i].data_ptr()) if i else C.c_void_p(t.data_ptr())
        rc = _lib.load().ccpol3b_eval_device(
            self._ctx, nm, ld, p(com, 0), p(com, 1), p(com, 2), p(site, 0), p(site, 1), p(site, 2),
            float(box), p(fsite, 0), p(fsite, 1), p(fsite, 2), p(fcom2), p(fcom3), p(energies),
            C.c_void_p(stream))
        _lib.check(rc, self._ctx, "ccpol3b_eval_device")
        return fsite, fcom2, fcom3, energies

    def eval_batch_device(self, mol_start, box_len, com, site, fsite=None, fcom2=None, fcom3=None,
                          energies=None, stream=None):
        """Many independent boxes packed along the molecule axis (see eval_device)."""
        import torch
        nbox = len(box_len)
        nm, ld = com.shape[1], site.shape[2]
        dev = com.device
        ms = np.ascontiguousarray(mol_start, dtype=np.int32)
        bl = np.ascontiguousarray(box_len, dtype=np.float64)
        assert ms[0] == 0 and ms[-1] == nm and len(ms) == nbox + 1
        if fsite is None:
            fsite = torch.zeros((3, nm, ld), dtype=torch.float64, device=dev)
        if fcom2 is None:
            fcom2 = torch.zeros((3, nm), dtype=torch.float64, device=dev)
        if fcom3 is None:
            fcom3 = torch.zeros((3, nm), dtype=torch.float64, device=dev)
        if energies is None:
            energies = torch.zeros((nbox, 3), dtype=torch.float64, device=dev)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        p = lambda t, *i: C.c_void_p(t[i].data_ptr()) if i else C.c_void_p(t.data_ptr())
        rc = _lib.load().ccpol3b_eval_batch_device(
            self._ctx, nbox, ms.ctypes.data_as(C.POINTER(C.c_int)), _dp(bl), ld,
            p(com, 0), p(com, 1), p(com, 2), p(site, 0), p(site, 1), p(site, 2),
            p(fsite, 0), p(fsite, 1), p(fsite, 2), p(fcom2), p(fcom3), p(energies), C.c_void_p(stream))
        _lib.check(rc, self._ctx, "ccpol3b_eval_batch_device")
        return fsite, fcom2, fcom3, energies

    def eval_batch_grad_device(self, mol_start, box_len, com, site, phi=None, stream=None):
        """eval_batch_device plus phi [nbox, 728] = d(qfs3 of the box)/d(cs3_s, cs3_l)
        (kcal/mol per unit coefficient).  Returns (fsite, fcom2, fcom3, energies, phi)."""
        import torch
        nbox = len(box_len)
        nm, ld = com.shape[1], site.shape[2]
        dev = com.device
        ms = np.ascontiguousarray(mol_start, dtype=np.int32)
        bl = np.ascontiguousarray(box_len, dtype=np.float64)
        assert ms[0] == 0 and ms[-1] == nm and len(ms) == nbox + 1
        fsite = torch.zeros((3, nm, ld), dtype=torch.float64, device=dev)
        fcom2 = torch.zeros((3, nm), dtype=torch.float64, device=dev)
        fcom3 = torch.zeros((3, nm), dtype=torch.float64, device=dev)
        energies = torch.zeros((nbox, 3), dtype=torch.float64, device=dev)
        if phi is None:
            phi = torch.empty((nbox, 2 * NLIN), dtype=torch.float64, device=dev)
        assert phi.is_contiguous() and phi.shape == (nbox, 2 * NLIN) and phi.dtype == torch.float64
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        p = lambda t, *i: C.c_void_p(t[i].data_ptr()) if i else C.c_void_p(t.data_ptr())
        rc = _lib.load().ccpol3b_eval_batch_grad_device(
            self._ctx, nbox, ms.ctypes.data_as(C.POINTER(C.c_int)), _dp(bl), ld,
            p(com, 0), p(com, 1), p(com, 2), p(site, 0), p(site, 1), p(site, 2),
            p(fsite, 0), p(fsite, 1), p(fsite, 2), p(fcom2), p(fcom3), p(energies), p(phi), C.c_void_p(stream))
        _lib.check(rc, self._ctx, "ccpol3b_eval_batch_grad_device")
        return fsite, fcom2, fcom3, energies, phi

    def set_coefficients(self, cs_sr, cs_lr):
        """Replace the 364 SR and 364 LR linear coefficients of this context."""
        a = np.ascontiguousarray(cs_sr, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(cs_lr, dtype=np.float64).reshape(-1)
        assert a.size == NLIN and b.size == NLIN
        _lib.check(_lib.load().ccpol3b_set_coefficients(self._ctx, _dp(a), _dp(b)), self._ctx,
                   "ccpol3b_set_coefficients")
        self._coef_ref, self._coef_version = None, None  # the last tensor loaded by ThreeBodyEnergy

    def gram(self, phi, stream=None):
        """G = phi^T phi (float64 [ncol, ncol]) of a float64 cuda matrix phi [nrow, ncol]
        on the FP64 matrix cores."""
        import torch
        assert phi.is_cuda and phi.dtype == torch.float64 and phi.dim() == 2
        phi = phi.contiguous()
        nrow, ncol = phi.shape
        g = torch.empty((ncol, ncol), dtype=torch.float64, device=phi.device)
        if stream is None:
            stream = torch.cuda.current_stream(phi.device).cuda_stream
        rc = _lib.load().ccpol3b_gram_device(self._ctx, nrow, ncol, C.c_void_p(phi.data_ptr()), ncol,
                                             C.c_void_p(g.data_ptr()), ncol, C.c_void_p(stream))
        _lib.check(rc, self._ctx, "ccpol3b_gram_device")
        return g

    # ------------------------------------------------------------------ launch shapes
    def get_tuning(self) -> dict:
        t = _lib.Tuning()
        _lib.check(_lib.load().ccpol3b_get_tuning(self._ctx, C.byref(t)), self._ctx)
        return t.as_dict()

    def set_tuning(self, **kw):
        t = _lib.Tuning(**{**self.get_tuning(), **kw})
        _lib.check(_lib.load().ccpol3b_set_tuning(self._ctx, C.byref(t)), self._ctx, "ccpol3b_set_tuning")

    def autotune(self, mol_start, box_len, com, site, reps=3, stream=None):
        """Time the candidate launch shapes on this batch and keep the fastest.
        Returns (tuning dict, ms per evaluation)."""
        import torch
        ms = np.ascontiguousarray(mol_start, dtype=np.int32)
        bl = np.ascontiguousarray(box_len, dtype=np.float64)
        nbox, ld = len(bl), site.shape[2]
        if stream is None:
            stream = torch.cuda.current_stream(com.device).cuda_stream
        p = lambda t, i: C.c_void_p(t[i].data_ptr())
        t = _lib.Tuning()
        best = C.c_double(0.0)
        rc = _lib.load().ccpol3b_autotune(self._ctx, nbox, ms.ctypes.data_as(C.POINTER(C.c_int)), _dp(bl), ld,
                                          p(com, 0), p(com, 1), p(com, 2), p(site, 0), p(site, 1), p(site, 2),
                                          int(reps), C.c_void_p(stream), C.byref(t), C.byref(best))
        _lib.check(rc, self._ctx, "ccpol3b_autotune")
        return t.as_dict(), best.value

    def graph(self, mol_start, box_len, com, site, phi=False, headroom=0.25, stream=None):
        """Record the batched evaluation of these tensors into a HIP graph (BatchGraph)."""
        return BatchGraph(self, mol_start, box_len, com, site, phi, headroom, stream)

    # ------------------------------------------------------------------ introspection
    def stats(self) -> dict:
        s = _lib.Stats()
        _lib.check(_lib.load().ccpol3b_last_stats(self._ctx, C.byref(s)), self._ctx)
        return s.as_dict()

    def fs3_list(self) -> np.ndarray:
        L = _lib.load()
        n = L.ccpol3b_get_fs3_list(self._ctx, None, 0)
        out = np.zeros((max(n, 1), 3), dtype=np.int32)
        L.ccpol3b_get_fs3_list(self._ctx, out.ctypes.data_as(C.POINTER(C.c_int)), n)
        return out[:n]

    def fs2_list(self) -> np.ndarray:
        L = _lib.load()
        n = L.ccpol3b_get_fs2_list(self._ctx, None, 0)
        out = np.zeros((max(n, 1), 3), dtype=np.int32)
        L.ccpol3b_get_fs2_list(self._ctx, out.ctypes.data_as(C.POINTER(C.c_int)), n)
        return out[:n]


class BatchGraph:
    """A batched evaluation (ThreeBody.eval_batch_device semantics) recorded into a
    HIP graph for fixed tensors; ``launch()`` replays it with one launch and no
    host synchronisation.  Owns its output tensors: fsite [3, nm, ld] (site
    forces, accumulated by every replay -- zero it between replays if needed),
    fcom2, fcom3 [3, nm], energies [nbox, 3] and, with phi=True, phi [nbox, 728]."""

    def __init__(self, tb, mol_start, box_len, com, site, phi, headroom, stream):
        import torch
        self.ms = np.ascontiguousarray(mol_start, dtype=np.int32)
        self.bl = np.ascontiguousarray(box_len, dtype=np.float64)
        nbox, nm, ld = len(self.bl), com.shape[1], site.shape[2]
        dev = com.device
        self.com, self.site = com, site  # inputs read by every replay
        self.fsite = torch.zeros((3, nm, ld), dtype=torch.float64, device=dev)
        self.fcom2 = torch.zeros((3, nm), dtype=torch.float64, device=dev)
        self.fcom3 = torch.zeros((3, nm), dtype=torch.float64, device=dev)
        self.energies = torch.zeros((nbox, 3), dtype=torch.float64, device=dev)
        self.phi = torch.zeros((nbox, 2 * NLIN), dtype=torch.float64, device=dev) if phi else None
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        p = lambda t, *i: C.c_void_p(t[i].data_ptr()) if i else C.c_void_p(t.data_ptr())
        self._g = C.c_void_p()
        self._L = _lib.load()
        rc = self._L.ccpol3b_graph_create(
            tb._ctx, nbox, self.ms.ctypes.data_as(C.POINTER(C.c_int)), _dp(self.bl), ld,
            p(com, 0), p(com, 1), p(com, 2), p(site, 0), p(site, 1), p(site, 2),
            p(self.fsite, 0), p(self.fsite, 1), p(self.fsite, 2), p(self.fcom2), p(self.fcom3), p(self.energies),
            p(self.phi) if phi else None, float(headroom), C.c_void_p(stream), C.byref(self._g))
        _lib.check(rc, tb._ctx, "ccpol3b_graph_create")

    def launch(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(self.com.device).cuda_stream
        _lib.check(self._L.ccpol3b_graph_launch(self._g, C.c_void_p(stream)), None, "ccpol3b_graph_launch")

    def status(self) -> int:
        """0, or ECAPACITY when the last replay outgrew the recorded capacities
        (read after waiting for the stream)."""
        return self._L.ccpol3b_graph_status(self._g)

    def close(self):
        if self._g:
            self._L.ccpol3b_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def orbit_index(is_, js, ks, ip, jp, kp) -> int:
    """ind_lin_2 lookup (1-based sites, powers 0/1) through the library."""
    return _lib.load().ccpol3b_orbit_index(is_, js, ks, ip, jp, kp)


def read_data3b_c(path):
    """The C reader of the library (same result as params.read_data3b)."""
    L = _lib.load()
    b, g, r = np.zeros(289), np.zeros(289), np.zeros(289)
    cs = np.zeros(5000)
    n = C.c_int(5000)
    _lib.check(L.ccpol3b_read_data3b(str(path).encode(), _dp(b), _dp(g), _dp(r), 17, _dp(cs), C.byref(n)))
    f = lambda a: a.reshape(17, 17).T.copy()  # Fortran (is,js) -> [is, js]
    return f(b), f(g), f(r), cs[:n.value].copy()
