"""Rigid-water NVE on the device: the library's velocity-Verlet + NO_SQUISH
driver (ccpol3b_md_*, the rigid-body steps of nve_1 in evolve.f90:621-1062 and
nosquish, :8107-8204) with the three-body forces, and the molecular constants
it needs (principal-frame sites, mass and principal moments from FIELD,
initialization.f90:710-845).

Units are DL_POLY's: A, ps, amu, 10 J/mol; energies reported in kcal/mol."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .geometry import principal_coords

BOLTZ = 0.831451115  # 10 J mol^-1 K^-1 (common_arrays.f90:29)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def water_constants(path=None):
    """(pcoord [ns,3] A, mass amu, pmi [3] amu A^2) of the rigid model in FIELD."""
    _, pc, mass = principal_coords(path)
    totm = float(mass.sum())
    r2 = (pc ** 2).sum(1)
    pmi = np.array([(mass * (r2 - pc[:, a] ** 2)).sum() for a in range(3)])
    return np.ascontiguousarray(pc), totm, pmi


def maxwell_state(nm: int, temperature: float, mass: float, pmi, seed: int):
    """COM velocities [nm,3] (A/ps, zero total momentum) and body-frame angular
    momenta [nm,3] (amu A^2/ps) drawn from the Maxwell-Boltzmann distribution."""
    rng = np.random.default_rng(seed)
    kT = BOLTZ * temperature
    v = rng.normal(size=(nm, 3)) * np.sqrt(kT / mass)
    v -= v.mean(0)
    j = rng.normal(size=(nm, 3)) * np.sqrt(kT * np.asarray(pmi))[None, :]
    return v, j


class RigidMD:
    """NVE trajectory of rigid waters under the three-body term of ``tb``
    (a ThreeBody; its precision is used)."""

    def __init__(self, tb, com, quat, box, vel, jbody, constants=None):
        pc, mass, pmi = constants if constants is not None else water_constants()
        self.tb, self.nm, self.box = tb, len(com), float(box)
        self.pc = np.ascontiguousarray(pc, dtype=np.float64)
        self.mass, self.pmi = float(mass), np.ascontiguousarray(pmi, dtype=np.float64)
        self._L = _lib.load()
        self._md = C.c_void_p()
        rc = self._L.ccpol3b_md_create(tb._ctx, self.nm, self.box, len(self.pc), _dp(self.pc), self.mass,
                                       _dp(self.pmi), C.byref(self._md))
        _lib.check(rc, tb._ctx, "ccpol3b_md_create")
        self.set_state(com, quat, vel, jbody)

    def set_state(self, com, quat, vel, jbody):
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (com, quat, vel, jbody)]
        _lib.check(self._L.ccpol3b_md_set_state(self._md, *[_dp(a) for a in arr]), None, "ccpol3b_md_set_state")

    def state(self) -> dict:
        out = {"com": np.zeros((self.nm, 3)), "quat": np.zeros((self.nm, 4)), "vel": np.zeros((self.nm, 3)),
               "jbody": np.zeros((self.nm, 3))}
        _lib.check(self._L.ccpol3b_md_get_state(self._md, _dp(out["com"]), _dp(out["quat"]), _dp(out["vel"]),
                                                _dp(out["jbody"])), None, "ccpol3b_md_get_state")
        return out

    def run(self, nsteps: int, dt: float) -> np.ndarray:
        """Advance nsteps of dt (ps); returns [nsteps, 4] = (E3, Kt, Kr, total) in kcal/mol."""
        rec = np.zeros((max(nsteps, 1), 4))
        _lib.check(self._L.ccpol3b_md_run(self._md, int(nsteps), float(dt), _dp(rec)), self.tb._ctx,
                   "ccpol3b_md_run")
        return rec[:nsteps]

    @classmethod
    def from_revcon(cls, tb, path, nm=None, constants=None):
        """Restart from a REVCON file as the reference does (initialization.f90:55-67):
        COM mapped into the box, velocities, body-frame angular momenta, quaternions."""
        from .formats import read_revcon
        r = read_revcon(path, nm=nm)
        return cls(tb, r.com, r.quat, r.box, r.vel, r.jbody, constants=constants)

    def write_revcon(self, path, step: int, time: float, chit: float = 0.0, chip: float = 0.0):
        """Write the current state in the reference's REVCON layout (md.f90:266-275)."""
        from .formats import Revcon, write_revcon
        st = self.state()
        write_revcon(path, Revcon(step=int(step), time=float(time), box=self.box, chit=chit, chip=chip,
                                  com=st["com"], vel=st["vel"], jbody=st["jbody"], quat=st["quat"]))

    def close(self):
        if self._md:
            self._L.ccpol3b_md_destroy(self._md)
            self._md = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
