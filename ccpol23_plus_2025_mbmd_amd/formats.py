"""Writers and readers of the reference's on-disk formats beyond the parameter
readers of ``params.py``, and a binary parameter cache.

* ``data3b_full*`` writing (the refit's output format, read back by
  initialization.f90:660-701): the same records as the shipped files, each
  value in the layout of the reference's list-directed ``write(*,*)``: a
  written file is byte-identical to a shipped one with the same content (the
  shipped files only lack the final newline).
* ``REVCON`` (md.f90:266-275 writes it, initialization.f90:55-67 reads it on
  restart): a title record; ``step t box chit chip``; then one record per
  molecule ``x y z vx vy vz jx jy jz q0 q1 q2 q3`` (A, A/ps, body-frame
  angular momentum in amu A^2/ps, quaternion).  On restart the reference maps
  each COM component into the box with ``x - box*nint(x/box)``.
* A binary cache of a parsed parameter set (``.npz``), keyed by the SHA-256 of
  the text files it came from, so a stale cache is never used.
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
from pathlib import Path

import numpy as np

from .params import NLIN, NS3, Data3B, ThreeBodyParams, _fortran_float, _tokens


# ---------------------------------------------------------------- list-directed layout
def _real_field(x: float) -> str:
    """One REAL(8) item of a list-directed record: a separating blank, then 25
    characters -- 17 significant digits, fixed notation right-aligned in 20
    columns plus 5 blanks when 0.1 <= |x| < 1e17, otherwise d.dddE+eee."""
    x = float(x)
    if x == 0.0:
        return " " + "0.0000000000000000".rjust(20) + " " * 5
    mant, exp = f"{x:.16e}".split("e")  # 17 significant digits, correctly rounded
    k = int(exp)                          # |x| rounded = m 10^k, 1 <= m < 10
    if -1 <= k <= 16:
        s = f"{x:.{16 - k}f}"             # k + 1 digits before the point, 17 in all
        if "." not in s:
            s += "."
        return " " + s.rjust(20) + " " * 5
    return " " + f"{mant}E{'-' if k < 0 else '+'}{abs(k):03d}".rjust(25)


def _int_field(n: int) -> str:
    """One default INTEGER item of a list-directed record (a blank and 11 columns)."""
    return " " + f"{int(n):d}".rjust(11)


def write_data3b(path: str | os.PathLike, data: Data3B, header: str = " beta        gamma       r0  ",
                 label: str = "linear parameters") -> None:
    """Write one ``data3b_full*`` table: the header record, 17*17 records
    ``beta gamma r0 is js`` (is-major, 1-based), the label record, ``nlin`` and
    ``nlin`` records ``index coefficient``."""
    cs = np.asarray(data.cs, dtype=np.float64)
    if cs.shape != (NLIN,):
        raise ValueError(f"expected {NLIN} linear coefficients, got {cs.shape}")
    out = [header]
    for is_ in range(NS3):
        for js in range(NS3):
            out.append(_real_field(data.beta[is_, js]) + _real_field(data.gamma[is_, js])
                       + _real_field(data.r0[is_, js]) + _int_field(is_ + 1) + _int_field(js + 1))
    out.append(label)
    out.append(_int_field(NLIN))
    for n in range(NLIN):
        out.append(_int_field(n + 1) + _real_field(cs[n]))
    Path(path).write_text("\n".join(out) + "\n")


# ---------------------------------------------------------------- REVCON
@dataclasses.dataclass
class Revcon:
    """Restart state of a rigid-water run (md.f90:266-275)."""

    step: int
    time: float      # ps
    box: float       # A
    chit: float      # thermostat variable (0 for NVE)
    chip: float      # barostat variable (0 for NVE/NVT)
    com: np.ndarray    # [nm, 3] A
    vel: np.ndarray    # [nm, 3] A/ps
    jbody: np.ndarray  # [nm, 3] amu A^2/ps, body frame
    quat: np.ndarray   # [nm, 4] (q0, q1, q2, q3)
    title: str = "Created by Ommair"

    @property
    def nm(self) -> int:
        return len(self.com)


def _nint(x: np.ndarray) -> np.ndarray:
    """Fortran NINT: to nearest, halves away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def read_revcon(path: str | os.PathLike, nm: int | None = None, wrap: bool = True) -> Revcon:
    """Read a REVCON file as the reference's restart does (initialization.f90:55-67):
    the title record is skipped; ``nm`` molecules are read (all records when
    None); with ``wrap`` each COM component becomes ``x - box*nint(x/box)``.
    List-directed input is accepted: blanks or commas between items, D or E
    exponents, items continued over several lines."""
    lines = Path(path).read_text().splitlines()
    items = []
    for ln in lines[1:]:
        items.extend(_tokens(ln))
    step = int(float(_fortran_float(items[0])))
    time, box, chit, chip = (_fortran_float(v) for v in items[1:5])
    vals = np.array([_fortran_float(v) for v in items[5:]], dtype=np.float64)
    if nm is None:
        if len(vals) % 13:
            raise ValueError(f"{path}: {len(vals)} values after the header are not whole records of 13")
        nm = len(vals) // 13
    if len(vals) < 13 * nm:
        raise ValueError(f"{path}: {nm} molecules need {13 * nm} values, found {len(vals)}")
    rec = vals[:13 * nm].reshape(nm, 13)
    com = rec[:, 0:3].copy()
    if wrap and box > 0.0:
        com = com - box * _nint(com / box)
    return Revcon(step=step, time=time, box=box, chit=chit, chip=chip, com=com, vel=rec[:, 3:6].copy(),
                  jbody=rec[:, 6:9].copy(), quat=rec[:, 9:13].copy(), title=lines[0].strip())


def write_revcon(path: str | os.PathLike, r: Revcon) -> None:
    """Write ``r`` in the reference's REVCON layout (list-directed records)."""
    out = [" " + r.title, _int_field(r.step) + "".join(_real_field(v) for v in (r.time, r.box, r.chit, r.chip))]
    for m in range(r.nm):
        row = np.concatenate([r.com[m], r.vel[m], r.jbody[m], r.quat[m]])
        out.append("".join(_real_field(v) for v in row))
    Path(path).write_text("\n".join(out) + "\n")


# ---------------------------------------------------------------- binary parameter cache
def _digest(paths) -> str:
    h = hashlib.sha256()
    for p in paths:
        h.update(Path(p).name.encode())
        h.update(Path(p).read_bytes())
    return h.hexdigest()


def variant_sources(data_dir: str | os.PathLike) -> list[Path]:
    d = Path(data_dir)
    return [d / "CONTROL", d / "data3b_full_SR_5p5", d / "data3b_full_LR_4p5"]


def save_params_cache(path: str | os.PathLike, p: ThreeBodyParams, sources) -> None:
    """Store a parsed parameter set with the SHA-256 of the files it was read from."""
    arrays = {f"{reg}_{k}": getattr(getattr(p, reg), k) for reg in ("sr", "lr") for k in ("beta", "gamma", "r0", "cs")}
    scal = np.array([p.alpha3, p.delta33, p.rcut3, p.r0_tilda], dtype=np.float64)
    flags = np.array([p.s2_term, p.s3_term, p.nonadditive_3b], dtype=np.int8)
    with open(path, "wb") as f:  # np.savez would append .npz to a str path without it
        np.savez(f, digest=np.array(_digest(sources)), scalars=scal, flags=flags, **arrays)


def load_params_cache(path: str | os.PathLike, sources) -> ThreeBodyParams | None:
    """The cached parameter set, or None when the cache is missing or was made
    from other file contents (no pickled objects are read)."""
    path = Path(path)
    if not path.exists():
        return None
    with np.load(path, allow_pickle=False) as z:
        if str(z["digest"]) != _digest(sources):
            return None
        regs = {reg: Data3B(beta=z[f"{reg}_beta"], gamma=z[f"{reg}_gamma"], r0=z[f"{reg}_r0"], cs=z[f"{reg}_cs"],
                            path=f"{path}:{reg}") for reg in ("sr", "lr")}
        alpha3, delta33, rcut3, r0_tilda = (float(v) for v in z["scalars"])
        s2, s3, nonadd = (bool(v) for v in z["flags"])
    return ThreeBodyParams(sr=regs["sr"], lr=regs["lr"], alpha3=alpha3, delta33=delta33, rcut3=rcut3,
                           r0_tilda=r0_tilda, s2_term=s2, s3_term=s3, nonadditive_3b=nonadd)


def load_variant_cached(variant: str, cache_dir: str | os.PathLike, data_dir: str | os.PathLike | None = None):
    """load_variant through a binary cache in ``cache_dir`` (written on a miss)."""
    from .params import DATA_DIR, load_variant
    d = Path(data_dir) if data_dir is not None else DATA_DIR / variant
    src = variant_sources(d)
    cache = Path(cache_dir) / f"ccpol3b_params_{variant}.npz"
    p = load_params_cache(cache, src)
    if p is None:
        p = load_variant(variant, d)
        cache.parent.mkdir(parents=True, exist_ok=True)
        save_params_cache(cache, p, src)
    return p
