"""Readers for the reference's on-disk three-body parameter formats.

* ``data3b_full*`` files (one per region, SR or LR):
  line 1 header; 17*17 records ``beta gamma r0 is js`` in ``is``-major order;
  one label record (its text is ignored, as the reference's ``read(40,*)``
  ignores it -- one shipped file spells it "linaer"); ``nlin0``; then
  ``nlin0`` records ``index coefficient`` whose coefficients are taken in
  record order (initialization.f90:660-701).
* ``CONTROL`` records 14-15: ``nonadditive_3B S3_term S2_term`` and
  ``rcut3 alpha3 r0_tilda delta33`` (initialization.f90:482-483).
* ``FIELD`` site table (initialization.f90:498-508).

The staging rule of the reference README (README.md:134-140) is kept: the SR
table the routine reads as ``data3b_full`` is the shipped
``data3b_full_SR_5p5``; the LR table is ``data3b_full_LR_4p5``.
"""
from __future__ import annotations

import dataclasses
import os
from pathlib import Path

import numpy as np

NS3 = 17
NLIN = 364

DATA_DIR = Path(__file__).resolve().parent / "data"


@dataclasses.dataclass
class Data3B:
    """One ``data3b_full*`` table: 17x17 (beta, gamma, r0) and the linear coefficients."""

    beta: np.ndarray   # [17,17] bohr^-1, index [is, js] (0-based)
    gamma: np.ndarray  # [17,17] bohr^-1
    r0: np.ndarray     # [17,17] bohr
    cs: np.ndarray     # [nlin] linear coefficients (hartree)
    path: str = ""


@dataclasses.dataclass
class ThreeBodyParams:
    """Everything the three-body routine reads from module globals
    (threebody_potential.f90:17-99 via common_arrays.f90:201-218)."""

    sr: Data3B
    lr: Data3B
    alpha3: float = 6.51292     # 1/A
    delta33: float = 10.0
    rcut3: float = 7.5          # A, FS2 COM cutoff
    r0_tilda: float = 7.40446   # A, read but not used by the routine
    s2_term: bool = True
    s3_term: bool = True
    nonadditive_3b: bool = True

    def replace(self, **kw) -> "ThreeBodyParams":
        return dataclasses.replace(self, **kw)


def _tokens(line: str) -> list[str]:
    return line.replace(",", " ").split()


def _fortran_float(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def read_data3b(path: str | os.PathLike) -> Data3B:
    """Read one ``data3b_full*`` file (initialization.f90:660-680)."""
    lines = [ln for ln in Path(path).read_text().splitlines()]
    pos = 1  # header record skipped
    beta = np.zeros((NS3, NS3))
    gamma = np.zeros((NS3, NS3))
    r0 = np.zeros((NS3, NS3))
    for is_ in range(NS3):
        for js in range(NS3):
            while not lines[pos].strip():
                pos += 1
            t = _tokens(lines[pos])
            pos += 1
            beta[is_, js] = _fortran_float(t[0])
            gamma[is_, js] = _fortran_float(t[1])
            r0[is_, js] = _fortran_float(t[2])
    pos += 1  # label record ("linear parameters"), content ignored
    while not lines[pos].strip():
        pos += 1
    nlin0 = int(_tokens(lines[pos])[0])
    pos += 1
    cs = np.zeros(nlin0)
    for n in range(nlin0):
        while not lines[pos].strip():
            pos += 1
        t = _tokens(lines[pos])
        pos += 1
        cs[n] = _fortran_float(t[1])
    if nlin0 != NLIN:
        raise ValueError(f"{path}: nlin0={nlin0}, the FS3 orbit map has {NLIN} coefficients")
    return Data3B(beta=beta, gamma=gamma, r0=r0, cs=cs, path=str(path))


def _logical(tok: str) -> bool:
    t = tok.strip().upper().lstrip(".")
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise ValueError(f"not a Fortran logical: {tok!r}")


def read_control_3b(path: str | os.PathLike) -> dict:
    """CONTROL records 14 and 15 (initialization.f90:482-483)."""
    recs = Path(path).read_text().splitlines()
    r14 = _tokens(recs[13])
    r15 = _tokens(recs[14])
    return dict(
        nonadditive_3b=_logical(r14[0]),
        s3_term=_logical(r14[1]),
        s2_term=_logical(r14[2]),
        rcut3=_fortran_float(r15[0]),
        alpha3=_fortran_float(r15[1]),
        r0_tilda=_fortran_float(r15[2]),
        delta33=_fortran_float(r15[3]),
    )


def read_field_sites(path: str | os.PathLike):
    """FIELD site table: names, coordinates (A), masses, charges (initialization.f90:498-508)."""
    recs = Path(path).read_text().splitlines()
    ns = int(_tokens(recs[2])[0])
    names, xyz, mass, charge = [], [], [], []
    for r in recs[5:5 + ns]:
        t = _tokens(r)
        names.append(t[0])
        xyz.append([_fortran_float(v) for v in t[1:4]])
        mass.append(_fortran_float(t[4]))
        charge.append(_fortran_float(t[5]))
    return names, np.array(xyz), np.array(mass), np.array(charge)


def load_variant(variant: str = "fs2fs3", data_dir: str | os.PathLike | None = None) -> ThreeBodyParams:
    """Parameters of one shipped fit, staged as the reference README prescribes.

    ``variant`` is ``"fs2fs3"`` (CONTROL switches T T T) or ``"fs3"`` (T T F).
    """
    d = Path(data_dir) if data_dir is not None else DATA_DIR / variant
    ctl = read_control_3b(d / "CONTROL")
    return ThreeBodyParams(
        sr=read_data3b(d / "data3b_full_SR_5p5"),
        lr=read_data3b(d / "data3b_full_LR_4p5"),
        alpha3=ctl["alpha3"], delta33=ctl["delta33"], rcut3=ctl["rcut3"],
        r0_tilda=ctl["r0_tilda"], s2_term=ctl["s2_term"] and ctl["nonadditive_3b"],
        s3_term=ctl["s3_term"] and ctl["nonadditive_3b"],
        nonadditive_3b=ctl["nonadditive_3b"],
    )


def field_path() -> Path:
    return DATA_DIR / "fs2fs3" / "FIELD"
