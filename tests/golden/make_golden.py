"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference ships no numeric three-body vectors: its only test
(source/tests/check_fs2_periodic.py) checks FS2 through invariances and finite
differences, and nothing pins FS3 values.  These fixtures are therefore frozen
outputs of the oracle (the C restatement under oracle/, one thread, the
reference's loop and accumulation order).  They pin the oracle against
regressions and give the GPU path a committed target that does not depend on
rebuilding the oracle.  Inputs are stored explicitly (COM positions, site
offsets of sites 1-17, box), so the fixtures do not depend on the geometry
generators either.

Cases:
  box64_fs3      64 waters, L = 12.42 A, FS3-only parameter set (switches T T F)
  box64_fs2fs3   64 waters, L = 12.42 A, FS2+FS3 parameter set
  trimers_fs2fs3 12 isolated trimers (open boundary): three equilateral ones at
                 R_avg = 4.0 / 5.0 / 5.9 A (SR, switch band, LR) and nine with
                 R_avg ~ U(2.6, 6.2) A (most inside both cutoffs)

Per case: com [nm,3], site [nm,17,3] (A), box (A; <= 0 open); the outputs
qfs2, qfs3 (kcal/mol), vir (10 J/mol), fsite [nm,17,3], fcom2, fcom3 [nm,3]
(10 J/mol/A); the FS3 triplet list (int16 rows i,j,k) and the FS2 item count
with the SHA-256 of its int32 (k,i,j) rows.  Trimer cases stack nt trimers
along a leading axis.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

NS = 17  # sites 1-17 carry the three-body terms


def flat_arrays(com, site, ld=NS):
    """(x, y, z, xs, ys, zs) in the reference layout (xs[is + ld*i])."""
    nm = len(com)
    buf = np.zeros((3, nm, ld))
    buf[:, :, :site.shape[1]] = site.transpose(2, 0, 1)
    return (np.ascontiguousarray(com[:, 0]), np.ascontiguousarray(com[:, 1]), np.ascontiguousarray(com[:, 2]),
            buf[0].reshape(-1).copy(), buf[1].reshape(-1).copy(), buf[2].reshape(-1).copy())


def list_digest(rows: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(rows, dtype=np.int32).tobytes()).hexdigest()


def oracle_case(O, p, com, site, box):
    x, y, z, xs, ys, zs = flat_arrays(com, site)
    r = O.threebody_term(p, x=x, y=y, z=z, xs=xs, ys=ys, zs=zs, box=box, ld=NS, nthreads=1)
    fsite = np.stack([r["fxs"], r["fys"], r["fzs"]], -1)  # [nm, ld, 3]
    out = {
        "qfs2": r["qfs2"], "qfs3": r["qfs3"], "vir": r["vir"], "fsite": fsite[:, :NS],
        "fcom2": np.stack([r["fxfs2"], r["fyfs2"], r["fzfs2"]], -1),
        "fcom3": np.stack([r["fxfs3"], r["fyfs3"], r["fzfs3"]], -1),
        "hist": np.array([r["n_fs3_sr"], r["n_fs3_sw"], r["n_fs3_lr"]], dtype=np.int64),
    }
    l3 = O.fs3_list(p, x, y, z, box)
    l2 = O.fs2_list(p, x, y, z, box) if p.s2_term else np.zeros((0, 3), np.int32)
    assert len(l3) == r["n_fs3"] and (not p.s2_term or len(l2) == r["n_fs2"])
    out["fs3_list"] = l3.astype(np.int16)
    out["n_fs2"] = np.int64(len(l2))
    out["fs2_sha256"] = np.array(list_digest(l2))
    return out


def box_case(O, p, wb):
    out = oracle_case(O, p, wb.com, wb.site[:, :NS], wb.box)
    out.update(com=wb.com, site=wb.site[:, :NS], box=np.float64(wb.box))
    return out


def trimer_case(O, p, trimers):
    per = [oracle_case(O, p, t.com, t.site[:, :NS], t.box) for t in trimers]
    out = {k: np.stack([q[k] for q in per]) for k in ("qfs2", "qfs3", "vir", "fsite", "fcom2", "fcom3", "hist",
                                                     "n_fs2")}
    out["fs3_count"] = np.array([len(q["fs3_list"]) for q in per])
    out["com"] = np.stack([t.com for t in trimers])
    out["site"] = np.stack([t.site[:, :NS] for t in trimers])
    out["box"] = np.zeros(len(trimers))
    return out


def equilateral_trimers(seed=23):
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox, random_quaternions
    rng = np.random.default_rng(seed)
    out = []
    for d in (4.0, 5.0, 5.9):
        tri = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0], [d / 2, d * np.sqrt(3) / 2, 0.0]])
        out.append(WaterBox(tri - tri.mean(0), random_quaternions(3, rng), 0.0))
    return out


def main():
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box, trimer_set
    from ccpol23_plus_2025_mbmd_amd.params import load_variant
    from oracle import oracle as O

    fs3 = load_variant("fs3").replace(s2_term=False)
    fs2fs3 = load_variant("fs2fs3")
    cases = {
        "box64_fs3": box_case(O, fs3, lattice_box(64, seed=1)),
        "box64_fs2fs3": box_case(O, fs2fs3, lattice_box(64, seed=5)),
        "trimers_fs2fs3": trimer_case(O, fs2fs3, equilateral_trimers()
                                      + trimer_set(9, seed=17, ravg_lo=2.6, ravg_hi=6.2)),
    }
    for name, arrays in cases.items():
        path = HERE / f"{name}.npz"
        np.savez_compressed(path, **arrays)
        print(f"{path.name}: {path.stat().st_size} bytes")


if __name__ == "__main__":
    main()
