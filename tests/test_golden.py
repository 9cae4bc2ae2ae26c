"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the FP64 oracle with one thread, in the reference's loop order).

CPU: the oracle, rebuilt from oracle/, still reproduces them (lists and list
digests exactly, energies to 1e-12 relative, forces to 1e-11 relative RMS --
the multi-threaded run only reorders sums).
GPU: the HIP path, through the C ABI, matches them at the parity bar: FP64
energies within 1e-10 relative (plus a 2e-12 kcal/mol floor), site forces and
the full force output within 1e-9 relative RMS, FS3 COM forces within 1e-7,
triplet lists bit-exact; the mixed path within 1e-5 relative force RMS."""
import hashlib
import os
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
NS = 17
NTHREADS = min(8, os.cpu_count() or 1)
BOXES = ["box64_fs3", "box64_fs2fs3"]
E_FLOOR = 2e-12


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def params_for(name):
    from ccpol23_plus_2025_mbmd_amd.params import load_variant
    if name.endswith("_fs3"):
        return load_variant("fs3").replace(s2_term=False)
    return load_variant("fs2fs3")


def flat_box(com, site, box):
    """An object with the WaterBox attributes ThreeBody.evaluate reads (ld = 17)."""
    nm = len(com)
    buf = np.ascontiguousarray(site.transpose(2, 0, 1)).reshape(3, -1)
    return SimpleNamespace(nm=nm, ld=NS, x=np.ascontiguousarray(com[:, 0]), y=np.ascontiguousarray(com[:, 1]),
                           z=np.ascontiguousarray(com[:, 2]), xs=buf[0].copy(), ys=buf[1].copy(),
                           zs=buf[2].copy(), box=float(box))


def rel_rms(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    den = np.sqrt((b ** 2).sum())
    return np.sqrt(((a - b) ** 2).sum()) / den if den > 0 else np.sqrt((a ** 2).sum())


def outputs(r):
    """Oracle- or ThreeBody-style result dict -> fixture-style arrays."""
    fsite = np.stack([np.asarray(r["fxs"]).reshape(-1, NS), np.asarray(r["fys"]).reshape(-1, NS),
                      np.asarray(r["fzs"]).reshape(-1, NS)], -1)
    return {"qfs2": r["qfs2"], "qfs3": r["qfs3"], "vir": r["vir"], "fsite": fsite,
            "fcom2": np.stack([r["fxfs2"], r["fyfs2"], r["fzfs2"]], -1),
            "fcom3": np.stack([r["fxfs3"], r["fyfs3"], r["fzfs3"]], -1)}


def check(got, ref, e_tol, f_tol, com_tol):
    print("errors: " + ", ".join(
        f"{k} {abs(got[k] - float(ref[k])) / max(abs(float(ref[k])), 1e-300):.2e}" for k in ("qfs2", "qfs3", "vir"))
        + f", fsite {rel_rms(got['fsite'], ref['fsite']):.2e}, fcom2 {rel_rms(got['fcom2'], ref['fcom2']):.2e}"
        + f", fcom3 {rel_rms(got['fcom3'], ref['fcom3']):.2e}", flush=True)
    for k in ("qfs2", "qfs3"):
        assert abs(got[k] - float(ref[k])) <= e_tol * abs(float(ref[k])) + E_FLOOR, (k, got[k], float(ref[k]))
    assert abs(got["vir"] - float(ref["vir"])) <= 10 * e_tol * abs(float(ref["vir"])) + 418.4 * E_FLOOR
    assert rel_rms(got["fsite"], ref["fsite"]) <= f_tol
    full = lambda o: np.concatenate([o["fsite"].ravel(), o["fcom2"].ravel(), o["fcom3"].ravel()])
    assert rel_rms(full(got), full(ref)) <= f_tol
    assert rel_rms(got["fcom3"], ref["fcom3"]) <= com_tol
    assert rel_rms(got["fcom2"], ref["fcom2"]) <= com_tol


def digest(rows):
    return hashlib.sha256(np.ascontiguousarray(rows, dtype=np.int32).tobytes()).hexdigest()


# ------------------------------------------------------------------ CPU: the oracle
@pytest.mark.parametrize("name", BOXES)
def test_oracle_reproduces_golden_box(oracle, name):
    g = load(name)
    p = params_for(name)
    b = flat_box(g["com"], g["site"], g["box"])
    r = oracle.threebody_term(p, b, nthreads=NTHREADS)
    check(outputs(r), g, e_tol=1e-12, f_tol=1e-11, com_tol=1e-11)
    np.testing.assert_array_equal(oracle.fs3_list(p, b.x, b.y, b.z, b.box), g["fs3_list"].astype(np.int32))
    assert [r["n_fs3_sr"], r["n_fs3_sw"], r["n_fs3_lr"]] == list(g["hist"])
    if p.s2_term:
        l2 = oracle.fs2_list(p, b.x, b.y, b.z, b.box)
        assert len(l2) == int(g["n_fs2"]) and digest(l2) == str(g["fs2_sha256"])


def test_oracle_reproduces_golden_trimers(oracle):
    g = load("trimers_fs2fs3")
    p = params_for("trimers_fs2fs3")
    assert set(np.nonzero(g["hist"].sum(0))[0]) == {0, 1, 2}  # SR, switch band and LR all present
    for t in range(len(g["com"])):
        r = oracle.threebody_term(p, flat_box(g["com"][t], g["site"][t], 0.0))
        ref = {k: g[k][t] for k in ("qfs2", "qfs3", "vir", "fsite", "fcom2", "fcom3")}
        check(outputs(r), ref, e_tol=1e-12, f_tol=1e-11, com_tol=1e-11)
        assert r["n_fs3"] == g["fs3_count"][t] and r["n_fs2"] == g["n_fs2"][t]


# ------------------------------------------------------------------ GPU: the HIP path
@pytest.fixture(scope="module")
def golden_tb():
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    made = {}

    def get(name, precision):
        key = (params_for(name).s2_term, precision)
        if key not in made:
            made[key] = ThreeBody(params_for(name), precision)
        return made[key]
    return get


TOLS = {"fp64": dict(e_tol=1e-10, f_tol=1e-9, com_tol=1e-9),
        # mixed: 1e-5 holds for the site forces and the full force output; the
        # energies and the COM forces derived from them (f_d' E + f_d w' E) do
        # not meet it -- the short-range triplet energies cancel ~1e4-fold, so
        # FP32 rounding costs ~1e-4 of them (tests/test_mixed_limits.py).  The
        # energy and COM tolerances are the measured envelope on these boxes
        # (COM 1.6e-3 on the 64-water box with the FS3-only parameter set).
        "mixed": dict(e_tol=3e-3, f_tol=1e-5, com_tol=4e-3)}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", BOXES)
def test_gpu_matches_golden_box(golden_tb, name, precision):
    g = load(name)
    print(f"golden {name} {precision}", flush=True)
    tb = golden_tb(name, precision)
    got = tb.evaluate(flat_box(g["com"], g["site"], g["box"]))
    check(outputs(got), g, **TOLS[precision])
    np.testing.assert_array_equal(tb.fs3_list(), g["fs3_list"].astype(np.int32))
    assert [got["n_fs3_sr"], got["n_fs3_switch"], got["n_fs3_lr"]] == list(g["hist"])
    if params_for(name).s2_term:
        l2 = tb.fs2_list()
        assert len(l2) == int(g["n_fs2"]) and digest(l2) == str(g["fs2_sha256"])


@pytest.mark.gpu
def test_gpu_matches_golden_trimers_fp64(golden_tb):
    g = load("trimers_fs2fs3")
    print("golden trimers fp64", flush=True)
    tb = golden_tb("trimers_fs2fs3", "fp64")
    for t in range(len(g["com"])):
        got = tb.evaluate(flat_box(g["com"][t], g["site"][t], 0.0))
        ref = {k: g[k][t] for k in ("qfs2", "qfs3", "vir", "fsite", "fcom2", "fcom3")}
        check(outputs(got), ref, **TOLS["fp64"])
        assert got["n_fs3"] == g["fs3_count"][t] and got["n_fs2_items"] == g["n_fs2"][t]


@pytest.mark.gpu
def test_gpu_matches_golden_trimers_mixed(golden_tb):
    """Mixed path on the trimer set, held to its bar on the force output of the
    whole set (1e-5 relative RMS) and to 3e-3 on the summed FS3 energy
    magnitudes.  The virial of one isolated trimer is not checked here: its
    terms r dE/dr cancel by up to ~1e5, so FP32 rounding of the terms moves it
    by up to ~10 % (the periodic boxes above check it)."""
    g = load("trimers_fs2fs3")
    print("golden trimers mixed", flush=True)
    tb = golden_tb("trimers_fs2fs3", "mixed")
    full = lambda o: np.concatenate([o["fsite"].ravel(), o["fcom2"].ravel(), o["fcom3"].ravel()])
    fg, fr, de, e3 = [], [], 0.0, 0.0
    for t in range(len(g["com"])):
        got = outputs(tb.evaluate(flat_box(g["com"][t], g["site"][t], 0.0)))
        ref = {k: g[k][t] for k in ("qfs2", "qfs3", "vir", "fsite", "fcom2", "fcom3")}
        fg.append(full(got))
        fr.append(full(ref))
        de += abs(got["qfs3"] - float(ref["qfs3"]))
        e3 += abs(float(ref["qfs3"]))
        print(f"trimer {t}: force {rel_rms(full(got), full(ref)):.2e} qfs3 {got['qfs3']:.6e} / "
              f"{float(ref['qfs3']):.6e} vir {got['vir']:.6e} / {float(ref['vir']):.6e}")
    err_f = rel_rms(np.concatenate(fg), np.concatenate(fr))
    print(f"set: force {err_f:.2e}, |dE3| / |E3| {de / e3:.2e}")
    assert err_f <= 1e-5
    assert de <= 3e-3 * e3
