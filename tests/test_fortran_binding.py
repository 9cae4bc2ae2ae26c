"""The Fortran binding (include/ccpol3b_f.f90) driven by a Fortran program,
the way the reference's MD loop would call the three-body term."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
from ccpol23_plus_2025_mbmd_amd.params import DATA_DIR, load_variant

HERE = Path(__file__).resolve().parent
DRIVER = HERE / "fortran" / "_build" / "drive_box"


def write_input(path, p, box, variant, precision=0):
    d = DATA_DIR / variant
    with open(path, "w") as f:
        g = lambda v: f"{float(v):.17g}"
        f.write(f"{box.nm} 50 {g(box.box)}\n")
        f.write(f"{g(p.alpha3)} {g(p.delta33)} {g(p.rcut3)} {int(p.s2_term)} {int(p.s3_term)} {precision}\n")
        f.write(f"{d / 'data3b_full_SR_5p5'}\n{d / 'data3b_full_LR_4p5'}\n")
        for i in range(box.nm):
            f.write(f"{g(box.com[i, 0])} {g(box.com[i, 1])} {g(box.com[i, 2])}\n")
        xs = box.xs.reshape(box.nm, box.ld)
        ys = box.ys.reshape(box.nm, box.ld)
        zs = box.zs.reshape(box.nm, box.ld)
        for i in range(box.nm):
            for s in range(17):
                f.write(f"{g(xs[i, s])} {g(ys[i, s])} {g(zs[i, s])}\n")


def ensure_driver():
    if not DRIVER.exists():
        if shutil.which("/opt/rocm/bin/amdflang") is None:
            pytest.skip("amdflang not available to build the Fortran driver")
        subprocess.run(["make", "-C", str(HERE / "fortran")], check=True, capture_output=True)
    return DRIVER


def test_fortran_probe(tmp_path):
    """Parameter reader, orbit map and create() through the Fortran interface."""
    import torch
    from ccpol23_plus_2025_mbmd_amd.threebody import orbit_index

    drv = ensure_driver()
    p = load_variant("fs2fs3")
    box = lattice_box(8, seed=1)
    inp = tmp_path / "in.txt"
    write_input(inp, p, box, "fs2fs3")
    r = subprocess.run([str(drv), "probe", str(inp)], capture_output=True, text=True, timeout=60)
    lines = dict(l.split(None, 1) for l in r.stdout.strip().splitlines() if " " in l)
    orb = [int(v) for v in lines["orbits"].split()]
    assert orb == [orbit_index(1, 1, 1, 0, 0, 0), orbit_index(17, 17, 17, 1, 1, 1), orbit_index(2, 3, 1, 0, 1, 0)]
    c = [float(v.replace("D", "E")) for v in lines["coef"].split()]
    assert c[0] == p.sr.cs[0] and c[1] == p.lr.cs[363]
    expect = 0 if torch.cuda.is_available() else -5  # CCPOL3B_ENODEV without a GPU
    assert int(lines["create"]) == expect


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["fs2fs3", "fs3"])
def test_fortran_drop_in_matches_python(tmp_path, variant):
    """A 64-water box evaluated from Fortran equals the Python entry point's result."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody

    assert DRIVER.exists(), "tests/fortran/_build/drive_box is built by __graft_entry__.build()"
    p = load_variant(variant)
    box = lattice_box(64, seed=5)
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    write_input(inp, p, box, variant)
    print(f"fortran driver: {variant}, 64 waters, seed 5", flush=True)
    r = subprocess.run([str(DRIVER), "run", str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout, r.stderr)
    vals = np.loadtxt(out, max_rows=1)
    f = np.loadtxt(out, skiprows=1, max_rows=64 * 17)
    com = np.loadtxt(out, skiprows=1 + 64 * 17)
    ref = ThreeBody(p, "fp64").evaluate(box)
    for got, key in ((vals[0], "qfs2"), (vals[1], "qfs3")):
        assert abs(got - ref[key]) <= 1e-12 * abs(ref[key]) + 1e-14, (key, got, ref[key])
    assert abs(vals[2] - ref["vir"]) <= 1e-12 * abs(ref["vir"]) + 1e-12
    fr = np.stack([ref["fxs"][:, :17].reshape(-1), ref["fys"][:, :17].reshape(-1), ref["fzs"][:, :17].reshape(-1)], 1)
    np.testing.assert_allclose(f, fr, rtol=1e-10, atol=1e-10 * np.abs(fr).max())
    cr = np.stack([ref[k] for k in ("fxfs2", "fyfs2", "fzfs2", "fxfs3", "fyfs3", "fzfs3")], 1)
    np.testing.assert_allclose(com, cr, rtol=1e-10, atol=1e-10 * np.abs(cr).max())


MD_DRIVER = HERE / "fortran" / "_build" / "drive_md"


@pytest.mark.gpu
def test_fortran_md_loop_matches_python(tmp_path):
    """Ten 0.5 fs NVE steps of 64 waters run from a Fortran program through the
    ccpol3b_md_* bindings equal the same run through the Python driver."""
    from ccpol23_plus_2025_mbmd_amd.md import RigidMD, maxwell_state, water_constants
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody

    assert MD_DRIVER.exists(), "tests/fortran/_build/drive_md is built by __graft_entry__.build()"
    p = load_variant("fs2fs3")
    box = lattice_box(64, seed=11)
    pc, mass, pmi = water_constants()
    v, j = maxwell_state(box.nm, 298.0, mass, pmi, seed=4)
    nsteps, dt = 10, 5e-4
    d = DATA_DIR / "fs2fs3"
    g = lambda a: " ".join(f"{float(x):.17g}" for x in np.ravel(a))
    inp, out = tmp_path / "md_in.txt", tmp_path / "md_out.txt"
    with open(inp, "w") as f:
        f.write(f"{box.nm} {g([box.box])}\n")
        f.write(f"{g([p.alpha3, p.delta33, p.rcut3])} {int(p.s2_term)} {int(p.s3_term)} 0\n")
        f.write(f"{d / 'data3b_full_SR_5p5'}\n{d / 'data3b_full_LR_4p5'}\n")
        for arr in (box.com, box.quats, v, j):
            for row in arr:
                f.write(g(row) + "\n")
        f.write(f"{len(pc)}\n")
        for row in pc:
            f.write(g(row) + "\n")
        f.write(f"{g([mass])}\n{g(pmi)}\n{nsteps} {g([dt])}\n")
    print(f"fortran md loop: 64 waters, {nsteps} steps", flush=True)
    r = subprocess.run([str(MD_DRIVER), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout, r.stderr)
    rec_f = np.loadtxt(out, max_rows=nsteps)
    rest = np.loadtxt(out, skiprows=nsteps, usecols=(0, 1, 2))
    quat_f = np.loadtxt(out, skiprows=nsteps + box.nm, max_rows=box.nm)
    com_f, vel_f, jb_f = rest[:box.nm], rest[2 * box.nm:3 * box.nm], rest[3 * box.nm:]

    md = RigidMD(ThreeBody(p, "fp64"), box.com, box.quats, box.box, v, j)
    rec = md.run(nsteps, dt)
    st = md.state()
    md.close()
    # same library and arithmetic; forces are summed with atomics (round-off order)
    np.testing.assert_allclose(rec_f, rec, rtol=1e-10, atol=1e-10)
    for got, key in ((com_f, "com"), (quat_f, "quat"), (vel_f, "vel"), (jb_f, "jbody")):
        np.testing.assert_allclose(got, st[key], rtol=1e-9, atol=1e-12 * np.abs(st[key]).max(), err_msg=key)


# ------------------------------------------------ the integration snippets at the reference's shapes
REF_DRIVER = HERE / "fortran" / "_build" / "drive_ref"
REF_SOURCE = HERE / "fortran" / "ref_dropin.f90"
INTEGRATION = HERE.parent / "INTEGRATION.md"


def integration_snippets():
    """The first two ```fortran blocks of the integration guide at the
    repository root (context creation and the new body of threebody_term)."""
    text = INTEGRATION.read_text()
    blocks, pos = [], 0
    while len(blocks) < 2:
        a = text.index("```fortran\n", pos) + len("```fortran\n")
        b = text.index("```", a)
        blocks.append([ln.strip() for ln in text[a:b].splitlines() if ln.strip()])
        pos = b + 3
    return blocks


def test_integration_snippets_are_the_tested_program():
    """Every line of the two snippets a maintainer copies from the integration
    guide appears, in the same order, in tests/fortran/ref_dropin.f90 (the drop-in
    that drive_ref and drive_nve link, which the GPU tests below run against
    the golden boxes and through 1000 MD steps); and the cutoff passed is
    rcut3 (A), not the squared rcutsd3."""
    src = [ln.strip() for ln in REF_SOURCE.read_text().splitlines() if ln.strip()]
    for block in integration_snippets():
        at = 0
        for line in block:
            assert line in src[at:], f"snippet line not in ref_dropin.f90 (in order): {line!r}"
            at = src.index(line, at) + 1
    flat = " ".join(" ".join(b) for b in integration_snippets())
    assert "tb_par%rcut3 = rcut3" in flat and "rcutsd3" not in flat


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box64_fs2fs3", "box64_fs3"])
def test_reference_shapes_drop_in_matches_golden(tmp_path, name):
    """The reference's globals at their declared shapes (nsite = 50, nom = 5000,
    bS3_s(50,50) read with ld_pair = 50), the integration snippets, one call
    of threebody_term: the golden 64-water box at the FP64 parity bar."""
    assert REF_DRIVER.exists(), "tests/fortran/_build/drive_ref is built by __graft_entry__.build()"
    with np.load(HERE / "golden" / f"{name}.npz", allow_pickle=False) as f:
        g = {k: f[k] for k in f.files}
    variant = "fs3" if name.endswith("_fs3") else "fs2fs3"
    p = load_variant(variant)
    s2 = 0 if variant == "fs3" else int(p.s2_term)
    nm = len(g["com"])
    d = DATA_DIR / variant
    fmt = lambda v: f"{float(v):.17g}"
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    with open(inp, "w") as fh:
        fh.write(f"{nm} {fmt(g['box'])}\n")
        fh.write(f"{fmt(p.alpha3)} {fmt(p.delta33)} {fmt(p.rcut3)} {s2} {int(p.s3_term)}\n")
        fh.write(f"{d / 'data3b_full_SR_5p5'}\n{d / 'data3b_full_LR_4p5'}\n")
        for row in g["com"]:
            fh.write(" ".join(fmt(v) for v in row) + "\n")
        for m in range(nm):
            for s in range(17):
                fh.write(" ".join(fmt(v) for v in g["site"][m, s]) + "\n")
    print(f"fortran reference-shape drop-in: {name}", flush=True)
    r = subprocess.run([str(REF_DRIVER), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout, r.stderr)
    vals = np.loadtxt(out, max_rows=1)
    fsite = np.loadtxt(out, skiprows=1, max_rows=nm * 17).reshape(nm, 17, 3)
    com = np.loadtxt(out, skiprows=1 + nm * 17)
    for got, key in ((vals[0], "qfs2"), (vals[1], "qfs3")):
        ref = float(g[key])
        assert abs(got - ref) <= 1e-10 * abs(ref) + 2e-12, (key, got, ref)
    assert abs(vals[2] - float(g["vir"])) <= 1e-9 * abs(float(g["vir"])) + 1e-9
    rel = lambda a, b: np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())
    assert rel(fsite, g["fsite"]) <= 1e-9
    if s2:
        assert rel(com[:, :3], g["fcom2"]) <= 1e-9
    assert rel(com[:, 3:], g["fcom3"]) <= 1e-9


# ------------------------------------------------ a Fortran MD loop calling the drop-in every step
NVE_DRIVER = HERE / "fortran" / "_build" / "drive_nve"


def read_nve(path, nm):
    """Energies per step and the snapshots of drive_nve's output."""
    lines = Path(path).read_text().splitlines()
    energies, snaps, k = [], {}, 0
    arr = lambda rows: np.array([[float(v) for v in ln.split()] for ln in rows])
    while k < len(lines):
        tag = lines[k].split()
        if tag[0] == "E":
            energies.append([float(v) for v in tag[2:]])
            k += 1
        else:
            step = int(tag[1])
            k += 1
            vals = [float(v) for v in lines[k].split()]
            k += 1
            snap = {"qfs2": vals[0], "qfs3": vals[1], "vir": vals[2]}
            snap["com"] = arr(lines[k:k + nm]); k += nm
            snap["site"] = arr(lines[k:k + 17 * nm]).reshape(nm, 17, 3); k += 17 * nm
            snap["fsite"] = arr(lines[k:k + 17 * nm]).reshape(nm, 17, 3); k += 17 * nm
            c = arr(lines[k:k + nm]); k += nm
            snap["fcom2"], snap["fcom3"] = c[:, :3], c[:, 3:]
            snaps[step] = snap
    return np.array(energies), snaps


@pytest.mark.gpu
def test_fortran_nve_loop_1000_steps(tmp_path, oracle):
    """256 waters (L = 19.71 A) from 298 K, FS2+FS3 in FP64: 1000 NVE steps of
    0.5 fs by a Fortran host loop (tests/fortran/drive_nve.f90) that calls the
    documented drop-in threebody_term at every step, as the reference's MD
    loop calls the original.  Every 100 steps: the snapshot's triplet and
    item lists bit-exact with the oracle's, and the energies, virial, site
    forces and FS2 and FS3 COM forces the Fortran loop received within the
    FP64 bar of the oracle's.  Over the first 20 steps the total energy stays
    put (the three-body term alone has no short-range repulsion, so the
    trajectory is unphysical after ~20 fs, but every snapshot is a valid
    geometry to compare on)."""
    from ccpol23_plus_2025_mbmd_amd.md import maxwell_state, water_constants
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    assert NVE_DRIVER.exists(), "tests/fortran/_build/drive_nve is built by __graft_entry__.build()"
    p = load_variant("fs2fs3")
    box = lattice_box(256, seed=17)
    pc, mass, pmi = water_constants()
    v, j = maxwell_state(box.nm, 298.0, mass, pmi, seed=5)
    fmt = lambda x: f"{float(x):.17g}"
    d = DATA_DIR / "fs2fs3"
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    with open(inp, "w") as fh:
        fh.write(f"{box.nm} {fmt(box.box)} 1000 100 {fmt(5e-4)}\n")
        fh.write(f"{fmt(p.alpha3)} {fmt(p.delta33)} {fmt(p.rcut3)} {int(p.s2_term)} {int(p.s3_term)}\n")
        fh.write(f"{d / 'data3b_full_SR_5p5'}\n{d / 'data3b_full_LR_4p5'}\n")
        fh.write(" ".join(fmt(x) for x in [mass, *pmi]) + "\n")
        for s in range(17):
            fh.write(" ".join(fmt(x) for x in pc[s]) + "\n")
        for i in range(box.nm):
            fh.write(" ".join(fmt(x) for x in [*box.com[i], *box.quats[i], *v[i], *j[i]]) + "\n")
    print("fortran NVE loop: 256 waters, 1000 steps of 0.5 fs, drop-in called every step", flush=True)
    r = subprocess.run([str(NVE_DRIVER), str(inp), str(out)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout[-2000:], r.stderr[-2000:])
    energies, snaps = read_nve(out, box.nm)
    assert energies.shape == (1001, 3) and sorted(snaps) == list(range(0, 1001, 100))
    etot = energies.sum(1)
    k = energies[:21, 1] + energies[:21, 2]
    drift = np.abs(etot[:21] - etot[0]).max()
    print(f"first 20 steps: E3 {energies[0, 0]:.4f} -> {energies[20, 0]:.4f}, max |dE| {drift:.2e} kcal/mol",
          flush=True)
    assert np.abs(energies[20, 0] - energies[0, 0]) > 10 * drift and drift <= 2e-3 * k.mean()
    rel = lambda a, b: float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    tb = ThreeBody(p, "fp64")
    for step, sn in sorted(snaps.items()):
        nm = box.nm
        x, y, z = (sn["com"][:, c].copy() for c in range(3))
        xs, ys, zs = (sn["site"][:, :, c].reshape(-1).copy() for c in range(3))
        ref = oracle.threebody_term(p, x=x, y=y, z=z, xs=xs, ys=ys, zs=zs, box=box.box, ld=17, nthreads=16)
        tb.threebody_term(x, y, z, xs, ys, zs, box.box, *(np.zeros(17 * nm) for _ in range(3)),
                          *(np.zeros(nm) for _ in range(6)), 0.0, ld=17)
        np.testing.assert_array_equal(tb.fs3_list(), oracle.fs3_list(p, x, y, z, box.box))
        np.testing.assert_array_equal(tb.fs2_list(), oracle.fs2_list(p, x, y, z, box.box))
        for key in ("qfs2", "qfs3"):
            assert abs(sn[key] - ref[key]) <= 1e-10 * abs(ref[key]) + 2e-12, (step, key, sn[key], ref[key])
        assert abs(sn["vir"] - ref["vir"]) <= 1e-9 * abs(ref["vir"]) + 1e-9, step
        fr = np.stack([ref[c].reshape(nm, 17) for c in ("fxs", "fys", "fzs")], -1)
        e_site = rel(sn["fsite"], fr)
        e_c2 = rel(sn["fcom2"], np.stack([ref[c] for c in ("fxfs2", "fyfs2", "fzfs2")], -1))
        e_c3 = rel(sn["fcom3"], np.stack([ref[c] for c in ("fxfs3", "fyfs3", "fzfs3")], -1))
        print(f"step {step}: E3 {sn['qfs2'] + sn['qfs3']:.6f} kcal/mol, site {e_site:.1e}, "
              f"COM FS2 {e_c2:.1e}, COM FS3 {e_c3:.1e}", flush=True)
        assert e_site <= 1e-9 and e_c2 <= 1e-9 and e_c3 <= 1e-9, step
