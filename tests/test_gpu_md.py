"""GPU tests of the device NVE driver against the numpy restatement of the
reference step (oracle/md_ref.py) with oracle forces, and its energy conservation."""
import numpy as np
import pytest

from oracle import md_ref

pytestmark = pytest.mark.gpu


def oracle_forces(oracle, params):
    def force_fn(c, site):
        ld = 28
        xs = np.zeros((len(c), ld, 3))
        xs[:, :site.shape[1]] = site
        r = oracle.threebody_term(params, x=c[:, 0].copy(), y=c[:, 1].copy(), z=c[:, 2].copy(),
                                  xs=xs[:, :, 0].reshape(-1).copy(), ys=xs[:, :, 1].reshape(-1).copy(),
                                  zs=xs[:, :, 2].reshape(-1).copy(), box=force_fn.box, ld=ld, nthreads=8)
        fsite = np.stack([r["fxs"], r["fys"], r["fzs"]], -1)[:, :site.shape[1]]
        fcom = np.stack([r["fxfs2"] + r["fxfs3"], r["fyfs2"] + r["fyfs3"], r["fzfs2"] + r["fzfs3"]], -1)
        return fsite, fcom, r["qfs2"] + r["qfs3"]
    return force_fn


@pytest.mark.parametrize("variant,nsteps", [("fs3", 100), ("fs2fs3", 10)])
def test_md_matches_reference_step(variant, nsteps, oracle, params_fs3, params_fs2fs3):
    """0.5 fs steps of 64 waters (L = 12.42 A) from 298 K Maxwell velocities:
    device driver vs the restated reference step with oracle forces; the FS3
    variant for 100 steps (the 64-water NVE run of the FS3 workload), FS2+FS3
    for 10."""
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    from ccpol23_plus_2025_mbmd_amd.md import RigidMD, maxwell_state, water_constants
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    p = params_fs3.replace(s2_term=False) if variant == "fs3" else params_fs2fs3
    box = lattice_box(64, seed=1)
    pc, mass, pmi = water_constants()
    v, j = maxwell_state(box.nm, 298.0, mass, pmi, seed=2)
    print(f"md vs reference step: {variant}, 64 waters, {nsteps} steps", flush=True)
    md = RigidMD(ThreeBody(p, "fp64"), box.com, box.quats, box.box, v, j)
    rec = md.run(nsteps, 5e-4)
    got = md.state()
    ff = oracle_forces(oracle, p)
    ff.box = box.box
    st = {"com": box.com.copy(), "quat": box.quats.copy(), "vel": v.copy(), "jbody": j.copy()}
    f = ff(st["com"], md_ref.sites(st["quat"], pc))[:2]
    e = []
    for _ in range(nsteps):
        st, f, en = md_ref.nve_step(st, f, ff, 5e-4, pc, mass, pmi, box.box)
        e.append(en)
    for key, tol in (("com", 1e-9), ("quat", 1e-9), ("vel", 1e-7), ("jbody", 1e-7)):
        a, b = got[key], st[key]
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        print(f"{key}: max relative difference {err:.2e} after {nsteps} steps", flush=True)
        assert err <= tol, (key, err)
    e = np.array(e)
    assert np.allclose(rec[:, 0], e[:, 0], rtol=1e-9, atol=1e-10)
    assert np.allclose(rec[:, 1], e[:, 1], rtol=1e-9)
    assert np.allclose(rec[:, 2], e[:, 2], rtol=1e-9)


def test_md_energy_conservation():
    """NVE with the three-body term alone, 20 steps of 0.5 fs, 256 waters from 298 K.
    The three-body term alone has no short-range repulsion (the two-body term
    the host adds is what keeps molecules apart), so a longer run collapses;
    over 10 fs the total energy must stay put up to the small jumps of the
    reference's hard per-pair cutoffs (6.3 A for FS3, rcut3 for FS2)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    from ccpol23_plus_2025_mbmd_amd.md import RigidMD, maxwell_state, water_constants
    from ccpol23_plus_2025_mbmd_amd.params import load_variant
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    box = lattice_box(256, seed=3)
    pc, mass, pmi = water_constants()
    v, j = maxwell_state(box.nm, 298.0, mass, pmi, seed=4)
    md = RigidMD(ThreeBody(load_variant("fs2fs3"), "fp64"), box.com, box.quats, box.box, v, j)
    rec = md.run(20, 5e-4)
    etot = rec[:, 3]
    k = rec[:, 1] + rec[:, 2]
    drift = np.abs(etot - etot[0]).max()
    print(f"E3 {rec[0, 0]:.4f} -> {rec[-1, 0]:.4f}, K {k[0]:.3f} -> {k[-1]:.3f}, max |dE| {drift:.3e} kcal/mol",
          flush=True)
    assert np.all(np.isfinite(rec))
    assert np.abs(rec[-1, 0] - rec[0, 0]) > 10 * drift  # energy flows between E3 and K
    assert drift <= 2e-3 * k.mean(), (drift, k.mean())


def test_md_revcon_restart_continues_the_run(tmp_path, params_fs2fs3):
    """10 steps, REVCON, restart, 10 more steps == 20 steps without the restart
    (the REVCON values round-trip exactly, and the COM wrap only shifts a
    molecule by a whole box, which no kernel can see)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    from ccpol23_plus_2025_mbmd_amd.md import RigidMD, maxwell_state, water_constants
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    box = lattice_box(64, seed=6)
    pc, mass, pmi = water_constants()
    v, j = maxwell_state(box.nm, 298.0, mass, pmi, seed=8)
    tb = ThreeBody(params_fs2fs3, "fp64")
    print("md revcon restart: 64 waters, 10 + 10 steps", flush=True)
    a = RigidMD(tb, box.com, box.quats, box.box, v, j)
    ra = a.run(20, 5e-4)
    sa = a.state()
    b = RigidMD(tb, box.com, box.quats, box.box, v, j)
    b.run(10, 5e-4)
    b.write_revcon(tmp_path / "REVCON", step=10, time=10 * 5e-4)
    b.close()
    c = RigidMD.from_revcon(tb, tmp_path / "REVCON")
    rc = c.run(10, 5e-4)
    sc = c.state()
    np.testing.assert_allclose(rc, ra[10:], rtol=1e-10, atol=1e-10)
    L = box.box
    d = sc["com"] - sa["com"]
    np.testing.assert_allclose(d - L * np.round(d / L), 0.0, atol=1e-9)
    for key in ("quat", "vel", "jbody"):
        np.testing.assert_allclose(sc[key], sa[key], rtol=1e-9, atol=1e-12 * np.abs(sa[key]).max(), err_msg=key)


def test_md_1000_steps_snapshot_parity(oracle, params_fs2fs3):
    """1000 NVE steps of 0.5 fs, 256 waters (L = 19.71 A), FS2+FS3 in FP64, as
    the device driver runs them; every 100 steps the snapshot's triplet and
    item lists are bit-exact with the oracle's and its energies, forces and
    virial are within the FP64 bar.  (With the three-body term as the only
    force the trajectory is not physical after ~20 fs -- there is no two-body
    repulsion -- but every snapshot is a valid geometry to compare on.)"""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox, lattice_box
    from ccpol23_plus_2025_mbmd_amd.md import RigidMD, maxwell_state, water_constants
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    box = lattice_box(256, seed=17)
    pc, mass, pmi = water_constants()
    v, j = maxwell_state(box.nm, 298.0, mass, pmi, seed=5)
    tb = ThreeBody(params_fs2fs3, "fp64")
    md = RigidMD(tb, box.com, box.quats, box.box, v, j)
    for block in range(10):
        rec = md.run(100, 5e-4)
        assert np.all(np.isfinite(rec)), block
        st = md.state()
        snap = WaterBox(st["com"], st["quat"], box.box)
        print(f"md snapshot at step {100 * (block + 1)}: E3 {rec[-1, 0]:.6f} kcal/mol", flush=True)
        got = tb.evaluate(snap)
        np.testing.assert_array_equal(tb.fs3_list(), oracle.fs3_list(params_fs2fs3, snap.x, snap.y, snap.z, snap.box))
        np.testing.assert_array_equal(tb.fs2_list(), oracle.fs2_list(params_fs2fs3, snap.x, snap.y, snap.z, snap.box))
        ref = oracle.threebody_term(params_fs2fs3, snap, nthreads=16)
        for k in ("qfs2", "qfs3"):
            assert abs(got[k] - ref[k]) <= 1e-10 * abs(ref[k]) + 2e-12, (block, k, got[k], ref[k])
        f = np.stack([got[k][:, :17] for k in ("fxs", "fys", "fzs")])
        fr = np.stack([ref[k][:, :17] for k in ("fxs", "fys", "fzs")])
        assert np.sqrt(((f - fr) ** 2).sum() / (fr ** 2).sum()) <= 1e-9, block
        assert abs(got["vir"] - ref["vir"]) <= 1e-9 * abs(ref["vir"]) + 1e-9, block
    md.close()
