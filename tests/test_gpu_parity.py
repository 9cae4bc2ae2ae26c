"""GPU parity of the HIP three-body path against the FP64 oracle, through the C ABI.

Bar: triplet lists bit-exact (same entries, same order); FP64 energies within
1e-10 relative, site and COM forces within 1e-9 relative RMS, virial within
1e-9 relative; the mixed-precision path within 1e-5 relative force RMS."""
import os

import numpy as np
import pytest

from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box, trimer_set, box_length

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


def rel_rms(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.sqrt((b ** 2).sum())
    return np.sqrt(((a - b) ** 2).sum()) / den if den > 0 else np.sqrt((a ** 2).sum())


def site_forces(r, ns=17):
    return np.stack([np.asarray(r["fxs"])[:, :ns], np.asarray(r["fys"])[:, :ns], np.asarray(r["fzs"])[:, :ns]])


def com_forces(r, term):
    return np.stack([r[f"fx{term}"], r[f"fy{term}"], r[f"fz{term}"]])


@pytest.fixture(scope="module")
def tb_fs2fs3(params_fs2fs3):
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    print("creating ThreeBody(fs2fs3, fp64) on cuda:0", flush=True)
    return ThreeBody(params_fs2fs3, "fp64")


@pytest.fixture(scope="module")
def tb_fs3(params_fs3):
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    return ThreeBody(params_fs3.replace(s2_term=False), "fp64")


def all_forces(r, s2, s3):
    parts = [site_forces(r).reshape(-1)]
    if s2:
        parts.append(com_forces(r, "fs2").reshape(-1))
    if s3:
        parts.append(com_forces(r, "fs3").reshape(-1))
    return np.concatenate(parts)


# COM-force arrays alone are held to 1e-9 too, except where FP64 itself cannot
# fix them that well: the FS3 COM force is f_d' E + f_d w' E over region
# energies that cancel up to ~1e6-fold, so for some isolated trimers the
# oracle's own result moves by more than 1e-9 when only the order of the
# molecules changes (tests/test_oracle_physics.py::test_fs3_com_conditioning).
# There the bar is four times that self-variation (oracle_com_floor).
COM_TOL = 1e-9


def oracle_com_floor(oracle, params, box):
    """Relative RMS change of the oracle's FS3 and FS2 COM forces when the
    molecules of `box` are evaluated in reverse order."""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox
    ref = oracle.threebody_term(params, box)
    perm = np.arange(box.nm)[::-1]
    rp = oracle.threebody_term(params, WaterBox(box.com[perm], box.quats[perm], box.box))
    return max(rel_rms(com_forces(rp, t)[:, np.argsort(perm)], com_forces(ref, t)) for t in ("fs2", "fs3"))
# Energies: 1e-10 relative plus an absolute floor.  The oracle's FS3 trimer
# energy itself moves by up to 8.4e-13 kcal/mol (3.4e-10 relative for a
# 2.4e-3 kcal/mol trimer) when only the molecule order is permuted, so the floor
# is 2e-12 kcal/mol; for periodic boxes the relative term dominates.
E_FLOOR = 2e-12


def assert_parity(got, ref, s2=True, s3=True, e_tol=1e-10, f_tol=1e-9, com_tol=COM_TOL):
    if s3:
        assert got["n_fs3"] == ref["n_fs3"]
        assert abs(got["qfs3"] - ref["qfs3"]) <= e_tol * abs(ref["qfs3"]) + E_FLOOR, (got["qfs3"], ref["qfs3"])
        assert rel_rms(com_forces(got, "fs3"), com_forces(ref, "fs3")) <= com_tol
    if s2:
        assert got["n_fs2_items"] == ref["n_fs2"]
        assert abs(got["qfs2"] - ref["qfs2"]) <= e_tol * abs(ref["qfs2"]) + E_FLOOR, (got["qfs2"], ref["qfs2"])
        assert rel_rms(com_forces(got, "fs2"), com_forces(ref, "fs2")) <= com_tol
    assert rel_rms(all_forces(got, s2, s3), all_forces(ref, s2, s3)) <= f_tol
    assert rel_rms(site_forces(got), site_forces(ref)) <= f_tol
    assert abs(got["vir"] - ref["vir"]) <= 10 * e_tol * abs(ref["vir"]) + 418.4 * E_FLOOR, (got["vir"], ref["vir"])


def test_fs3_list_longer_than_first_grid(oracle, params_fs3):
    """A per-call FS3 evaluation launches one block per triplet for 96 triplets
    per molecule, or the last call's count if that is more.  125 waters at 1.4x
    water density hold 156 per molecule: the first call on a fresh context is
    flagged, writes nothing, and is evaluated again with a grid that covers the
    list.  Its outputs match the oracle, and equal bit for bit those of the next
    call, whose grid comes from the first call's count."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    n = 125
    box = lattice_box(n, seed=9, box=box_length(n) * (1 / 1.4) ** (1 / 3))
    print("fs3 list longer than the first grid: 125 waters at 1.4x density", flush=True)
    p = params_fs3.replace(s2_term=False)
    tb = ThreeBody(p, "fp64")
    first = tb.evaluate(box)
    assert first["n_fs3"] > 96 * n
    second = tb.evaluate(box)
    for k in ("qfs3", "vir", "fxs", "fys", "fzs", "fxfs3", "fyfs3", "fzfs3"):
        assert np.array_equal(first[k], second[k]), k
    np.testing.assert_array_equal(tb.fs3_list(), oracle.fs3_list(p, box.x, box.y, box.z, box.box))
    assert_parity(first, oracle.threebody_term(p, box), s2=False)


@pytest.mark.parametrize("n,seed", [(64, 1), (64, 2), (256, 3)])
def test_lists_bit_exact(tb_fs2fs3, oracle, params_fs2fs3, n, seed):
    box = lattice_box(n, seed=seed)
    print(f"lists: n={n} seed={seed}", flush=True)
    tb_fs2fs3.evaluate(box)
    l3 = tb_fs2fs3.fs3_list()
    l2 = tb_fs2fs3.fs2_list()
    np.testing.assert_array_equal(l3, oracle.fs3_list(params_fs2fs3, box.x, box.y, box.z, box.box))
    np.testing.assert_array_equal(l2, oracle.fs2_list(params_fs2fs3, box.x, box.y, box.z, box.box))


def test_fs2_dense_box(oracle, params_fs2fs3):
    """FS2 alone on 216 waters compressed to 1.3x water density (~77 neighbours
    within 7.5 A, ~3,000 candidates per spectator): the spectator pass then
    tests its candidates in several rounds and evaluates its queue more than
    once per molecule.  FS2 list bit-exact; energy, forces and virial against
    the oracle."""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    p = params_fs2fs3.replace(s3_term=False)
    base = lattice_box(216, seed=51)
    s = (1 / 1.3) ** (1 / 3)
    box = WaterBox(base.com * s, base.quats, base.box * s)
    print("fs2 dense box 216 waters", flush=True)
    tb = ThreeBody(p, "fp64")
    got = tb.evaluate(box)
    ref = oracle.threebody_term(p, box, nthreads=NTHREADS)
    assert got["n_fs2_items"] > 216 * 1200
    assert_parity(got, ref, s3=False)
    np.testing.assert_array_equal(tb.fs2_list(), oracle.fs2_list(p, box.x, box.y, box.z, box.box))


def test_fs3_variant_64_box(tb_fs3, oracle, params_fs3):
    """FS3-only variant (switches T T F) on the 64-water box, L = 12.42 A."""
    box = lattice_box(64, seed=1)
    print("fs3 variant 64 waters", flush=True)
    got = tb_fs3.evaluate(box)
    ref = oracle.threebody_term(params_fs3.replace(s2_term=False), box, nthreads=NTHREADS)
    assert_parity(got, ref, s2=False)
    assert got["n_fs3_sr"] == ref["n_fs3_sr"] and got["n_fs3_switch"] == ref["n_fs3_sw"]
    assert got["n_fs3_lr"] == ref["n_fs3_lr"]


@pytest.mark.parametrize("n,seed", [(64, 5), (256, 7)])
def test_fs2fs3_box(tb_fs2fs3, oracle, params_fs2fs3, n, seed):
    box = lattice_box(n, seed=seed)
    print(f"fs2fs3 box n={n}", flush=True)
    got = tb_fs2fs3.evaluate(box)
    ref = oracle.threebody_term(params_fs2fs3, box, nthreads=NTHREADS)
    assert_parity(got, ref)


def test_benchmark_box_against_oracle(tb_fs2fs3, oracle, params_fs2fs3):
    """The box bench.py times (512 waters, seed 2025: 31,290 triplets, 450,861
    FS2 items) in FP64 against the oracle: both lists bit for bit, energies,
    forces and virial to the parity bars."""
    box = lattice_box(512, seed=2025)
    print("benchmark box: 512 waters, seed 2025", flush=True)
    got = tb_fs2fs3.evaluate(box)
    l3, l2 = tb_fs2fs3.fs3_list(), tb_fs2fs3.fs2_list()
    ref = oracle.threebody_term(params_fs2fs3, box, nthreads=NTHREADS)
    assert (ref["n_fs3"], ref["n_fs2"]) == (31290, 450861)
    np.testing.assert_array_equal(l3, oracle.fs3_list(params_fs2fs3, box.x, box.y, box.z, box.box))
    np.testing.assert_array_equal(l2, oracle.fs2_list(params_fs2fs3, box.x, box.y, box.z, box.box))
    assert_parity(got, ref)


def region_trimers(seed=23):
    """Equilateral trimers at R_avg = 4.0 (SR), 5.0 (switch) and 5.9 A (LR)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox, random_quaternions
    rng = np.random.default_rng(seed)
    out = []
    for d in (4.0, 5.0, 5.9):
        tri = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0], [d / 2, d * np.sqrt(3) / 2, 0.0]])
        out.append(WaterBox(tri - tri.mean(0), random_quaternions(3, rng), 0.0))
    return out


def test_trimers_all_regions(tb_fs2fs3, oracle, params_fs2fs3):
    trimers = trimer_set(40, seed=17) + region_trimers()
    seen = set()
    for t in trimers:
        got = tb_fs2fs3.evaluate(t)
        ref = oracle.threebody_term(params_fs2fs3, t)
        assert_parity(got, ref, com_tol=max(COM_TOL, 4 * oracle_com_floor(oracle, params_fs2fs3, t)))
        for key in ("n_fs3_sr", "n_fs3_switch", "n_fs3_lr"):
            if got[key]:
                seen.add(key)
    assert seen == {"n_fs3_sr", "n_fs3_switch", "n_fs3_lr"}


def test_drop_in_semantics(tb_fs2fs3, params_fs2fs3):
    """Site forces and virial accumulate, COM forces are overwritten, torques zeroed."""
    box = lattice_box(27, seed=4, box=9.3)
    nm, ld = box.nm, box.ld
    base = tb_fs2fs3.evaluate(box)
    rng = np.random.default_rng(0)
    f0 = [rng.normal(size=nm * ld) for _ in range(3)]
    fx, fy, fz = (a.copy() for a in f0)
    c = [rng.normal(size=nm) for _ in range(6)]
    t = [rng.normal(size=nm) for _ in range(6)]
    q2, q3, vir = tb_fs2fs3.threebody_term(box.x, box.y, box.z, box.xs, box.ys, box.zs, box.box, fx, fy, fz,
                                           *c, vir=5.0, txfs2=t[0], tyfs2=t[1], tzfs2=t[2], txfs3=t[3],
                                           tyfs3=t[4], tzfs3=t[5], ld=ld)
    # forces are summed with FP64 atomics: reproducible to round-off, not bitwise
    scale = np.abs(base["fxs"]).max()
    np.testing.assert_allclose(fx - f0[0], base["fxs"].reshape(-1), rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(c[0], base["fxfs2"], rtol=1e-12, atol=1e-12 * np.abs(base["fxfs2"]).max())
    np.testing.assert_allclose(c[3], base["fxfs3"], rtol=1e-12, atol=1e-12 * np.abs(base["fxfs3"]).max())
    for a in t:
        assert not a.any()
    assert abs(vir - 5.0 - base["vir"]) < 1e-12 * max(1.0, abs(base["vir"]))
    assert q2 == base["qfs2"] and q3 == base["qfs3"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_repeatable(params_fs2fs3, precision):
    """Energies and virial use fixed reduction trees, forces fixed-point sums:
    every output is the same bits run to run (the reference merges its forces
    in a fixed order, threebody_potential.f90:1964-1977)."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb = ThreeBody(params_fs2fs3, precision)
    box = lattice_box(256, seed=9)
    a = tb.evaluate(box)
    for _ in range(3):
        b = tb.evaluate(box)
        for k in ("qfs2", "qfs3", "vir"):
            assert a[k] == b[k]
        for k in ("fxs", "fys", "fzs", "fxfs2", "fyfs2", "fzfs2", "fxfs3", "fyfs3", "fzfs3"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_overlapping_sites_fail_cleanly(params_fs2fs3):
    """A spectator's oxygen exactly on a pair member's centre of mass makes the
    FS2 exchange-charge term 1/|r_ks - R_i| infinite (the reference divides by
    zero there).  The call fails with CCPOL3B_EINVAL, names the cause, and
    leaves every output untouched."""
    from ccpol23_plus_2025_mbmd_amd import _lib
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    base = lattice_box(8, seed=2)
    com = base.com.copy()
    com[2] = com[0] - base.site[2, 0]  # O site of molecule 2 on the COM of molecule 0
    wb = WaterBox(com, base.quats, 0.0)
    tb = ThreeBody(params_fs2fs3, "fp64")
    nm, ld = wb.nm, wb.ld
    outs = [np.full(nm * ld, 3.0) for _ in range(3)] + [np.full(nm, 5.0) for _ in range(6)]
    keep = [o.copy() for o in outs]
    with pytest.raises(_lib.CcpolError, match="fixed-point"):
        tb.threebody_term(wb.x, wb.y, wb.z, wb.xs, wb.ys, wb.zs, 0.0, *outs, vir=1.0, ld=ld)
    for o, k in zip(outs, keep):
        np.testing.assert_array_equal(o, k)


def test_box_description_changes_between_calls(params_fs2fs3):
    """One context, calls whose box description (box length, molecule count,
    open boundary) changes or repeats: each result equals a fresh context's.
    The context copies the box description to the device only when it differs
    from the last copy."""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    a = lattice_box(64, seed=21)
    b = WaterBox(a.com * 1.02, a.quats, a.box * 1.02)   # same molecules, longer box
    c = lattice_box(100, seed=22)
    d = WaterBox(a.com, a.quats, 0.0)                   # open boundary
    tb = ThreeBody(params_fs2fs3, "fp64")
    for box in (a, b, a, c, d, a, a):
        got = tb.evaluate(box)
        ref = ThreeBody(params_fs2fs3, "fp64").evaluate(box)
        for k in ("qfs2", "qfs3", "vir", "n_fs3", "n_fs2_items"):
            assert got[k] == ref[k], k
        for k in ("fxs", "fys", "fzs", "fxfs2", "fxfs3"):
            np.testing.assert_allclose(got[k], ref[k], rtol=0, atol=1e-13 * np.abs(ref[k]).max())


def test_mixed_precision_error_512(params_fs2fs3, oracle):
    """Mixed (FP32 exponential/polynomial work in the shifted pair-vector basis)
    vs the FP64 path at 512 waters: site forces, COM forces, and what an
    integrator consumes (molecular force = sites + both COM terms, torque)."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    box = lattice_box(512, seed=3)
    print("mixed 512", flush=True)
    tb = ThreeBody(params_fs2fs3, "fp64")
    ref = tb.evaluate(box)
    tb.set_precision("mixed")
    got = tb.evaluate(box)
    err_f = rel_rms(site_forces(got), site_forces(ref))
    err_all = rel_rms(all_forces(got, True, True), all_forces(ref, True, True))
    err_c = rel_rms(com_forces(got, "fs3"), com_forces(ref, "fs3"))
    err_e = abs(got["qfs3"] - ref["qfs3"]) / abs(ref["qfs3"])

    def molecular(r):
        return site_forces(r).sum(2) + com_forces(r, "fs2") + com_forces(r, "fs3")

    def torque(r):
        f = site_forces(r).transpose(1, 2, 0)
        return np.cross(box.site[:, :17], f).sum(1)

    err_m = rel_rms(molecular(got), molecular(ref))
    err_t = rel_rms(torque(got), torque(ref))
    vir_err = abs(got["vir"] - ref["vir"]) / abs(ref["vir"])
    err_e2 = abs(got["qfs2"] - ref["qfs2"]) / abs(ref["qfs2"])
    err_c2 = rel_rms(com_forces(got, "fs2"), com_forces(ref, "fs2"))
    print(f"mixed vs fp64: site force rms {err_f:.3e}, all forces {err_all:.3e}, FS3 COM {err_c:.3e}, "
          f"molecular force {err_m:.3e}, torque {err_t:.3e}, FS3 energy {err_e:.3e}, virial {vir_err:.3e}, "
          f"FS2 energy {err_e2:.3e}, FS2 COM {err_c2:.3e}")
    # The 1e-5 force bar holds for the site forces (and the full output vector,
    # which they dominate).
    assert err_f <= 1e-5
    assert err_all <= 1e-5
    # It does NOT hold for the FS3 COM forces, the molecular forces and the
    # energy: their envelopes below are measured, not a bar (this box: FS3 COM
    # 1.4e-3 -- 8e-5 for the FS2 + FS3 COM vector, whose FP64 FS2 part
    # dominates -- molecular 1.4e-4, torque 6.5e-5, energy 6.5e-4).  The
    # short-range triplet energies cancel ~1e4-fold in the shifted basis, and
    # FP32 rounding of the coefficients alone costs 6e-4 of them
    # (tests/test_mixed_limits.py), so the benchmark's headline is the FP64 path.
    assert err_e <= 2e-3
    assert err_c <= 3e-3
    assert err_m <= 4e-4
    assert err_t <= 2e-4
    # the virial sums r dE/dr over site pairs and COM vectors; in a periodic
    # 512-water box its FP32 error stays at the energy's level
    assert vir_err <= 5e-3
    # FS2 in mixed precision (FP32 charge-site distances, cancellation-free
    # pair differences): measured energy 7.8e-6, COM 8.7e-7 on this box
    assert err_e2 <= 1e-4 and err_c2 <= 1e-5


def test_fs2_mixed_against_oracle(params_fs2fs3, oracle):
    """FS2 alone (FS3 off) in mixed precision against the oracle at 256 and 512
    waters: energy, site forces (sites 1-5), COM forces and virial; the FP32
    charge-site distances enter T_k through cancellation-free pair differences.
    Measured: energy 4.5e-6 / 6.0e-5, site forces 3.2e-7 / 3.5e-7, COM forces
    7.1e-7 / 8.5e-7, virial 6.0e-6 / 1.3e-5 (256 / 512 waters)."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    p = params_fs2fs3.replace(s3_term=False)
    tb = ThreeBody(p, "mixed")
    for n, seed in ((256, 61), (512, 62)):
        box = lattice_box(n, seed=seed)
        got = tb.evaluate(box)
        ref = oracle.threebody_term(p, box, nthreads=NTHREADS)
        err_e = abs(got["qfs2"] - ref["qfs2"]) / abs(ref["qfs2"])
        err_f = rel_rms(site_forces(got, 5), site_forces(ref, 5))
        err_c = rel_rms(com_forces(got, "fs2"), com_forces(ref, "fs2"))
        err_v = abs(got["vir"] - ref["vir"]) / abs(ref["vir"])
        print(f"FS2 mixed {n} waters: energy {err_e:.2e}, site forces {err_f:.2e}, COM {err_c:.2e}, "
              f"virial {err_v:.2e}", flush=True)
        assert got["n_fs2_items"] == ref["n_fs2"]
        assert err_f <= 2e-6 and err_c <= 5e-6 and err_e <= 2e-4 and err_v <= 1e-4


def test_batch_matches_single(params_fs2fs3):
    import torch
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb = ThreeBody(params_fs2fs3, "fp64")
    boxes = [lattice_box(n, seed=s) for n, s in ((40, 1), (64, 2), (100, 3))]
    dev = torch.device("cuda:0")

    def pack(bs):
        com = torch.tensor(np.concatenate([b.com for b in bs]).T.copy(), dtype=torch.float64, device=dev)
        site = torch.tensor(np.concatenate([np.stack([b.xs.reshape(b.nm, b.ld), b.ys.reshape(b.nm, b.ld),
                                                      b.zs.reshape(b.nm, b.ld)]) for b in bs], axis=1).copy(),
                            dtype=torch.float64, device=dev)
        return com.contiguous(), site.contiguous()

    com, site = pack(boxes)
    starts = np.cumsum([0] + [b.nm for b in boxes])
    fs, c2, c3, en = tb.eval_batch_device(starts, [b.box for b in boxes], com, site)
    torch.cuda.synchronize()
    for bi, b in enumerate(boxes):
        cb, sb = pack([b])
        fs1, c21, c31, en1 = tb.eval_device(cb, sb, b.box)
        torch.cuda.synchronize()
        sl = slice(starts[bi], starts[bi + 1])
        # forces: fixed-point sums; energies: fixed trees -- both bitwise
        for a, b_ in ((fs[:, sl], fs1), (c2[:, sl], c21), (c3[:, sl], c31)):
            assert torch.equal(a, b_)
        assert torch.equal(en[bi], en1)


@pytest.mark.parametrize("terms", ["fs2fs3", "fs2", "fs3"])
def test_large_batch_matches_single(params_fs2fs3, terms):
    """6,100 molecules in six boxes: the offset scans of the list build then
    take several tiles (2048 molecules per tile with both terms, 4096 with one
    of FS2, 8192 with FS3 alone), against one box per call, where each takes
    one tile."""
    import torch
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    p = params_fs2fs3
    if terms == "fs2":
        p = p.replace(s3_term=False)
    elif terms == "fs3":
        p = p.replace(s2_term=False)
    tb = ThreeBody(p, "fp64")
    boxes = [lattice_box(n, seed=s) for n, s in ((700, 41), (1000, 42), (900, 43), (1200, 44), (800, 45),
                                                 (1500, 46))]
    print("large batch", terms, sum(b.nm for b in boxes), "molecules", flush=True)
    dev = torch.device("cuda:0")

    def pack(bs):
        com = torch.tensor(np.concatenate([b.com for b in bs]).T.copy(), dtype=torch.float64, device=dev)
        site = torch.tensor(np.concatenate([np.stack([b.xs.reshape(b.nm, b.ld), b.ys.reshape(b.nm, b.ld),
                                                      b.zs.reshape(b.nm, b.ld)]) for b in bs], axis=1).copy(),
                            dtype=torch.float64, device=dev)
        return com.contiguous(), site.contiguous()

    com, site = pack(boxes)
    starts = np.cumsum([0] + [b.nm for b in boxes])
    fs, c2, c3, en = tb.eval_batch_device(starts, [b.box for b in boxes], com, site)
    torch.cuda.synchronize()
    st = tb.stats()
    n3 = n2 = 0
    for bi, b in enumerate(boxes):
        cb, sb = pack([b])
        fs1, c21, c31, en1 = tb.eval_device(cb, sb, b.box)
        torch.cuda.synchronize()
        n3 += tb.stats()["n_fs3"]
        n2 += tb.stats()["n_fs2_items"]
        sl = slice(starts[bi], starts[bi + 1])
        for a, b_ in ((fs[:, sl], fs1), (c2[:, sl], c21), (c3[:, sl], c31)):
            assert torch.equal(a, b_)
        assert torch.equal(en[bi], en1)
    assert (st["n_fs3"], st["n_fs2_items"]) == (n3, n2)
    assert (n3 > 0) == (terms != "fs2") and (n2 > 0) == (terms != "fs3")


def test_empty_and_tiny_inputs(tb_fs2fs3, oracle, params_fs2fs3):
    """nm = 0, 1, 2 and three far-apart molecules: no triplets or items, zero
    energies, accumulated outputs untouched, vir unchanged."""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox, random_quaternions
    rng = np.random.default_rng(2)
    z = np.zeros(0)
    q2, q3, vir = tb_fs2fs3.threebody_term(z, z, z, z, z, z, 10.0, z, z, z, z, z, z, z, z, z, 3.5, ld=28)
    assert (q2, q3, vir) == (0.0, 0.0, 3.5)
    for com in ([[0.0, 0.0, 0.0]], [[0.0, 0.0, 0.0], [2.9, 0.0, 0.0]],
                [[0.0, 0.0, 0.0], [9.0, 0.0, 0.0], [0.0, 9.0, 0.0]]):
        com = np.array(com)
        box = WaterBox(com, random_quaternions(len(com), rng), 0.0)
        print(f"tiny input: {len(com)} molecules", flush=True)
        got = tb_fs2fs3.evaluate(box)
        ref = oracle.threebody_term(params_fs2fs3, box)
        assert got["n_fs3"] == ref["n_fs3"] == 0 and got["n_fs2_items"] == ref["n_fs2"] == 0
        assert got["qfs2"] == 0.0 and got["qfs3"] == 0.0 and got["vir"] == 0.0
        for k in ("fxs", "fys", "fzs", "fxfs2", "fxfs3"):
            assert not np.asarray(got[k]).any(), k


def test_open_cluster(tb_fs2fs3, oracle, params_fs2fs3):
    """A 40-water cluster without periodic images (box <= 0)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox
    b = lattice_box(40, seed=12)
    box = WaterBox(b.com - b.com.mean(0), b.quats, 0.0)
    print("open cluster 40 waters", flush=True)
    got = tb_fs2fs3.evaluate(box)
    ref = oracle.threebody_term(params_fs2fs3, box, nthreads=NTHREADS)
    assert got["n_fs3"] > 0 and got["n_fs2_items"] > 0
    assert_parity(got, ref)


def _triangles(com_a, box_a, cut_a):
    """Triangles of the graph 'COM distance < cut' (minimum image, the list
    kernels' FP64 arithmetic in bohr): the FS3 triplet count when the cutoff is
    below half the box, and a third of the FS2 item count."""
    b2a = 0.529177249
    x = com_a / b2a
    L = box_a / b2a
    d = x[:, None, :] - x[None, :, :]
    d = d - L * (np.sign(d / L) * np.floor(np.abs(d / L) + 0.5))
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    a = (r2 < (cut_a / b2a) ** 2).astype(np.float32)
    np.fill_diagonal(a, 0.0)
    return int(round(float(((a @ a) * a).sum(dtype=np.float64)) / 6))


def test_large_box_lists_and_invariances(tb_fs2fs3, params_fs2fs3):
    """2048 waters (L = 39.4 A, both cutoffs below L/2): triplet and item counts
    against the triangles of the cutoff graphs; energies invariant and forces
    equivariant under a permutation of the molecules; total force zero
    (translation invariance of E)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox
    box = lattice_box(2048, seed=31)
    print("large box 2048 waters", flush=True)
    got = tb_fs2fs3.evaluate(box)
    assert got["n_fs3"] == _triangles(box.com, box.box, 6.3)
    assert got["n_fs2_items"] == 3 * _triangles(box.com, box.box, params_fs2fs3.rcut3)
    l3 = tb_fs2fs3.fs3_list()
    assert len(l3) == got["n_fs3"] and np.all(np.diff(l3[:, 0]) >= 0)
    assert np.all((l3[:, 0] < l3[:, 1]) & (l3[:, 1] < l3[:, 2]))
    perm = np.random.default_rng(4).permutation(box.nm)
    pb = WaterBox(box.com[perm], box.quats[perm], box.box)
    gp = tb_fs2fs3.evaluate(pb)
    assert gp["n_fs3"] == got["n_fs3"] and gp["n_fs2_items"] == got["n_fs2_items"]
    for k in ("qfs2", "qfs3"):
        assert abs(gp[k] - got[k]) <= 1e-11 * abs(got[k]), k
    assert abs(gp["vir"] - got["vir"]) <= 1e-10 * abs(got["vir"])
    fs, fsp = site_forces(got), site_forces(gp)
    assert rel_rms(fsp, fs[:, perm]) <= 1e-11
    for t in ("fs2", "fs3"):
        assert rel_rms(com_forces(gp, t), com_forces(got, t)[:, perm]) <= 1e-9
    total = fs.sum(axis=(1, 2)) + com_forces(got, "fs2").sum(1) + com_forces(got, "fs3").sum(1)
    assert np.all(np.abs(total) <= 1e-10 * np.abs(fs).sum())
