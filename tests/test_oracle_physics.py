"""Pins the oracle's FS2/FS3 physics by what the mathematics fixes: central
finite differences of the energy (forces, torques, rigid virial), symmetries
(periodic shift, translation, rotation, molecule permutation), Newton's third
law, region special cases, switch continuity and linearity in the fitted
coefficients.  The FS2 block restates the reference's own check
(source/tests/check_fs2_periodic.py) on its three fixture trimers."""
import numpy as np
import pytest

from helpers import (BOHR2A, INTERNAL_PER_KCAL, System, energy, fd_force, fd_torque,
                     molecular_forces, run_oracle, torques, trimer)

CONV = BOHR2A / INTERNAL_PER_KCAL  # internal force units -> kcal/mol/bohr (check_fs2_periodic.py:26)


def fs3_only(p):
    return p.replace(s2_term=False, s3_term=True)


def fs2_only(p):
    return p.replace(s2_term=True, s3_term=False)


# ---------------------------------------------------------------- FS2 (reference test)
def fs2_ref_system(geometry):
    """check_fs2_periodic.py:32-40 -- COM rows scaled by sqrt(geometry), 5 sites, bohr."""
    g = np.sqrt(geometry)
    com = np.array([[0.0, 0.0, 0.0], [5.5, 0.3, 0.2], [2.0, 5.0, 0.7]]) * g
    sx = [0.0, 1.43, -1.43, 0.3, -0.3]
    sy = [-0.1, 1.0, 1.0, -0.5, -0.5]
    sz = [0.0, 0.0, 0.0, 0.5, -0.5]
    site = np.repeat(np.stack([sx, sy, sz], 1)[None], 3, axis=0)
    return com, site


@pytest.fixture(scope="module")
def fs2_ref_params(params_fs2fs3):
    # check_fs2_periodic.py:14-18: alpha3=6.51292, rcutsd3=6.1**2, delta33=10
    return params_fs2fs3.replace(s2_term=True, s3_term=False, rcut3=6.1, alpha3=6.51292, delta33=10.0)


def _calc_bohr(O, p, com_b, site_b, box_b):
    """Call the oracle on bohr inputs (the reference test calls FS2_term in bohr)."""
    s = System(com_b * BOHR2A, site_b * BOHR2A, box_b * BOHR2A, ld=5)
    r = run_oracle(O, p, s)
    return r


def near(actual, expected, tol):
    assert abs(actual - expected) <= tol * max(1.0, abs(expected)), (actual, expected)


@pytest.mark.parametrize("geometry", [1, 2, 3])
def test_fs2_reference_check(oracle, fs2_ref_params, geometry):
    p = fs2_ref_params
    com, site = fs2_ref_system(geometry)
    base = _calc_bohr(oracle, p, com, site, 40.0)
    eb, vb = base["qfs2"], base["vir"]
    assert eb != 0.0
    fb = molecular_forces(base, 5)
    sb = np.stack([base["fxs"], base["fys"], base["fzs"]])
    for i in range(3):
        for a in range(3):
            for sg in (-1, 1):
                c2 = com.copy()
                c2[i, a] += sg * 40.0
                r = _calc_bohr(oracle, p, c2, site, 40.0)
                near(r["qfs2"], eb, 1e-10)
                near(r["vir"], vb, 1e-7)
                near(np.abs(molecular_forces(r, 5) - fb).max(), 0.0, 1e-7)
                near(np.abs(np.stack([r["fxs"], r["fys"], r["fzs"]]) - sb).max(), 0.0, 1e-7)
    # derivatives in a representation crossing periodic boundaries
    c = com.copy()
    c[0, 0] += 40.0
    c[1, 1] -= 40.0
    r = _calc_bohr(oracle, p, c, site, 40.0)
    fs = np.stack([r["fxs"], r["fys"], r["fzs"]], -1)[:, :5]
    vcom = (fs * site).sum() * CONV
    v = r["vir"] / INTERNAL_PER_KCAL + vcom
    f = molecular_forces(r, 5) * CONV
    h = 1e-5
    for i in range(3):
        for a in range(3):
            cp, cm = c.copy(), c.copy()
            cp[i, a] += h
            cm[i, a] -= h
            fd = -(_calc_bohr(oracle, p, cp, site, 40.0)["qfs2"] - _calc_bohr(oracle, p, cm, site, 40.0)["qfs2"]) / (2 * h)
            near(f[i, a], fd, 1e-7)
    ep = _calc_bohr(oracle, p, c * (1 + h), site, 40.0 * (1 + h))["qfs2"]
    em = _calc_bohr(oracle, p, c * (1 - h), site, 40.0 * (1 - h))["qfs2"]
    near(v, (ep - em) / (2 * h), 1e-7)


# ---------------------------------------------------------------- FS3 regions
REGION_CASES = [("sr", 3.9), ("switch", 5.0), ("lr", 5.75)]


@pytest.mark.parametrize("region,d", REGION_CASES)
@pytest.mark.parametrize("which", ["fs3", "fs2"])
def test_fd_forces_and_torques(oracle, params_fs2fs3, region, d, which):
    p = fs3_only(params_fs2fs3) if which == "fs3" else fs2_only(params_fs2fs3)
    s = trimer(d, seed=11, shape=(1.0, 1.04, 0.97))
    r = run_oracle(oracle, p, s)
    if which == "fs3":
        hist = {"sr": r["n_fs3_sr"], "switch": r["n_fs3_sw"], "lr": r["n_fs3_lr"]}
        assert hist[region] == 1, hist
    E = lambda t: energy(run_oracle(oracle, p, t))
    F = molecular_forces(r, 17) / INTERNAL_PER_KCAL
    T = torques(r, s, 17) / INTERNAL_PER_KCAL
    scale = max(np.abs(F).max(), np.abs(T).max())
    assert scale > 0
    for m in range(3):
        for a in range(3):
            assert abs(fd_force(E, s, m, a) - F[m, a]) <= 1e-6 * scale
            assert abs(fd_torque(E, s, m, a) - T[m, a]) <= 1e-6 * scale


@pytest.mark.parametrize("region,d", REGION_CASES)
def test_newton_third_law(oracle, params_fs2fs3, region, d):
    s = trimer(d, seed=5, shape=(1.0, 0.96, 1.05))
    r = run_oracle(oracle, params_fs2fs3, s)
    F = molecular_forces(r, 17)
    T = torques(r, s, 17)
    scale = np.abs(F).max()
    assert np.abs(F.sum(0)).max() < 1e-10 * scale
    L = (np.cross(s.com, F) + T).sum(0)
    assert np.abs(L).max() < 1e-10 * scale * np.abs(s.com).max()


@pytest.mark.parametrize("region,d", REGION_CASES)
def test_translation_rotation_permutation(oracle, params_fs2fs3, region, d):
    p = params_fs2fs3
    s = trimer(d, seed=7, shape=(1.03, 1.0, 0.98))
    r0 = run_oracle(oracle, p, s)
    e0 = energy(r0)
    F0 = molecular_forces(r0, 17)
    # translation
    t = s.copy()
    t.com += np.array([3.3, -1.7, 11.2])
    assert abs(energy(run_oracle(oracle, p, t)) - e0) <= 1e-10 * abs(e0) + 1e-12
    # global rotation
    from helpers import rotvec_matrix
    R = rotvec_matrix([0.3, -1.0, 0.4], 0.77)
    t = s.copy()
    t.com = t.com @ R.T
    t.site = t.site @ R.T
    r1 = run_oracle(oracle, p, t)
    assert abs(energy(r1) - e0) <= 1e-10 * abs(e0) + 1e-12
    np.testing.assert_allclose(molecular_forces(r1, 17), F0 @ R.T, rtol=1e-8, atol=1e-9 * np.abs(F0).max())
    # every molecule permutation
    import itertools
    for perm in itertools.permutations(range(3)):
        t = System(s.com[list(perm)], s.site[list(perm)], s.box, s.ld)
        rp = run_oracle(oracle, p, t)
        assert abs(energy(rp) - e0) <= 1e-10 * abs(e0) + 1e-12
        np.testing.assert_allclose(molecular_forces(rp, 17), F0[list(perm)], rtol=1e-8,
                                   atol=1e-9 * np.abs(F0).max())


def test_region_special_cases(oracle, params_fs2fs3):
    p = fs3_only(params_fs2fs3)
    rng = np.random.default_rng(3)
    noisy_lr = p.lr.__class__(beta=p.lr.beta, gamma=p.lr.gamma, r0=p.lr.r0,
                             cs=p.lr.cs + rng.normal(size=364))
    noisy_sr = p.sr.__class__(beta=p.sr.beta, gamma=p.sr.gamma, r0=p.sr.r0,
                             cs=p.sr.cs + rng.normal(size=364))
    s_sr = trimer(4.2, seed=1)
    s_lr = trimer(5.8, seed=1)
    e_sr = energy(run_oracle(oracle, p, s_sr))
    assert energy(run_oracle(oracle, p.replace(lr=noisy_lr), s_sr)) == e_sr
    e_lr = energy(run_oracle(oracle, p, s_lr))
    assert energy(run_oracle(oracle, p.replace(sr=noisy_sr), s_lr)) == e_lr
    s_sw = trimer(5.0, seed=1)
    e_sw = energy(run_oracle(oracle, p, s_sw))
    assert energy(run_oracle(oracle, p.replace(sr=noisy_sr), s_sw)) != e_sw
    assert energy(run_oracle(oracle, p.replace(lr=noisy_lr), s_sw)) != e_sw


def test_switch_continuity(oracle, params_fs2fs3):
    """E is continuous (C1 switch) across R_avg = 4.5 and 5.5 A."""
    p = fs3_only(params_fs2fs3)
    for edge in (4.5, 5.5):
        es, ds = [], []
        for eps in (-2e-6, -1e-6, 1e-6, 2e-6):
            s = trimer(edge + eps, seed=21)
            es.append(energy(run_oracle(oracle, p, s)))
        # the jump across the edge is as small as a step of the same size inside
        jump = abs(es[2] - es[1])
        inside = max(abs(es[1] - es[0]), abs(es[3] - es[2]))
        assert jump <= 3 * inside + 1e-14, (edge, es)


def test_linearity_in_coefficients(oracle, params_fs2fs3):
    p = fs3_only(params_fs2fs3)
    s = trimer(5.0, seed=9)
    rng = np.random.default_rng(4)
    c1, c2 = rng.normal(size=364), rng.normal(size=364)

    def e(csr, clr):
        q = p.replace(sr=p.sr.__class__(p.sr.beta, p.sr.gamma, p.sr.r0, csr),
                      lr=p.lr.__class__(p.lr.beta, p.lr.gamma, p.lr.r0, clr))
        return energy(run_oracle(oracle, q, s))

    z = np.zeros(364)
    assert e(z, z) == 0.0
    a = e(c1, c2)
    b = e(2.5 * c1, 2.5 * c2)
    assert abs(b - 2.5 * a) <= 1e-12 * abs(b)
    assert abs(e(c1 + c2, z) - e(c1, z) - e(c2, z)) <= 1e-11 * abs(e(c1 + c2, z))


def test_fd_rigid_virial_fs3_box(oracle, params_fs2fs3):
    """Rigid-molecule virial: dE/dlambda under COM + box scaling equals
    vir/418.4 + sum_s f_s.s/418.4 (the reference's criterion,
    check_fs2_periodic.py:82-84; initialization.f90:1407-1432)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    wb = lattice_box(16, seed=2, box=9.2)
    s = System.from_waterbox(wb)
    p = params_fs2fs3
    r = run_oracle(oracle, p, s)
    assert r["n_fs3"] > 0 and r["n_fs2"] > 0
    fs = np.stack([r["fxs"], r["fys"], r["fzs"]], -1)[:, :17]
    v = (r["vir"] + (fs * s.site[:, :17]).sum()) / INTERNAL_PER_KCAL
    h = 1e-6
    sp, sm = s.copy(), s.copy()
    sp.com *= 1 + h
    sp.box *= 1 + h
    sm.com *= 1 - h
    sm.box *= 1 - h
    dE = (energy(run_oracle(oracle, p, sp)) - energy(run_oracle(oracle, p, sm))) / (2 * h)
    assert abs(v - dE) <= 1e-6 * max(abs(dE), 1e-3), (v, dE)


def test_periodic_shift_box(oracle, params_fs2fs3):
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    wb = lattice_box(16, seed=2, box=9.2)
    s = System.from_waterbox(wb)
    r0 = run_oracle(oracle, params_fs2fs3, s)
    t = s.copy()
    t.com[3, 0] += s.box
    t.com[7, 2] -= 2 * s.box
    r1 = run_oracle(oracle, params_fs2fs3, t)
    assert r1["n_fs3"] == r0["n_fs3"] and r1["n_fs2"] == r0["n_fs2"]
    assert abs(energy(r1) - energy(r0)) <= 1e-10 * abs(energy(r0))
    np.testing.assert_allclose(molecular_forces(r1, 17), molecular_forces(r0, 17), rtol=1e-8,
                               atol=1e-10 * np.abs(molecular_forces(r0, 17)).max())


def test_fs3_com_conditioning(oracle, params_fs2fs3):
    """The FS3 COM force of a trimer is f_d' E + f_d w' E over region energies
    that cancel up to ~1e6-fold, so FP64 fixes it only to ~1e-16 times that
    cancellation.  Measured on the oracle itself: reversing the order of the
    molecules (the same physics, another summation order) moves the FS3 COM
    forces of the 40 trimers of the GPU parity set by more than 1e-9 relative
    for some of them, while their site forces stay within 1e-12.  This is the
    floor the GPU COM-force parity bar allows for (tests/test_gpu_parity.py)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox, trimer_set
    worst_com, worst_site = 0.0, 0.0
    for t in trimer_set(40, seed=17):
        ref = oracle.threebody_term(params_fs2fs3, t)
        perm = np.array([2, 1, 0])
        rp = oracle.threebody_term(params_fs2fs3, WaterBox(t.com[perm], t.quats[perm], t.box))
        inv = np.argsort(perm)
        c = np.stack([ref[k] for k in ("fxfs3", "fyfs3", "fzfs3")])
        cp = np.stack([rp[k] for k in ("fxfs3", "fyfs3", "fzfs3")])[:, inv]
        if np.abs(c).max() > 0:
            worst_com = max(worst_com, np.sqrt(((cp - c) ** 2).sum() / (c ** 2).sum()))
        f = np.stack([ref[k][:, :17] for k in ("fxs", "fys", "fzs")])
        fp = np.stack([rp[k][:, :17] for k in ("fxs", "fys", "fzs")])[:, inv]
        if np.abs(f).max() > 0:
            worst_site = max(worst_site, np.sqrt(((fp - f) ** 2).sum() / (f ** 2).sum()))
    print(f"oracle self-variation under molecule reversal: FS3 COM {worst_com:.2e}, site {worst_site:.2e}")
    assert worst_com > 1e-9
    assert worst_site < 1e-11
