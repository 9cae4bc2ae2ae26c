"""Independent pins of the oracle's FS2 value and of the FS3 LR and switch-band values.

The other oracle tests (finite differences, invariances, the reference's own
check_fs2_periodic.py) accept any self-consistent energy: a wrong sign of the
exchange-charge shift, a wrong charge placement or a wrong constant would pass
them.  This file fixes the values themselves:

* ``fs2_brute`` is a numpy restatement of the FS2 energy written directly from
  source/threebody_potential.f90:297-438 (item loop :297-336, damping :339-349,
  exchange charges :356-372, spectator vectors :385-418, energy :436-438, times
  h2kcal at :604).  It does not share code with the C oracle and takes the shift
  g as an argument.
* g3 is a default-REAL(4) literal in the reference (:177, no kind suffix; the
  Makefile passes no -fdefault-real-8), so its value is the float 0.4; the
  oracle must match the brute force at that value and miss it at the double 0.4
  by the first-order amount dE/dg * (0.4f - 0.4).
* Closed forms: g -> 0 gives E_FS2 = 0 and E is linear in g at small g.  The
  four exchange charges of a pair (+1 at R_i, R_j; -1 at R_i + g u_ij and
  R_j - g u_ij) carry no charge and no dipole; far away they act as a linear
  quadrupole -2 g (d + g) P2 / r^3 about the pair midpoint.  With the damping
  switched off (delta33 = 0) and a spectator far from a close pair, the oracle's
  energy must approach S_ij times the spectator's five charges in that
  potential: the sign fixes the direction of the shift and the value fixes which
  sites carry which charge.
* FS3: one-hot coefficient vectors make the energy an orbit sum that a
  pure-Python brute force evaluates; here on an LR trimer (the LR tables alone)
  and a switch-band trimer (both tables, weighted by the cubic switch written
  from its closed form, :1420-1423).
"""
import itertools

import numpy as np
import pytest

from helpers import BOHR2A, H2KCAL, System, run_oracle, trimer

# threebody_potential.f90:175-177, 197-201
ETA3 = 0.48056977605478400
BETA3 = 0.85836606579445000
G3_REF = float(np.float32(0.4))  # `0.400000000000000` is a default REAL(4) literal
CHARGES = np.array([0.258504607653328389, 0.564050362569283648, 0.564050362569283648,
                    -0.693302666395947953, -0.693302666395947953])


def _nint(v):
    """Fortran NINT: round half away from zero."""
    return np.where(v >= 0, np.floor(v + 0.5), -np.floor(-v + 0.5))


def fs2_brute(com_a, site_a, box_a, alpha3, delta33, rcut3, g):
    """E_FS2 in kcal/mol of one box, restated from threebody_potential.f90:297-438.
    com_a [nm,3], site_a [nm,>=5,3] in A; box_a <= 0 means no periodic images."""
    xx = np.asarray(com_a, float) / BOHR2A
    xs = np.asarray(site_a, float)[:, :5] / BOHR2A
    nm = len(xx)
    boxi = box_a / BOHR2A if box_a > 0 else 0.0
    alp3 = alpha3 * BOHR2A
    rc2 = rcut3 ** 2 * BOHR2A ** (-2.0)
    r0dmp = 7.40446 / BOHR2A
    k, i, j = np.meshgrid(np.arange(nm), np.arange(nm), np.arange(nm), indexing="ij")
    keep = (k != i) & (j > i) & (j != k)
    k, i, j = k[keep], i[keep], j[keep]

    def mi(d):
        return d - boxi * _nint(d / boxi) if boxi > 0 else d
    xij = mi(xx[i] - xx[j])
    xki = mi(xx[k] - xx[i])
    xkj = xki + xij  # the same periodic image for all three vectors (:325-328)
    dij2 = (xij ** 2).sum(-1)
    dki2 = (xki ** 2).sum(-1)
    dkj2 = (xkj ** 2).sum(-1)
    ok = (dij2 < rc2) & (dki2 < rc2) & (dkj2 < rc2)
    k, i, j = k[ok], i[ok], j[ok]
    xij, xki, xkj, dij2, dki2, dkj2 = xij[ok], xki[ok], xkj[ok], dij2[ok], dki2[ok], dkj2[ok]
    flrki = (1.0 / (1.0 + np.exp(alp3 * (np.sqrt(dki2) - r0dmp)))) ** delta33
    flrkj = (1.0 / (1.0 + np.exp(alp3 * (np.sqrt(dkj2) - r0dmp)))) ** delta33
    # exchange overlap of the O,H,H sites of i and j
    S = np.zeros(len(k))
    for a in range(3):
        for b in range(3):
            v = xij + (xs[i, a] - xs[j, b])
            S += ETA3 * np.exp(-BETA3 * np.sqrt((v ** 2).sum(-1)))
    u = xij / np.sqrt(dij2)[:, None]
    T = np.zeros(len(k))
    for ks in range(5):
        kc_i = xki + xs[k, ks]            # site ks of k minus R_i
        kc_j = xkj + xs[k, ks]            # site ks of k minus R_j
        kq_i = kc_i - g * u               # minus the exchange charge i+ = R_i + g u_ij
        kq_j = kc_j + g * u               # minus the exchange charge j+ = R_j - g u_ij
        inv = lambda w: 1.0 / np.sqrt((w ** 2).sum(-1))
        T += CHARGES[ks] * (inv(kc_i) + inv(kc_j) - inv(kq_i) - inv(kq_j))
    return float((S * T * flrki * flrkj).sum() * H2KCAL), len(k)


def fs2_only(p, **kw):
    return p.replace(s2_term=True, s3_term=False, **kw)


def brute_of(p, s: System, g=G3_REF):
    return fs2_brute(s.com, s.site, s.box, p.alpha3, p.delta33, p.rcut3, g)


def fs2_ref_trimer(geometry):
    """The three trimers of check_fs2_periodic.py:32-40 (bohr inputs, box 40 bohr)."""
    gs = np.sqrt(geometry)
    com = np.array([[0.0, 0.0, 0.0], [5.5, 0.3, 0.2], [2.0, 5.0, 0.7]]) * gs
    sx = [0.0, 1.43, -1.43, 0.3, -0.3]
    sy = [-0.1, 1.0, 1.0, -0.5, -0.5]
    sz = [0.0, 0.0, 0.0, 0.5, -0.5]
    site = np.repeat(np.stack([sx, sy, sz], 1)[None], 3, axis=0)
    return System(com * BOHR2A, site * BOHR2A, 40.0 * BOHR2A, ld=5)


def test_g3_is_the_single_precision_literal():
    assert G3_REF == 0.4000000059604645
    assert G3_REF != 0.4


@pytest.mark.parametrize("geometry", [1, 2, 3])
def test_fs2_value_reference_trimers(oracle, params_fs2fs3, geometry):
    # check_fs2_periodic.py:14-18: alpha3 = 6.51292, rcutsd3 = 6.1**2, delta33 = 10
    p = fs2_only(params_fs2fs3, rcut3=6.1, alpha3=6.51292, delta33=10.0)
    s = fs2_ref_trimer(geometry)
    e_or = run_oracle(oracle, p, s)["qfs2"]
    e_bf, n = brute_of(p, s)
    assert n > 0 and e_bf != 0.0
    assert abs(e_or - e_bf) <= 1e-12 * abs(e_bf), (e_or, e_bf)


def test_fs2_value_box64(oracle, params_fs2fs3):
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    p = fs2_only(params_fs2fs3)
    wb = lattice_box(64, seed=5)
    s = System.from_waterbox(wb)
    r = run_oracle(oracle, p, s)
    e_bf, n = brute_of(p, s)
    assert n == r["n_fs2"]
    assert abs(r["qfs2"] - e_bf) <= 1e-11 * abs(e_bf), (r["qfs2"], e_bf)


@pytest.mark.parametrize("seed", [2, 9, 31])
def test_fs2_value_random_trimers(oracle, params_fs2fs3, seed):
    p = fs2_only(params_fs2fs3)
    s = trimer(4.3 + 0.4 * (seed % 3), seed=seed, shape=(1.0, 1.08, 0.93))
    e_or = run_oracle(oracle, p, s)["qfs2"]
    e_bf, n = brute_of(p, s)
    assert n == 3 and e_bf != 0.0
    assert abs(e_or - e_bf) <= 1e-12 * abs(e_bf), (e_or, e_bf)


def test_fs2_g3_value_is_pinned(oracle, params_fs2fs3):
    """The oracle sits on the brute force at g = 0.4f, and off it at g = 0.4 by
    the first-order shift dE/dg * (0.4f - 0.4) (about 1.5e-8 relative)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    p = fs2_only(params_fs2fs3)
    s = System.from_waterbox(lattice_box(64, seed=5))
    e_or = run_oracle(oracle, p, s)["qfs2"]
    e_f = brute_of(p, s, G3_REF)[0]
    e_d = brute_of(p, s, 0.4)[0]
    h = 1e-4
    dedg = (brute_of(p, s, 0.4 + h)[0] - brute_of(p, s, 0.4 - h)[0]) / (2 * h)
    predicted = dedg * (G3_REF - 0.4)
    assert abs((e_f - e_d) - predicted) <= 1e-3 * abs(predicted)
    assert abs(e_or - e_f) <= 0.02 * abs(predicted)   # on the float-literal value
    assert abs(e_or - e_d) >= 0.9 * abs(predicted)    # and measurably off the double one


def test_fs2_vanishes_without_shift(params_fs2fs3):
    s = trimer(4.6, seed=4, shape=(1.0, 1.05, 0.96))
    p = fs2_only(params_fs2fs3)
    e0, n = brute_of(p, s, 0.0)
    e3 = brute_of(p, s, G3_REF)[0]
    # ((a + b) - a) - b of the bracket is zero up to its rounding
    assert n == 3 and e3 != 0.0 and abs(e0) <= 1e-12 * abs(e3)
    # and E is linear in g at small g
    e_small = brute_of(p, s, 1e-6)[0]
    assert abs(e_small / 1e-6 - (brute_of(p, s, 2e-6)[0] - e_small) / 1e-6) <= 1e-4 * abs(e_small / 1e-6)


def quadrupole_potential(g, d, u, r):
    """Potential at r (relative to the pair midpoint) of the four charges +1 at
    +-(d/2) u and -1 at +-(d/2 + g) u, to leading order in 1/|r|: the set has
    no charge and no dipole, and its linear quadrupole sum_q q z^2 = -2 g (d + g)
    gives phi = -2 g (d + g) P2(u . r^) / |r|^3 (all orders in g)."""
    rn = np.linalg.norm(r, axis=-1)
    cth = (r @ u) / rn
    return -2.0 * g * (d + g) * 0.5 * (3.0 * cth ** 2 - 1.0) / rn ** 3


@pytest.mark.parametrize("dist_a,tol", [(160.0, 0.02), (320.0, 0.01)])
def test_fs2_distant_spectator_closed_form(oracle, params_fs2fs3, dist_a, tol):
    """Damping off (delta33 = 0), cutoff far away: with molecule k at distance D
    from a close pair (i, j), every item but (k; i, j) has a vanishing exchange
    overlap, and E_FS2/h2kcal -> S_ij sum_ks q_ks phi(r_ks - M) with phi the
    potential of the outward-shifted exchange charges (i+ = R_i + g u_ij,
    j+ = R_j - g u_ij) and M the pair midpoint.  The next multipole is smaller
    by (d/D)^2 per site, but the neutral spectator cancels its sum to ~1e-3 of
    sum |q phi|, so the closed form approaches the oracle as 1/D (3.5 % at
    40 A, 0.7 % at 160 A, 0.25 % at 320 A).  An inward shift flips the sign, a
    charge on the wrong site or a wrong g changes the value by O(1)."""
    rng = np.random.default_rng(7)
    base = trimer(3.0, seed=12, shape=(1.0, 1.0, 1.0))
    com = base.com.copy()
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    com[2] = 0.5 * (com[0] + com[1]) + dist_a * direction
    s = System(com, base.site, 0.0, ld=base.ld)
    p = fs2_only(params_fs2fs3, rcut3=1000.0, delta33=0.0)
    e_or = run_oracle(oracle, p, s)["qfs2"] / H2KCAL
    xb = s.com / BOHR2A
    st = s.site[:, :5] / BOHR2A
    xij = xb[0] - xb[1]
    d = np.linalg.norm(xij)
    u = xij / d
    S = sum(ETA3 * np.exp(-BETA3 * np.linalg.norm(xij + st[0, a] - st[1, b]))
            for a in range(3) for b in range(3))
    mid = 0.5 * (xb[0] + xb[1])
    closed = S * (CHARGES * quadrupole_potential(G3_REF, d, u, xb[2] + st[2] - mid)).sum()
    assert closed != 0.0
    assert abs(e_or - closed) <= tol * abs(closed), (e_or, closed)
    # the inward shift gives the opposite sign
    inward = S * (CHARGES * quadrupole_potential(-G3_REF, d, u, xb[2] + st[2] - mid)).sum()
    assert np.sign(inward) != np.sign(e_or)


# ------------------------------------------------------------- FS3 LR and switch band
NTYPE = [1, 2, 2, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6]


def orbit_of_labels():
    seen = {}
    for iat, jat, kat in itertools.product(range(17), repeat=3):
        for ip, jp, kp in itertools.product(range(2), repeat=3):
            key = tuple(sorted([(NTYPE[iat], ip), (NTYPE[jat], jp), (NTYPE[kat], kp)]))
            seen.setdefault(key, len(seen))
    return seen


def orbit_sum(tab, s: System, n_orbit):
    """Sum over all site triples and powers whose label multiset is orbit n of
    e(ab) e(ac) e(bc) r_bc^p r_ac^q r_ab^s (bohr), e from the pair table `tab`."""
    keys = orbit_of_labels()
    target = {v: k for k, v in keys.items()}[n_orbit]
    pos = s.com[:, None, :] / BOHR2A + s.site[:, :17] / BOHR2A
    dist = lambda m1, m2: np.linalg.norm(pos[m1][:, None, :] - pos[m2][None, :, :], axis=-1)
    rab, rac, rbc = dist(0, 1), dist(0, 2), dist(1, 2)
    B, G, R0 = tab.beta, tab.gamma, tab.r0
    e = lambda r, a, b: np.exp(-B[a, b] * r) / (1 + np.exp(-G[a, b] * (r - R0[a, b])))
    tot = 0.0
    for a, b, c in itertools.product(range(17), repeat=3):
        for pa, pb, pc in itertools.product(range(2), repeat=3):
            if tuple(sorted([(NTYPE[a], pa), (NTYPE[b], pb), (NTYPE[c], pc)])) != target:
                continue
            tot += (e(rab[a, b], a, b) * e(rac[a, c], a, c) * e(rbc[b, c], b, c)
                    * rbc[b, c] ** pa * rac[a, c] ** pb * rab[a, b] ** pc)
    return tot


def damping_and_switch(p, s: System):
    """f_d(R_avg) (:894-904) and the SR weight of the cubic switch (:1420-1423),
    from their closed forms: f_d = (1 + exp(alpha (R_avg - 6.20446 A)))^-delta,
    w_SR = 1 + 2c^3 - 3c^2 with c = (R_avg - 4.5 A) / 1 A, clamped to [0, 1]."""
    R = np.linalg.norm(s.com[:, None] - s.com[None], axis=-1)
    ravg_a = (R[0, 1] + R[1, 2] + R[0, 2]) / 3
    fd = (1.0 / (1.0 + np.exp(p.alpha3 * (ravg_a - 6.20446)))) ** p.delta33
    c = min(max(ravg_a - 4.5, 0.0), 1.0)
    return ravg_a, fd, 1.0 + 2 * c ** 3 - 3 * c ** 2


def onehot(n):
    v = np.zeros(364)
    v[n] = 1.0
    return v


@pytest.mark.parametrize("n_orbit", [0, 57, 200, 363])
def test_one_hot_lr_region(oracle, params_fs2fs3, n_orbit):
    p = params_fs2fs3.replace(s2_term=False)
    s = trimer(5.8, seed=21, shape=(1.0, 1.03, 0.98))
    ravg, fd, w_sr = damping_and_switch(p, s)
    assert ravg >= 5.5 and w_sr == 0.0
    Tab = p.sr.__class__
    q = p.replace(sr=Tab(p.sr.beta, p.sr.gamma, p.sr.r0, np.zeros(364)),
                  lr=Tab(p.lr.beta, p.lr.gamma, p.lr.r0, onehot(n_orbit)))
    r = run_oracle(oracle, q, s)
    assert r["n_fs3_lr"] == 1
    ref = fd * orbit_sum(p.lr, s, n_orbit)
    assert ref != 0.0
    assert abs(r["qfs3"] / H2KCAL - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("n_orbit", [3, 140, 301])
def test_one_hot_switch_band(oracle, params_fs2fs3, n_orbit):
    p = params_fs2fs3.replace(s2_term=False)
    s = trimer(5.05, seed=8, shape=(1.0, 1.04, 0.97))
    ravg, fd, w_sr = damping_and_switch(p, s)
    assert 4.5 < ravg < 5.5 and 0.0 < w_sr < 1.0
    Tab = p.sr.__class__
    only_sr = p.replace(sr=Tab(p.sr.beta, p.sr.gamma, p.sr.r0, onehot(n_orbit)),
                        lr=Tab(p.lr.beta, p.lr.gamma, p.lr.r0, np.zeros(364)))
    only_lr = p.replace(sr=Tab(p.sr.beta, p.sr.gamma, p.sr.r0, np.zeros(364)),
                        lr=Tab(p.lr.beta, p.lr.gamma, p.lr.r0, onehot(n_orbit)))
    e_sr = run_oracle(oracle, only_sr, s)
    e_lr = run_oracle(oracle, only_lr, s)
    assert e_sr["n_fs3_sw"] == 1
    ref_sr = fd * w_sr * orbit_sum(p.sr, s, n_orbit)
    ref_lr = fd * (1.0 - w_sr) * orbit_sum(p.lr, s, n_orbit)
    assert abs(e_sr["qfs3"] / H2KCAL - ref_sr) <= 1e-12 * abs(ref_sr)
    assert abs(e_lr["qfs3"] / H2KCAL - ref_lr) <= 1e-12 * abs(ref_lr)
