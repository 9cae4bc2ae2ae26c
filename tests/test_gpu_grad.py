"""GPU tests of the gradient path: per-box coefficient gradients, the
torch.autograd.Function, and the matrix-core Gram matrix.

Pinned by: linearity of E in the coefficients (phi . c = qfs3); the oracle's
energy with one-hot coefficient vectors (phi_n = E(c = e_n)); central finite
differences of the GPU energy in COM positions (h = 1e-5 A), orientations
(h = 1e-5 rad) and coefficients; and an extended-precision Gram reference."""
import numpy as np
import pytest
import torch

from ccpol23_plus_2025_mbmd_amd.geometry import principal_coords, trimer_set

pytestmark = pytest.mark.gpu

NLIN = 364


def pack(boxes):
    com = np.concatenate([b.com for b in boxes])
    site = np.concatenate([b.site for b in boxes])
    ms = np.cumsum([0] + [b.nm for b in boxes]).astype(np.int32)
    bl = np.array([b.box for b in boxes], float)
    return com, site, ms, bl


def coef_vector(p):
    return np.concatenate([p.sr.cs, p.lr.cs])


def away_from_cutoffs(b, h=0.01):
    d = np.linalg.norm(b.com[:, None] - b.com[None], axis=-1)[np.triu_indices(3, 1)]
    return np.all(np.abs(d - 6.3) > h) and np.all(np.abs(d - 7.5) > h)


@pytest.fixture(scope="module")
def trimers():
    ts = [b for b in trimer_set(96, seed=44) if away_from_cutoffs(b)]
    return ts[:64]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_phi_linear_in_coefficients(params_fs2fs3, trimers, precision):
    """sum_n phi[b, n] c_n equals the FS3 energy of every box."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    print(f"phi linearity: {len(trimers)} trimers, {precision}", flush=True)
    tb = ThreeBody(params_fs2fs3, precision)
    com, site, ms, bl = pack(trimers)
    dev = torch.device("cuda:0")
    com_t = torch.tensor(com.T.copy(), device=dev)
    site_t = torch.tensor(site.transpose(2, 0, 1).copy(), device=dev)
    _, _, _, en, phi = tb.eval_batch_grad_device(ms, bl, com_t, site_t)
    c = torch.tensor(coef_vector(params_fs2fs3), device=dev)
    e3 = en[:, 1]
    lin = phi @ c
    tol = 1e-10 if precision == "fp64" else 1e-4
    scale = (phi.abs() @ c.abs())
    assert torch.all((lin - e3).abs() <= tol * scale + 1e-12), (lin - e3).abs().max().item()
    # the plain batch path gives the same energies
    _, _, _, en2 = tb.eval_batch_device(ms, bl, com_t, site_t)
    assert torch.allclose(en2, en, rtol=1e-12, atol=1e-13)


def test_phi_matches_oracle_one_hot(params_fs2fs3, oracle):
    """phi[b, n] = qfs3 of the oracle with the coefficient vector e_n (SR and LR)."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    from helpers import System, run_oracle
    ts = [b for b in trimer_set(400, seed=7) if away_from_cutoffs(b)]
    # one trimer per region: SR, switch band, LR (R_avg from the COM triangle)
    def ravg(b):
        return np.linalg.norm(b.com[:, None] - b.com[None], axis=-1)[np.triu_indices(3, 1)].mean()
    def dmax(b):
        return np.linalg.norm(b.com[:, None] - b.com[None], axis=-1).max()
    pick = [next(b for b in ts if ravg(b) < 4.4), next(b for b in ts if 4.6 < ravg(b) < 5.4),
            next(b for b in ts if 5.6 < ravg(b) < 6.0 and dmax(b) < 6.25)]
    tb = ThreeBody(params_fs2fs3, "fp64")
    com, site, ms, bl = pack(pick)
    dev = torch.device("cuda:0")
    _, _, _, en, phi = tb.eval_batch_grad_device(ms, bl, torch.tensor(com.T.copy(), device=dev),
                                                 torch.tensor(site.transpose(2, 0, 1).copy(), device=dev))
    phi = phi.cpu().numpy()
    p = params_fs2fs3.replace(s2_term=False)
    for bi, b in enumerate(pick):
        s = System(b.com, b.site, 0.0, ld=28)
        for col in (0, 5, 130, 363, NLIN + 0, NLIN + 77, NLIN + 363):
            onehot = np.zeros(2 * NLIN)
            onehot[col] = 1.0
            q = p.replace(sr=p.sr.__class__(p.sr.beta, p.sr.gamma, p.sr.r0, onehot[:NLIN]),
                          lr=p.lr.__class__(p.lr.beta, p.lr.gamma, p.lr.r0, onehot[NLIN:]))
            ref = run_oracle(oracle, q, s)["qfs3"]
            assert abs(phi[bi, col] - ref) <= 1e-10 * abs(ref) + 1e-15, (bi, col, phi[bi, col], ref)


def rotate_quat(q, axis, ang):
    """q' such that R(q') = R(axis, ang) R(q) (lab-frame rotation)."""
    axis = axis / np.linalg.norm(axis)
    r = np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis])
    w1, v1 = r[0], r[1:]
    w2, v2 = q[0], q[1:]
    return np.concatenate([[w1 * w2 - v1 @ v2], w1 * v2 + w2 * v1 + np.cross(v1, v2)])


def test_autograd_matches_finite_differences(params_fs2fs3, trimers):
    """dE/d(COM), dE/d(orientation) and dE/dc through torch.autograd against
    central differences of the GPU FP64 energy."""
    from ccpol23_plus_2025_mbmd_amd.autograd import threebody_energy
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb = ThreeBody(params_fs2fs3, "fp64")
    dev = torch.device("cuda:0")
    sel = trimers[:12]
    com0 = np.concatenate([b.com for b in sel])
    q0 = np.concatenate([b.quats for b in sel])
    ms = np.cumsum([0] + [3] * len(sel)).astype(np.int32)
    bl = np.zeros(len(sel))
    pc = torch.tensor(principal_coords()[1], device=dev)
    c0 = torch.tensor(coef_vector(params_fs2fs3), device=dev)
    print(f"autograd FD: {len(sel)} trimers", flush=True)

    def E(com, quat, coef):
        return threebody_energy(tb, com, quat, coef, pc, ms, bl)

    com = torch.tensor(com0, device=dev, requires_grad=True)
    quat = torch.tensor(q0, device=dev, requires_grad=True)
    coef = c0.clone().requires_grad_(True)
    e = E(com, quat, coef)
    w = torch.linspace(0.5, 1.5, len(sel), device=dev, dtype=torch.float64)
    (w * e).sum().backward()
    g_com, g_quat, g_coef = com.grad.cpu().numpy(), quat.grad.cpu().numpy(), coef.grad.cpu().numpy()

    def L(comn, quatn, coefn=c0):
        with torch.no_grad():
            return float((w * E(torch.tensor(comn, device=dev), torch.tensor(quatn, device=dev), coefn)).sum())

    h = 1e-5
    scale = np.abs(g_com).max()
    for m in (0, 4, 11, 20, 33):
        for a in range(3):
            cp, cm = com0.copy(), com0.copy()
            cp[m, a] += h
            cm[m, a] -= h
            fd = (L(cp, q0) - L(cm, q0)) / (2 * h)
            assert abs(fd - g_com[m, a]) <= 1e-5 * scale + 1e-8, (m, a, fd, g_com[m, a])
    # orientation: rotation by +-h about a lab axis; dE/dtheta = sum over sites of (e x s) . dE/ds
    for m in (1, 7, 30):
        for a in range(3):
            e_ax = np.eye(3)[a]
            qp, qm = q0.copy(), q0.copy()
            qp[m] = rotate_quat(q0[m], e_ax, h)
            qm[m] = rotate_quat(q0[m], e_ax, -h)
            fd = (L(com0, qp) - L(com0, qm)) / (2 * h)
            # chain rule through the quaternion: d q / d theta for the lab rotation
            dq = (rotate_quat(q0[m], e_ax, 1e-7) - rotate_quat(q0[m], e_ax, -1e-7)) / 2e-7
            an = float(g_quat[m] @ dq)
            assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3) + 1e-8, (m, a, fd, an)
    # coefficients: E is linear in c, so central differences are exact up to round-off
    for col in (3, 200, NLIN + 10, NLIN + 300):
        cp, cm = c0.clone(), c0.clone()
        cp[col] += 1.0
        cm[col] -= 1.0
        fd = (L(com0, q0, cp) - L(com0, q0, cm)) / 2.0
        assert abs(fd - g_coef[col]) <= 1e-9 * abs(fd) + 1e-13 * max(1.0, abs(L(com0, q0))), (col, fd, g_coef[col])


def test_gram_matches_extended_precision(params_fs2fs3):
    """FP64 matrix-core Gram against a long-double reference (small, ragged
    shapes) and FP64 BLAS (the 4096 x 728 refit shape), within the FP64
    dot-product bound |fl(sum a b) - sum a b| <= n u sum |a| |b|."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb = ThreeBody(params_fs2fs3, "fp64")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(3)
    u = 2.0 ** -53
    # columns of very different scales and signs, like SR/LR coefficient gradients;
    # shapes that are not multiples of the 64 x 64 tiles or of the 16-row stages
    for nrow, ncol in ((300, 70), (1, 5), (1037, 129)):
        a = rng.normal(size=(nrow, ncol)) * np.exp(rng.uniform(-20, 5, size=ncol))
        a[:, min(7, ncol - 1)] = 0.0
        print(f"gram {nrow} x {ncol}", flush=True)
        g = tb.gram(torch.tensor(a, device=dev)).cpu().numpy()
        al = a.astype(np.longdouble)
        ref = (al.T @ al).astype(np.float64)
        bound = (nrow + 2) * u * (np.abs(a).T @ np.abs(a)) + 1e-300
        assert np.all(np.abs(g - ref) <= bound), np.max(np.abs(g - ref) / bound)
        assert np.array_equal(g, g.T)
    # no rows: the zero matrix
    assert not tb.gram(torch.zeros((0, 9), dtype=torch.float64, device=dev)).any()
    # the config-4 shape: 4096 x 728
    b = rng.normal(size=(4096, 728)) * np.exp(rng.uniform(-10, 3, size=728))
    bt = torch.tensor(b, device=dev)
    g = tb.gram(bt)
    ref = bt.T @ bt
    m = bt.abs().max(0).values
    bound = 1e-12 * (bt.abs().T @ bt.abs()) + 16 * 2.0 ** -53 * 4096 * torch.outer(m, m)
    err = ((g - ref).abs() / bound).max().item()
    assert err <= 1.0, err


def test_refit_reproduces_energies_and_writes_tables(tmp_path, params_fs2fs3):
    """Fit the 728 coefficients to FS3 energies of 1024 trimers made with the
    shipped coefficients; the fit reproduces them, and the data3b_full tables
    written from it, read back into a new context, give the fitted energies."""
    import shutil
    from ccpol23_plus_2025_mbmd_amd.params import DATA_DIR, load_variant
    from ccpol23_plus_2025_mbmd_amd.refit import design_matrix, solve_coefficients, write_fit
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    ts = [b for b in trimer_set(1400, seed=12) if away_from_cutoffs(b)][:1024]
    com, site, ms, bl = pack(ts)
    dev = torch.device("cuda:0")
    com_t = torch.tensor(com.T.copy(), device=dev)
    site_t = torch.tensor(site.transpose(2, 0, 1).copy(), device=dev)
    print(f"refit: {len(ts)} trimers", flush=True)
    tb = ThreeBody(params_fs2fs3, "fp64")
    phi, y = design_matrix(tb, ms, bl, com_t, site_t)
    c, res = solve_coefficients(tb, phi, y)
    scale = float(torch.sqrt(torch.mean(y ** 2)))
    print(f"refit residual {res:.3e} kcal/mol, energy RMS {scale:.3e}", flush=True)
    assert res <= 2e-5 * scale, (res, scale)
    # weights only reweight the residual; an exact target set stays reproduced
    w = torch.linspace(0.5, 2.0, len(ts), dtype=torch.float64, device=dev)
    _, res_w = solve_coefficients(tb, phi, y, weights=w)
    assert res_w <= 2e-5 * scale
    # tables written from the fit, loaded as a parameter set of their own
    d = tmp_path / "fit"
    write_fit(d, params_fs2fs3, c)
    shutil.copy(DATA_DIR / "fs2fs3" / "CONTROL", d / "CONTROL")
    q = load_variant("fs2fs3", d)
    assert np.array_equal(np.concatenate([q.sr.cs, q.lr.cs]), c.cpu().numpy())
    tb2 = ThreeBody(q, "fp64")
    _, _, _, en2 = tb2.eval_batch_device(ms, bl, com_t, site_t)
    fitted = phi @ c
    assert torch.allclose(en2[:, 1], fitted, rtol=1e-9, atol=1e-9 * scale)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autograd_matches_reference_analytic_forces(params_fs2fs3, trimers, oracle, precision):
    """dE/d(COM) through torch.autograd equals minus the oracle's analytic
    molecular force (sites + FS2 and FS3 COM terms) of every trimer: FP64 within
    1e-9 relative RMS; mixed within its molecular-force error, 1e-3."""
    from ccpol23_plus_2025_mbmd_amd.autograd import KCAL_PER_FORCE_UNIT, threebody_energy
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    from helpers import System, molecular_forces, run_oracle
    tb = ThreeBody(params_fs2fs3, precision)
    dev = torch.device("cuda:0")
    sel = trimers[:24]
    ms = np.cumsum([0] + [3] * len(sel)).astype(np.int32)
    bl = np.zeros(len(sel))
    pc = torch.tensor(principal_coords()[1], device=dev)
    com = torch.tensor(np.concatenate([b.com for b in sel]), device=dev, requires_grad=True)
    quat = torch.tensor(np.concatenate([b.quats for b in sel]), device=dev)
    coef = torch.tensor(coef_vector(params_fs2fs3), device=dev)
    print(f"autograd vs oracle forces: {len(sel)} trimers, {precision}", flush=True)
    threebody_energy(tb, com, quat, coef, pc, ms, bl).sum().backward()
    got = com.grad.cpu().numpy()
    ref = np.concatenate([-molecular_forces(run_oracle(oracle, params_fs2fs3, System(b.com, b.site, 0.0, ld=28)), 17)
                          for b in sel]) * KCAL_PER_FORCE_UNIT
    err = np.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum())
    assert err <= (1e-9 if precision == "fp64" else 1e-3), err


def rotation_residual(com, g_com, quat, g_quat):
    """Per box of 3 molecules and lab axis e: sum_m (e x x_m) . dE/dx_m + dE/dq_m . dq_m/dtheta,
    the derivative of E under a rigid rotation of the whole trimer (zero for an
    isolated trimer), with dq/dtheta = (0, e) (x) q / 2; and the sum of the
    magnitudes of its terms."""
    nb = len(com) // 3
    res = np.zeros((nb, 3))
    mag = np.zeros((nb, 3))
    for a in range(3):
        e = np.eye(3)[a]
        t1 = (np.cross(e, com) * g_com).sum(1)
        q0, qv = quat[:, 0], quat[:, 1:]
        dq = 0.5 * np.concatenate([(-(qv @ e))[:, None], q0[:, None] * e + np.cross(e, qv)], 1)
        t2 = (g_quat * dq).sum(1)
        res[:, a] = (t1 + t2).reshape(nb, 3).sum(1)
        mag[:, a] = (np.abs(t1) + np.abs(t2)).reshape(nb, 3).sum(1)
    return res, mag


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_4096_trimers_gradient_path(params_fs2fs3, precision):
    """The differentiation workload at full size: 4096 isolated trimers, R_avg
    uniform in 2.5-9.0 A (SR, switch and LR regions), monomer separations
    >= 2.3 A, random orientations, through torch.autograd in one batched call
    (energies, dE/d(COM), dE/d(orientation) via the quaternions, dE/dc), with
    size-independent properties: every region present; E_FS3 = phi . c for
    each trimer (E is linear in the coefficients) and dE/dc = sum_t w_t phi_t;
    the COM gradients of each trimer sum to zero (translation invariance) and
    the derivative under a rigid rotation of each trimer vanishes (rotation
    invariance), to FP64 rounding (mixed: to FP32 rounding); the Gram matrix of the 4096 x 728 coefficient gradients on the
    matrix cores against FP64 torch.  Mixed precision also against the FP64
    gradients (measured 6.4e-5 COM, 5.2e-5 orientation, 2.7e-8 coefficients;
    asserted at 2e-4, 2e-4, 1e-7)."""
    from ccpol23_plus_2025_mbmd_amd.autograd import threebody_energy
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    tri = trimer_set(4096, seed=2025)
    com0, site0, ms, bl = pack(tri)
    q0 = np.concatenate([b.quats for b in tri])
    pc = torch.tensor(principal_coords()[1], device=dev)
    c0 = coef_vector(params_fs2fs3)
    w = torch.linspace(0.5, 1.5, len(tri), device=dev, dtype=torch.float64)
    print(f"4096 trimers through the gradient path, {precision}", flush=True)

    def grads(prec):
        tb = ThreeBody(params_fs2fs3, prec)
        com = torch.tensor(com0, device=dev, requires_grad=True)
        quat = torch.tensor(q0, device=dev, requires_grad=True)
        coef = torch.tensor(c0, device=dev, requires_grad=True)
        e = threebody_energy(tb, com, quat, coef, pc, ms, bl)
        (w * e).sum().backward()
        st = tb.stats()
        com_t = torch.tensor(com0.T.copy(), device=dev).contiguous()
        site_t = torch.tensor(site0.transpose(2, 0, 1).copy(), device=dev).contiguous()
        _, _, _, en, phi = tb.eval_batch_grad_device(ms, bl, com_t, site_t)
        return (e.detach().cpu().numpy(), com.grad.cpu().numpy(), quat.grad.cpu().numpy(),
                coef.grad.cpu().numpy(), en.cpu().numpy(), phi, st, tb)

    e, g_com, g_quat, g_coef, en, phi, st, tb = grads(precision)
    print(f"FS3 triplets: SR {st['n_fs3_sr']}, switch {st['n_fs3_switch']}, LR {st['n_fs3_lr']} "
          f"(the other trimers have a pair beyond the 6.3 A FS3 cutoff)", flush=True)
    assert st["n_fs3_sr"] > 50 and st["n_fs3_switch"] > 50 and st["n_fs3_lr"] > 50, st
    phi_h = phi.cpu().numpy()
    lin = np.abs(phi_h @ c0 - en[:, 1])
    scale_e = np.abs(phi_h) @ np.abs(c0)
    assert np.all(lin <= (1e-12 if precision == "fp64" else 1e-5) * scale_e + 1e-15), (lin / scale_e).max()
    wg = w.cpu().numpy() @ phi_h
    assert np.all(np.abs(g_coef - wg) <= 1e-12 * (w.cpu().numpy() @ np.abs(phi_h)) + 1e-300)
    # FP64: invariance to rounding.  Mixed: FS2's site forces on the spectator and
    # COM forces on the pair come from FP32 derivatives formed by different
    # threads, so their sum vanishes to FP32 rounding of the terms (measured
    # 1.0e-5 of the summed magnitudes)
    inv_tol = 1e-9 if precision == "fp64" else 1e-4
    tsum = np.abs(g_com.reshape(-1, 3, 3).sum(1))
    tmag = np.abs(g_com.reshape(-1, 3, 3)).sum(1)
    print(f"translation residual max {(tsum / (tmag + 1e-300)).max():.1e}", flush=True)
    assert np.all(tsum <= inv_tol * tmag + 1e-15)
    res, mag = rotation_residual(com0, g_com, q0, g_quat)
    print(f"rotation residual max {(np.abs(res) / (mag + 1e-300)).max():.1e}", flush=True)
    assert np.all(np.abs(res) <= inv_tol * mag + 1e-15)
    # the Gram of this phi on the matrix cores
    g = tb.gram(phi)
    ref = phi.T @ phi
    m = phi.abs().max(0).values
    bound = 1e-12 * (phi.abs().T @ phi.abs()) + 16 * 2.0 ** -53 * len(tri) * torch.outer(m, m)
    assert ((g - ref).abs() <= bound).all(), ((g - ref).abs() / bound).max().item()
    if precision == "mixed":
        _, g64, q64, c64, *_ = grads("fp64")
        rel = lambda a, b: float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
        errs = {"com": rel(g_com, g64), "quat": rel(g_quat, q64), "coef": rel(g_coef, c64)}
        print(f"mixed vs fp64 gradients: {errs}", flush=True)
        assert errs["com"] <= 2e-4 and errs["quat"] <= 2e-4 and errs["coef"] <= 1e-7, errs


def test_mixed_gradient_against_fd_of_fp64_energy(params_fs2fs3, trimers):
    """The mixed-precision dE/d(COM) and dE/d(orientation) against central
    differences (h = 1e-5 A, 1e-5 rad) of the FP64 energy: the mixed gradient is
    the derivative of the reference energy up to the mixed path's force error
    (measured 8.2e-5 of the largest gradient component; asserted at 3e-4)."""
    from ccpol23_plus_2025_mbmd_amd.autograd import threebody_energy
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb_m = ThreeBody(params_fs2fs3, "mixed")
    tb_d = ThreeBody(params_fs2fs3, "fp64")
    dev = torch.device("cuda:0")
    sel = trimers[:12]
    com0 = np.concatenate([b.com for b in sel])
    q0 = np.concatenate([b.quats for b in sel])
    ms = np.cumsum([0] + [3] * len(sel)).astype(np.int32)
    bl = np.zeros(len(sel))
    pc = torch.tensor(principal_coords()[1], device=dev)
    c0 = torch.tensor(coef_vector(params_fs2fs3), device=dev)
    print(f"mixed gradient vs FD of the FP64 energy: {len(sel)} trimers", flush=True)
    com = torch.tensor(com0, device=dev, requires_grad=True)
    quat = torch.tensor(q0, device=dev, requires_grad=True)
    threebody_energy(tb_m, com, quat, c0, pc, ms, bl).sum().backward()
    g_com, g_quat = com.grad.cpu().numpy(), quat.grad.cpu().numpy()

    def L(comn, quatn):
        with torch.no_grad():
            return float(threebody_energy(tb_d, torch.tensor(comn, device=dev), torch.tensor(quatn, device=dev),
                                          c0, pc, ms, bl).sum())

    h = 1e-5
    scale = np.abs(g_com).max()
    worst = 0.0
    for m in (0, 4, 11, 20, 33):
        for a in range(3):
            cp, cm = com0.copy(), com0.copy()
            cp[m, a] += h
            cm[m, a] -= h
            fd = (L(cp, q0) - L(cm, q0)) / (2 * h)
            worst = max(worst, abs(fd - g_com[m, a]) / scale)
    for m in (1, 7, 30):
        for a in range(3):
            e_ax = np.eye(3)[a]
            qp, qm = q0.copy(), q0.copy()
            qp[m] = rotate_quat(q0[m], e_ax, h)
            qm[m] = rotate_quat(q0[m], e_ax, -h)
            fd = (L(com0, qp) - L(com0, qm)) / (2 * h)
            dq = (rotate_quat(q0[m], e_ax, 1e-7) - rotate_quat(q0[m], e_ax, -1e-7)) / 2e-7
            worst = max(worst, abs(fd - float(g_quat[m] @ dq)) / scale)
    print(f"largest difference / largest component: {worst:.2e}", flush=True)
    assert worst <= 3e-4, worst


def test_autograd_reuses_loaded_coefficients(params_fs2fs3, trimers):
    """The same coefficient tensor in consecutive forwards is loaded once (no
    host copy); a modified or different tensor is loaded again and changes the
    energy; a second derivative raises (the adjoint is first order only)."""
    from ccpol23_plus_2025_mbmd_amd.autograd import threebody_energy
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb = ThreeBody(params_fs2fs3, "fp64")
    dev = torch.device("cuda:0")
    sel = trimers[:6]
    com = torch.tensor(np.concatenate([b.com for b in sel]), device=dev, requires_grad=True)
    quat = torch.tensor(np.concatenate([b.quats for b in sel]), device=dev)
    ms = np.cumsum([0] + [3] * len(sel)).astype(np.int32)
    bl = np.zeros(len(sel))
    pc = torch.tensor(principal_coords()[1], device=dev)
    coef = torch.tensor(coef_vector(params_fs2fs3), device=dev)
    print("autograd coefficient reuse", flush=True)
    e1 = threebody_energy(tb, com, quat, coef, pc, ms, bl)
    calls = []
    orig = tb.set_coefficients
    tb.set_coefficients = lambda a, b: (calls.append(1), orig(a, b))
    e2 = threebody_energy(tb, com, quat, coef, pc, ms, bl)
    assert not calls and torch.equal(e1, e2)
    with torch.no_grad():
        coef[:NLIN] *= 2.0
    e3 = threebody_energy(tb, com, quat, coef, pc, ms, bl)
    assert len(calls) == 1 and not torch.equal(e3, e1)
    g, = torch.autograd.grad(e3.sum(), com, create_graph=True)
    with pytest.raises(RuntimeError):
        torch.autograd.grad(g.sum(), com)
