"""On-disk formats: data3b_full writing (refit output), REVCON restart files,
and the binary parameter cache.  CPU only.

Pinned by: byte identity with the four shipped data3b_full files (the reference
wrote them with list-directed output); exact read-back of written values;
the reference's restart semantics for REVCON (initialization.f90:55-67:
title skipped, COM mapped with x - box*nint(x/box), NINT halves away from 0)."""
import shutil

import numpy as np
import pytest

from ccpol23_plus_2025_mbmd_amd.formats import (Revcon, _real_field, load_params_cache, load_variant_cached,
                                                read_revcon, save_params_cache, variant_sources, write_data3b,
                                                write_revcon)
from ccpol23_plus_2025_mbmd_amd.params import DATA_DIR, Data3B, load_variant, read_data3b

SHIPPED = [(v, f) for v in ("fs2fs3", "fs3") for f in ("data3b_full_SR_5p5", "data3b_full_LR_4p5")]


@pytest.mark.parametrize("variant,name", SHIPPED)
def test_data3b_writer_reproduces_shipped_files(tmp_path, variant, name):
    src = DATA_DIR / variant / name
    text = src.read_text()
    lines = text.splitlines()
    d = read_data3b(src)
    out = tmp_path / name
    write_data3b(out, d, header=lines[0], label=lines[290])
    # identical up to the final newline, which the shipped copies lack
    assert out.read_text().rstrip("\n") == text.rstrip("\n")


def test_data3b_round_trip_is_exact(tmp_path):
    d = read_data3b(DATA_DIR / "fs2fs3" / "data3b_full_LR_4p5")
    rng = np.random.default_rng(5)
    cs = d.cs * np.exp(rng.uniform(-40, 40, d.cs.shape)) * rng.choice([-1, 1], d.cs.shape)
    cs[:8] = [0.0, 0.1, -0.1, 0.09999999999999999, 1e17, 9.999999999999998e16, 1e-300, -4.0e5]
    beta = d.beta * (1 + rng.uniform(-0.5, 0.5, d.beta.shape))
    fit = Data3B(beta=beta, gamma=d.gamma, r0=d.r0, cs=cs)
    out = tmp_path / "data3b_full"
    write_data3b(out, fit)
    back = read_data3b(out)
    for k in ("beta", "gamma", "r0", "cs"):
        assert np.array_equal(getattr(back, k), getattr(fit, k)), k


def test_real_field_layout():
    # fixed notation in 20 columns + 5 blanks, or d.dddE+eee in 25, after a separating blank
    assert _real_field(0.34551522041334698) == "  0.34551522041334698     "
    assert _real_field(-16.313258739257499) == "  -16.313258739257499     "
    assert _real_field(-4.1602711135645502e-2) == "  -4.1602711135645502E-002"
    assert _real_field(1.1836449333911699e-2) == "   1.1836449333911699E-002"
    assert all(len(_real_field(x)) == 26 for x in (0.0, 1e-300, 5e17, -0.25, 123456.789))


def revcon_state(nm, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(nm, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return Revcon(step=1234, time=0.617, box=12.417180486121673, chit=0.0, chip=0.0,
                  com=rng.uniform(-6, 6, (nm, 3)), vel=rng.normal(size=(nm, 3)), jbody=rng.normal(size=(nm, 3)),
                  quat=q)


def test_revcon_round_trip(tmp_path):
    r = revcon_state(64, 1)
    path = tmp_path / "REVCON"
    write_revcon(path, r)
    back = read_revcon(path)
    assert back.nm == 64 and back.step == 1234 and back.time == r.time and back.box == r.box
    for k in ("com", "vel", "jbody", "quat"):
        assert np.array_equal(getattr(back, k), getattr(r, k)), k
    # the reference reads exactly nm records and ignores anything after them
    assert read_revcon(path, nm=10).nm == 10


def test_revcon_restart_semantics(tmp_path):
    """Title skipped; list-directed items (commas, D exponents, split records);
    COM mapped into the box with NINT rounding halves away from zero."""
    box = 10.0
    text = ("  any title at all\n"
            "  77, 1.5D-01 10.0 0.0 0.0\n"
            "  5.0 -5.0 17.5 1 2 3 4 5 6\n   1.0 0.0 0.0 0.0\n"
            " -5.0 5.000000000000001 -14.9,0.5D0,0,0,0,0,0,0.5,0.5,0.5,0.5\n")
    path = tmp_path / "REVCON"
    path.write_text(text)
    r = read_revcon(path)
    assert r.step == 77 and r.time == 0.15 and r.box == box and r.nm == 2
    np.testing.assert_array_equal(r.com[0], [5.0 - 10.0, -5.0 + 10.0, 17.5 - 20.0])
    np.testing.assert_array_equal(r.com[1], [-5.0 + 10.0, 5.000000000000001 - 10.0, -14.9 + 10.0])
    np.testing.assert_array_equal(r.vel[1], [0.5, 0.0, 0.0])
    np.testing.assert_array_equal(r.quat[1], [0.5] * 4)
    raw = read_revcon(path, wrap=False)
    assert raw.com[0, 2] == 17.5
    with pytest.raises(ValueError):
        read_revcon(path, nm=3)


def test_params_cache(tmp_path):
    d = tmp_path / "fs2fs3"
    shutil.copytree(DATA_DIR / "fs2fs3", d)
    src = variant_sources(d)
    assert load_params_cache(tmp_path / "none.npz", src) is None
    p = load_variant("fs2fs3", d)
    cache = tmp_path / "p.npz"
    save_params_cache(cache, p, src)
    q = load_params_cache(cache, src)
    assert q is not None
    for reg in ("sr", "lr"):
        for k in ("beta", "gamma", "r0", "cs"):
            assert np.array_equal(getattr(getattr(q, reg), k), getattr(getattr(p, reg), k))
    assert (q.alpha3, q.delta33, q.rcut3, q.s2_term, q.s3_term) == (p.alpha3, p.delta33, p.rcut3, p.s2_term,
                                                                     p.s3_term)
    # a changed source file makes the cache stale
    f = d / "data3b_full_LR_4p5"
    f.write_text(f.read_text().replace("linear parameters", "linear  parameters"))
    assert load_params_cache(cache, src) is None
    # the cached loader rebuilds it
    r = load_variant_cached("fs2fs3", tmp_path / "cache", d)
    assert np.array_equal(r.lr.cs, p.lr.cs)
    assert (tmp_path / "cache" / "ccpol3b_params_fs2fs3.npz").exists()
    assert np.array_equal(load_variant_cached("fs2fs3", tmp_path / "cache", d).sr.cs, p.sr.cs)
