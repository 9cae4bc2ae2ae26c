"""Host-side checks of libccpol3b.so that need no GPU: the library loads,
exports every entry point include/ccpol3b.h declares, and its host helpers
(orbit map, data3b reader) agree with the oracle and the Python reader."""
import re
from pathlib import Path

import numpy as np
import pytest

from ccpol23_plus_2025_mbmd_amd import _lib
from ccpol23_plus_2025_mbmd_amd.params import DATA_DIR, read_data3b

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "ccpol3b.h").read_text()
    return sorted(set(re.findall(r"\b(ccpol3b_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    L = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTED) == syms


def test_orbit_index_matches_oracle(oracle):
    from ccpol23_plus_2025_mbmd_amd.threebody import orbit_index
    t, _ = oracle.ind_lin_2()
    rng = np.random.default_rng(1)
    for _ in range(3000):
        s = rng.integers(1, 18, 3)
        p = rng.integers(0, 2, 3)
        idx = p[2] + 2 * p[1] + 4 * p[0] + 8 * (s[2] - 1) + 136 * (s[1] - 1) + 2312 * (s[0] - 1)
        assert orbit_index(int(s[0]), int(s[1]), int(s[2]), int(p[0]), int(p[1]), int(p[2])) == t[idx]
    assert orbit_index(0, 1, 1, 0, 0, 0) == _lib.EINVAL


@pytest.mark.parametrize("variant", ["fs3", "fs2fs3"])
@pytest.mark.parametrize("name", ["data3b_full_SR_5p5", "data3b_full_LR_4p5"])
def test_c_reader_matches_python_reader(variant, name):
    from ccpol23_plus_2025_mbmd_amd.threebody import read_data3b_c
    b, g, r, cs = read_data3b_c(DATA_DIR / variant / name)
    ref = read_data3b(DATA_DIR / variant / name)
    np.testing.assert_array_equal(b, ref.beta)
    np.testing.assert_array_equal(g, ref.gamma)
    np.testing.assert_array_equal(r, ref.r0)
    np.testing.assert_array_equal(cs, ref.cs)


def test_create_rejects_bad_arguments(params_fs2fs3):
    import ctypes as C
    L = _lib.load()
    p = _lib.Params()
    p.nlin = 363  # wrong coefficient count is rejected before any device call
    p.ld_pair = 17
    ctx = C.c_void_p()
    assert L.ccpol3b_create(C.byref(p), C.byref(ctx)) == _lib.EINVAL
    assert not ctx.value


def test_no_device_fails_loudly(params_fs2fs3):
    """Without a usable gfx950 device the context cannot be created; there is no
    CPU fallback."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    with pytest.raises(_lib.CcpolError):
        ThreeBody(params_fs2fs3)
