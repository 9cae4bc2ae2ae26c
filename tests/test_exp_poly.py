"""The FP64 exponential of the FS3 pair stage (exp_cw in csrc/devcommon.h):
its constants, read from the header, and its algorithm restated here in
numpy (fused multiply-adds emulated in 80-bit arithmetic) against numpy's
extended-precision exp.  Pins the Cody-Waite split (the head of ln 2 has 32
significant bits, head + tail = ln 2 to 1e-19) and the interpolation
polynomial (within 2 ulp over the pair-stage argument range and beyond)."""
import re
from pathlib import Path

import numpy as np

HDR = Path(__file__).resolve().parents[1] / "ccpol23_plus_2025_mbmd_amd" / "csrc" / "devcommon.h"
LD = np.longdouble


def constants():
    src = HDR.read_text()
    body = src[src.index("double exp_cw(double x)"):]
    body = body[:body.index("\n}\n")]
    hexnum = r"(-?0x[0-9a-fA-F.]+p[-+]?\d+)"
    hi = float.fromhex(re.search(r"LN2_HI = " + hexnum, body).group(1))
    lo = float.fromhex(re.search(r"LN2_LO = " + hexnum, body).group(1))
    block = body[body.index("C[12] = {"):body.index("};")]
    c = [float.fromhex(v) for v in re.findall(hexnum, block)]
    return hi, lo, np.array(c)


def fma(a, b, c):
    return (LD(a) * LD(b) + LD(c)).astype(np.float64)


def exp_cw(x):
    hi, lo, c = constants()
    n = np.rint(x * 1.4426950408889634)
    r = fma(-n, lo, fma(-n, hi, x))
    p = np.full_like(x, c[11])
    for k in range(10, -1, -1):
        p = fma(p, r, c[k])
    return np.ldexp(p, n.astype(np.int64))


def test_constants():
    hi, lo, c = constants()
    ln2 = LD("0.693147180559945309417232121458176568")
    assert len(c) == 12 and c[0] == 1.0 and c[1] == 1.0
    m, _ = np.frexp(hi)
    assert float(m * 2.0 ** 32) == int(m * 2.0 ** 32)  # 32 significant bits: n * hi exact for |n| < 2^21
    assert abs(float(LD(hi) + LD(lo) - ln2)) < 1e-19
    # close to the Taylor coefficients 1/k! (an interpolant of e^r, not a different function)
    fact = np.cumprod([1.0] + list(range(1, 12)))
    assert np.allclose(c * fact, 1.0, rtol=5e-3)


def test_accuracy():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-120, 60, 400_000), rng.uniform(-700, 700, 100_000),
                        np.linspace(-0.36, 0.36, 10_001)])
    y = exp_cw(x)
    ref = np.exp(x.astype(LD))
    ulp = np.abs((y.astype(LD) - ref) / ref) / 2.0 ** -53
    assert float(ulp.max()) <= 2.0, float(ulp.max())
    assert exp_cw(np.array([0.0]))[0] == 1.0
