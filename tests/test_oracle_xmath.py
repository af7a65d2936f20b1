"""The oracle's fp32 transcendentals of the direction samplers (xsincos ..
xpow in oracle/oracle.c, restated in the kernels by rzk_math.h) against the
exact functions (float64 numpy), in ulp of the float32 result, on dense grids
of the arguments the samplers pass: angles within two turns, the whole range
of atan / atan2, [-1, 1] for acos, the 24-bit uniforms and beyond for log, the
exponents pow sees (1 / (n + 1) of Phong exponents).  A dropped series term, a
wrong reduction constant or a swapped quadrant moves some point by far more
than the bound."""
import ctypes as C

import numpy as np
import pytest

import oracle


def _run(fn, x, y=None):
    L = oracle.lib()
    L.orc_xmath.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_xmath.restype = None
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    yy = None if y is None else np.ascontiguousarray(y, np.float32)
    L.orc_xmath(fn, len(x), x.ctypes.data, None if yy is None else yy.ctypes.data, out.ctypes.data)
    return out


def _ulps(got, exact):
    exact = np.asarray(exact, np.float64)
    ulp = np.spacing(np.maximum(np.abs(exact).astype(np.float32), np.float32(1.2e-38))).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / ulp


RNG = np.random.default_rng(0)
ANGLES = np.concatenate([np.linspace(-2 * np.pi, 2 * np.pi, 400_001), RNG.uniform(-8, 8, 100_000),
                         np.arange(-8, 9) * (np.pi / 4)]).astype(np.float32)


@pytest.mark.parametrize("fn,f,bound", [(0, np.sin, 1.5), (1, np.cos, 1.5)])
def test_sin_cos(fn, f, bound):
    e = _ulps(_run(fn, ANGLES), f(ANGLES.astype(np.float64)))
    assert e.max() <= bound
    # 0 and the quarter turns land exactly on the polynomials' ends
    assert _run(0, np.zeros(1))[0] == 0.0 and _run(1, np.zeros(1))[0] == 1.0


def test_tan_away_from_poles():
    t = ANGLES[np.abs(np.cos(ANGLES.astype(np.float64))) > 1e-3]
    assert _ulps(_run(2, t), np.tan(t.astype(np.float64))).max() <= 3.5


def test_atan_whole_range():
    a = np.concatenate([np.linspace(-3, 3, 400_001),
                        np.sign(RNG.uniform(-1, 1, 100_000)) * 10 ** RNG.uniform(-8, 6, 100_000)]).astype(np.float32)
    assert _ulps(_run(3, a), np.arctan(a.astype(np.float64))).max() <= 2.5
    assert _run(3, np.array([np.inf, -np.inf], np.float32)).tolist() == pytest.approx([np.pi / 2, -np.pi / 2], rel=1e-7)


def test_atan2_quadrants_and_zeros():
    y = RNG.uniform(-1, 1, 300_000).astype(np.float32)
    x = RNG.uniform(-1, 1, 300_000).astype(np.float32)
    assert _ulps(_run(4, y, x), np.arctan2(y.astype(np.float64), x.astype(np.float64))).max() <= 2.5
    zy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0], np.float32)
    zx = np.array([0.0, 0.0, -0.0, -0.0, 0.0, -0.0, -1.0], np.float32)
    got = _run(4, zy, zx)
    want = np.arctan2(zy.astype(np.float64), zx.astype(np.float64))
    assert np.array_equal(np.signbit(got), np.signbit(want))
    assert _ulps(got, want).max() <= 1.0


def test_acos():
    c = np.linspace(-1, 1, 1_000_001).astype(np.float32)
    assert _ulps(_run(5, c), np.arccos(c.astype(np.float64))).max() <= 1.5
    assert np.isnan(_run(5, np.array([1.0001, -2.0], np.float32))).all()


def test_log_uniforms_and_beyond():
    u = np.concatenate([np.arange(1, 1 << 24, 7) / 2 ** 24, 10 ** RNG.uniform(-44, 38, 100_000)]).astype(np.float32)
    assert _ulps(_run(6, u), np.log(u.astype(np.float64))).max() <= 2.0
    assert _run(6, np.zeros(1))[0] == -np.inf and np.isnan(_run(6, -np.ones(1))[0])


def test_exp_and_pow():
    t = RNG.uniform(-100, 80, 300_000).astype(np.float32)
    assert _ulps(_run(7, t), np.exp(t.astype(np.float64))).max() <= 1.0
    x = (np.arange(1, 1 << 24, 61) / 2 ** 24).astype(np.float32)
    y = (1.0 / (RNG.integers(1, 2600, len(x)) + 1.0)).astype(np.float32)
    assert _ulps(_run(8, x, y), np.power(x.astype(np.float64), y.astype(np.float64))).max() <= 2.0
    assert _run(8, np.array([0.0, 1.0], np.float32), np.array([0.5, 0.3], np.float32)).tolist() == [0.0, 1.0]
