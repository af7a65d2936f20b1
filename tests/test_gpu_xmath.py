"""The kernels' fp32 transcendentals (rzk_math.h xsincos .. xpow, evaluated on
the device through libprobe.so) give the oracle's bits for every argument of
dense grids over the samplers' domains and for the special cases (signed zeros,
subnormals, the turn's quarter points, arguments outside the domains), so the
host restatement samples the same directions as the kernels."""
import ctypes as C
import os

import numpy as np
import pytest

from test_oracle_xmath import _run

pytestmark = pytest.mark.gpu
PROBE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "renzoku_amd", "libprobe.so")


def _dev(fn, x, y=None):
    lib = C.CDLL(PROBE)
    lib.rzk_probe_xmath.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.rzk_probe_xmath.restype = C.c_int
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    yy = None if y is None else np.ascontiguousarray(y, np.float32)
    rc = lib.rzk_probe_xmath(fn, len(x), x.ctypes.data, None if yy is None else yy.ctypes.data, out.ctypes.data)
    assert rc == 0, f"rzk_probe_xmath: hip error {rc}"
    return out


def _same(a, b):
    # bit-equal, NaNs of either sign counting as equal
    nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(a.view(np.uint32)[~nan], b.view(np.uint32)[~nan]) and np.array_equal(np.isnan(a), np.isnan(b))


RNG = np.random.default_rng(1)
SPECIAL = np.array([0.0, -0.0, 1e-45, -1e-45, 1.17e-38, 1.0, -1.0, 0.5, -0.5, np.pi / 4, np.pi / 2, np.pi,
                    2 * np.pi, -np.pi, 1e4, 2e4, np.inf, -np.inf, np.nan], np.float32)


@pytest.mark.parametrize("fn", [0, 1, 2, 3, 5, 6, 7])
def test_unary_device_equals_oracle(fn):
    dom = {0: (-7, 7), 1: (-7, 7), 2: (-7, 7), 3: (-50, 50), 5: (-1, 1), 6: (0, 4), 7: (-105, 90)}[fn]
    x = np.concatenate([np.linspace(*dom, 1_000_003), RNG.uniform(*dom, 200_000), SPECIAL]).astype(np.float32)
    if fn == 6:
        x = np.concatenate([x, (np.arange(1, 1 << 24, 13) / 2 ** 24).astype(np.float32)])
    print(f"fn {fn}: {len(x)} arguments", flush=True)
    assert _same(_dev(fn, x), _run(fn, x))


@pytest.mark.parametrize("fn", [4, 8])
def test_binary_device_equals_oracle(fn):
    n = 500_000
    if fn == 4:
        y = np.concatenate([RNG.uniform(-1, 1, n), SPECIAL, np.zeros(len(SPECIAL))]).astype(np.float32)
        x = np.concatenate([RNG.uniform(-1, 1, n), np.zeros(len(SPECIAL)), SPECIAL]).astype(np.float32)
    else:
        x = np.concatenate([(np.arange(1, 1 << 24, 31) / 2 ** 24), [0.0, 1.0, 0.5]]).astype(np.float32)
        y = (1.0 / (RNG.integers(1, 2600, len(x)) + 1.0)).astype(np.float32)
    print(f"fn {fn}: {len(x)} argument pairs", flush=True)
    assert _same(_dev(fn, x, y), _run(fn, x, y))
