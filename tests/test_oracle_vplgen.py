"""The oracle's bounce-1 VPL generation (oracle.c orc_generate_vpls,
vpl_generator.cc:127-181) pinned by what the sampling fixes:

* sincos_turn, the engine's cos/sin of 2 pi u, against the C library in float64
  (absolute error below 2e-7 over a fine grid and random points);
* every VPL's power is Le * A_light * pi / pdf_light: the cosine-weighted
  direction's pdf cancels |n_q . wo| and the uniform point's pdf is 1 / A;
* emission cosines about the light's shape normal average 2/3 with variance
  1/18 (cosine-weighted hemisphere), in a closed room where every sample stores
  a VPL (in an open scene the rays that escape are not a random subset);
* samples are taken in order: a smaller request is a prefix of a larger one,
  and the same seed gives the same set.
"""
import math

import numpy as np

import oracle
from renzoku_amd import scenes


def test_sincos_turn_matches_libm():
    rng = np.random.default_rng(1)
    us = np.concatenate([np.arange(0, 1, 1 / 8192), rng.random(20000)]).astype(np.float32)
    worst = 0.0
    for u in us:
        s, c = oracle.sincos_turn(float(u))
        th = 2.0 * math.pi * float(u)
        worst = max(worst, abs(s - math.sin(th)), abs(c - math.cos(th)))
    assert worst < 2e-7
    assert oracle.sincos_turn(0.0) == (0.0, 1.0)
    s, c = oracle.sincos_turn(0.25)
    assert s == 1.0 and abs(c) == 0.0


def _cornell():
    sc = scenes.cornell_box(64, 64)
    return sc, oracle.OrcScene(sc)


def _light_area_normal(sc):
    lt = sc.light_tris
    P = sc.positions[lt].reshape(-1, 3, 3).astype(np.float64)
    cr = np.cross(P[:, 0] - P[:, 1], P[:, 0] - P[:, 2])
    return 0.5 * np.linalg.norm(cr, axis=1).sum(), cr[0] / np.linalg.norm(cr[0])


def test_vpl_power_is_the_closed_form():
    sc, osc = _cornell()
    vpl, vmat, used = osc.generate_vpls(6000, seed=7)
    assert used >= 6000 and (vmat >= 0).all()
    area, _ = _light_area_normal(sc)
    le = sc.light_radiance[0].astype(np.float64)        # one light: pdf_light = 1
    assert np.allclose(vpl[:, 18:21], (le * area * math.pi)[None, :], rtol=2e-5)
    assert np.allclose(np.linalg.norm(vpl[:, 15:18], axis=1), 1.0, atol=1e-6)


def test_emission_is_cosine_weighted():
    """in a closed room every light ray stores a VPL, so the VPLs' emission
    cosines are the sampled ones: mean 2/3, variance 1/18"""
    sc = scenes.room(64, 64, seed=3, target_tris=3000)
    vpl, vmat, used = oracle.OrcScene(sc).generate_vpls(8000, seed=5)
    assert used == 8000                                 # nothing missed, nothing reached a light
    _, n = _light_area_normal(sc)
    cos = -vpl[:, 15:18].astype(np.float64) @ n
    assert cos.min() > 0.0
    assert abs(cos.mean() - 2.0 / 3.0) < 5 * math.sqrt(1.0 / 18.0 / len(cos))
    assert abs(cos.var() - 1.0 / 18.0) < 0.005


def test_vpls_are_taken_in_sample_order():
    sc, osc = _cornell()
    a, am, au = osc.generate_vpls(3000, seed=11)
    b, bm, bu = osc.generate_vpls(1000, seed=11)
    assert np.array_equal(a[:1000].view(np.uint32), b.view(np.uint32)) and np.array_equal(am[:1000], bm)
    assert bu <= au
    c, _, cu = osc.generate_vpls(3000, seed=11)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32)) and cu == au
    d, _, _ = osc.generate_vpls(3000, seed=12)
    assert not np.array_equal(a, d)
