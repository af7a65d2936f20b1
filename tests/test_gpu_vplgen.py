"""rzk_generate_vpls (bounce-1 VPL generation on the device) against the
oracle's orc_generate_vpls on the same light samples: the same number of light
samples used, the same materials and the same VPL records, bit for bit (the
light sampling, the cosine-weighted direction through sincos_turn, the closest
hit on the reference tree and the receiver fill are the same fp32 operations in
the same order on both sides).  Requests that span several batches of light
samples, and bad requests."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _scene(name):
    from renzoku_amd import scenes
    if name == "cornell":
        return scenes.cornell_box(64, 64)
    return scenes.room(64, 64, seed=5, target_tris=20_000)


@pytest.mark.parametrize("name,n,seed", [("cornell", 4096, 7), ("room", 20_000, 3)])
def test_generate_vpls_matches_oracle(ctx, name, n, seed):
    import renzoku_amd as rz
    sc = _scene(name)
    gs = rz.Scene(ctx, sc)
    vpl, vmat, used = rz.generate_vpls(ctx, gs, n, seed=seed)
    ov, om, ou = oracle.OrcScene(sc).generate_vpls(n, seed=seed)
    print(f"{name}: {n} VPLs from {used} light samples", flush=True)
    assert used == ou
    assert np.array_equal(vmat, om)
    assert np.array_equal(vpl.view(np.uint32), ov.view(np.uint32))


def test_generate_vpls_over_several_batches(ctx):
    import renzoku_amd as rz
    sc = _scene("cornell")
    gs = rz.Scene(ctx, sc)
    n = 230_000                                   # more than one batch of 2^18 light samples
    vpl, vmat, used, work = rz.generate_vpls(ctx, gs, n, seed=21, count_work=True)
    ov, om, ou = oracle.OrcScene(sc).generate_vpls(n, seed=21)
    assert used == ou and used > (1 << 18)
    assert np.array_equal(vmat, om)
    assert np.array_equal(vpl.view(np.uint32), ov.view(np.uint32))
    # every sample used was traced; a batch may trace a few past the last one used
    assert used <= work["closest_rays"] < used + (1 << 18) and work["nodes"] > used and work["tris"] > 0
    # a smaller request is the prefix of the larger one
    v2, m2, u2 = rz.generate_vpls(ctx, gs, 1000, seed=21)
    assert np.array_equal(v2.view(np.uint32), vpl[:1000].view(np.uint32)) and u2 <= used


def test_generate_vpls_rejects_bad_input(ctx):
    import renzoku_amd as rz
    gs = rz.Scene(ctx, _scene("cornell"))
    with pytest.raises(rz.RzkError):
        rz.generate_vpls(ctx, gs, 0)
    with pytest.raises(rz.RzkError):
        rz.generate_vpls(ctx, gs, 1000, max_samples=10)     # too few light samples for the request
