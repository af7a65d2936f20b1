"""The C++ drop-in adapter (renzoku_amd/adapters/aggregate_gpu.h): a compiled
program calls the three Aggregate::hit overloads through a base-class pointer
(closest, any, segment; aggregate.h:20-26) and the batched forms; answers are
compared with the oracle (hit ids and t bit-exact, booleans exact), and the
closest hit's whole HitRecord (normals, uvn frame, tangent, surface and
material bindings) with the oracle's receiver of the same hit."""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "adapter_test")


def test_aggregate_adapter_matches_oracle(tmp_path, small_room):
    sc = small_room
    assert os.path.exists(BIN), "tests/cpp/adapter_test not built (make -C tests/cpp)"
    rng = np.random.default_rng(21)
    n, m = 3000, 2000
    lo, hi = sc.positions.reshape(-1, 3).min(0), sc.positions.reshape(-1, 3).max(0)
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    p = (lo + (hi - lo) * rng.random((m, 3))).astype(np.float32)
    q = (lo + (hi - lo) * rng.random((m, 3))).astype(np.float32)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([sc.ntri, len(sc.surf_material), len(sc.materials), len(sc.light_radiance), n, m],
                 np.int32).tofile(f)
        for a, dt in [(sc.positions, np.float32), (sc.tangents, np.float32), (sc.tri_surface, np.int32),
                      (sc.surf_material, np.int32), (sc.surf_light, np.int32), (sc.surf_flags, np.int32),
                      (sc.materials, np.float32), (sc.light_radiance, np.float32), (rays, np.float32),
                      (p, np.float32), (q, np.float32)]:
            np.ascontiguousarray(a, dt).tofile(f)
    r = subprocess.run([BIN, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    t = np.frombuffer(raw, np.float32, n, 0)
    tri = np.frombuffer(raw, np.int32, n, 12 * n)
    anyh = np.frombuffer(raw, np.uint8, n, 16 * n)
    blocked = np.frombuffer(raw, np.uint8, m, 17 * n)
    osc = oracle.OrcScene(sc)
    orays = np.stack([oracle.ray_make(o[i], d[i]) for i in range(n)])
    ot, ob, og, otri = osc.trace_closest(orays)
    assert np.array_equal(tri, otri)
    hit = otri >= 0
    assert np.array_equal(t[hit].view(np.uint32), ot[hit].view(np.uint32))
    assert np.array_equal(anyh, osc.trace_any(orays))
    assert np.array_equal(blocked, osc.trace_segments(p, q))
    # the whole record: normals, uvn frame and tangent (triangle.cc:277-296, surface.cc:62-96) are the
    # oracle's receiver fields bit for bit; surface and material pointers are the hit surface's bindings
    off = 17 * n + m
    frame = np.frombuffer(raw, np.float32, 18 * n, off).reshape(n, 18)
    sidx = np.frombuffer(raw, np.int32, n, off + 72 * n)
    midx = np.frombuffer(raw, np.int32, n, off + 76 * n)
    recv, om, _ = osc.fill_receivers(orays, ot, ob, og, otri)
    assert np.array_equal(frame[hit, 0:6].view(np.uint32), recv[hit, 3:9].view(np.uint32))      # n, n_shading
    assert np.array_equal(frame[hit, 6:15].view(np.uint32), recv[hit, 12:21].view(np.uint32))   # uvn
    assert np.array_equal(frame[hit, 15:18], np.asarray(sc.tangents, np.float32).reshape(-1, 3)[otri[hit]])
    surf = np.asarray(sc.tri_surface)[otri[hit]]
    assert np.array_equal(sidx[hit], surf)
    assert np.array_equal(midx[hit], om[hit])
    assert (sidx[~hit] == -1).all() and (midx[~hit] == -1).all()
