"""Pins for the oracle functions behind the LightSlice guide: the per-cluster
upper-hemisphere flag and direction (lightslice.cc:978-1018), the per-slice
cluster pdf (lightslice.cc:1114-1176), incident radiance and the reduced matrix
(vpl.cc:411-440, lightslice.cc:721-748) and cluster_by_sampling's draw order
(cluster_builder.cc:82-273).

Each check is built so that a plausible slip fails it: the geometric normal in
place of the shading normal (or the reverse), a dropped non-black or weight
term, a missing horizon zero, a transposed matrix index, centres drawn in
proportion to the column norm instead of alpha, representatives drawn before
the centres or out of cluster order.  Expected values come from float64 closed
forms and brute force, not from the oracle's own code."""
import math

import numpy as np
import pytest

import oracle

VPL_F = 21


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _onb9(n):
    """an orthonormal frame (u, v, n) with n as given"""
    n = _unit(n)
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    u = _unit(np.cross(n, a))
    v = np.cross(n, u)
    return np.concatenate([u, v, n]).astype(np.float32)


def _receiver(p, n_geom, n_shade, wo):
    r = np.zeros(21, np.float32)
    r[0:3] = p
    r[3:6] = _unit(n_geom)
    r[6:9] = _unit(n_shade)
    r[9:12] = _unit(wo)
    r[12:21] = _onb9(n_shade)
    return r


def _vpl(q, qn, power, wi_stored=(0, 0, 1)):
    v = np.zeros(VPL_F, np.float32)
    v[0:3] = q
    v[3:6] = _unit(qn)
    v[6:15] = _onb9(qn)
    v[15:18] = _unit(wi_stored)
    v[18:21] = power
    return v


def _unit_f32(d):
    """unit_vector as vec3_cpu.h:406-413 computes it: reciprocal of the length, then multiply"""
    d = np.asarray(d, np.float32)
    s = np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])
    s = np.float32(s + np.float32(d[2] * d[2]))
    inv = np.float32(1.0) / np.sqrt(s)
    return (d * inv).astype(np.float32)


def _lambert_scene(cornell):
    """the Cornell scene with one extra all-Lambert material for the VPL BRDF"""
    return oracle.OrcScene(cornell)


# ---------------------------------------------------------- slice cluster info -------
def test_slice_cluster_info_upper_flags(cornell):
    """lightslice.cc:1008-1015: upper = visible && dot(shading_n, wi) > 1e-6 &&
    incident radiance != black; slice_wi = unit_vector(q - p)."""
    osc = _lambert_scene(cornell)
    lam = int(np.nonzero(cornell.materials[:, 0] == 0)[0][0])
    p = np.array([10.0, 20.0, 30.0], np.float32)
    n_g = np.array([0.0, 0.0, 1.0])
    n_s = np.array([0.0, 0.6, 0.8])
    recv = _receiver(p, n_g, n_s, [0, 0, 1])
    white = (1.0, 1.0, 1.0)

    def toward(dirn, dist=50.0):
        return (p + dist * _unit(dirn)).astype(np.float32)

    cases = []   # (vpl, vmat, Vrow value, expected upper, what)
    q = toward([0, 0, 1])
    cases.append((_vpl(q, -_unit([0, 0, 1]), white), -1, 1.0, 1, "above both normals"))
    # below the shading normal, above the geometric one: the shading normal decides
    d = _unit([0.0, -0.9, 0.436])
    assert d @ _unit(n_s) < -0.1 and d @ n_g > 0.1
    cases.append((_vpl(toward(d), -d, white), -1, 1.0, 0, "below shading, above geometric"))
    # above the shading normal, below the geometric one
    d = _unit([0.0, 0.9, -0.2])
    assert d @ _unit(n_s) > 0.1 and d @ n_g < -0.1
    cases.append((_vpl(toward(d), -d, white), -1, 1.0, 1, "above shading, below geometric"))
    # black incident radiance: zero power
    cases.append((_vpl(q, [0, 0, -1], (0.0, 0.0, 0.0)), -1, 1.0, 0, "zero power"))
    # black: the VPL normal is perpendicular to pq, so G = 0
    cases.append((_vpl(q, [1, 0, 0], white), -1, 1.0, 0, "G = 0"))
    # black: a Lambertian VPL facing away from the receiver (its BRDF toward -wi is 0)
    cases.append((_vpl(q, [0, 0, 1], white, wi_stored=(0, 0, 1)), lam, 1.0, 0, "VPL BRDF zero"))
    # the same VPL turned toward the receiver is lit (BRDF kd / pi)
    cases.append((_vpl(q, [0, 0, -1], white, wi_stored=(0, 0, -1)), lam, 1.0, 1, "VPL BRDF non-zero"))
    # invisible
    cases.append((_vpl(q, [0, 0, -1], white), -1, 0.0, 0, "invisible"))
    # exactly in the shading tangent plane: dot = 0, not > 1e-6
    d = _unit(np.cross(n_s, [1.0, 0, 0]))
    assert abs(d @ _unit(n_s)) < 1e-12
    cases.append((_vpl(toward(d), -d + np.array([0, 0, -0.5]), white), -1, 1.0, 0, "in the tangent plane"))

    vpl = np.stack([c[0] for c in cases])
    vmat = np.array([c[1] for c in cases], np.int32)
    Vrow = np.array([c[2] for c in cases], np.float32)
    reps = np.arange(len(cases), dtype=np.int32)
    wi, up = osc.slice_cluster_info(recv, reps, vpl, vmat, Vrow, 1, 0.01)
    for c, case in enumerate(cases):
        assert up[c] == case[3], case[4]
        exp = _unit_f32(vpl[c, :3] - p)
        assert np.array_equal(wi[c].view(np.uint32), exp.view(np.uint32)), case[4]
    # the reps array indexes the VPL set: reversed reps give reversed rows
    wi2, up2 = osc.slice_cluster_info(recv, reps[::-1].copy(), vpl, vmat, Vrow, 1, 0.01)
    assert np.array_equal(up2, up[::-1]) and np.array_equal(wi2, wi[::-1])


# ---------------------------------------------------------- slice cluster pdf --------
def test_slice_cluster_pdf_bruteforce():
    """lightslice.cc:1114-1176: sqrt(sum over neighbour rows of |R(row, rep)|^2) x weight,
    0 where the cluster is not in the upper hemisphere."""
    rng = np.random.default_rng(11)
    S, V = 37, 53                                    # rows != cols: a transposed index reads other data
    R = rng.gamma(0.7, 1.0, (V, S, 3)).astype(np.float32)
    nb = rng.choice(S, 9, replace=False).astype(np.int32)
    reps = rng.choice(V, 20, replace=False).astype(np.int32)
    w = rng.uniform(0.5, 40.0, 20).astype(np.float32)
    upper = (rng.random(20) < 0.7).astype(np.int32)
    upper[:2] = [0, 1]
    got = oracle.slice_cluster_pdf(R, nb, reps, w, upper)
    R64 = R.astype(np.float64)
    for c in range(20):
        norm = math.sqrt(sum((R64[reps[c], r] ** 2).sum() for r in nb))
        exp = norm * w[c] if upper[c] else 0.0
        if upper[c]:
            assert got[c] == pytest.approx(exp, rel=2e-6), c
            assert abs(got[c] - norm) > 1e-3 * norm or abs(w[c] - 1) < 1e-3   # the weight matters
        else:
            assert got[c] == 0.0 and norm > 0, c
    # the pdf reads only the neighbour rows: changing another row changes nothing
    other = np.setdiff1d(np.arange(S), nb)[0]
    R2 = R.copy()
    R2[:, other] *= 7.0
    assert np.array_equal(oracle.slice_cluster_pdf(R2, nb, reps, w, upper), got)


# ---------------------------------------------------------- incident radiance --------
def _clamp64(mode, thr, power, fs, fl, G):
    """vpl_interface.h:50-123 closed forms (Rgb::value = L2 norm, rgb.h:85)"""
    power, fs, fl = (np.asarray(x, np.float64) for x in (power, fs, fl))
    if mode == 1:
        return fl * fs * min(G, thr) * power
    if mode == 2:
        refl = fs * G * fl
        lum = np.linalg.norm(refl)
        return (refl * (thr / lum) if lum > thr else refl) * power
    if mode == 3:
        rad = power * G * fs * fl
        b = np.linalg.norm(rad)
        return rad * (thr / b) if b > thr else rad
    return power * G * fs * fl


def _G64(p, n, q, qn):
    pq = np.asarray(q, np.float64) - np.asarray(p, np.float64)
    return abs((np.asarray(n, np.float64) @ pq) * (-(np.asarray(qn, np.float64) @ pq)) / (pq @ pq) ** 2)


def _hemisphere_cases(cornell, seed, n):
    """receivers whose geometric and shading normals differ by 20-60 degrees,
    and VPLs on both sides of either normal (material VPLs and light VPLs)"""
    rng = np.random.default_rng(seed)
    lam = np.nonzero(cornell.materials[:, 0] == 0)[0]
    out = []
    for _ in range(n):
        p = rng.uniform(100, 400, 3).astype(np.float32)
        n_g = _unit(rng.normal(size=3))
        tilt = math.radians(rng.uniform(20, 60))
        t = _unit(np.cross(n_g, rng.normal(size=3)))
        n_s = math.cos(tilt) * n_g + math.sin(tilt) * t
        recv = _receiver(p, n_g, n_s, n_s)
        d = _unit(rng.normal(size=3))
        q = (p + rng.uniform(30, 300) * d).astype(np.float32)
        qn = _unit(-d + 0.6 * rng.normal(size=3))
        vm = -1 if rng.random() < 0.3 else int(rng.choice(lam))
        vpl = _vpl(q, qn, rng.uniform(1e3, 1e5, 3).astype(np.float32), wi_stored=_unit(qn + 0.3 * rng.normal(size=3)))
        out.append((recv, vpl, vm))
    return out


@pytest.mark.parametrize("mode,thr", [(0, 0.0), (1, 1e-5), (2, 1.3e-5), (3, 0.25)])
def test_incident_radiance_uses_geometric_normal(cornell, mode, thr):
    """vpl.cc:411-440: G from the receiver's geometric normal r.n (not dg.n),
    surface BRDF white, the VPL's BRDF toward -wi, the incident clamp."""
    osc = _lambert_scene(cornell)
    cases = _hemisphere_cases(cornell, 3 + mode, 60)
    clamped = free = 0
    separated = 0
    for recv, vpl, vm in cases:
        got = osc.incident(recv, vpl, vm, mode, thr)
        p, q = recv[:3], vpl[:3]
        wi = _unit(q.astype(np.float64) - p)
        fl = np.ones(3) if vm < 0 else oracle.bsdf_eval(cornell.materials[vm], vpl[6:15], vpl[15:18],
                                                        (-wi).astype(np.float32)).astype(np.float64)
        G = _G64(p, recv[3:6], q, vpl[3:6])
        exp = _clamp64(mode, thr, vpl[18:21], np.ones(3), fl, G)
        assert np.allclose(got, exp, rtol=2e-5, atol=1e-30), (mode, got, exp)
        # the shading normal would give another value (H14): the case can tell them apart
        Gs = _G64(p, recv[6:9], q, vpl[3:6])
        alt = _clamp64(mode, thr, vpl[18:21], np.ones(3), fl, Gs)
        if np.abs(exp).max() > 0 and not np.allclose(alt, exp, rtol=1e-3):
            separated += 1
        unclamped = _clamp64(0, 0.0, vpl[18:21], np.ones(3), fl, G)
        if np.abs(exp).max() > 0:
            if np.allclose(unclamped, exp, rtol=1e-6):
                free += 1
            else:
                clamped += 1
    # a clamped REFLECTIVITY / RADIANCE term no longer depends on |G|, so only
    # the cases below the threshold (about half) separate the two normals
    assert separated >= 15, separated
    if mode != 0:
        assert clamped >= 5 and free >= 5, (clamped, free)   # both branches of the clamp taken


def test_reduced_matrix_composition(cornell):
    """lightslice.cc:721-748: R(r, k) = vis x incident radiance (geometric normal),
    V(r, k) = vis; rows without a material stay zero; column-major Rgb."""
    from renzoku_amd import scenes, vplgen
    osc = oracle.OrcScene(cornell)
    px, py = scenes.pixel_grid(256, 256)
    sel = np.arange(0, 65536, 2111)
    rays = oracle.camera_rays(cornell.camera, px[sel], py[sel])
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    # tilt the shading normals of half the rows so geometric and shading normals differ
    for i in range(0, len(recv), 2):
        if m[i] >= 0:
            n_s = _unit(recv[i, 3:6] + np.array([0.3, -0.2, 0.25]))
            recv[i] = _receiver(recv[i, :3], recv[i, 3:6], n_s, recv[i, 9:12])
    assert (m < 0).any() and (m >= 0).sum() > 10
    vpl, vmat, _ = osc.generate_vpls(40, seed=5)
    for mode, thr in ((1, 0.01), (3, 5.0)):
        R, V = osc.reduced_matrix(recv, m, vpl, vmat, mode, thr, 1e-3)
        assert R.shape == (40, len(sel), 3) and V.shape == (40, len(sel))
        nonzero = 0
        for r in range(len(sel)):
            if m[r] < 0:
                assert not R[:, r].any() and not V[:, r].any()
                continue
            blocked = osc.trace_segments(np.repeat(recv[r:r + 1, :3], 40, 0), vpl[:, :3])
            for k in range(40):
                vis = 0.0 if blocked[k] else 1.0
                assert V[k, r] == vis
                wi = _unit(vpl[k, :3].astype(np.float64) - recv[r, :3])
                fl = np.ones(3) if vmat[k] < 0 else oracle.bsdf_eval(cornell.materials[vmat[k]], vpl[k, 6:15],
                                                                     vpl[k, 15:18], (-wi).astype(np.float32))
                G = _G64(recv[r, :3], recv[r, 3:6], vpl[k, :3], vpl[k, 3:6])
                exp = vis * _clamp64(mode, thr, vpl[k, 18:21], np.ones(3), fl, G)
                assert np.allclose(R[k, r], exp, rtol=2e-5, atol=1e-30), (mode, r, k)
                nonzero += bool(exp.any())
        assert nonzero > 100


# ---------------------------------------------------------- cluster_by_sampling -------
def _floats(seed, n):
    return (oracle.mt_stream(seed, n) & 0xFFFFFF).astype(np.float64) / 2 ** 24


def _alpha64(R):
    M = np.sqrt((R.astype(np.float64) ** 2).sum(2))          # (cols, rows) entry magnitudes
    norms = np.sqrt((M ** 2).sum(1))
    return norms * norms.sum() - M @ M.sum(0)


def _draw(cdf, u):
    """DiscretePdf::sample (pdf.cc:59-77): the first entry whose cdf exceeds u"""
    return int(np.searchsorted(cdf, u, side="right"))


def _collinear_R(seed):
    """columns 0-3 share one direction and have large norms (small alpha);
    the rest point anywhere with small norms: alpha and norm orders disagree"""
    rng = np.random.default_rng(seed)
    rows, cols = 6, 14
    R = rng.uniform(0.0, 1.0, (cols, rows, 3)).astype(np.float32)
    base = rng.uniform(0.5, 1.0, (rows, 3)).astype(np.float32)
    for k, s in enumerate((9.0, 8.0, 7.0, 6.0)):
        R[k] = base * s
    return R


def test_first_centre_is_drawn_in_proportion_to_alpha():
    R = _collinear_R(1)
    a = _alpha64(R)
    mags = np.sqrt((R.astype(np.float64) ** 2).reshape(len(R), -1).sum(1))
    pa, pm = a / a.sum(), mags / mags.sum()
    assert np.abs(pa - pm).sum() > 0.4           # the two laws are far apart
    n = 3000
    counts = np.zeros(len(R))
    for s in range(n):
        counts[oracle.cluster_by_sampling(R, 1, 1000 + s)["centers"][0]] += 1
    chi_a = ((counts - n * pa) ** 2 / (n * pa)).sum()
    chi_m = ((counts - n * pm) ** 2 / (n * pm)).sum()
    assert chi_a < 40.0, chi_a                   # 13 dof: p ~ 1e-4 at 40
    assert chi_m > 500.0, chi_m


def test_representative_is_drawn_in_proportion_to_column_norm():
    # one cluster: every column is its member, the representative follows |m_i|
    R = _collinear_R(2)
    mags = np.sqrt((R.astype(np.float64) ** 2).reshape(len(R), -1).sum(1))
    pm = mags / mags.sum()
    a = _alpha64(R)
    pa = a / a.sum()
    n = 3000
    counts = np.zeros(len(R))
    for s in range(n):
        out = oracle.cluster_by_sampling(R, 1, 5000 + s)
        assert (out["assign"] == 0).all()
        rep = out["reps"][0]
        counts[rep] += 1
        assert out["rep_w"][0] == pytest.approx(1.0 / pm[rep], rel=1e-5)
    chi_m = ((counts - n * pm) ** 2 / (n * pm)).sum()
    chi_a = ((counts - n * pa) ** 2 / (n * pa)).sum()
    chi_u = ((counts - n / len(R)) ** 2 / (n / len(R))).sum()
    assert chi_m < 40.0, chi_m
    assert chi_a > 500.0 and chi_u > 500.0, (chi_a, chi_u)


@pytest.mark.parametrize("seed", [2013, 7, 99, 12345])
def test_cluster_by_sampling_draw_order(seed):
    """cluster_builder.cc:121-156 then :242-262 on one MT stream: centre draws
    until C distinct centres (repeats add 1/p to the weight), then exactly one
    draw per cluster in cluster order for its representative; the stream is left
    right after the last representative draw."""
    rng = np.random.default_rng(seed)
    rows, cols, C = 5, 60, 12
    R = rng.gamma(0.6, 1.0, (cols, rows, 3)).astype(np.float32)
    R[:5] *= 20.0                                        # a few heavy columns, so centres repeat
    a = _alpha64(R)
    cdf_a = np.cumsum(a / a.sum())
    cdf_a[-1] = 1.0
    u = _floats(seed, 4000)
    # centres from the spec, in float64
    centres, weights, t = [], [], 0
    while len(centres) < C:
        k = _draw(cdf_a, u[t])
        margin = np.abs(cdf_a - u[t]).min()
        assert margin > 1e-5, "a draw sits on a cdf boundary; pick another seed"
        p = a[k] / a.sum()
        if k in centres:
            weights[centres.index(k)] += 1.0 / p
        else:
            centres.append(k)
            weights.append(1.0 / p)
        t += 1
    repeats = t - C
    out = oracle.cluster_by_sampling(R, C, seed)
    assert out["centers"].tolist() == centres
    assert np.allclose(out["center_w"], weights, rtol=1e-4)
    # assignment: brute-force argmin of | |K_j| |m_i| - m_i . K_j |, K_j = weight_j x column
    M = R.reshape(cols, -1).astype(np.float64)
    K = M[centres] * np.array(weights)[:, None]
    dist = np.abs(np.linalg.norm(K, axis=1)[None, :] * np.linalg.norm(M, axis=1)[:, None] - M @ K.T)
    srt = np.sort(dist, 1)
    clear = (srt[:, 1] - srt[:, 0]) > 1e-4 * srt[:, 1]
    assert np.array_equal(out["assign"][clear], dist.argmin(1)[clear])
    assign = out["assign"]
    mags = np.linalg.norm(M, axis=1)
    # representatives: cluster c takes draw t + c over its members in increasing index order
    reps = []
    for c in range(C):
        mem = np.nonzero(assign == c)[0]
        if len(mem) == 0:
            continue
        cdf_m = np.cumsum(mags[mem] / mags[mem].sum())
        cdf_m[-1] = 1.0
        assert np.abs(cdf_m - u[t + c]).min() > 1e-5
        reps.append(int(mem[_draw(cdf_m, u[t + c])]))
    assert out["reps"].tolist() == reps
    # the stream continues right after the C representative draws
    nxt = oracle.lib().orc_mt_rand()
    assert nxt == pytest.approx(u[t + C], abs=0)
    assert repeats >= 0


# ---------------------------------------------------------- gather_clusters ----------
def test_gather_clusters_weights_after_the_clamp(cornell):
    """lightslice.cc:1325-1339: radiance += weight * Lo, Lo already clamped inside
    gather_point.  Unit weights give the plain gather bit for bit; other weights give
    the weighted float64 sum of the clamped terms, which differs from folding the
    weight into the VPL power once the RADIANCE clamp (not linear in power) engages."""
    from renzoku_amd import scenes, vplgen
    osc = oracle.OrcScene(cornell)
    px, py = scenes.pixel_grid(256, 256)
    sel = np.arange(0, 65536, 1499)
    rays = oracle.camera_rays(cornell.camera, px[sel], py[sel])
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    vpl, vmat, ne = osc.generate_vpls(48, seed=3)
    w = np.random.default_rng(2).uniform(0.5, 30.0, 48).astype(np.float32)
    for mode, thr in ((1, 3e-6), (3, 0.04)):
        ones = osc.gather_clusters(recv, m, vpl, vmat, np.ones(48, np.float32), mode, thr, 1e-3, ne)
        assert np.array_equal(ones.view(np.uint32), osc.gather(recv, m, vpl, vmat, mode, thr, 1e-3, ne).view(np.uint32))
        got = osc.gather_clusters(recv, m, vpl, vmat, w, mode, thr, 1e-3, ne)
        exp = np.zeros((len(sel), 3))
        for i in range(len(sel)):
            for k in range(48):
                lo, vis = osc.gather_point(recv[i], m[i], vpl[k], vmat[k], mode, thr, 1e-3)
                if vis:
                    exp[i] += float(w[k]) * lo.astype(np.float64)
        exp /= ne
        assert np.allclose(got, exp, rtol=1e-5, atol=1e-30)
        folded = vpl.copy()
        folded[:, 18:21] *= w[:, None]
        alt = osc.gather(recv, m, folded, vmat, mode, thr, 1e-3, ne)
        if mode == 3:
            assert not np.allclose(alt, exp, rtol=1e-3)       # the clamp is not linear in power
        else:
            assert np.allclose(alt, exp, rtol=1e-5, atol=1e-30)
