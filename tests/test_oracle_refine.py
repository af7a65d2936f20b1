"""The oracle's LightSlice per-slice refinement (oracle/refine.cpp,
lightslice.cc:837-902) pinned by what the algorithm fixes:

* it is a refinement: every slice's clusters partition the kept global
  clusters' members, only ever splitting a cluster, and each representative
  lies in its own cluster;
* its draw count is exactly 2 per split of a cluster with more than two
  members plus 1 per new representative, and the MT stream continues there;
* the slice cost is the pairwise closed form sum_ij (|c_i||c_j| - c_i.c_j)
  (float64 check of the float32 restatement);
* with max_slice_clusters equal to the number of global clusters nothing
  changes and no draw is taken;
* on small tie-free inputs an independent step-by-step Python restatement
  (float32 arithmetic in the reference's order, Python's sort and a scan for
  the largest cost, which agree with std::sort / priority_queue when keys are
  distinct) gives the same clusters, representatives and weights bit for bit.
"""
import numpy as np
import pytest

import oracle


def _inputs(V, S, nnb, seed, zero_frac=0.4):
    rng = np.random.default_rng(seed)
    R = rng.uniform(0, 1, (V, S, 3)).astype(np.float32) * (rng.uniform(size=(V, S, 1)) > zero_frac)
    nb = np.array([[(s + d) % S for d in range(nnb)] for s in range(S)], np.int32)
    return R.astype(np.float32), nb


def test_refinement_partitions_the_global_clusters():
    R, nb = _inputs(400, 16, 4, 1, zero_frac=0.7)        # many columns zero over a slice: ties
    o = oracle.cluster_and_refine(R, 12, 30, nb, 2013, with_perm=True)
    ncl = len(o["reps"])
    members = set(o["mem"].tolist())
    gl = [set(o["mem"][o["off"][c]:o["off"][c + 1]].tolist()) for c in range(ncl)]
    for s in range(16):
        n = o["count"][s]
        assert ncl <= n <= 30
        sizes = o["size"][s, :n]
        assert sizes.min() >= 1 and sizes.sum() == len(o["mem"])
        perm = o["perm"][s, :sizes.sum()]
        assert sorted(perm.tolist()) == sorted(members)
        pos = 0
        for c in range(n):
            part = set(perm[pos:pos + sizes[c]].tolist())
            assert o["rep"][s, c] in part
            assert any(part <= g for g in gl)                     # a piece of one global cluster
            pos += sizes[c]


def test_refinement_draw_accounting_and_stream_continuation():
    R, nb = _inputs(300, 12, 4, 2)
    o = oracle.cluster_and_refine(R, 10, 24, nb, 77)
    L = oracle.lib()
    after = L.orc_mt_rand()
    # the same stream: clustering, then exactly sum(draws) numbers, then `after`
    oracle.cluster_by_sampling(R, 10, 77)
    for _ in range(int(o["draws"].sum())):
        L.orc_mt_rand()
    assert L.orc_mt_rand() == after
    for s in range(12):
        splits = o["count"][s] - len(o["reps"])
        assert 2 * splits <= o["draws"][s] <= 4 * splits


def test_slice_cost_closed_form():
    rng = np.random.default_rng(3)
    R = rng.uniform(-1, 1, (50, 9, 3)).astype(np.float32)
    rows = np.array([0, 4, 7], np.int32)
    idx = rng.choice(50, 17, replace=False).astype(np.int32)
    C = R[idx][:, rows, :].reshape(len(idx), -1).astype(np.float64)   # columns c_i (flattened slice rows)
    n = np.linalg.norm(C, axis=1)
    ref = (np.outer(n, n) - C @ C.T).sum()
    got = oracle.slice_cost(R, rows, idx)
    assert got == pytest.approx(ref, rel=1e-4, abs=1e-3)
    assert oracle.slice_cost(R, rows, idx[:1]) == pytest.approx(0.0, abs=1e-5)


def test_no_refinement_when_max_equals_global_count():
    R, nb = _inputs(200, 8, 4, 4)
    base = oracle.cluster_by_sampling(R, 9, 5)
    o = oracle.cluster_and_refine(R, 9, len(base["reps"]), nb, 5)
    for s in range(8):
        assert o["count"][s] == len(base["reps"]) and o["draws"][s] == 0
        assert np.array_equal(o["rep"][s, :o["count"][s]], base["reps"])
        assert np.array_equal(o["w"][s, :o["count"][s]].view(np.uint32), base["rep_w"].view(np.uint32))


# ---- an independent restatement for tie-free inputs ----------------------------
def _f32sum(vals):
    acc = np.float32(0.0)
    for v in vals:
        acc = np.float32(acc + np.float32(v))
    return acc


def _pdf(arr):
    s = _f32sum(arr)
    n = len(arr)
    if n == 0 or s == np.float32(0):
        return np.zeros(n, np.float32), np.zeros(n, np.float32)
    pdf = np.array([np.float32(a) / s for a in arr], np.float32)
    cdf = np.zeros(n, np.float32)
    cdf[0] = pdf[0]
    for i in range(1, n):
        cdf[i] = np.float32(cdf[i - 1] + pdf[i])
    cdf[-1] = np.float32(1.0)
    return pdf, cdf


def _sample(pdf, cdf, u):
    for i in range(len(cdf)):
        if u < cdf[i]:
            return i, pdf[i]
    return -1, np.float32(0)


def _cost(X):
    """X: (n, M) float32 columns of the slice matrix"""
    n, M = X.shape
    norms = [np.float32(np.sqrt(_f32sum([X[j, r] * X[j, r] for r in range(M)]))) for j in range(n)]
    u = [np.float32(0)] * M
    for j in range(n):
        for r in range(M):
            u[r] = np.float32(u[r] + X[j, r])
    v = [np.float32(0)] * n
    for r in range(M):
        for j in range(n):
            v[j] = np.float32(v[j] + np.float32(X[j, r] * u[r]))
    c = _f32sum(norms)
    cost = np.float32(0)
    for j in range(n):
        cost = np.float32(cost + np.float32(np.float32(norms[j] * c) - v[j]))
    return cost


def _py_refine(R, nb, glob, mags, norms_R, max_c, draw):
    S = R.shape[1]
    out = []
    for s in range(S):
        X = R[:, nb[s], :].reshape(R.shape[0], -1).astype(np.float32)     # column k -> 3*nnb values
        cls = [dict(idx=list(g["idx"]), pdf=_pdf([mags[k] for k in g["idx"]]), rep=g["rep"], w=g["w"]) for g in glob]
        costs = [_cost(X[c["idx"]]) for c in cls]
        alive = list(range(len(cls)))
        while alive and len(cls) < max_c:
            top = max(alive, key=lambda c: costs[c])                    # distinct costs: the heap's top
            cl = cls[top]
            n = len(cl["idx"])
            if n <= 1:
                break
            if n == 2:
                i1, i2 = [cl["idx"][0]], [cl["idx"][1]]
            else:
                i, _ = _sample(*cl["pdf"], draw())
                j, _ = _sample(*cl["pdf"], draw())
                if i < 0 or j < 0:
                    i1, i2 = cl["idx"][:n // 2], cl["idx"][n // 2:]
                else:
                    a, b = X[cl["idx"][i]], X[cl["idx"][j]]
                    ab = (a - b).astype(np.float32)
                    line = []
                    for k in range(n):
                        if k in (i, j):
                            continue
                        ac = (a - X[cl["idx"][k]]).astype(np.float32)
                        line.append((k, _f32sum([np.float32(ac[r] * ab[r]) for r in range(len(ab))])))
                    line.sort(key=lambda p: p[1])
                    gaps = [np.float32(line[k][1] - line[k - 1][1]) for k in range(1, len(line))]
                    best, cut = np.float32(0), -1
                    for k, g in enumerate(gaps, start=1):
                        if g > best:
                            best, cut = g, k
                    if cut < 0:
                        i1, i2 = cl["idx"][:n // 2], cl["idx"][n // 2:]
                    else:
                        i1 = [cl["idx"][i]] + [cl["idx"][p[0]] for p in line[:cut]]
                        i2 = [cl["idx"][j]] + [cl["idx"][p[0]] for p in line[cut:]]
            new = []
            for part in (i1, i2):
                c = dict(idx=part, pdf=_pdf([norms_R[k] for k in part]))
                k, pk = _sample(*c["pdf"], draw())
                c["rep"], c["w"] = (part[len(part) // 2], np.float32(0)) if k < 0 else (part[k], np.float32(1) / pk)
                new.append((c, _cost(X[part])))
            cls[top], costs[top] = new[0]
            cls.append(new[1][0])
            costs.append(new[1][1])
            alive.append(len(cls) - 1)
        out.append(cls)
    return out


@pytest.mark.parametrize("seed", [5, 6])
def test_refinement_matches_independent_restatement(seed):
    V, S = 90, 5
    R, nb = _inputs(V, S, 3, seed, zero_frac=0.0)       # continuous values: no equal keys or costs
    o = oracle.cluster_and_refine(R, 6, 11, nb, seed)
    # replay: clustering's draws, then this restatement draws from the same stream
    oracle.cluster_by_sampling(R, 6, seed)
    L = oracle.lib()
    mags = oracle.column_magnitudes(R)
    norms_R = np.array([np.sqrt(_f32sum([np.float32(np.float32(np.float32(e[0] * e[0]) + np.float32(e[1] * e[1]))
                                                    + np.float32(e[2] * e[2])) for e in R[k]])) for k in range(V)],
                       np.float32)
    glob = [dict(idx=o["mem"][o["off"][c]:o["off"][c + 1]].tolist(), rep=o["reps"][c], w=o["rep_w"][c])
            for c in range(len(o["reps"]))]
    py = _py_refine(R, nb, glob, mags, norms_R, 11, lambda: np.float32(L.orc_mt_rand()))
    for s in range(S):
        assert o["count"][s] == len(py[s])
        for c, cl in enumerate(py[s]):
            assert o["size"][s, c] == len(cl["idx"])
            assert o["rep"][s, c] == cl["rep"]
            assert np.float32(o["w"][s, c]).view(np.uint32) == np.float32(cl["w"]).view(np.uint32)
