"""CPU pins of the oracle's restatement of the reference's GPU-variant method
(ReconstructionCudaKernels.cu:242-868): each part against what the
mathematics fixes, independently of the restated code.

- checkTriangles (.cu:586-632) is "some triangle of the points with edges
  shorter than 2 x dispersion contains the output sample": a float64 brute
  force over all triples with a barycentric containment test;
- the variant's reprojection (.cu:419-437) against the closed form of a
  point moving linearly in homogeneous space, seen through lens (u, v);
- the nearest search (.cu:476-579) against a brute-force minimum over the
  samples of every leaf whose bounds reach the output sample;
- the 1R first surface (.cu:332-473) on a single static plane against the
  tent-weighted mean of every sample within R (float64);
- the 2R surfaces (.cu:706-867) on a single static plane against the tent
  mean of every sample within the dispersion (float64);
- front-to-back order: with a near and a far plane both present everywhere,
  the first surface is the near one (a reversed order gives the far colour).
No GPU."""
import numpy as np
import pytest


def _queries(orc, W, H, M, n, seed, margin=0):
    pat = orc.output_pattern(M)
    L = orc.lib()
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x = int(rng.integers(margin, W - margin))
        y = int(rng.integers(margin, H - margin))
        i = int(rng.integers(M))
        p = L.orc_hash_bits(x, y) % 64
        t = pat[p * M + i]
        out.append((x, y, i, np.array([t[0] + x, t[1] + y, t[2], t[3], t[4]], np.float32)))
    return out


def _contains(a, b, c):
    """Origin strictly inside triangle abc (float64), or None when it is
    within rounding of an edge."""
    o1 = a[0] * b[1] - b[0] * a[1]
    o2 = b[0] * c[1] - c[0] * b[1]
    o3 = c[0] * a[1] - a[0] * c[1]
    if min(abs(o1), abs(o2), abs(o3)) < 1e-9:
        return None
    return (o1 > 0 and o2 > 0 and o3 > 0) or (o1 < 0 and o2 < 0 and o3 < 0)


def test_triangles_brute_force(orc):
    rng = np.random.default_rng(5)
    disp = 0.5
    dcomp = 4.0 * disp * disp
    seen = {True: 0, False: 0}
    for trial in range(3000):
        n = int(rng.integers(0, 11))
        spread = [0.4, 0.8, 1.2][trial % 3]
        pts = rng.uniform(-spread, spread, (n, 2)).astype(np.float32)
        if trial % 5 == 0 and n:
            pts[:, 1] = np.abs(pts[:, 1])  # all on one side: no triangle around the origin
        P = pts.astype(np.float64)
        expect, ambiguous = False, False
        for i in range(n):
            for j in range(i + 1, n):
                for k in range(j + 1, n):
                    d = [((P[a] - P[b]) ** 2).sum() for a, b in ((i, j), (j, k), (i, k))]
                    if any(abs(x - dcomp) < 1e-6 for x in d):
                        ambiguous = True
                        continue
                    if max(d) >= dcomp:
                        continue
                    c = _contains(P[i], P[j], P[k])
                    if c is None:
                        ambiguous = True
                    elif c:
                        expect = True
        if ambiguous and not expect:
            continue
        got = orc.refgpu_triangles(pts, disp)
        assert got == expect, (trial, pts.tolist())
        seen[expect] += 1
    assert seen[True] > 300 and seen[False] > 300, seen


@pytest.fixture(scope="module")
def motion(tlr, orc):
    W, H, spp = 64, 48, 16
    sb = tlr.make_scene(2, W, H, spp, seed=31)
    o = orc.Oracle(W, H, spp, sb.coc[0], sb.coc[1], sb.records)
    bi = o.build(threads=4)
    rep, src = o.reprojected()
    pos = {s: k for k, s in enumerate(src.tolist())}
    P = rep[[pos[s] for s in o.layout().tolist()]].astype(np.float64)
    # samples of the leaves the hierarchy holds (invalid leaves are not in it)
    s01, _ = o.leaves()
    intree = np.zeros(len(P), bool)
    for lf in o.hierarchy()[0]:
        intree[s01[lf][0]:s01[lf][1]] = True
    return sb, o, bi, P, intree


def test_reprojection_closed_form(orc, motion):
    sb, o, bi, P, _ = motion
    cc0, cc1 = float(bi.coc_adj[0]), float(bi.coc_adj[1])
    for x, y, i, q in _queries(orc, sb.width, sb.height, 16, 40, 3):
        got = o.refgpu_reproject(q).astype(np.float64)
        dt = float(q[4]) - P[:, 3]
        Wt = P[:, 2] + dt * P[:, 6]
        X = (P[:, 0] * P[:, 2] + dt * P[:, 4]) / Wt + (cc0 / Wt + cc1) * float(q[2])
        Y = (P[:, 1] * P[:, 2] + dt * P[:, 5]) / Wt + (cc0 / Wt + cc1) * float(q[3])
        tol = 2e-5 * (np.abs(X) + np.abs(cc0 / Wt) + abs(cc1) + abs(float(q[0]))) + 1e-5
        assert np.all(np.abs(got[:, 0] - (X - float(q[0]))) <= tol)
        assert np.all(np.abs(got[:, 1] - (Y - float(q[1]))) <= tol)
        assert np.allclose(got[:, 2], got[:, 0] ** 2 + got[:, 1] ** 2, rtol=1e-6, atol=1e-12)


def test_nearest_brute_force(orc, motion):
    """The walk shrinks its radius to each nearer sample, so the answer k
    is nearest among the samples of every leaf whose bounds reach the output
    sample at radius |k| (those leaves were all visited): a brute force over
    the leaves that a bounds test alone selects."""
    sb, o, bi, P, intree = motion
    s01, _ = o.leaves()
    disp = 0.5  # 16 spp
    found = 0
    for x, y, i, q in _queries(orc, sb.width, sb.height, 16, 300, 4):
        d2 = o.refgpu_reproject(q)[:, 2].astype(np.float64)
        k, col = o.refgpu_nearest(q)
        r = np.sqrt(d2[k]) if k >= 0 else 16 * disp
        lv = o.leaves_bruteforce(q, float(r) * (1 - 1e-6))
        idx = np.concatenate([np.arange(*s01[l]) for l in lv]) if len(lv) else np.zeros(0, int)
        assert intree[idx].all()
        if k < 0:
            assert not col.any()
            assert np.all(d2[idx] > (16 * disp) ** 2 * (1 - 1e-6))
            continue
        assert intree[k]
        assert np.all(d2[idx] >= d2[k] * (1 - 1e-6)), (x, y, i)
        assert np.array_equal(col[:3], P[k, 7:10].astype(np.float32)) and col[3] == 1.0
        found += 1
    assert found > 250


def test_first_surface_single_plane(tlr, orc):
    """One static plane with defocus: every 1R leaf is the same surface, so a
    stage-0 colour is the tent-weighted mean of every sample within R."""
    W, H, spp, M = 48, 40, 16, 8
    sb = tlr.make_scene(1, W, H, spp, seed=12, num_quads=0)
    o = orc.Oracle(W, H, spp, sb.coc[0], sb.coc[1], sb.records)
    o.build(threads=4)
    rep, src = o.reprojected()
    pos = {s: k for k, s in enumerate(src.tolist())}
    P = rep[[pos[s] for s in o.layout().tolist()]].astype(np.float64)
    R = 0.5
    n = 0
    for x, y, i, q in _queries(orc, W, H, M, 300, 6, margin=8):
        col, stage = o.refgpu_query(q)
        if stage != 0:
            continue
        d = np.sqrt(o.refgpu_reproject(q)[:, 2].astype(np.float64))
        w = np.maximum(0.0, 1.0 - d / R)
        sel = d <= R
        expect = (w[sel, None] * P[sel, 7:10]).sum(0) / w[sel].sum()
        assert np.allclose(col[:3], expect, atol=3e-6), (x, y, i)
        assert abs(col[3] - 1.0) < 1e-6
        n += 1
    assert n > 200


def test_2r_single_plane(tlr, orc):
    """failKernel on a single static plane: one surface holds every 2R leaf,
    and the colour is the tent mean (weights 1 - d / dispersion, positive
    only) of the samples within the dispersion."""
    W, H, spp, M = 48, 40, 16, 8
    sb = tlr.make_scene(1, W, H, spp, seed=13, num_quads=0)
    o = orc.Oracle(W, H, spp, sb.coc[0], sb.coc[1], sb.records)
    o.build(threads=4)
    rep, src = o.reprojected()
    pos = {s: k for k, s in enumerate(src.tolist())}
    P = rep[[pos[s] for s in o.layout().tolist()]].astype(np.float64)
    disp = 0.5
    n = 0
    for x, y, i, q in _queries(orc, W, H, M, 200, 7, margin=8):
        ok, col = o.refgpu_failed(q)
        d = np.sqrt(o.refgpu_reproject(q)[:, 2].astype(np.float64))
        w = 1.0 - d / disp
        sel = w > 0
        assert ok == bool(sel.any())
        if not ok:
            continue
        expect = (w[sel, None] * P[sel, 7:10]).sum(0) / w[sel].sum()
        assert np.allclose(col[:3], expect, atol=3e-6), (x, y, i)
        n += 1
    assert n > 150


def test_first_surface_is_the_near_plane(tlr, orc):
    """Half of each pixel's samples from a red plane at depth 2, half from a
    green plane at depth 20: the first surface is red.  Leaves group samples
    by position, so some hold samples of both planes and let a little green
    into the first surface; a back-to-front order would make it green."""
    W, H, spp = 40, 32, 16
    near = tlr.make_scene(1, W, H, spp, seed=1, num_quads=0, w_max=1.6)
    far = tlr.make_scene(1, W, H, spp, seed=2, num_quads=0, w_max=16.0)
    rec = near.records.reshape(H, W, spp, 16).copy()
    rec[..., :, 7:10] = [1.0, 0.0, 0.0]
    f = far.records.reshape(H, W, spp, 16).copy()
    f[..., 7:10] = [0.0, 1.0, 0.0]
    rec[:, :, spp // 2:] = f[:, :, spp // 2:]
    o = orc.Oracle(W, H, spp, near.coc[0], near.coc[1], rec.reshape(-1, 16))
    o.build(threads=4)
    img, q, stg = o.reconstruct_refgpu(out_spp=8, want_queries=True, threads=4)
    inner = (slice(6, H - 6), slice(6, W - 6))
    q0, s0 = q[inner], stg[inner]
    first = s0 == 0
    assert first.mean() > 0.5
    assert q0[first][:, 0].mean() > 0.9 and q0[first][:, 1].mean() < 0.1
