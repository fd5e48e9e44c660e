"""CPU pins of the oracle's build: reprojection closed forms, grid and leaf
partition invariants, bounds containment/merge invariants, hierarchy equivalence
with brute force.  No GPU."""
import numpy as np
import pytest


def _scene(tlr, preset, W, H, spp, seed):
    return tlr.make_scene(preset, W, H, spp, seed=seed)


@pytest.fixture(scope="module")
def built(tlr, orc):
    out = {}
    for name, (preset, W, H, spp) in {"defocus": (1, 96, 80, 4), "motion": (2, 96, 72, 16),
                                      "strong": (3, 120, 90, 8)}.items():
        sb = _scene(tlr, preset, W, H, spp, 1000 * preset + 5)
        o = orc.Oracle(W, H, spp, sb.coc[0], sb.coc[1], sb.records)
        o.build(threads=4)
        out[name] = (sb, o)
    return out


@pytest.mark.parametrize("name", ["defocus", "motion", "strong"])
def test_reprojection_round_trip(built, name):
    # Reprojecting the (u,v,t)-centre sample back to its own (u,v,t) returns
    # its input position (Reconstruction.hpp:93-121 inverted in float64).
    sb, o = built[name]
    rep, src = o.reprojected()
    rec = sb.records[src].astype(np.float64)
    x, y, w, t = (rep[:, k].astype(np.float64) for k in range(4))
    mv = rec[:, 11:14]
    C0, C1 = sb.coc
    P = np.stack([x * w, y * w, w], 1) + (rec[:, 6] - t)[:, None] * mv
    xr = P[:, 0] / P[:, 2] + (C0 / P[:, 2] + C1) * rec[:, 4]
    yr = P[:, 1] / P[:, 2] + (C0 / P[:, 2] + C1) * rec[:, 5]
    assert np.abs(xr - rec[:, 0]).max() < 2e-3
    assert np.abs(yr - rec[:, 1]).max() < 2e-3
    assert np.abs(P[:, 2] - rec[:, 3]).max() < 1e-3 * rec[:, 3].max()
    # t is snapped to 0.5 unless that puts the sample behind the camera
    assert set(np.unique(t)).issubset({0.0, 0.5, 1.0})


@pytest.mark.parametrize("name", ["defocus", "motion", "strong"])
def test_grid_and_partition(built, name):
    sb, o = built[name]
    bi = o.info()
    rep, src = o.reprojected()
    cells = o.cells()
    # cell = floor(scale*xy) - bbmin_int (Reconstruction.cpp:617-618)
    sx = np.floor(np.float32(bi.scale) * rep[:, 0]).astype(np.int64) - bi.bbmin_int[0]
    sy = np.floor(np.float32(bi.scale) * rep[:, 1]).astype(np.int64) - bi.bbmin_int[1]
    assert np.array_equal(cells, sy * bi.grid_w + sx)
    # the scale targets <= ~48 samples per screen pixel (Reconstruction.cpp:590)
    assert bi.scale == np.float32(np.sqrt(np.float32(bi.max_hits) / np.float32(48)))
    layout = o.layout()
    assert sorted(layout.tolist()) == sorted(src.tolist())  # a permutation
    il = o.initial_leaves()
    s01, b = o.leaves()
    assert s01[0, 0] == 0 and s01[-1, 1] == len(layout)
    assert np.array_equal(s01[1:, 0], s01[:-1, 1])  # leaves tile the layout
    # every leaf lies inside one initial leaf and holds one t
    starts = il[:, 4]
    owner = np.searchsorted(starts, s01[:, 0], side="right") - 1
    assert np.all(s01[:, 1] <= starts[owner] + il[owner, 5])
    rec_t = {s: t for s, t in zip(src.tolist(), rep[:, 3].tolist())}
    rec_w = {s: w for s, w in zip(src.tolist(), rep[:, 2].tolist())}
    tl = np.array([rec_t[s] for s in layout.tolist()])
    wl = np.array([rec_w[s] for s in layout.tolist()])
    for a, e in s01:
        assert np.all(tl[a:e] == tl[a])
    # inside each initial leaf the samples are sorted by (t, w)
    for x0, y0, x1, y1, off, cnt in il:
        if cnt > 1:
            key = tl[off:off + cnt] * 1e6 + wl[off:off + cnt]
            tt, ww = tl[off:off + cnt], wl[off:off + cnt]
            assert np.all((tt[1:] > tt[:-1]) | ((tt[1:] == tt[:-1]) & (ww[1:] >= ww[:-1])))
    # each initial leaf's cells hold exactly its samples
    cell_of = dict(zip(src.tolist(), cells.tolist()))
    for x0, y0, x1, y1, off, cnt in il[:200]:
        for s in layout[off:off + cnt].tolist():
            c = cell_of[s]
            cx, cy = c % bi.grid_w, c // bi.grid_w
            assert x0 <= cx < x1 and y0 <= cy < y1


def test_bounds_contain_trajectory(orc):
    # computeTUXPlanes bounds are exact at ta, tb, tmid for u = +-1 (the
    # constant term is their min/max), so the trajectory lies inside there.
    L = orc.lib()
    rng = np.random.default_rng(4)
    out = np.zeros(14, np.float32)
    cc = np.array([-15.24, 19.05], np.float32)
    nvalid = 0
    for _ in range(400):
        w0 = rng.uniform(1, 20)
        x0 = rng.uniform(0, 200)
        y0 = rng.uniform(0, 150)
        mot = np.array([rng.uniform(-30, 30) * w0, rng.uniform(-30, 30) * w0, rng.uniform(-0.2, 0.2) * w0],
                       np.float32)
        pos = np.array([x0 * w0, y0 * w0, w0], np.float32)
        if not L.orc_tux_planes(pos, mot, 200.0, 150.0, cc, out):
            continue
        nvalid += 1
        p = out[:12].reshape(4, 3).astype(np.float64)
        ta, tb = float(out[12]), float(out[13])
        for t in (ta, tb, 0.5 * (ta + tb)):
            P = pos.astype(np.float64) + t * mot.astype(np.float64)
            ex, ey = P[0] / P[2], P[1] / P[2]
            coc = cc[0] / P[2] + cc[1]
            for u in (-1.0, 1.0):
                xv, yv = ex + coc * u, ey + coc * u
                tol = 1e-3 * (1 + abs(xv) + abs(yv))
                assert p[0, 0] * t + p[0, 1] * u + p[0, 2] <= xv + tol
                assert p[1, 0] * t + p[1, 1] * u + p[1, 2] >= xv - tol
                assert p[2, 0] * t + p[2, 1] * u + p[2, 2] <= yv + tol
                assert p[3, 0] * t + p[3, 1] * u + p[3, 2] >= yv - tol
    assert nvalid > 300


def test_bounds_static_closed_form(orc):
    # A static sample's planes are exactly x = x0/w0 + coc(w0)*u
    # (TimeBounds.cpp:231-241 and the constant-term fold).
    L = orc.lib()
    out = np.zeros(14, np.float32)
    cc = np.array([-15.24, 19.05], np.float32)
    for w0, x0, y0 in ((2.0, 10.5, 20.25), (7.5, 99.0, 3.0), (15.0, 0.5, 149.5)):
        pos = np.array([x0 * w0, y0 * w0, w0], np.float32)
        assert L.orc_tux_planes(pos, np.zeros(3, np.float32), 200.0, 150.0, cc, out)
        coc = np.float32(cc[0] / np.float32(w0) + cc[1])
        assert out[12] == 0 and out[13] == 1
        for k, c in ((0, x0), (1, x0), (2, y0), (3, y0)):
            assert out[3 * k] == 0
            assert abs(out[3 * k + 1] - coc) < 1e-5
            assert abs(out[3 * k + 2] - c) < 1e-4 * (1 + c)


def test_merge_is_conservative(orc):
    L = orc.lib()
    rng = np.random.default_rng(8)
    o = np.zeros(14, np.float32)
    for _ in range(500):
        a = rng.normal(0, 20, 14).astype(np.float32)
        b = rng.normal(0, 20, 14).astype(np.float32)
        a[12:] = [0.0, 1.0]
        b[12:] = [0.2, 0.9]
        L.orc_merge_bounds(a, b, o)
        assert o[12] == 0.0 and o[13] == 1.0
        for t in (0.0, 1.0):
            for u in (-1.0, 1.0):
                ev = lambda p, k: p[3 * k] * t + p[3 * k + 1] * u + p[3 * k + 2]
                tol = 1e-3 * (1 + np.abs(a).max() + np.abs(b).max())
                for k in (1, 3):  # upper planes bound from above
                    assert ev(o, k) >= max(ev(a, k), ev(b, k)) - tol
                for k in (0, 2):  # lower planes bound from below
                    assert ev(o, k) <= min(ev(a, k), ev(b, k)) + tol
    # an invalid operand is skipped; merging equal planes reproduces them
    inv = np.zeros(14, np.float32)
    L.orc_merge_bounds(a, inv, o)
    assert np.array_equal(o, a)
    a = np.array([0.5, 3.0, 10.0] * 4 + [0.0, 1.0], np.float32)
    L.orc_merge_bounds(a, a, o)
    assert np.allclose(o, a, atol=1e-5)


@pytest.mark.parametrize("name", ["defocus", "motion", "strong"])
def test_hierarchy_finds_every_intersecting_leaf(built, name):
    sb, o = built[name]
    rng = np.random.default_rng(12)
    disp = {1: 2.0, 2: 1.34, 4: 1.10, 8: 0.80, 16: 0.50}[o.info().spp_pow2]
    mism = 0
    for _ in range(300):
        q = np.array([rng.uniform(0, sb.width), rng.uniform(0, sb.height), rng.uniform(-1, 1),
                      rng.uniform(-1, 1), rng.uniform(0, 1)], np.float32)
        for R in (disp, 2 * disp):
            lv, _, _ = o.gather(q, R)
            bf = o.leaves_bruteforce(q, R)
            mism += set(lv.tolist()) != set(bf.tolist())
    assert mism == 0


@pytest.mark.parametrize("name", ["defocus", "motion", "strong"])
def test_hierarchy_is_a_balanced_median_tree(built, name):
    """Structure of the median-split hierarchy: every valid leaf appears once;
    internal node g splits slots [lo, hi) at mid = (lo+hi)//2 with g = mid-1,
    so the depth is ceil(log2 n); each internal bound holds its children at
    the corners of the (t, u) domain."""
    sb, o = built[name]
    order, ch, b = o.hierarchy()
    s01, lb = o.leaves()
    valid = np.nonzero(lb[:, 12] < lb[:, 13])[0]
    n = order.size
    assert sorted(order.tolist()) == valid.tolist()
    seen = []
    depth = 0
    stack = [(0, n, 0)]
    while stack:
        lo, hi, d = stack.pop()
        depth = max(depth, d)
        if hi - lo == 1:
            seen.append(lo)
            continue
        mid = (lo + hi) // 2
        g = mid - 1
        for k, (a, z) in enumerate(((lo, mid), (mid, hi))):
            c = int(ch[g, k])
            if z - a == 1:
                assert c < 0  # leaf id tagged with bit 31
                assert (c & 0x7FFFFFFF) == order[a]
                cb = lb[order[a]]
            else:
                assert c == (a + z) // 2 - 1
                cb = b[c]
            for t in (0.0, 1.0):
                for u in (-1.0, 1.0):
                    ev = lambda p, j: p[3 * j] * t + p[3 * j + 1] * u + p[3 * j + 2]
                    tol = 1e-3 * (1 + np.abs(cb[:12]).max())
                    assert ev(b[g], 1) >= ev(cb, 1) - tol and ev(b[g], 3) >= ev(cb, 3) - tol
                    assert ev(b[g], 0) <= ev(cb, 0) + tol and ev(b[g], 2) <= ev(cb, 2) + tol
            stack.append((a, z, d + 1))
    assert sorted(seen) == list(range(n))
    assert depth == int(np.ceil(np.log2(n)))


def test_error_paths(orc):
    # every sample behind the camera at t = 0.5, 0 and 1 -> all discarded
    rec = np.zeros((4, 4, 2, 16), np.float32)
    rec[..., 3] = -1.0
    o = orc.Oracle(4, 4, 2, -15.24, 19.05, rec)
    with pytest.raises(RuntimeError, match="\\(3\\)"):
        o.build()
