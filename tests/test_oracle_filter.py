"""CPU pins of the oracle's filter: gather against an independent float64
brute force, the tent filter against a direct computation on a single
pinhole plane, colour invariance.  No GPU."""
import numpy as np
import pytest


def _queries(orc, W, H, M, n, seed):
    pat = orc.output_pattern(M)
    L = orc.lib()
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x, y, i = int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(M))
        p = L.orc_hash_bits(x, y) % 64
        t = pat[p * M + i]
        out.append((x, y, i, np.array([t[0] + x, t[1] + y, t[2], t[3], t[4]], np.float32)))
    return out


def test_gather_matches_float64_bruteforce(tlr, orc):
    sb = tlr.make_scene(2, 80, 64, 16, seed=77)
    o = orc.Oracle(80, 64, 16, sb.coc[0], sb.coc[1], sb.records)
    bi = o.build(threads=4)
    rep, src = o.reprojected()
    lay = o.layout()
    pos = {s: k for k, s in enumerate(src.tolist())}
    P = rep[[pos[s] for s in lay.tolist()]].astype(np.float64)
    s01, _ = o.leaves()
    cc0, cc1 = float(bi.coc_adj[0]), float(bi.coc_adj[1])
    checked = 0
    for x, y, i, q in _queries(orc, 80, 64, 16, 300, 1):
        R = 0.5
        lv, sm, sf = o.gather(q, R)
        bf = o.leaves_bruteforce(q, R)
        assert set(lv.tolist()) == set(bf.tolist())
        idx = np.concatenate([np.arange(*s01[l]) for l in bf]) if len(bf) else np.zeros(0, int)
        px, py, pw, pt = P[idx, 0], P[idx, 1], P[idx, 2], P[idx, 3]
        mv = P[idx, 4:7]
        dt = float(q[4]) - pt
        Xw = px * pw + dt * mv[:, 0]
        Yw = py * pw + dt * mv[:, 1]
        Ww = pw + dt * mv[:, 2]
        coc = cc0 / Ww + cc1
        rx = Xw / Ww + coc * q[2]
        ry = Yw / Ww + coc * q[3]
        d = np.hypot(rx - q[0], ry - q[1])
        inside = set(idx[d < R - 1e-4].tolist())
        outside = set(idx[d > R + 1e-4].tolist())
        got = set(sm.tolist())
        assert inside <= got and not (outside & got)
        # surfaces are contiguous runs in gather order
        assert np.all(np.diff(sf) >= 0)
        checked += len(got)
    assert checked > 1000


def test_tent_filter_single_pinhole_plane(tlr, orc):
    # One static plane, pinhole (C0 = C1 = 0): every sample reprojects to its
    # own position and all samples form one surface, so each output sample
    # with samples inside R is their tent-weighted mean (ReconstructionOur.cpp:268-297).
    W, H, spp, M = 48, 40, 16, 8
    sb = tlr.make_scene(1, W, H, spp, seed=3, num_quads=0, coc_c0=0.0, coc_c1=0.0)
    o = orc.Oracle(W, H, spp, 0.0, 0.0, sb.records)
    o.build(threads=4)
    img, q, st = o.reconstruct(out_spp=M, want_queries=True, threads=4)
    # positions as reprojected to the lens centre (x*w*(1/w) in FP32)
    rep, src = o.reprojected()
    rec = sb.records[src].astype(np.float64)
    px, py = rep[:, 0].astype(np.float64), rep[:, 1].astype(np.float64)
    R = 0.5
    n = 0
    for x, y, i, qq in _queries(orc, W, H, M, 400, 2):
        d = np.hypot(px - qq[0], py - qq[1])
        w = 1 - d / R
        sel = (d <= R) & (w > 0)
        if not sel.any():
            continue
        expect = (w[sel, None] * rec[sel, 7:10]).sum(0) / w[sel].sum()
        assert np.allclose(q[y, x, i, :3], expect, atol=2e-6), (x, y, i)
        n += 1
    assert n > 300


def test_constant_colour_is_preserved(tlr, orc):
    sb = tlr.make_scene(2, 64, 48, 8, seed=9)
    rec = sb.records.copy()
    rec[:, 7:10] = [0.25, 0.5, 0.75]
    o = orc.Oracle(64, 48, 8, sb.coc[0], sb.coc[1], rec)
    o.build(threads=4)
    img, _, st = o.reconstruct(out_spp=8, threads=4)
    assert np.allclose(img[..., :3], [0.25, 0.5, 0.75], atol=1e-5)
    assert np.allclose(img[..., 3], 1.0, atol=1e-5)
    assert st["count"] == 64 * 48 * 8


def test_pattern_layout(orc):
    # 64 patterns x M; positions inside the pixel, lens on the unit disk,
    # t stratified into M strata per pattern (Reconstruction.cpp:459-511)
    for M in (1, 8, 16, 32):
        p = orc.output_pattern(M)
        assert p.shape == (64 * M, 6)
        assert np.all((p[:, 0] >= 0) & (p[:, 0] < 1) & (p[:, 1] >= 0) & (p[:, 1] < 1))
        assert np.all(p[:, 2] ** 2 + p[:, 3] ** 2 <= 1 + 1e-5)
        for j in range(64):
            blk = p[j * M:(j + 1) * M]
            assert np.all(np.diff(blk[:, 5]) >= 0)  # sorted by key
            if M > 1:
                strata = np.sort(np.floor(blk[:, 4] * M).astype(int))
                assert np.array_equal(strata, np.arange(M))
    p = orc.output_pattern(16, override_uvt=(0.25, -0.5, 0.75))
    assert p.shape == (64, 6)
    assert np.all(p[:, :2] == 0.5) and np.all(p[:, 2:5] == np.float32([0.25, -0.5, 0.75]))
