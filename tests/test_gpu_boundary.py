"""GPU: boundary entry points against the oracle.

- tlr_copy_query_gather: the general kernel's gathered sets (leaves front to
  back, gathered tree-layout sample indices and their surfaces) of 240 output
  samples at R and 2R equal the oracle's collectInputSamples2
  (ReconstructionOur.cpp:412-511) element for element;
- the debug image (#surfaces, #samples within 1R of each pixel's last output
  sample, ReconstructionOur.cpp:312-314) equals the oracle's, through both
  paths and through the TreeGather mirror;
- overrideRefocusDistance (Reconstruction.cpp:88-95) through explicit CoC
  coefficients: per-output-sample colours equal the oracle's built with the
  same coefficients, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATS = ("leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched", "surfaces", "num_2r",
         "surfaces_merged", "at_least_one", "count")


@pytest.fixture(scope="module")
def pair(tlr, orc):
    sb = tlr.make_scene(3, 96, 72, 16, seed=3301)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build()
    o = orc.Oracle(sb.width, sb.height, sb.spp, sb.coc[0], sb.coc[1], sb.records)
    o.build()
    return sb, ctx, o


def _osamples(tlr, orc, sb, M, n, seed):
    pat = tlr.output_pattern(M)
    rng = np.random.default_rng(seed)
    xs, ys, ii = rng.integers(0, sb.width, n), rng.integers(0, sb.height, n), rng.integers(0, M, n)
    o5 = np.zeros((n, 5), np.float32)
    for j, (x, y, i) in enumerate(zip(xs, ys, ii)):
        t = pat[(orc.lib().orc_hash_bits(int(x), int(y)) % 64) * M + i]
        o5[j] = (np.float32(t[0]) + np.float32(x), np.float32(t[1]) + np.float32(y), t[2], t[3], t[4])
    return o5


@pytest.mark.parametrize("scale", [1, 2])
def test_query_gather_matches_oracle(tlr, orc, pair, scale):
    sb, ctx, o = pair
    R = 0.5 * scale   # 16 spp: dispersion 0.5
    o5 = _osamples(tlr, orc, sb, 8, 240, 11 + scale)
    print("gather of %d output samples at R=%.2f" % (len(o5), R), flush=True)
    got = ctx.query_gather(o5, R)
    nonempty = 0
    for j in range(len(o5)):
        lv, sm, sf = o.gather(o5[j], R)
        assert got[j] is not None, "output sample %d overflowed" % j
        glv, gsm, gsf = got[j]
        assert np.array_equal(glv, lv), "output sample %d: leaves differ" % j
        assert np.array_equal(gsm, sm), "output sample %d: gathered samples differ" % j
        assert np.array_equal(gsf, sf), "output sample %d: surfaces differ" % j
        nonempty += len(sm) > 0
    assert nonempty > 200


@pytest.mark.parametrize("path", [0, 1])
def test_debug_image_matches_oracle(tlr, pair, path):
    sb, ctx, o = pair
    _, st = ctx.reconstruct(out_spp=8, path=path, debug_image=True)
    dbg = ctx.debug_image()
    _, _, rst, rdbg = o.reconstruct(out_spp=8, want_debug=True)
    assert np.array_equal(dbg, rdbg), "debug image differs at %d pixels" % int(np.any(dbg != rdbg, axis=-1).sum())
    assert dbg[..., 0].max() >= 2 and dbg[..., 1].max() > 0


def test_debug_image_through_tree_gather(tlr, orc):
    sb = tlr.make_scene(2, 48, 40, 8, seed=77)
    tg = tlr.TreeGather(sb, tlr.CameraParams(out_spp=4))
    img = np.zeros((40, 48, 4), np.float32)
    dbg = np.zeros((40, 48, 4), np.float32)
    tg.reconstruct_dof_motion(img, dbg)
    o = orc.Oracle(48, 40, 8, sb.coc[0], sb.coc[1], sb.records)
    o.build()
    ref, _, _, rdbg = o.reconstruct(out_spp=4, want_debug=True)
    assert np.array_equal(dbg, rdbg)
    assert np.abs(img - ref).max() <= 2e-3


def test_refocus_override_matches_oracle(tlr, orc):
    sb = tlr.make_scene(2, 96, 72, 16, seed=3302)
    # a perspective projection, window 96 px wide, lens radius 0.2, refocused to 4
    proj = np.array([[1.5, 0, 0, 0], [0, 2.0, 0, 0], [0, 0, -1.002, -0.2002], [0, 0, -1, 0]], np.float32)
    cam = tlr.CameraParams(out_spp=8, override_refocus_distance=4.0, projection=proj, window_width=96,
                           aperture_radius=0.2)
    cc = cam.refocus_coc()
    print("refocused CoC coefficients", cc, "header", sb.coc, flush=True)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build(coc_override=cc)
    assert list(ctx.build_info().coc_adj) == [float(cc[0]), float(cc[1])]
    img, st = ctx.reconstruct(out_spp=8, collect_stats=True)
    q = ctx.query_colors(8)
    o = orc.Oracle(96, 72, 16, sb.coc[0], sb.coc[1], sb.records)
    o.build(coc_override=cc)
    ref, qref, rst = o.reconstruct(out_spp=8, want_queries=True)
    assert np.array_equal(q.view(np.uint32), qref.view(np.uint32))
    for k in STATS:
        assert st[k] == rst[k], k
    # the mirror class builds with the same coefficients
    tg = tlr.TreeGather(sb, cam)
    assert np.array_equal(tg.reconstruct_dof_motion(), img)
    ctx.build()
    plain, _ = ctx.reconstruct(out_spp=8)
    assert np.abs(plain - img).max() > 1e-2
