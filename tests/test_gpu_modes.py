"""GPU parity of the reconstruction modes around the default one (the
reference's pinhole, refocus, uv-sweep and Gaussian-filter variants): the
same kernels with other output templates, CoC coefficients or weights.

Pinhole: App::reconstructPinhole (App.cpp:454-460) reconstructs with
overrideUVT = (0,0,1): one output sample per pixel at the pixel centre
(Reconstruction.cpp:454-455,493-500).  Refocus: the refocus sweep
(App.cpp:252-254) builds a new TreeGather with aperture / focal-distance
adjustments (Reconstruction.cpp:77-87).  uv sweep: overrideUVT with other
(u, v, t).  Gaussian lens / time filters: Reconstruction.hpp:378 weights."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(img, st, ref, rst, q, qref, what):
    qerr = np.abs(q - qref).max(axis=-1)
    nbad = int(np.count_nonzero(qerr > 1e-4))
    err = np.abs(img - ref)
    print("%s: max-abs %.3g mean-abs %.3g, queries differing %d / %d" % (what, err.max(), err.mean(), nbad,
                                                                           qerr.size), flush=True)
    assert nbad == 0
    assert err.max() <= 2e-3 and err.mean() <= 1e-4
    for k in ("leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched", "surfaces",
              "num_2r", "count"):
        assert st[k] == rst[k], (what, k, st[k], rst[k])


@pytest.fixture(scope="module")
def scene(tlr, orc):
    sb = tlr.make_scene(2, 112, 80, 16, seed=4242)
    return sb


@pytest.mark.parametrize("uvt", [(0.0, 0.0, 1.0), (0.5, -0.25, 0.3), (-0.9, 0.8, 0.0)])
def test_override_uvt(tlr, orc, scene, uvt):
    sb = scene
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build()
    img, st = ctx.reconstruct(out_spp=1, override_uvt=uvt, collect_stats=True)
    q = ctx.query_colors(1)
    o = orc.Oracle(sb.width, sb.height, sb.spp, sb.coc[0], sb.coc[1], sb.records)
    o.build()
    ref, qref, rst = o.reconstruct(out_spp=1, override_uvt=uvt, want_queries=True)
    _check(img, st, ref, rst, q, qref, "uvt %s" % (uvt,))


@pytest.mark.parametrize("adj", [(0.5, 1.0), (1.0, 1.3), (2.0, 0.8)])
def test_refocus_adjust(tlr, orc, scene, adj):
    sb = scene
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build(aperture_adjust=adj[0], focal_adjust=adj[1])
    o = orc.Oracle(sb.width, sb.height, sb.spp, sb.coc[0], sb.coc[1], sb.records)
    o.build(aperture_adjust=adj[0], focal_adjust=adj[1])
    g, r = ctx.build_info(), o.info()
    assert list(g.coc_adj) == list(r.coc_adj)
    assert g.num_leaves == r.num_leaves
    img, st = ctx.reconstruct(out_spp=8, collect_stats=True)
    q = ctx.query_colors(8)
    ref, qref, rst = o.reconstruct(out_spp=8, want_queries=True)
    _check(img, st, ref, rst, q, qref, "adjust %s" % (adj,))


@pytest.mark.parametrize("lens,time", [(True, False), (False, True), (True, True)])
def test_gaussian_filters(tlr, orc, scene, lens, time):
    sb = scene
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build()
    img, st = ctx.reconstruct(out_spp=8, lens_gaussian=lens, time_gaussian=time, collect_stats=True)
    q = ctx.query_colors(8)
    o = orc.Oracle(sb.width, sb.height, sb.spp, sb.coc[0], sb.coc[1], sb.records)
    o.build()
    ref, qref, rst = o.reconstruct(out_spp=8, lens_gaussian=int(lens), time_gaussian=int(time), want_queries=True)
    _check(img, st, ref, rst, q, qref, "lens %s time %s" % (lens, time))
