"""GPU parity: every build stage and the final image of the HIP path against
the CPU oracle on the same seeded inputs.

Bit-exact: reprojection, bucketing layout, initial leaves, leaf membership and
leaf bounds, per-query results when no FP reordering is involved.  Images:
max-abs <= 2e-3 and mean-abs <= 1e-4 linear RGB (the stated tolerance).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # (name, preset, W, H, spp, out_spp)
    ("defocus_cfg1", 1, 128, 128, 4, 16),
    ("motion_cfg2_small", 2, 128, 96, 16, 8),
    ("strong_cfg3_small", 3, 160, 120, 16, 8),
]

_cache = {}


def _pair(tlr, orc, case):
    if case in _cache:
        return _cache[case]
    name, preset, W, H, spp, out_spp = case
    print("case", name, flush=True)
    sbuf = tlr.make_scene(preset, W, H, spp, seed=1000 * preset + 11)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sbuf)
    ctx.build()
    o = orc.Oracle(W, H, spp, sbuf.coc[0], sbuf.coc[1], sbuf.records)
    o.build()
    _cache[case] = (sbuf, ctx, o)
    return _cache[case]


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_build_info(tlr, orc, case):
    sbuf, ctx, o = _pair(tlr, orc, case)
    g, r = ctx.build_info(), o.info()
    assert g.num_kept == r.num_kept
    assert g.num_bucketed == r.num_bucketed
    assert g.spp_pow2 == r.spp_pow2
    assert list(g.coc_adj) == list(r.coc_adj)
    assert list(g.bbmin) == list(r.bbmin) and list(g.bbmax) == list(r.bbmax)
    assert g.scale == r.scale
    assert list(g.bbmin_int) == list(r.bbmin_int)
    assert (g.grid_w, g.grid_h) == (r.grid_w, r.grid_h)
    assert g.max_hits == r.max_hits
    assert g.num_initial_leaves == r.num_initial_leaves
    assert g.num_leaves == r.num_leaves


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_reprojection_bitexact(tlr, orc, case):
    sbuf, ctx, o = _pair(tlr, orc, case)
    gx, gy, gw, gt = ctx.reprojected()
    ref, src = o.reprojected()
    assert np.count_nonzero(gw > 0) == len(src)
    for k, a in enumerate((gx, gy, gw, gt)):
        assert np.array_equal(_bits(a[src]), _bits(ref[:, k])), "field %d differs" % k


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_layout_and_initial_leaves(tlr, orc, case):
    sbuf, ctx, o = _pair(tlr, orc, case)
    assert np.array_equal(ctx.layout(), o.layout())
    gi = ctx.initial_leaves()
    ri = o.initial_leaves()
    gi = gi[np.lexsort((gi[:, 1], gi[:, 0], gi[:, 4]))]
    ri = ri[np.lexsort((ri[:, 1], ri[:, 0], ri[:, 4]))]
    assert np.array_equal(gi, ri)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_leaves_bitexact(tlr, orc, case):
    sbuf, ctx, o = _pair(tlr, orc, case)
    gs, gb = ctx.leaves()
    rs, rb = o.leaves()
    assert np.array_equal(gs, rs), "leaf membership differs"
    bad = np.nonzero(np.any(_bits(gb) != _bits(rb), axis=1))[0]
    assert bad.size == 0, "leaf bounds differ at %d leaves, first %s" % (bad.size, bad[:5])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_hierarchy_bitexact(tlr, orc, case):
    """Median-split hierarchy: same leaf order, same children, same merged
    bounds as the oracle's recursive build."""
    sbuf, ctx, o = _pair(tlr, orc, case)
    go, gc, gb = ctx.hierarchy()
    ro, rc, rb = o.hierarchy()
    assert go.size == ro.size == ctx.build_info().num_valid_leaves
    assert np.array_equal(go, ro), "leaf order differs at %s" % np.nonzero(go != ro)[0][:5]
    assert np.array_equal(gc, rc), "children differ"
    assert np.array_equal(_bits(gb), _bits(rb)), "internal bounds differ"


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_image_parity(tlr, orc, case):
    sbuf, ctx, o = _pair(tlr, orc, case)
    out_spp = case[5]
    img, st = ctx.reconstruct(out_spp=out_spp, collect_stats=True)
    ref, qref, rst = o.reconstruct(out_spp=out_spp, want_queries=True)
    q = ctx.query_colors(out_spp)
    qerr = np.abs(q - qref).max(axis=-1)
    nbad = int(np.count_nonzero(qerr > 1e-4))
    err = np.abs(img - ref)
    print("%s: max-abs %.3g mean-abs %.3g, queries differing %d / %d, gpu %.2f ms" % (
        case[0], err.max(), err.mean(), nbad, qerr.size, st["query_ms"]))
    for k in ("leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched",
              "surfaces", "num_2r", "surfaces_merged", "at_least_one", "count"):
        assert st[k] == rst[k], "stat %s: gpu %s oracle %s" % (k, st[k], rst[k])
    assert st["overflow"] == 0
    assert nbad == 0, "per-query colours differ at %d output samples" % nbad
    assert err.max() <= 2e-3 and err.mean() <= 1e-4


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fast_path_equals_general(tlr, orc, case):
    """The 1R lane kernel (+ general kernel for the rest), with and without
    stats, gives bit-identical output samples to the general kernel alone."""
    sbuf, ctx, o = _pair(tlr, orc, case)
    M = case[5]
    _, st_g = ctx.reconstruct(out_spp=M, path=1, collect_stats=True)
    qg = ctx.query_colors(M)
    _, st_f = ctx.reconstruct(out_spp=M, path=0, collect_stats=True)
    qf = ctx.query_colors(M)
    _, st_n = ctx.reconstruct(out_spp=M, path=0, collect_stats=False)
    qn = ctx.query_colors(M)
    print("%s: general %.2f ms; fast %.2f + general %.2f ms for %d of %d queries" % (
        case[0], st_g["query_ms"], st_n["fast_ms"], st_n["general_ms"], st_n["general_queries"],
        st_n["count"] if st_n["count"] else qg.size // 4))
    assert np.array_equal(_bits(qg), _bits(qf))
    assert np.array_equal(_bits(qg), _bits(qn))
    assert st_n["general_queries"] == st_f["general_queries"]
    for k in ("leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched",
              "surfaces", "num_2r", "surfaces_merged", "at_least_one", "count"):
        assert st_g[k] == st_f[k], k
