"""GPU: the reference's edge cases against the oracle.

- samples with w <= 0 at t = 0.5 that the reprojection moves to t = 0 or 1,
  samples crossing w = 0 inside the shutter (time-span clipping), and samples
  discarded outright (Reconstruction.hpp:93-121, TimeBounds.cpp:247-260,
  Reconstruction.cpp:561-587): every build stage and every output sample
  bit for bit;
- input spp that is not a power of two (3, 12: rounded up for the filter
  radius, Reconstruction.cpp:65-73): bit for bit;
- the error paths: spp > 256 (TLR_E_UNSUPPORTED_SPP, ReconstructionOur.cpp:66),
  every sample discarded (TLR_E_ALL_DISCARDED, Reconstruction.cpp:587), an
  output image smaller than the buffer (TLR_E_IMAGE_TOO_SMALL,
  Reconstruction.cpp:223), a NULL device destination (TLR_E_INVALID), and a
  query beyond the per-query capacity (TLR_E_CAPACITY, through the test build
  lib/libtlr_tiny_cap.so whose general kernel holds 8 leaves).
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import edge_scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TINY_LIB = os.path.join(ROOT, "temporal-lightfield-reconstruction_amd", "lib", "libtlr_tiny_cap.so")
STATS = ("leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched", "surfaces", "num_2r",
         "surfaces_merged", "at_least_one", "count")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _compare(tlr, orc, sb, M):
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build()
    o = orc.Oracle(sb.width, sb.height, sb.spp, sb.coc[0], sb.coc[1], sb.records)
    o.build()
    g, r = ctx.build_info(), o.info()
    assert (g.num_kept, g.num_bucketed, g.spp_pow2, g.num_leaves) == \
        (r.num_kept, r.num_bucketed, r.spp_pow2, r.num_leaves)
    gx, gy, gw, gt = ctx.reprojected()
    ref, src = o.reprojected()
    assert np.count_nonzero(gw > 0) == len(src)
    for k, a in enumerate((gx, gy, gw, gt)):
        assert np.array_equal(_bits(a[src]), _bits(ref[:, k])), "reprojected field %d differs" % k
    assert np.array_equal(ctx.layout(), o.layout())
    gs, gb = ctx.leaves()
    rs, rb = o.leaves()
    assert np.array_equal(gs, rs), "leaf membership differs"
    assert np.array_equal(_bits(gb), _bits(rb)), "leaf bounds differ"
    _, qref, rst = o.reconstruct(out_spp=M, want_queries=True)
    for path in (0, 1):
        _, st = ctx.reconstruct(out_spp=M, path=path, collect_stats=True)
        q = ctx.query_colors(M)
        bad = np.any(_bits(q) != _bits(qref), axis=-1)
        assert not bad.any(), "path %d: %d output samples differ" % (path, int(bad.sum()))
        for k in STATS:
            assert st[k] == rst[k], "path %d stat %s: gpu %s oracle %s" % (path, k, st[k], rst[k])
    return ctx, o


def test_w_crossings_and_partial_discards(tlr, orc):
    sb = edge_scenes.wcross_scene(tlr)
    cat, crosses = edge_scenes.classify(sb.records)
    counts = np.bincount(cat, minlength=4)
    print("reprojection branches (0.5, 0, 1, discarded): %s; crossing w=0 in the shutter: %d" % (
        counts.tolist(), int(crosses.sum())), flush=True)
    assert counts[1] > 0 and counts[2] > 0 and counts[3] > 0 and crosses.sum() > 0
    ctx, o = _compare(tlr, orc, sb, 8)
    assert ctx.build_info().num_kept == int(np.count_nonzero(cat < 3))


@pytest.mark.parametrize("spp", [3, 12])
def test_non_power_of_two_spp(tlr, orc, spp):
    print("spp", spp, flush=True)
    sb = tlr.make_scene(2, 56, 40, spp, seed=90 + spp)
    ctx, _ = _compare(tlr, orc, sb, 8)
    assert ctx.build_info().spp_pow2 == (4 if spp == 3 else 16)


def test_unsupported_spp(tlr, orc):
    sb = tlr.make_scene(2, 4, 3, 300, seed=5)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    with pytest.raises(tlr.TlrError) as e:
        ctx.build()
    assert e.value.code == 4
    o = orc.Oracle(4, 3, 300, sb.coc[0], sb.coc[1], sb.records)
    with pytest.raises(RuntimeError, match="\\(4\\)"):
        o.build()


def test_all_discarded(tlr, orc):
    rec = edge_scenes.all_discarded_records()
    H, W, spp = rec.shape[:3]
    sb = tlr.UVTSampleBuffer(W, H, spp, -15.24, 19.05, rec)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    with pytest.raises(tlr.TlrError) as e:
        ctx.build()
    assert e.value.code == 3
    o = orc.Oracle(W, H, spp, -15.24, 19.05, rec)
    with pytest.raises(RuntimeError, match="\\(3\\)"):
        o.build()
    # the context stays usable
    good = tlr.make_scene(1, 16, 12, 2, seed=3)
    ctx.set_samples(good)
    ctx.build()
    img, _ = ctx.reconstruct(out_spp=4)
    assert np.isfinite(img).all()


def test_image_too_small(tlr):
    sb = tlr.make_scene(1, 24, 16, 2, seed=4)
    tg = tlr.TreeGather(sb, tlr.CameraParams(out_spp=4))
    with pytest.raises(tlr.TlrError) as e:
        tg.reconstruct_dof_motion(np.zeros((15, 24, 4), np.float32))
    assert e.value.code == 5
    img = tg.reconstruct_dof_motion(np.zeros((20, 30, 4), np.float32))
    assert img.shape == (20, 30, 4)


def test_null_device_destination(tlr):
    sb = tlr.make_scene(1, 16, 12, 2, seed=6)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build()
    p = tlr.ReconParams()
    p.out_spp = 4
    rc = tlr.lib().tlr_reconstruct(ctx.h, ctypes.byref(p), None, 1, None)
    assert rc == 1, "expected TLR_E_INVALID, got %d" % rc
    img, _ = ctx.reconstruct(out_spp=4)
    assert np.isfinite(img).all()


def _tiny_child():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    tlr = ge.load_package()
    assert os.path.samefile(tlr.TLR_LIB, TINY_LIB)
    print("case: 1920x1080-type scene at 96x72, 16 spp, tiny leaf capacity", flush=True)
    sb = tlr.make_scene(3, 96, 72, 16, seed=77)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build()
    try:
        ctx.reconstruct(out_spp=4, path=1)
    except tlr.TlrError as e:
        print("error:", e, flush=True)
        sys.exit(0 if e.code == 7 else 1)
    print("no capacity error", flush=True)
    sys.exit(1)


def test_capacity_error():
    assert os.path.exists(TINY_LIB), "libtlr_tiny_cap.so is not built"
    env = dict(os.environ, TLR_LIBRARY=TINY_LIB)
    r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--tiny-child"], env=env, timeout=240,
                       capture_output=True, text=True)
    print(r.stdout + r.stderr)
    assert r.returncode == 0


if __name__ == "__main__" and "--tiny-child" in sys.argv:
    _tiny_child()
