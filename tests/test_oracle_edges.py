"""CPU: the edge-case buffers used by the GPU edge tests exercise what they
are meant to (every reprojection branch, w = 0 crossings, discards), and the
oracle's handling of them agrees with closed forms: the kept count is the
float64 count of samples with a valid centre time, discarded samples are
absent from the layout, and every sample kept at t' = 0 or 1 carries that
time after reprojection (Reconstruction.hpp:93-121)."""
import numpy as np
import pytest

import edge_scenes


def test_wcross_scene_takes_every_branch(tlr, orc):
    sb = edge_scenes.wcross_scene(tlr)
    cat, crosses = edge_scenes.classify(sb.records)
    counts = np.bincount(cat, minlength=4)
    assert counts[1] > 20 and counts[2] > 20 and counts[3] > 50 and crosses.sum() > 100
    o = orc.Oracle(sb.width, sb.height, sb.spp, sb.coc[0], sb.coc[1], sb.records)
    bi = o.build()
    assert bi.num_kept == int(np.count_nonzero(cat < 3))
    ref, src = o.reprojected()
    assert set(src.tolist()).isdisjoint(np.nonzero(cat == 3)[0].tolist())
    # reprojected t is the chosen centre time (ref columns: x, y, w, t, ...)
    want_t = np.array([0.5, 0.0, 1.0], np.float32)[cat[src]]
    assert np.array_equal(ref[:, 3], want_t)
    assert np.all(ref[:, 2] > 0)
    img, _, st = o.reconstruct(out_spp=4, row_stride=4)
    assert np.isfinite(img).all()


def test_all_discarded_is_an_error(orc):
    rec = edge_scenes.all_discarded_records()
    H, W, spp = rec.shape[:3]
    o = orc.Oracle(W, H, spp, -15.24, 19.05, rec)
    with pytest.raises(RuntimeError, match="\\(3\\)"):
        o.build()


def test_spp_above_256_is_an_error(orc):
    rec = np.zeros((2, 2, 300, 16), np.float32)
    rec[..., 3] = 2.0
    o = orc.Oracle(2, 2, 300, -15.24, 19.05, rec)
    with pytest.raises(RuntimeError, match="\\(4\\)"):
        o.build()
