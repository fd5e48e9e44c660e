"""Sample buffers for the reference's edge cases, built on the host from a
seeded synthetic scene (no reference data exists for them).

The reprojection to the (u,v,t) centre tries t' = 0.5, then 0, then 1, and
keeps the first with P0.w > 0 (Reconstruction.hpp:93-121); a sample with no
such t' is discarded (Reconstruction.cpp:561-587).  The tree bounds clip a
sample's time span where w(t) crosses 0 inside the shutter
(TimeBounds.cpp:247-260).  `wcross_scene` gives part of the image motion in w
strong enough that every one of those branches is taken by some sample, and
discards a few samples outright.
"""
import numpy as np


def classify(rec):
    """Per record, which centre time the reprojection picks (float64 model of
    P0.w = w - t*mv_w + t'*mv_w): 0 -> 0.5, 1 -> 0, 2 -> 1, 3 -> discarded;
    and whether w(t) crosses 0 inside [0, 1] at the chosen t'."""
    r = rec.reshape(-1, 16).astype(np.float64)
    w, t, mvw = r[:, 3], r[:, 6], r[:, 13]
    base = w - t * mvw
    cat = np.full(len(r), 3, np.int32)
    for k, tp in reversed(list(enumerate((0.5, 0.0, 1.0)))):
        cat[base + tp * mvw > 0] = k
    w0, w1 = base, base + mvw
    crosses = (cat < 3) & ((w0 <= 0) | (w1 <= 0))
    return cat, crosses


def wcross_scene(tlr, W=64, H=48, spp=8, seed=41, block=(20, 14, 10, 8)):
    """A 960x540-type scene whose pixels in `block` (x0, y0, width, height)
    move in w at up to 4x their depth per shutter, so samples take each
    reprojection branch and some cross w = 0 inside the shutter; 1% of the
    samples sit behind the camera with no motion (discarded at every t').
    Samples whose chosen P0.w would fall below 0.2 w keep their original
    motion, so reprojected positions stay within a few image widths."""
    sb = tlr.make_scene(2, W, H, spp, seed=seed)
    rec = sb.records.reshape(H, W, spp, 16).copy()
    rng = np.random.default_rng(seed)
    sel = np.zeros((H, W, spp), bool)
    x0, y0, bw, bh = block
    sel[y0:y0 + bh, x0:x0 + bw] = True
    k = rng.uniform(2.5, 4.0, size=sel.shape) * rng.choice([-1.0, 1.0], size=sel.shape)
    w = rec[..., 3].astype(np.float64)
    t = rec[..., 6].astype(np.float64)
    mvw_new = (k * w).astype(np.float32)
    base = w - t * mvw_new
    chosen = np.full(sel.shape, np.nan)
    for tp in (1.0, 0.0, 0.5):  # the first valid in (0.5, 0, 1) wins
        v = base + tp * mvw_new
        chosen = np.where(v > 0, v, chosen)
    ok = np.isfinite(chosen) & (chosen >= 0.2 * w)
    take = sel & ok
    rec[..., 13] = np.where(take, mvw_new, rec[..., 13])
    # discarded outright: behind the camera, no motion
    behind = rng.random(sel.shape) < 0.01
    rec[..., 3] = np.where(behind, -np.abs(rec[..., 3]), rec[..., 3])
    for f in (11, 12, 13):
        rec[..., f] = np.where(behind, 0.0, rec[..., f])
    return tlr.UVTSampleBuffer(W, H, spp, sb.coc[0], sb.coc[1], rec)


def all_discarded_records(W=8, H=6, spp=2):
    rec = np.zeros((H, W, spp, 16), np.float32)
    rec[..., 0] = np.arange(W)[None, :, None] + 0.5
    rec[..., 1] = np.arange(H)[:, None, None] + 0.5
    rec[..., 3] = -1.0
    rec[..., 7:11] = 0.5
    return rec
