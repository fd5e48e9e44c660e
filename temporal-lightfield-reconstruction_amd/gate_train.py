"""Training and calibration driver for the learned gate.

The gate is a 24-64-64-1 MLP (BF16 operands, FP32 accumulation on MFMA)
that scores each output sample from features the direct filter computes
(kernel k_query_direct).  Its labels come from the plain method itself: 1
where the output sample has exactly one surface at 1R and no 2R fetch, so
the direct filter (every sample within R, one tent sum) equals the plain
result up to summation order.  Everything here runs on the library's own
kernels: data collection (plain + direct passes), training (forward,
backward and Adam kernels), inference (forward kernel).

Typical use::

    ctx = tlr.Ctx(0)
    params, losses = train_gate(tlr, ctx, scenes)            # on the GPU
    thr, curve = calibrate_threshold(tlr, ctx, val_scene, 1e-3)
    ctx.gate_set_threshold(thr)
    img, st = ctx.reconstruct(out_spp=16, gate=True)
"""
import concurrent.futures as _cf

import numpy as np

GATE_FEATURES = 24


def init_params(seed=0):
    """He-style initialisation of the 5,825 parameters (layout of gate.hpp)."""
    F, H = 24, 64
    rng = np.random.default_rng(seed)
    p = np.zeros(F * H + H + H * H + H + H + 1, np.float32)
    o = 0
    p[o:o + F * H] = rng.normal(0, np.sqrt(2.0 / F), F * H)
    o += F * H + H
    p[o:o + H * H] = rng.normal(0, np.sqrt(2.0 / H), H * H)
    o += H * H + H
    p[o:o + H] = rng.normal(0, np.sqrt(1.0 / H), H)
    return p


def _scenes(tlr, specs, workers=8):
    """Generates sample buffers (host, in parallel) in the order of specs:
    (preset, W, H, spp, seed)."""
    def gen(s):
        return tlr.make_scene(s[0], s[1], s[2], s[3], seed=s[4])
    with _cf.ThreadPoolExecutor(workers) as ex:
        for sb in ex.map(gen, specs):
            yield sb


def train_gate(tlr, ctx, specs, out_spp=16, stride=16, epochs=4, batch=4096, lr=2e-3, pos_weight=1.0,
               seed=0, params=None, log=None):
    """Collects features and labels from every scene of `specs` (stride: keep
    every stride-th output sample) and trains the gate on the GPU.  Returns
    (trained parameters, per-epoch mean loss)."""
    ctx.gate_set_params(init_params(seed) if params is None else params)
    ctx.gate_clear_training()
    for k, sb in enumerate(_scenes(tlr, specs)):
        ctx.set_samples(sb)
        ctx.build()
        n = ctx.gate_append_training(out_spp=out_spp, stride=stride)
        if log:
            log("gate data: scene %d/%d, %d rows" % (k + 1, len(specs), n))
    losses = ctx.gate_train(epochs=epochs, batch=batch, lr=lr, pos_weight=pos_weight, seed=seed)
    if log:
        log("gate training losses: %s" % np.array2string(losses, precision=4))
    return ctx.gate_get_params(), losses


def image_from_samples(cols, W, H, M):
    """Per-pixel combine of per-output-sample colours (the k_combine rule):
    sum in order, normalise by alpha, red where nothing was found."""
    c = cols.reshape(H, W, M, 4).astype(np.float64).sum(axis=2)
    a = c[..., 3:4]
    img = np.where(a != 0, c / np.where(a != 0, a, 1.0), np.array([1.0, 0.0, 0.0, 1.0]))
    return img.astype(np.float32)


def gated_samples(X, logits, direct, plain, thr):
    """Per-output-sample colours of the gated path at threshold thr."""
    acc = (logits >= thr) & (X[:, 0] > 0)
    return np.where(acc[:, None], direct, plain), acc


def calibrate_threshold(tlr, ctx, sbuf, budget, out_spp=16, quantiles=64):
    """Lowest logit threshold whose gated image stays within `budget`
    mean-abs RGB of the plain image on `sbuf` (with half the budget kept in
    reserve).  Returns (threshold, [(thr, accepted fraction, mean-abs)])."""
    ctx.set_samples(sbuf)
    ctx.build()
    X, y, direct, plain = ctx.gate_collect(out_spp=out_spp, stride=1)
    z = ctx.gate_forward(X)
    W, H = sbuf.width, sbuf.height
    ref = image_from_samples(plain, W, H, out_spp)
    finite = np.isfinite(z)
    cands = np.unique(np.quantile(z[finite], np.linspace(0.0, 1.0, quantiles)))
    curve = []
    best = float("inf")
    for thr in cands[::-1]:
        q, acc = gated_samples(X, np.where(finite, z, -np.inf), direct, plain, thr)
        err = float(np.abs(image_from_samples(q, W, H, out_spp) - ref)[..., :3].mean())
        curve.append((float(thr), float(acc.mean()), err))
        if err <= 0.5 * budget:
            best = float(thr)
        else:
            break
    return best, curve
