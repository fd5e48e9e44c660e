"""Training and calibration driver for the learned gate.

The plain method's 1R lane kernel finishes most output samples and hands the
rest to the general kernel (2R fetch, several surfaces, coverage test).  The
gate is a 24-64-64-1 MLP (BF16 operands, FP32 accumulation on MFMA) that
scores each hand-off from features the lane kernel computes on the way
(kernel k_query_fast_gate); a hand-off it accepts keeps the lane kernel's
provisional colour (its first 1R surface, tent-filtered) and skips the
general kernel.  Labels come from the plain method itself: 1 where the
provisional colour is within 0.001 of the plain method's final colour in every
channel.  Everything runs on the library's own kernels: data collection,
training (forward, backward and Adam kernels), inference (forward kernel).

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


def train_gate(tlr, ctx, specs, out_spp=16, stride=1, epochs=4, batch=4096, lr=2e-3, pos_weight=1.0,
               seed=0, params=None, log=None):
    """Collects features and labels of the hand-offs of every scene of
    `specs` (stride: keep every stride-th hand-off) and trains the gate on
    the GPU.  Returns (trained parameters, per-epoch mean loss)."""
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
    c = cols.reshape(H * W, M, 4).astype(np.float64).sum(axis=1)
    a = c[:, 3:4]
    img = np.where(a != 0, c / np.where(a != 0, a, 1.0), np.array([1.0, 0.0, 0.0, 1.0]))
    return img.reshape(H, W, 4)


class GateErrorModel:
    """Image error of gating one build at any threshold, without running it:
    the gated image differs from the plain one only where accepted hand-offs
    replace their plain colour by the provisional one."""

    def __init__(self, ctx, W, H, M):
        X, y, prov, plain, q = ctx.gate_collect(out_spp=M, stride=1)
        self.X, self.y, self.prov, self.plain, self.q = X, y, prov, plain, q.astype(np.int64)
        self.z = ctx.gate_forward(X) if len(X) else np.zeros(0, np.float32)
        cols = ctx.query_colors(M)  # plain per-output-sample colours (the collect ran the plain method)
        self.S = cols.reshape(H * W, M, 4).astype(np.float64).sum(axis=1)
        self.W, self.H, self.M = W, H, M
        self.nq = W * H * M
        self.has = X[:, 0] >= 0 if len(X) else np.zeros(0, bool)

    def _img(self, S):
        a = S[:, 3:4]
        return np.where(a != 0, S / np.where(a != 0, a, 1.0), np.array([1.0, 0.0, 0.0, 1.0]))

    def evaluate(self, thr):
        """(accepted hand-offs, their fraction of all hand-offs, mean-abs RGB
        of the gated image against the plain one)."""
        acc = self.has & (self.z >= thr)
        n = int(acc.sum())
        if n == 0:
            return 0, 0.0, 0.0
        pix = self.q[acc] // self.M
        D = np.zeros((self.H * self.W, 4), np.float64)
        np.add.at(D, pix, (self.prov[acc] - self.plain[acc]).astype(np.float64))
        touched = np.unique(pix)
        err = np.abs(self._img(self.S[touched] + D[touched]) - self._img(self.S[touched]))[:, :3].sum()
        return n, n / max(len(self.X), 1), float(err / (self.H * self.W * 3))


def calibrate_threshold(tlr, ctx, sbuf, budget, out_spp=16, quantiles=64, use=0.6):
    """Lowest logit threshold whose gated image stays within `use` of
    `budget` mean-abs RGB of the plain image on `sbuf` (the rest kept in
    reserve for unseen scenes).  Returns (threshold, [(thr, accepted fraction
    of hand-offs, mean-abs)])."""
    ctx.set_samples(sbuf)
    ctx.build()
    m = GateErrorModel(ctx, sbuf.width, sbuf.height, out_spp)
    z = m.z[m.has]
    z = z[np.isfinite(z)]
    if z.size == 0:
        return float("inf"), []
    cands = np.unique(np.quantile(z, np.linspace(0.0, 1.0, quantiles)))
    curve = []
    best = float("inf")
    for thr in cands[::-1]:
        _, frac, err = m.evaluate(thr)
        curve.append((float(thr), float(frac), err))
        if err <= use * budget:
            best = float(thr)
        else:
            break
    return best, curve
