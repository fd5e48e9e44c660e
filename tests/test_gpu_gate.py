"""GPU: learned-gate MLP kernels (MFMA BF16, FP32 accumulation) against the
CPU reference (oracle/gate_ref.py); the gate's training rows are the 1R lane
kernel's hand-offs with the plain method's colours and correct labels; the
gated path equals the plain one when the gate accepts nothing and the
provisional colours where it accepts everything; a trained gate beats an
untrained one and stays within its error budget on a held-out scene."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import gate_ref  # noqa: E402


@pytest.fixture(scope="module")
def ctx(tlr):
    c = tlr.Ctx(0)
    c.gate_set_params(gate_ref.init_params(1))
    return c


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, gate_ref.F)).astype(np.float32)
    y = (rng.random(n) < 0.6).astype(np.float32)
    return X, y


@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000, 4096])
def test_gate_forward_matches_cpu(ctx, n):
    X, _ = _rows(n, n)
    print("gate forward n=%d" % n, flush=True)
    z = ctx.gate_forward(X)
    zr, _, _ = gate_ref.forward(gate_ref.init_params(1), X)
    err = np.abs(z - zr) / (1 + np.abs(zr))
    assert err.max() < 2e-3, err.max()


@pytest.mark.parametrize("n,pw", [(32, 1.0), (500, 1.0), (4096, 2.5)])
def test_gate_gradient_matches_cpu(ctx, n, pw):
    X, y = _rows(n, 100 + n)
    print("gate grad n=%d" % n, flush=True)
    g, loss = ctx.gate_grad(X, y, pos_weight=pw)
    gr, lr = gate_ref.grad(gate_ref.init_params(1), X, y, pos_weight=pw)
    assert abs(loss - lr) <= 1e-4 * abs(lr) + 1e-3
    for name, a, b in (("W1", gate_ref.OFF_W1, gate_ref.OFF_B1), ("b1", gate_ref.OFF_B1, gate_ref.OFF_W2),
                       ("W2", gate_ref.OFF_W2, gate_ref.OFF_B2), ("b2", gate_ref.OFF_B2, gate_ref.OFF_W3),
                       ("w3", gate_ref.OFF_W3, gate_ref.OFF_B3), ("b3", gate_ref.OFF_B3, gate_ref.NPARAM)):
        rel = np.linalg.norm(g[a:b] - gr[a:b]) / (np.linalg.norm(gr[a:b]) + 1e-12)
        assert rel < 5e-3, (name, rel)


def test_gate_training_reduces_loss(tlr, ctx):
    c = tlr.Ctx(0)
    c.gate_set_params(gate_ref.init_params(2))
    sb = tlr.make_scene(2, 96, 72, 16, seed=5)
    c.set_samples(sb)
    c.build()
    _, st = c.reconstruct(out_spp=8)
    n = c.gate_append_training(out_spp=8, stride=1)
    # one row per hand-off of the first 1R lane pass (the entries the plain
    # path's second lane pass takes)
    assert n == st["wide_queries"] > 0
    losses = c.gate_train(epochs=6, batch=512, lr=2e-3, seed=1)
    print("training losses", losses, flush=True)
    assert losses[-1] < losses[0]


@pytest.fixture(scope="module")
def collected(tlr):
    sb = tlr.make_scene(2, 96, 72, 16, seed=21)
    c = tlr.Ctx(0)
    c.gate_set_params(gate_ref.init_params(0))
    c.set_samples(sb)
    c.build()
    plain, st = c.reconstruct(out_spp=8)
    qplain = c.query_colors(8).reshape(-1, 4).copy()
    X, y, prov, pl, q = c.gate_collect(out_spp=8, stride=1)
    return sb, c, plain, st, qplain, (X, y, prov, pl, q)


def test_collected_rows_are_the_hand_offs(collected):
    """One row per hand-off; the plain colour of each row is the plain
    method's colour of that output sample, bit for bit; the label says
    whether the provisional colour is within 0.001 of it in every channel."""
    sb, c, plain, st, qplain, (X, y, prov, pl, q) = collected
    # the hand-offs of the first 1R lane pass (the plain path's second lane
    # pass takes the same entries)
    assert len(X) == st["wide_queries"] > 0
    assert len(np.unique(q)) == len(q) and q.max() < qplain.shape[0]
    assert np.array_equal(pl.view(np.uint32), qplain[q].view(np.uint32))
    has = X[:, 0] >= 0
    print("hand-offs %d, with a provisional colour %d, labelled 1: %d" % (len(X), has.sum(), (y > 0.5).sum()),
          flush=True)
    assert has.mean() > 0.3 and np.all(np.isfinite(X)) and np.all(np.abs(X) <= 16)
    want = has & (np.abs(prov - pl).max(axis=1) <= 1e-3)
    assert np.array_equal(y > 0.5, want)
    assert np.all(X[~has, 0] == -1) and np.all(X[~has, 1:] == 0)


def test_gate_thresholds_bound_the_gated_path(tlr, collected):
    """Threshold +inf: the gated path is the plain one, bit for bit, and
    every hand-off goes to the general kernel.  Threshold -inf: every
    hand-off with a provisional colour keeps it (the per-output-sample colours
    are the plain ones with those replaced), the rest run the general kernel."""
    sb, c, plain, st, qplain, (X, y, prov, pl, q) = collected
    c.gate_set_threshold(float("inf"))
    g, sg = c.reconstruct(out_spp=8, gate=True)
    assert np.array_equal(g, plain)
    assert sg["gate_candidates"] == st["wide_queries"] == sg["gate_rejected"]
    c.gate_set_threshold(-float("inf"))
    g, sg = c.reconstruct(out_spp=8, gate=True)
    has = X[:, 0] >= 0
    assert sg["gate_rejected"] == int((~has).sum())
    want = qplain.copy()
    want[q[has]] = prov[has]
    got = c.query_colors(8).reshape(-1, 4)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_trained_gate_gains_over_untrained(tlr):
    """What training buys: on a held-out scene, the trained gate lets far more
    hand-offs keep their provisional colour than the untrained network at the
    same image-error budget, and a held-out gated image stays within it."""
    import importlib
    gt = importlib.import_module(tlr.__name__ + ".gate_train")
    ctx = tlr.Ctx(0)
    train = [(2, 240, 136, 16, 700 + k) for k in range(6)]
    params, losses = gt.train_gate(tlr, ctx, train, out_spp=8, stride=1, epochs=8, batch=1024, lr=2e-3, seed=3)
    print("losses", losses, flush=True)
    assert losses[-1] < 0.8 * losses[0]
    val = tlr.make_scene(2, 240, 136, 16, seed=800)
    budget = 1e-3
    thr, curve = gt.calibrate_threshold(tlr, ctx, val, budget, out_spp=8, use=0.5)
    acc_trained = max([a for t, a, e in curve if e <= 0.5 * budget] + [0.0])
    c0 = tlr.Ctx(0)
    c0.gate_set_params(gt.init_params(3))
    thr0, curve0 = gt.calibrate_threshold(tlr, c0, val, budget, out_spp=8, use=0.5)
    acc_untrained = max([a for t, a, e in curve0 if e <= 0.5 * budget] + [0.0])
    print("hand-offs accepted at %.0e: trained %.3f untrained %.3f" % (budget, acc_trained, acc_untrained),
          flush=True)
    assert acc_trained > acc_untrained + 0.1
    # held-out image through the gated pipeline
    test = tlr.make_scene(2, 240, 136, 16, seed=900)
    ctx.set_samples(test)
    ctx.build()
    plain, _ = ctx.reconstruct(out_spp=8)
    ctx.gate_set_threshold(thr)
    gated, st = ctx.reconstruct(out_spp=8, gate=True)
    err = np.abs(gated - plain)[..., :3].mean()
    print("held-out: hand-offs %d, rejected %d, mean-abs %.2e" % (st["gate_candidates"], st["gate_rejected"], err),
          flush=True)
    assert st["gate_rejected"] < st["gate_candidates"]
    assert err <= budget
