"""GPU: learned-gate MLP kernels (MFMA BF16, FP32 accumulation) against the
CPU reference (oracle/gate_ref.py); direct filter equals the plain method on
single-surface output samples; gated reconstruction stays within its error
budget of the plain output."""
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
    # separable toy problem: label = feature 0 > 0
    rng = np.random.default_rng(3)
    c = tlr.Ctx(0)
    c.gate_set_params(gate_ref.init_params(2))
    sb = tlr.make_scene(2, 64, 48, 8, seed=5)
    c.set_samples(sb)
    c.build()
    n = c.gate_append_training(out_spp=8, stride=1)
    assert n == 64 * 48 * 8
    losses = c.gate_train(epochs=4, batch=2048, lr=2e-3, seed=1)
    print("training losses", losses, flush=True)
    assert losses[-1] < losses[0]


def test_direct_filter_equals_plain_on_single_surface(tlr):
    sb = tlr.make_scene(2, 96, 72, 16, seed=21)
    c = tlr.Ctx(0)
    c.gate_set_params(gate_ref.init_params(0))
    c.set_samples(sb)
    c.build()
    X, y, direct, plain = c.gate_collect(out_spp=8, stride=1)
    single = y > 0.5
    print("single-surface fraction %.3f" % single.mean(), flush=True)
    assert single.mean() > 0.3
    # same samples, same weights, only the summation order differs
    assert np.abs(direct[single] - plain[single]).max() < 1e-5
    assert np.all(np.isfinite(X))


def test_gated_reconstruction_error_budget(tlr):
    """With a threshold that accepts nothing the gated path equals the plain
    one; accepting everything gives the direct filter."""
    sb = tlr.make_scene(2, 96, 72, 16, seed=22)
    c = tlr.Ctx(0)
    c.gate_set_params(gate_ref.init_params(0))
    c.set_samples(sb)
    c.build()
    plain, _ = c.reconstruct(out_spp=8)
    c.gate_set_threshold(1e30)
    g, st = c.reconstruct(out_spp=8, gate=True, collect_stats=True)
    assert np.array_equal(g, plain)
    assert st["gate_rejected"] == 96 * 72 * 8


def test_trained_gate_gains_over_untrained(tlr):
    """What training buys: on a held-out scene, the trained gate routes far
    more output samples to the direct filter than the untrained network at
    the same image-error budget, and the gated image stays within it."""
    import importlib
    gt = importlib.import_module(tlr.__name__ + ".gate_train")
    ctx = tlr.Ctx(0)
    train = [(2, 240, 136, 16, 700 + k) for k in range(4)]
    params, losses = gt.train_gate(tlr, ctx, train, out_spp=8, stride=2, epochs=6, batch=2048, lr=2e-3, seed=3)
    print("losses", losses, flush=True)
    assert losses[-1] < 0.5 * losses[0]
    val = tlr.make_scene(2, 240, 136, 16, seed=800)
    budget = 1e-3
    # calibrated at half the budget: the held-out image must stay within it
    thr, curve = gt.calibrate_threshold(tlr, ctx, val, 0.5 * budget, out_spp=8)
    acc_trained = max([a for t, a, e in curve if e <= 0.25 * budget] + [0.0])
    c0 = tlr.Ctx(0)
    c0.gate_set_params(gt.init_params(3))
    thr0, curve0 = gt.calibrate_threshold(tlr, c0, val, 0.5 * budget, out_spp=8)
    acc_untrained = max([a for t, a, e in curve0 if e <= 0.25 * budget] + [0.0])
    print("accepted at %.0e: trained %.3f untrained %.3f" % (budget, acc_trained, acc_untrained), flush=True)
    assert acc_trained > acc_untrained + 0.2
    # held-out image through the gated pipeline
    test = tlr.make_scene(2, 240, 136, 16, seed=900)
    ctx.set_samples(test)
    ctx.build()
    plain, _ = ctx.reconstruct(out_spp=8)
    ctx.gate_set_threshold(thr)
    gated, st = ctx.reconstruct(out_spp=8, gate=True)
    err = np.abs(gated - plain)[..., :3].mean()
    print("held-out: rejected %.3f, mean-abs %.2e" % (st["gate_rejected"] / (240 * 136 * 8), err), flush=True)
    assert st["gate_rejected"] < 0.8 * 240 * 136 * 8
    assert err <= budget
