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
