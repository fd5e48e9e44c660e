"""CPU pin of the gate reference: with BF16 rounding switched off, its
analytic gradient equals central finite differences in float64."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import gate_ref  # noqa: E402


def test_gradient_matches_finite_differences(monkeypatch):
    monkeypatch.setattr(gate_ref, "bf16", lambda x: np.asarray(x, np.float64))
    rng = np.random.default_rng(0)
    p = gate_ref.init_params(3).astype(np.float64)
    X = rng.normal(0, 1, (40, gate_ref.F))
    y = (rng.random(40) < 0.5).astype(np.float64)

    def loss(q):
        # float64 forward without the kernel's float32 casts
        W1, b1, W2, b2, w3, b3 = (q[gate_ref.OFF_W1:gate_ref.OFF_B1].reshape(24, 64), q[gate_ref.OFF_B1:gate_ref.OFF_W2],
                                  q[gate_ref.OFF_W2:gate_ref.OFF_B2].reshape(64, 64), q[gate_ref.OFF_B2:gate_ref.OFF_W3],
                                  q[gate_ref.OFF_W3:gate_ref.OFF_B3], q[gate_ref.OFF_B3])
        h1 = np.maximum(X @ W1 + b1, 0)
        h2 = np.maximum(h1 @ W2 + b2, 0)
        z = h2 @ w3 + b3
        w = np.where(y > 0.5, 1.7, 1.0)
        return np.sum(w * (np.logaddexp(0, z) - y * z))

    # analytic gradient of the reference (float32 casts inside are exact enough
    # for a 1e-4 relative check)
    g, _ = gate_ref.grad(p.astype(np.float32), X.astype(np.float32), y, pos_weight=1.7)
    idx = rng.choice(gate_ref.NPARAM, 60, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-4
        fd = (loss(p + e) - loss(p - e)) / 2e-4
        assert abs(fd - g[i]) <= 2e-3 * (1 + abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5], np.float32)
    r = gate_ref.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0 + 2 ** -7                # representable
    assert r[3] == 1.0 + 2 ** -6                # tie to even (up)
    assert r[4] == -2.5
