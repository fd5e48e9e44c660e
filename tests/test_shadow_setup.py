"""CPU: the library's host-side shadow set-up (tlr_shadow_setup: the
matrices of Reconstruction.cpp:365-388 in FW's float arithmetic) equals the
oracle's restatement bit for bit, for random cameras and lights."""
import numpy as np
import pytest


def perspective(fovy, aspect, near, far):
    f = 1.0 / np.tan(fovy / 2)
    return np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                     [0, 0, (far + near) / (near - far), 2 * far * near / (near - far)], [0, 0, -1, 0]])


@pytest.mark.parametrize("seed", range(8))
def test_shadow_setup_matches_oracle(tlr, orc, seed):
    rng = np.random.default_rng(seed)
    P = perspective(rng.uniform(0.6, 1.4), rng.uniform(1.0, 2.0), rng.uniform(0.05, 1.0), rng.uniform(20, 500))
    PL = perspective(rng.uniform(0.6, 1.4), rng.uniform(1.0, 2.0), rng.uniform(0.05, 1.0), rng.uniform(20, 500))
    C2W = np.eye(4)
    C2W[:3, :3] = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    C2W[:3, 3] = rng.normal(size=3) * 5
    W2L = np.eye(4)
    W2L[:3, :3] = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    W2L[:3, 3] = rng.normal(size=3) * 5
    w, h = int(rng.integers(16, 2000)), int(rng.integers(16, 1200))
    cc, lc = rng.normal(size=2) * 10, rng.normal(size=2) * 10
    a = tlr.shadow_setup(P, C2W, PL, W2L, w, h, cc, lc, 100.0)
    b = orc.shadow_setup(P, C2W, PL, W2L, w, h, cc, lc, 100.0)
    assert bytes(a) == bytes(b)
