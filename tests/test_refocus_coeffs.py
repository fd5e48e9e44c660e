"""CPU: the refocus coefficients (CameraParams::getCocCoeffs,
CameraParams.hpp:97-118, as the facades compute them for
overrideRefocusDistance) against their closed form.  With D = getDofMatrix of
lens point (0.5, 0), projection * D - projection has m02 = P00 * 0.5 A / f and
m03 = P00 * 0.5 A, so (C0, C1) = (-0.5 A P00 W, 0.5 A P00 W / f): the CoC
vanishes at the focus distance (C0 / f + C1 = 0) and grows with the aperture."""
import numpy as np
import pytest


@pytest.mark.parametrize("A,f,W", [(0.2, 4.0, 96), (0.05, 1.5, 1920), (1.0, 10.0, 640)])
def test_coc_coeffs_closed_form(tlr, A, f, W):
    proj = np.array([[1.5, 0.1, 0.2, 0], [0, 2.0, 0.1, 0], [0, 0, -1.002, -0.2002], [0, 0, -1, 0]], np.float32)
    cc = tlr.coc_coeffs(proj, W, A, f)
    P00 = float(proj[0, 0])
    want = np.array([-0.5 * A * P00 * W, 0.5 * A * P00 * W / f])
    assert np.allclose(cc, want, rtol=2e-6, atol=1e-6)
    assert abs(cc[0] / np.float32(f) + cc[1]) <= 1e-5 * abs(cc[1])
