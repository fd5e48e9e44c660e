"""CPU pins of the oracle's shadow reconstruction (TreeGather::reconstructShadows
and reconstructDofMotionShadows, Reconstruction.cpp:302-433; the filter's
shadow branches, ReconstructionOur.cpp:78-80, 283-291, 325-347), each against
an outcome fixed by hand, not a retyped formula:

- the per-sample light test (o.w - 0.01 <= sample w at t + dot(o.xy -
  sample xy, sample w-gradient)) on a constructed light-space plane: a query
  on the plane is lit (the plain tent mean), one 0.02 behind it is black, one
  0.009 behind it still lit, and with a w-gradient exactly the samples on one
  side of the query light it;
- the dispersion scaled by sqrt(density): density 4 gives the tent mean at
  twice the radius;
- the nearest-sample fallback takes the light test too;
- the extended result is the nearest sample's position, w, w-gradient and
  density;
- FW's matrix inverse against float64, lightSpaceDot against the geometric
  cosine between a surface's normal (rebuilt from its w-gradients) and the
  direction to the light, and the whole per-pixel shading of
  reconstructDofMotionShadows on a pinhole scene lit from the camera, with
  and without an occluder in the light's buffer.

tests/test_oracle_mutants.py breaks the bias sign, the density scaling and
the fallback's light test and checks that these pins catch each."""
import numpy as np
import pytest

import pin_scenes as ps


def plane_colour(x, y):
    """Not linear in position, so tent means at different radii differ."""
    return (0.5 + 0.5 * np.sin(3.0 * x), 0.5 + 0.5 * np.cos(2.3 * y), 0.25)


def plane(x, y, px, py, k):
    return (5.0, plane_colour(x, y))


_built = {}


def _oracle(orc, name, fn, wg=None, density=None):
    if name not in _built:
        rec = ps.records(fn)
        if wg is not None:
            rec[..., 14] = wg[0]
            rec[..., 15] = wg[1]
        o = orc.Oracle(ps.W, ps.H, ps.SPP, ps.C0, ps.C1, rec)
        if density is not None:
            o.set_density(density(rec))
        o.build()
        _built[name] = (rec, o)
    return _built[name]


def _q(x, y, w, density=1.0):
    return np.array([[x, y, 0.0, 0.0, 0.5, w, density, 0.0]], np.float32)


O = (12.5, 10.5)


def test_query_on_the_plane_is_lit(orc):
    rec, o = _oracle(orc, "plane", plane)
    col, ext, _ = o.reconstruct_queries(_q(*O, 5.0), shadow=True)
    assert np.allclose(col[0, :3], ps.tent_mean(rec, O), atol=1e-5) and col[0, 3] == 1.0


def test_query_behind_the_plane_is_black(orc):
    rec, o = _oracle(orc, "plane", plane)
    col, _, _ = o.reconstruct_queries(_q(*O, 5.02), shadow=True)
    assert np.array_equal(col[0], np.array([0, 0, 0, 1], np.float32))


def test_bias_keeps_a_query_just_behind_lit(orc):
    """0.009 behind the plane: o.w - 0.01 is still in front of it."""
    rec, o = _oracle(orc, "plane", plane)
    col, _, _ = o.reconstruct_queries(_q(*O, 5.009), shadow=True)
    assert np.allclose(col[0, :3], ps.tent_mean(rec, O), atol=1e-5)


def test_w_gradient_lights_one_side(orc):
    """Every sample with w-gradient (0.2, 0) and the query at w = 5.01: a
    sample lights it exactly when o.x >= its x (5.0 <= 5 + 0.2 dx); the tent
    weights of the others still count in alpha."""
    rec, o = _oracle(orc, "plane_wg", plane, wg=(0.2, 0.0))
    col, _, _ = o.reconstruct_queries(_q(*O, 5.01), shadow=True)
    r = rec.reshape(-1, 16).astype(np.float64)
    d = np.hypot(r[:, 0] - O[0], r[:, 1] - O[1])
    sel = d < ps.R
    wgt = 1.0 - d[sel] / ps.R
    lit = (O[0] - r[sel, 0]) >= 0
    want = (wgt[:, None] * r[sel, 7:10] * lit[:, None]).sum(0) / wgt.sum()
    assert 0 < lit.sum() < lit.size
    assert np.allclose(col[0, :3], want, atol=1e-5) and abs(col[0, 3] - 1.0) < 1e-6


def test_density_scales_the_dispersion(orc):
    """Density 4: radius sqrt(4) * 0.5 = 1.0, so the tent mean over 1 px."""
    rec, o = _oracle(orc, "plane", plane)
    col, _, _ = o.reconstruct_queries(_q(*O, 5.0, density=4.0), shadow=True)
    want = ps.tent_mean(rec, O, radius=1.0)
    assert not np.allclose(want, ps.tent_mean(rec, O), atol=1e-4)
    assert np.allclose(col[0, :3], want, atol=1e-5)


def test_nearest_fallback_takes_the_light_test(orc):
    """No sample within 1.6 px of the hole's centre: the colour is the
    nearest sample's (ReconstructionOur.cpp:336-347) when the query is lit,
    black when it lies behind the plane."""
    rec, o = _oracle(orc, "hole", ps.hole)
    want, _ = ps.nearest_colour(rec, ps.HOLE_CENTRE)
    col, _, _ = o.reconstruct_queries(_q(*ps.HOLE_CENTRE, 5.0), shadow=True)
    assert np.allclose(col[0, :3], want, atol=1e-6)
    col, _, _ = o.reconstruct_queries(_q(*ps.HOLE_CENTRE, 5.02), shadow=True)
    assert np.array_equal(col[0], np.array([0, 0, 0, 1], np.float32))


def test_extended_result_is_the_nearest_sample(orc):
    def dens(rec):
        return (0.5 + (np.arange(rec[..., 0].size) % 7) / 4.0).astype(np.float32)

    rec, o = _oracle(orc, "plane_dens", plane, wg=(0.05, -0.03), density=dens)
    _, ext, _ = o.reconstruct_queries(_q(*O, 5.0), shadow=True)
    r = rec.reshape(-1, 16)
    d = np.hypot(r[:, 0].astype(np.float64) - O[0], r[:, 1] - O[1])
    k = int(np.argmin(d))
    assert np.allclose(ext[0, :2], r[k, :2], atol=1e-6)
    assert ext[0, 2] == np.float32(5.0)
    assert np.allclose(ext[0, 4:6], [0.05, -0.03])
    assert ext[0, 3] == np.float32(0.5 + (k % 7) / 4.0)


# ----------------------------------------------------------------- matrices
def test_fw_inverse_matches_float64(orc):
    rng = np.random.default_rng(0)
    for _ in range(50):
        m = rng.normal(size=(4, 4)) + 3 * np.eye(4)
        inv = orc.mat4_inverse(m.astype(np.float32))
        want = np.linalg.inv(m.astype(np.float32).astype(np.float64))
        assert np.allclose(inv, want, rtol=1e-4, atol=1e-5)


def perspective(fovy, aspect, near, far):
    f = 1.0 / np.tan(fovy / 2)
    return np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                     [0, 0, (far + near) / (near - far), 2 * far * near / (near - far)], [0, 0, -1, 0]])


def look(eye, target, up=(0, 1, 0)):
    """World-to-camera (camera looks down -z)."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    zc = eye - target
    zc /= np.linalg.norm(zc)
    xc = np.cross(up, zc)
    xc /= np.linalg.norm(xc)
    yc = np.cross(zc, xc)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = xc, yc, zc
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def test_light_space_dot_is_the_cosine_to_the_light(orc):
    """A tilted plane seen by a camera at the origin, lit by a point light
    elsewhere: the normal rebuilt from the plane's w-gradients (finite
    differences of w over window pixels) and the light-space dot give the
    geometric normal and cos(normal, direction to the light)."""
    W, H = 64, 48
    P = perspective(np.radians(60), W / H, 0.1, 100.0)
    C2W = np.eye(4)
    W2L = look((3.0, 2.0, 1.0), (0.0, 0.0, -6.0))
    PL = perspective(np.radians(50), W / H, 0.1, 100.0)
    sp = orc.shadow_setup(P, C2W, PL, W2L, W, H, (0, 0), (0, 0), 100.0)
    n_world = np.array([0.3, 0.2, 1.0])
    n_world /= np.linalg.norm(n_world)
    p0 = np.array([0.0, 0.0, -6.0])

    def w_at(xw, yw):
        ndc = np.array([2 * xw / W - 1, 2 * yw / H - 1])
        d = np.array([ndc[0] / P[0, 0], ndc[1] / P[1, 1], -1.0])    # camera ray at w = 1
        t = n_world @ p0 / (n_world @ d)
        return t, t * d                                               # w, camera-space point

    light_pos = np.linalg.inv(W2L)[:3, 3]
    for xw, yw in [(32.0, 24.0), (20.5, 30.25), (45.0, 10.0)]:
        w, pt = w_at(xw, yw)
        wg = ((w_at(xw + 0.5, yw)[0] - w_at(xw - 0.5, yw)[0]), (w_at(xw, yw + 0.5)[0] - w_at(xw, yw - 0.5)[0]))
        d, n = orc.light_space_dot([xw * w, yw * w, 1.0, w], wg, sp)
        assert np.allclose(n, n_world, atol=2e-3), (n, n_world)
        to_light = (light_pos - pt) / np.linalg.norm(light_pos - pt)
        assert abs(d - to_light @ n_world) < 2e-3


def _pinhole_scene(W, H, spp, w_of):
    """Records of a plane seen through a pinhole (CoC 0): a 4x4 grid per pixel
    in window coordinates, w from w_of(x, y), one colour."""
    g = int(np.sqrt(spp))
    ys, xs, k = np.meshgrid(np.arange(H), np.arange(W), np.arange(spp), indexing="ij")
    x = xs + ((k % g) + 0.5) / g + 0.001 * ((k * 7) % 5)
    y = ys + ((k // g) + 0.5) / g + 0.001 * ((k * 3) % 5)
    rec = np.zeros((H, W, spp, 16), np.float32)
    rec[..., 0], rec[..., 1] = x, y
    rec[..., 3] = w_of(x, y)
    rec[..., 6] = 0.5
    rec[..., 7:10] = (0.6, 0.4, 0.2)
    rec[..., 10] = 1.0
    return rec


def test_dof_motion_shadows_lit_from_the_camera(orc):
    """Camera and light coincide; a fronto-parallel plane at w = 5 in both
    buffers: every shadow ray is lit, the normal faces the camera (world y =
    0, ambient 0.5), so a pixel is colour * (1.5 cos + 0.25) with cos =
    5 / |P| for the camera-space point P.  With an occluder at w = 3 over the
    left half of the light's buffer, the left pixels keep the ambient term
    only: colour * 0.25."""
    W, H, spp = 32, 24, 16
    P = perspective(np.radians(60), W / H, 0.1, 100.0)
    sp = orc.shadow_setup(P, np.eye(4), P, np.eye(4), W, H, (0, 0), (0, 0), 100.0)
    cam_rec = _pinhole_scene(W, H, spp, lambda x, y: np.full_like(x, 5.0))
    occ_rec = _pinhole_scene(W, H, spp, lambda x, y: np.where(x < 16, 3.0, 5.0))
    cam = orc.Oracle(W, H, spp, 0.0, 0.0, cam_rec)
    cam.build()
    c = np.array([0.6, 0.4, 0.2])
    for light_rec, shadowed in ((cam_rec, False), (occ_rec, True)):
        light = orc.Oracle(W, H, spp, 0.0, 0.0, light_rec)
        light.build()
        img = cam.reconstruct_dof_motion_shadows(light, sp, out_spp=16)
        for px, py in [(24, 12), (20, 8), (26, 16)] + ([(6, 12), (10, 8)] if shadowed else [(8, 12)]):
            X = (2 * (px + 0.5) / W - 1) * 5 / P[0, 0]
            Y = (2 * (py + 0.5) / H - 1) * 5 / P[1, 1]
            cosv = 5 / np.sqrt(X * X + Y * Y + 25)
            f = 0.25 if (shadowed and px < 16) else 1.5 * cosv + 0.25
            assert np.allclose(img[py, px, :3], c * f, atol=2e-3), (shadowed, px, py, img[py, px], c * f)
            assert abs(img[py, px, 3] - 1.0) < 1e-6


# -------------------------------------------------------- irregular buffers
@pytest.mark.parametrize("keep,spp_pow2", [(1.0, 16), (0.55, 8), (0.3, 4), (0.05, 1)])
def test_irregular_buffer_spp_is_the_average(orc, keep, spp_pow2):
    """An irregular buffer (spp 0, Reconstruction.cpp:65-71) takes the
    average samples per pixel, integer division, rounded up to a power of
    two: 16 -> 16, 8.8 -> 8, 4.8 -> 4, 0.8 -> 0 -> 1."""
    rec = ps.records(plane).reshape(-1, 16)
    n = int(round(keep * rec.shape[0]))
    sub = rec[np.random.default_rng(0).permutation(rec.shape[0])[:n]]
    o = orc.Oracle.irregular(ps.W, ps.H, ps.C0, ps.C1, sub)
    info = o.build()
    assert info.spp_pow2 == spp_pow2
    assert info.num_input == n
