"""GPU: shadow reconstruction against the oracle, bit for bit.

- explicit shadow queries (tlr_reconstruct_queries with shadow, as
  TreeGather::reconstructShadows drives the filter): colours and extended
  results (nearest sample's xy, w, density, w-gradient) on a buffer with
  random w-gradients and densities, queries at random positions, lens
  positions, times, w around the scene's and densities 0.5-4;
- the same queries through the plain filter, with extended results;
- reconstructDofMotionShadows (camera pass, shadow rays, light pass,
  shading) on two seeded scenes with a light turned a few degrees from the
  camera, and on the pinhole scene lit from the camera whose image the
  oracle pins in closed form;
- error cases."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _pair(tlr, orc, sb, seed):
    rng = np.random.default_rng(seed)
    rec = sb.records
    rec[:, 14:16] = rng.normal(0, 0.05, (rec.shape[0], 2)).astype(np.float32)
    dens = rng.uniform(0.5, 2.0, rec.shape[0]).astype(np.float32)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.set_shadow_inputs(rec[:, 14:16], dens)
    ctx.build()
    o = orc.Oracle(sb.width, sb.height, sb.spp, sb.coc[0], sb.coc[1], rec)
    o.set_density(dens)
    o.build()
    return ctx, o, rec


def _queries(rng, sb, rec, n):
    q = np.zeros((n, 8), np.float32)
    q[:, 0] = rng.uniform(2, sb.width - 2, n)
    q[:, 1] = rng.uniform(2, sb.height - 2, n)
    q[:, 2:4] = rng.uniform(-1, 1, (n, 2))
    q[:, 4] = rng.uniform(0, 1, n)
    w = rec[rng.integers(0, rec.shape[0], n), 3]
    q[:, 5] = w * rng.choice([0.98, 1.0, 1.0005, 1.02, 1.2], n)
    q[:, 6] = rng.choice([0.5, 1.0, 2.0, 4.0], n)
    return q


@pytest.mark.parametrize("shadow", [True, False])
def test_queries_match_oracle(tlr, orc, shadow):
    sb = tlr.make_scene(2, 96, 72, 16, seed=41)
    ctx, o, rec = _pair(tlr, orc, sb, 5)
    q = _queries(np.random.default_rng(7), sb, rec, 20000)
    print("queries, shadow=%s" % shadow, flush=True)
    col, ext, st = ctx.reconstruct_queries(q, shadow=shadow)
    ocol, oext, ost = o.reconstruct_queries(q, shadow=shadow)
    assert np.array_equal(_bits(col), _bits(ocol))
    assert np.array_equal(_bits(ext[:, :6]), _bits(oext[:, :6]))
    for k in ("leaf_nodes", "samples_in_leaf_nodes", "samples_fetched", "surfaces", "num_2r"):
        assert st[k] == ost[k], k
    if shadow:
        black = np.all(col[:, :3] == 0, axis=1) & (col[:, 3] > 0)
        print("black %d of %d" % (black.sum(), len(q)), flush=True)
        assert 0 < black.sum() < len(q)


def _persp(fovy, aspect, near, far):
    f = 1.0 / np.tan(fovy / 2)
    return np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                     [0, 0, (far + near) / (near - far), 2 * far * near / (near - far)], [0, 0, -1, 0]])


def test_dof_motion_shadows_matches_oracle(tlr, orc):
    W, H = 64, 48
    cam_sb = tlr.make_scene(2, W, H, 16, seed=51)
    light_sb = tlr.make_scene(2, W, H, 16, seed=52)
    cam, ocam, _ = _pair(tlr, orc, cam_sb, 8)
    light, olight, _ = _pair(tlr, orc, light_sb, 9)
    P = _persp(np.radians(60), W / H, 0.1, 100.0)
    a = np.radians(4.0)
    W2L = np.eye(4)
    W2L[:3, :3] = [[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]]
    W2L[:3, 3] = [0.3, -0.1, 0.0]
    sp = tlr.shadow_setup(P, np.eye(4), P, W2L, W, H, cam_sb.coc, light_sb.coc, 100.0)
    osp = orc.shadow_setup(P, np.eye(4), P, W2L, W, H, cam_sb.coc, light_sb.coc, 100.0)
    assert bytes(sp) == bytes(osp)
    print("dof motion shadows %dx%d" % (W, H), flush=True)
    img, st = cam.reconstruct_dof_motion_shadows(light, sp, out_spp=8)
    ref = ocam.reconstruct_dof_motion_shadows(olight, osp, out_spp=8)
    print("light queries %d of %d" % (st["light_queries"], W * H * 8), flush=True)
    assert st["light_queries"] > 0.3 * W * H * 8
    assert np.array_equal(_bits(img), _bits(ref))


def test_dof_motion_shadows_pinhole_scene(tlr, orc):
    """The closed-form scene of the oracle's pins (camera = light, plane at
    w = 5, an occluder at w = 3 over the left half of the light's buffer)."""
    W, H, spp = 32, 24, 16
    g = 4
    ys, xs, k = np.meshgrid(np.arange(H), np.arange(W), np.arange(spp), indexing="ij")
    x = xs + ((k % g) + 0.5) / g + 0.001 * ((k * 7) % 5)
    y = ys + ((k // g) + 0.5) / g + 0.001 * ((k * 3) % 5)

    def records(w):
        rec = np.zeros((H, W, spp, 16), np.float32)
        rec[..., 0], rec[..., 1], rec[..., 3] = x, y, w
        rec[..., 6] = 0.5
        rec[..., 7:10] = (0.6, 0.4, 0.2)
        rec[..., 10] = 1.0
        return rec.reshape(-1, 16)

    P = _persp(np.radians(60), W / H, 0.1, 100.0)
    sp = tlr.shadow_setup(P, np.eye(4), P, np.eye(4), W, H, (0, 0), (0, 0), 100.0)
    osp = orc.shadow_setup(P, np.eye(4), P, np.eye(4), W, H, (0, 0), (0, 0), 100.0)
    cam_rec = records(np.full(x.shape, 5.0))
    light_rec = records(np.where(x < 16, 3.0, 5.0))
    ctxs, oracles = [], []
    for rec in (cam_rec, light_rec):
        sb = tlr.UVTSampleBuffer(W, H, spp, 0.0, 0.0, rec)
        c = tlr.Ctx(0)
        c.set_samples(sb)
        c.build()
        o = orc.Oracle(W, H, spp, 0.0, 0.0, rec)
        o.build()
        ctxs.append(c)
        oracles.append(o)
    img, _ = ctxs[0].reconstruct_dof_motion_shadows(ctxs[1], sp, out_spp=16)
    ref = oracles[0].reconstruct_dof_motion_shadows(oracles[1], osp, out_spp=16)
    assert np.array_equal(_bits(img), _bits(ref))
    assert np.allclose(img[12, 6, :3], np.array([0.6, 0.4, 0.2]) * 0.25, atol=2e-3)


def test_shadow_errors(tlr):
    sb = tlr.make_scene(2, 32, 24, 4, seed=3)
    c = tlr.Ctx(0)
    c.set_samples(sb)
    with pytest.raises(tlr.TlrError):
        c.reconstruct_queries(np.zeros((1, 8), np.float32))   # not built
    with pytest.raises(tlr.TlrError):
        c._chk(tlr.lib().tlr_set_shadow_inputs(c.h, None, None, 5))               # wrong sample count
    c.build()
    other = tlr.Ctx(0)
    other.set_samples(tlr.make_scene(2, 40, 24, 4, seed=4))
    other.build()
    P = _persp(1.0, 32 / 24, 0.1, 100.0)
    sp = tlr.shadow_setup(P, np.eye(4), P, np.eye(4), 32, 24, (0, 0), (0, 0), 100.0)
    with pytest.raises(tlr.TlrError):
        c.reconstruct_dof_motion_shadows(other, sp)   # sizes differ
    col, ext, _ = c.reconstruct_queries(np.zeros((0, 8), np.float32))
    assert col.shape == (0, 4)


def test_irregular_buffer_matches_oracle(tlr, orc):
    """An irregular buffer (spp 0: 70 % of a scene's samples in shuffled
    order) built and reconstructed on the GPU, and queried in shadow mode,
    equals the oracle bit for bit."""
    sb = tlr.make_scene(2, 64, 48, 8, seed=71)
    rng = np.random.default_rng(4)
    rec = sb.records[rng.permutation(sb.records.shape[0])[: int(0.7 * sb.records.shape[0])]].copy()
    rec[:, 14:16] = rng.normal(0, 0.05, (rec.shape[0], 2)).astype(np.float32)
    irr = tlr.UVTSampleBuffer(64, 48, 0, sb.coc[0], sb.coc[1], rec)
    ctx = tlr.Ctx(0)
    ctx.set_samples(irr)
    ctx.set_shadow_inputs(rec[:, 14:16], None)
    ctx.build()
    o = orc.Oracle.irregular(64, 48, sb.coc[0], sb.coc[1], rec)
    info = o.build()
    assert info.spp_pow2 == 8    # 5.6 per pixel -> 5 -> 8
    img, _ = ctx.reconstruct(out_spp=8)
    ref, _, _ = o.reconstruct(out_spp=8)
    assert np.array_equal(_bits(img), _bits(ref))
    q = _queries(np.random.default_rng(9), irr, rec, 5000)
    col, ext, _ = ctx.reconstruct_queries(q, shadow=True)
    ocol, oext, _ = o.reconstruct_queries(q, shadow=True)
    assert np.array_equal(_bits(col), _bits(ocol)) and np.array_equal(_bits(ext[:, :6]), _bits(oext[:, :6]))
