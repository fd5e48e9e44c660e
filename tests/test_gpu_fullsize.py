"""GPU parity at the full sizes of the benchmark workloads: 960x540 at 16 spp
(16 output samples per pixel) and 1920x1080 at 16 spp (32 output samples per
pixel, 400 quads + 5000 occluders).  The whole frame runs on the GPU; the
oracle filters every k-th scanline (its build is complete), and the output
samples and pixels of those rows must be equal bit for bit.  The build's leaf
count and hierarchy size must match exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# name: (preset, W, H, spp, out_spp, row stride, seed)
FULL = {
    "960x540": (2, 960, 540, 16, 16, 12, 2011),
    "1920x1080": (3, 1920, 1080, 16, 32, 60, 3011),
}


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def frames(tlr, orc):
    out = {}

    def get(name):
        if name not in out:
            preset, W, H, spp, M, stride, seed = FULL[name]
            print("scene %s" % name, flush=True)
            sb = tlr.make_scene(preset, W, H, spp, seed=seed)
            ctx = tlr.Ctx(0)
            ctx.set_samples(sb)
            ctx.build()
            o = orc.Oracle(W, H, spp, sb.coc[0], sb.coc[1], sb.records)
            o.build()
            out[name] = (sb, ctx, o)
        return out[name]
    return get


def _check_rows(what, img, q, ref, qref, stride):
    rows = slice(0, None, stride)
    nbad = int(np.count_nonzero(np.any(_bits(q[rows]) != _bits(qref[rows]), axis=-1)))
    print("%s: %d rows, output samples differing %d / %d" % (what, q[rows].shape[0], nbad,
                                                             q[rows].size // 4), flush=True)
    assert nbad == 0
    assert np.array_equal(_bits(img[rows]), _bits(ref[rows]))


@pytest.mark.parametrize("name", list(FULL))
def test_fullsize_plain(tlr, orc, frames, name):
    preset, W, H, spp, M, stride, seed = FULL[name]
    sb, ctx, o = frames(name)
    g, r = ctx.build_info(), o.info()
    assert g.num_leaves == r.num_leaves and g.num_valid_leaves == r.num_internal + 1
    img, st = ctx.reconstruct(out_spp=M)
    q = ctx.query_colors(M)
    ref, qref, _ = o.reconstruct(out_spp=M, want_queries=True, row_stride=stride)
    print("%s: GPU build %.1f ms, query %.1f ms" % (name, st["build_ms"], st["query_ms"]), flush=True)
    _check_rows(name, img, q, ref, qref, stride)


def test_fullsize_refgpu(tlr, orc, frames):
    preset, W, H, spp, M, stride, seed = FULL["960x540"]
    sb, ctx, o = frames("960x540")
    img, st = ctx.reconstruct(out_spp=M, path=2)
    q = ctx.query_colors(M)
    stg = ctx.query_stages(M)
    ref, qref, sref = o.reconstruct_refgpu(out_spp=M, want_queries=True, row_stride=stride)
    rows = slice(0, None, stride)
    assert np.array_equal(stg[rows], sref[rows])
    _check_rows("960x540 path 2", img, q, ref, qref, stride)
