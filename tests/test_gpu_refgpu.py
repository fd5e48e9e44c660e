"""GPU parity of the reference's GPU-variant method (path 2: recKernel,
failKernel and nearestKernel of ReconstructionCudaKernels.cu:242-868, driven
as in CudaReconstruction::reconstructGPU, ReconstructionCuda.cpp:327-465)
against the oracle's restatement: per output sample colour and stage bit for
bit, image bit for bit.  Small list capacities (list_cap) send output samples
through the overflow launches; the results must not change."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCENES = {
    "defocus": (1, 96, 80, 4, 1005),
    "motion": (2, 96, 72, 16, 2005),
    "strong": (3, 120, 90, 8, 3005),
    "motion_wide": (2, 240, 135, 16, 2007),
}


@pytest.fixture(scope="module")
def cases(tlr, orc):
    out = {}
    for name, (preset, W, H, spp, seed) in SCENES.items():
        sb = tlr.make_scene(preset, W, H, spp, seed=seed)
        o = orc.Oracle(W, H, spp, sb.coc[0], sb.coc[1], sb.records)
        o.build()
        out[name] = (sb, o)
    return out


def _run(tlr, sb, out_spp, **kw):
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build()
    img, st = ctx.reconstruct(out_spp=out_spp, path=2, **kw)
    M = 1 if kw.get("override_uvt") is not None else out_spp
    return img, st, ctx.query_colors(M), ctx.query_stages(M)


def _compare(what, img, q, stg, ref, qref, sref):
    nbad = int(np.count_nonzero(np.any(q != qref, axis=-1)))
    nstage = int(np.count_nonzero(stg != sref))
    frac = [float(np.mean(sref == k)) for k in range(3)]
    print("%s: stages 1R/2R/nearest %.3f/%.3f/%.3f, colours differing %d, stages differing %d, image max-abs %.3g"
          % (what, frac[0], frac[1], frac[2], nbad, nstage, float(np.abs(img - ref).max())), flush=True)
    assert nstage == 0
    assert nbad == 0
    assert np.array_equal(img, ref)


@pytest.mark.parametrize("name", list(SCENES))
def test_refgpu_bitexact(tlr, orc, cases, name):
    sb, o = cases[name]
    img, st, q, stg = _run(tlr, sb, 16)
    ref, qref, sref = o.reconstruct_refgpu(out_spp=16, want_queries=True)
    _compare(name, img, q, stg, ref, qref, sref)
    assert st["num_2r"] >= float(np.count_nonzero(sref == 1))


@pytest.mark.parametrize("cap", [1, 3, 8])
def test_refgpu_overflow_launches(tlr, orc, cases, cap):
    sb, o = cases["strong"]
    img, _, q, stg = _run(tlr, sb, 16, list_cap=cap)
    ref, qref, sref = o.reconstruct_refgpu(out_spp=16, want_queries=True)
    _compare("strong, list cap %d" % cap, img, q, stg, ref, qref, sref)


@pytest.mark.parametrize("out_spp,uvt", [(1, None), (4, None), (1, (0.0, 0.0, 1.0)), (1, (0.4, -0.7, 0.2))])
def test_refgpu_patterns(tlr, orc, cases, out_spp, uvt):
    sb, o = cases["motion"]
    img, _, q, stg = _run(tlr, sb, out_spp, override_uvt=uvt)
    ref, qref, sref = o.reconstruct_refgpu(out_spp=out_spp, want_queries=True, override_uvt=uvt)
    _compare("motion, out spp %d, uvt %s" % (out_spp, uvt), img, q, stg, ref, qref, sref)


def test_refgpu_tiny(tlr, orc):
    """A 16x12 buffer at 4 spp: few leaves, many output samples left to the
    nearest search."""
    sb = tlr.make_scene(2, 16, 12, 4, seed=7)
    o = orc.Oracle(16, 12, 4, sb.coc[0], sb.coc[1], sb.records)
    o.build()
    img, _, q, stg = _run(tlr, sb, 16)
    ref, qref, sref = o.reconstruct_refgpu(out_spp=16, want_queries=True)
    _compare("tiny", img, q, stg, ref, qref, sref)


def test_refgpu_rejects_gate(tlr):
    sb = tlr.make_scene(2, 16, 12, 4, seed=7)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sb)
    ctx.build()
    with pytest.raises(tlr.TlrError):
        ctx.reconstruct(out_spp=16, path=2, gate=True)
    with pytest.raises(tlr.TlrError):
        ctx.reconstruct(out_spp=16, path=3)
