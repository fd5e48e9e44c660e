"""GPU: batched reconstruction (tlr_reconstruct_batch).  Every image of a
batch of uneven buffers equals the one the single-buffer path makes from the
same buffer, bit for bit (same per-query arithmetic), in both forms of the
query phase (merged launches, chained per-buffer launches); the second call
of the same batch replays the captured graph; one buffer is also checked
against the CPU oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BUFS = [
    # (preset, W, H, spp, out_spp)
    (1, 64, 48, 4, 8),
    (2, 96, 72, 16, 16),
    (3, 80, 64, 8, 4),
    (2, 40, 30, 1, 8),
    (2, 128, 40, 32, 2),
]


@pytest.fixture(scope="module")
def batch(tlr):
    print("batch: uploading %d buffers" % len(BUFS), flush=True)
    sbufs = [tlr.make_scene(p, W, H, spp, seed=500 + k) for k, (p, W, H, spp, m) in enumerate(BUFS)]
    ctx = tlr.Ctx(0)
    for k, sb in enumerate(sbufs):
        ctx.batch_set_samples(k, sb)
    jobs = [(k, BUFS[k][4]) for k in range(len(BUFS))]
    imgs1, bms1, st1 = ctx.reconstruct_batch(jobs)
    imgs2, bms2, st2 = ctx.reconstruct_batch(jobs)
    return sbufs, ctx, jobs, (imgs1, st1), (imgs2, st2)


def test_batch_equals_single_buffer_path(tlr, batch):
    sbufs, ctx, jobs, (imgs, st), _ = batch
    print("batch stats", st, flush=True)
    assert st["autotuned"] == 1
    assert st["output_samples"] == sum(W * H * m for (_, W, H, _, m) in BUFS)
    single = tlr.Ctx(0)
    for k, sb in enumerate(sbufs):
        single.set_samples(sb)
        single.build()
        ref, _ = single.reconstruct(out_spp=BUFS[k][4])
        assert np.array_equal(imgs[k], ref), "buffer %d differs" % k


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_forms_equal_single_buffer_path(tlr, batch, mode):
    sbufs, _, jobs, _, _ = batch
    ctx = tlr.Ctx(0)
    ctx.batch_set_mode(mode)
    for k, sb in enumerate(sbufs):
        ctx.batch_set_samples(k, sb)
    print("batch form %d" % mode, flush=True)
    imgs, _, st = ctx.reconstruct_batch(jobs)
    assert st["chained"] == mode and st["autotuned"] == 1
    imgs2, _, st2 = ctx.reconstruct_batch(jobs)
    assert st2["graph_reused"] == 1 and st2["chained"] == mode
    single = tlr.Ctx(0)
    handed = 0
    for k, sb in enumerate(sbufs):
        single.set_samples(sb)
        single.build()
        ref, sst = single.reconstruct(out_spp=BUFS[k][4])
        # the first 1R lane pass's hand-offs (the entries the single-buffer
        # path's second lane pass takes)
        handed += sst["wide_queries"]
        assert np.array_equal(imgs[k], ref), "form %d: buffer %d differs" % (mode, k)
        assert np.array_equal(imgs2[k], ref), "form %d replay: buffer %d differs" % (mode, k)
    if mode == 1:
        # same fast kernel as the single-buffer path, so the same hand-offs
        # (the merged form's tuned variant may keep fewer leaves per lane)
        assert st["general_queries"] == handed


def test_batch_graph_replay(batch):
    sbufs, ctx, jobs, (imgs1, st1), (imgs2, st2) = batch
    assert st2["graph_reused"] == 1 and st2["autotuned"] == 0
    for a, b in zip(imgs1, imgs2):
        assert np.array_equal(a, b)


def test_batch_against_oracle(tlr, orc, batch):
    sbufs, ctx, jobs, (imgs, st), _ = batch
    k = 1
    sb = sbufs[k]
    o = orc.Oracle(sb.width, sb.height, sb.spp, sb.coc[0], sb.coc[1], sb.records)
    o.build()
    ref, _, _ = o.reconstruct(out_spp=BUFS[k][4])
    err = np.abs(imgs[k] - ref)
    assert err.max() <= 2e-3 and err.mean() <= 1e-4


def test_batch_shape_change_recaptures(tlr, batch):
    sbufs, ctx, jobs, _, _ = batch
    sub = [jobs[2], jobs[0]]
    imgs, _, st = ctx.reconstruct_batch(sub)
    assert st["graph_reused"] == 0
    single = tlr.Ctx(0)
    single.set_samples(sbufs[2])
    single.build()
    ref, _ = single.reconstruct(out_spp=BUFS[2][4])
    assert np.array_equal(imgs[0], ref)


def test_batch_errors(tlr):
    ctx = tlr.Ctx(0)
    with pytest.raises(tlr.TlrError):
        ctx.reconstruct_batch([(3, 8)])  # empty slot
    with pytest.raises(tlr.TlrError):
        ctx.batch_set_mode(2)


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_new_parameters_are_not_replayed_stale(tlr, mode):
    """A later call on the same slots with other CoC adjustments, and after a
    slot is re-uploaded with a different scene of the same size, gives the
    single-buffer path's images bit for bit (the captured graph must not
    replay the previous call's parameters)."""
    sbufs = [tlr.make_scene(p, W, H, spp, seed=900 + k) for k, (p, W, H, spp, m) in enumerate(BUFS[:3])]
    ctx = tlr.Ctx(0)
    ctx.batch_set_mode(mode)
    for k, sb in enumerate(sbufs):
        ctx.batch_set_samples(k, sb)
    single = tlr.Ctx(0)

    def check(jobs, bufs, label):
        imgs, _, st = ctx.reconstruct_batch(jobs)
        for k, job in enumerate(jobs):
            single.set_samples(bufs[job[0]])
            single.build(job[2] if len(job) > 2 else 1.0, job[3] if len(job) > 3 else 1.0)
            ref, _ = single.reconstruct(out_spp=job[1])
            assert np.array_equal(imgs[k], ref), "%s: item %d differs" % (label, k)
        return st

    print("form %d: plain call" % mode, flush=True)
    jobs = [(k, BUFS[k][4]) for k in range(3)]
    check(jobs, sbufs, "first call")
    print("form %d: aperture 1.3 / focal 0.8 on two items" % mode, flush=True)
    jobs2 = [(0, BUFS[0][4], 1.3, 1.0), (1, BUFS[1][4], 1.0, 0.8), (2, BUFS[2][4])]
    check(jobs2, sbufs, "new adjustments")
    print("form %d: slot 1 re-uploaded with another seed" % mode, flush=True)
    p, W, H, spp, m = BUFS[1]
    sbufs[1] = tlr.make_scene(p, W, H, spp, seed=1234)
    ctx.batch_set_samples(1, sbufs[1])
    check(jobs, sbufs, "re-uploaded slot")


def _committed_gate(tlr):
    import json
    import os
    d = os.path.dirname(tlr.__file__)
    params = np.fromfile(os.path.join(d, "gate_params.f32"), dtype="<f4")
    with open(os.path.join(d, "gate_params.json")) as f:
        return params, json.load(f)["threshold"]


def test_gated_batch_equals_gated_single_buffer(tlr, batch):
    """Items with the learned gate on (committed parameters) give the gated
    single-buffer path's images bit for bit, the others the plain ones; the
    batch runs the chained form, counts the hand-offs the gate accepted as the
    single-buffer calls do, replays its graph, and recaptures it (no stale
    threshold) when the threshold changes.  Completion times grow in item
    order."""
    sbufs, _, jobs, _, _ = batch
    params, thr = _committed_gate(tlr)
    flags = [True, False, True, True, False]
    ctx = tlr.Ctx(0)
    ctx.gate_set_params(params)
    single = tlr.Ctx(0)
    single.gate_set_params(params)
    for k, sb in enumerate(sbufs):
        ctx.batch_set_samples(k, sb)
    for t in (thr, thr - 1.0):
        ctx.gate_set_threshold(t)
        single.gate_set_threshold(t)
        print("gated batch, threshold %.4f" % t, flush=True)
        done = []
        imgs, _, st = ctx.reconstruct_batch(jobs, gate=flags, done=done)
        imgs2, _, st2 = ctx.reconstruct_batch(jobs, gate=flags)
        assert st["chained"] == 1 and st2["graph_reused"] == 1
        accepted = 0.0
        for k, sb in enumerate(sbufs):
            single.set_samples(sb)
            single.build()
            ref, sst = single.reconstruct(out_spp=BUFS[k][4], gate=flags[k])
            if flags[k]:
                accepted += sst["gate_candidates"] - sst["gate_rejected"]
            assert np.array_equal(imgs[k], ref), "threshold %g: buffer %d differs" % (t, k)
            assert np.array_equal(imgs2[k], ref), "threshold %g replay: buffer %d differs" % (t, k)
        assert st["gate_accepted"] == accepted == st2["gate_accepted"]
        assert all(d > 0 for d in done) and all(b >= a for a, b in zip(done, done[1:]))


def test_gated_batch_needs_parameters(tlr, batch):
    sbufs, _, jobs, _, _ = batch
    ctx = tlr.Ctx(0)
    ctx.batch_set_samples(0, sbufs[0])
    with pytest.raises(tlr.TlrError):
        ctx.reconstruct_batch([jobs[0]], gate=True)
