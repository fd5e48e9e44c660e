"""GPU: tlr_reconstruct_noalloc (the allocation-free entry point) against
tlr_reconstruct on the same build, through the C ABI: images bit for bit and
Stats equal, on two calls in a row, the second with a smaller buffer.  The
first buffer spans three row bands of the fixed scratch (the last one
partial, the width not a multiple of 8); the second fits one band.  Device
memory for the output comes from the HIP runtime through ctypes."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATS = ("leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched", "surfaces", "num_2r",
         "surfaces_merged", "at_least_one", "count", "overflow")


class DeviceBuffer:
    """A device allocation made by the test (the entry point makes none)."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        # poison, so that a region the entry point fails to write shows
        assert self.hip.hipMemset(self.ptr, 0x7F, ctypes.c_size_t(nbytes)) == 0

    def to_host(self, shape):
        out = np.empty(shape, np.float32)
        assert out.nbytes == self.nbytes
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes),
                                  2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


@pytest.mark.parametrize("collect_stats", [True, False])
def test_noalloc_equals_reconstruct(tlr, collect_stats):
    ctx = tlr.Ctx(0)
    # (preset, W, H, spp, out_spp): 520x264x16 output samples = 3 bands of
    # 2^20 at most (120 + 120 + 24 rows); then a one-band buffer
    cases = [(2, 520, 264, 8, 16), (3, 72, 40, 16, 8)]
    for preset, W, H, spp, M in cases:
        print("case %dx%d spp %d out %d stats %d" % (W, H, spp, M, collect_stats), flush=True)
        sb = tlr.make_scene(preset, W, H, spp, seed=4000 + W)
        ctx.set_samples(sb)
        ctx.build()
        ref, st_ref = ctx.reconstruct(out_spp=M, collect_stats=collect_stats)
        buf = DeviceBuffer(W * H * 16)
        try:
            st = ctx.reconstruct_noalloc(buf.ptr.value, out_spp=M, collect_stats=collect_stats)
            img = buf.to_host((H, W, 4))
        finally:
            buf.free()
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), \
            "%dx%d: %d pixels differ" % (W, H, int(np.count_nonzero(np.any(img != ref, axis=-1))))
        if collect_stats:
            for k in STATS:
                assert st[k] == st_ref[k], "%s: noalloc %s reconstruct %s" % (k, st[k], st_ref[k])


def test_noalloc_rejects_what_it_cannot_hold(tlr):
    ctx = tlr.Ctx(0)
    sb = tlr.make_scene(1, 16, 12, 2, seed=8)
    ctx.set_samples(sb)
    ctx.build()
    buf = DeviceBuffer(16 * 12 * 16)
    try:
        with pytest.raises(tlr.TlrError) as e:
            ctx.reconstruct_noalloc(buf.ptr.value, out_spp=300)
        assert e.value.code == 1 and "at most 256" in str(e.value)
        with pytest.raises(tlr.TlrError) as e:
            ctx.reconstruct_noalloc(buf.ptr.value, out_spp=4, path=2)
        assert e.value.code == 1
        with pytest.raises(tlr.TlrError) as e:
            ctx.reconstruct_noalloc(0, out_spp=4)
        assert e.value.code == 1
        ctx.reconstruct_noalloc(buf.ptr.value, out_spp=4)
    finally:
        buf.free()
