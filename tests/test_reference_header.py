"""CPU: the sample-file loader reads the reference's own header layout.

tests/golden/butterflies-1spp.bin.header is the reference's data file
(data/butterflies-1spp.bin.header: v1.3, 962x543, 1 spp, "Motion model:
affine", CoC -15.242497,19.053122, binary encoding); its 962x543x1 binary
payload is missing from the reference (.MISSING_LARGE_BLOBS), so a synthetic
payload of that shape stands in.  The loader must take width, height, spp,
the CoC coefficients and the encoding from the header, skip the motion-model
line (SampleBuffer.cpp:255-256), and force alpha to 1 (:372)."""
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "golden", "butterflies-1spp.bin.header")


def test_loads_the_reference_header_layout(tlr, tmp_path):
    W, H, spp = 962, 543, 1
    rng = np.random.default_rng(0)
    rec = rng.random((H * W * spp, 16), dtype=np.float32)
    rec[:, 10] = 0.25          # alpha in the file; the loader forces 1
    path = str(tmp_path / "butterflies-1spp.bin")
    rec.tofile(path)
    shutil.copy(HEADER, path + ".header")
    sb = tlr.UVTSampleBuffer.load(path)
    assert (sb.width, sb.height, sb.spp) == (W, H, spp)
    assert sb.coc == (float(np.float32(-15.242497)), float(np.float32(19.053122)))
    want = rec.copy()
    want[:, 10] = 1.0
    assert np.array_equal(sb.records, want)
