"""C-ABI checks that need no GPU: the library loads and exports every symbol
its headers declare; host-only entry points (sample files, output pattern)
behave like the reference's."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tlr_[a-z0-9_]+)\s*\(", src)))


def test_tlr_exports_every_declared_symbol(tlr):
    L = tlr.lib()
    names = _declared("tlr.h")
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(tlr.EXPORTED), set(names) ^ set(tlr.EXPORTED)


def test_scene_exports(tlr):
    L = tlr.scene_lib()
    for n in _declared("tlr_scene.h"):
        assert hasattr(L, n), n


def test_create_without_gpu_fails_cleanly(tlr):
    import ctypes
    try:
        import subprocess
        has_gpu = subprocess.run(["rocm-smi"], capture_output=True).returncode == 0 and \
            os.path.exists("/dev/kfd")
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = tlr.lib().tlr_create(ctypes.byref(h), 0)
    assert rc != 0  # no silent host fallback


@pytest.mark.parametrize("binary", [False, True])
def test_sample_file_round_trip(tlr, tmp_path, binary):
    sb = tlr.make_scene(1, 12, 10, 2, seed=4)
    path = str(tmp_path / ("buf.bin" if binary else "buf.txt"))
    sb.save(path, binary=binary, separate_header=binary)
    back = tlr.UVTSampleBuffer.load(path)
    assert (back.width, back.height, back.spp) == (12, 10, 2)
    assert np.allclose(back.coc, sb.coc, atol=1e-5)
    r0, r1 = sb.records, back.records
    cols = [c for c in range(16) if c != 10]
    if binary:
        assert np.array_equal(r1[:, cols], r0[:, cols])
    else:
        assert np.allclose(r1[:, cols], r0[:, cols], atol=1e-6 + 1e-6 * np.abs(r0[:, cols]).max())
    # alpha is forced to 1 on load (SampleBuffer.cpp:343,372), even though the
    # writer stores b there (SampleBuffer.cpp:498,542)
    assert np.all(r1[:, 10] == 1.0)


def test_binary_needs_separate_header(tlr, tmp_path):
    sb = tlr.make_scene(1, 4, 4, 1, seed=1)
    with pytest.raises(tlr.TlrError):
        sb.save(str(tmp_path / "x.bin"), binary=True, separate_header=False)


def test_missing_file(tlr):
    with pytest.raises(tlr.TlrError, match="TLR_E_IO"):
        tlr.UVTSampleBuffer.load("/nonexistent/file.bin")


@pytest.mark.parametrize("M", [1, 8, 16, 32])
def test_product_pattern_matches_oracle(tlr, orc, M):
    a = tlr.output_pattern(M)
    b = orc.output_pattern(M)
    assert np.array_equal(a, b[:, :5])
    a = tlr.output_pattern(M, override_uvt=(0.1, 0.2, 0.3))
    b = orc.output_pattern(M, override_uvt=(0.1, 0.2, 0.3))
    assert np.array_equal(a, b[:, :5])


def _fw_parse_float32(s):
    """FW::parseFloat (src/framework/base/String.cpp:428-487) on a plain
    number, each step an IEEE single-precision operation (numpy float32)."""
    f = np.float32
    t, neg = 0, False
    if s[:1] == "+":
        t = 1
    elif s[:1] == "-":
        t, neg = 1, True
    v = f(0)
    while t < len(s) and s[t].isdigit():
        v = f(f(v * f(10)) + f(int(s[t])))
        t += 1
    if t < len(s) and s[t] == ".":
        t += 1
        sc = f(1)
        while t < len(s) and s[t].isdigit():
            sc = f(sc * f(0.1))
            v = f(v + f(sc * f(int(s[t]))))
            t += 1
    if t < len(s) and s[t] in "eE":
        e = int(s[t + 1:])
        if e:
            v = f(v * np.power(f(10), f(e)))
    return f(-v) if neg else v


def test_text_records_parse_as_fw_parsefloat(tlr, tmp_path):
    """Text records go through the reference's float parser, not a correctly
    rounded one: digits accumulate in float, fraction digits as scale * d with
    scale *= 0.1f, exponents as powf(10, e), and the #INF / #IND / #QNAN /
    #SNAN literals give fixed bit patterns.  Three of the fields below round
    differently from the decimal value (strtof), so a correctly rounded parser
    fails this test."""
    sb = tlr.make_scene(1, 1, 1, 1, seed=3)
    path = tmp_path / "one.txt"
    sb.save(str(path), binary=False)
    lines = path.read_text().splitlines()
    fields = ["0.123456", "123.456789", "-7.654321", "1.5e2", "2.5E-3", "16777217", "1#INF", "-1#IND",
              "3#QNAN", "0.1", "9", "+42", "-0.000001", "3.14159", "2#SNAN", "1e1"]
    lines[-1] = ",".join(fields)
    path.write_text("\n".join(lines) + "\n")
    got = tlr.UVTSampleBuffer.load(str(path)).records[0].view(np.uint32)
    special = {6: 0x7F800000, 7: 0xFFC00000, 8: 0x7FC00001, 14: 0x7F800001}
    differ = 0
    for k, s in enumerate(fields):
        if k == 10:
            assert got[k] == np.float32(1).view(np.uint32)   # alpha forced to 1
        elif k in special:
            assert got[k] == special[k], (s, hex(got[k]))
        else:
            want = _fw_parse_float32(s).view(np.uint32)
            assert got[k] == want, (s, hex(got[k]), hex(want))
            differ += int(want != np.float32(float(s)).view(np.uint32))
    assert differ == 3
