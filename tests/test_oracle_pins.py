"""CPU pins of the oracle's decision rules, each against an outcome the
reference's own rules fix by hand (not a retyped formula):

- sameSurface (Reconstruction.hpp:428-525) on hand-built leaf bounds:
  parallel planes within and beyond the 0.1 threshold, planes crossing
  inside the (u,v) box, a one-sided difference, planes crossing in time;
- the leaf surface scan (ReconstructionTreeBuilder.cpp:222-340) on hand-built
  screen boxes: the time split, the 2.0 conflict threshold, crossing motion,
  and a CoC sequence on which the reference's cocboundst0[0] typo (:334,336)
  decides the split;
- Filterer::reconstruct's control flow (ReconstructionOur.cpp:50-359) on
  constructed static scenes (tests/pin_scenes.py): fetch2R and the triangle
  coverage (a front surface that does not cover gives the back colour, one
  that does gives its own), no fetch2R when the first surface holds all four
  quadrants, the emergency pass ending in the nearest-sample colour, and a
  3-sample front surface merged into the one behind it;
- generateOutputSamples (Reconstruction.cpp:452-512) recomposed here from the
  separately pinned primitives;
- the GPU-variant 2R stage's surface split at 4 points
  (ReconstructionCudaKernels.cu:634-868) on the 3-sample front surface.

tests/test_oracle_mutants.py rebuilds the oracle with each rule broken and
checks that one of these tests fails.
"""
import numpy as np
import pytest

import pin_scenes as ps

O5 = np.array([10.5, 10.5, 0.0, 0.0, 0.5], np.float32)
RANGE_U = 1.0 / np.sqrt(128.0)


def bounds(xmin=(0, 0, 10), xmax=(0, 0, 12), ymin=(0, 0, 10), ymax=(0, 0, 12), ts=(0.0, 1.0)):
    """14 floats: (t, uv, 1) coefficients of the xmin, xmax, ymin, ymax planes
    (TimeLensBounds::evaluate, Reconstruction.hpp:162), then the time span."""
    return np.array(list(xmin) + list(xmax) + list(ymin) + list(ymax) + list(ts), np.float32)


# ---------------------------------------------------------------- sameSurface
def test_same_surface_identical_and_parallel(orc):
    b = bounds()
    assert orc.same_surface(b, b, O5)
    # every plane shifted by 0.05 (< 0.1): equal within the threshold
    b2 = b.copy()
    b2[[2, 5, 8, 11]] += 0.05
    assert orc.same_surface(b, b2, O5)
    # parallel but 5 px apart: the ordering is the same at every corner
    b3 = b.copy()
    b3[[2, 5, 8, 11]] += 5.0
    assert orc.same_surface(b, b3, O5)


def test_same_surface_crossing_in_uv(orc):
    """xmin planes of slopes 3.81 and 18.29 through the same point at u = 0:
    at u = -/+ 1/sqrt(128) they differ by +-1.28 (> 0.1), opposite signs."""
    a = bounds(xmin=(0, 3.81, 10))
    b = bounds(xmin=(0, 18.29, 10))
    assert not orc.same_surface(a, b, O5)
    # slopes 1.0 apart: +-0.088 at the corners, inside the threshold
    c = bounds(xmin=(0, 4.81, 10))
    assert orc.same_surface(a, c, O5)


def test_same_surface_one_sided_difference(orc):
    """A difference of 0.3 at one u corner and 0.05 at the other compares as
    (+1, 0): ternaryEqual accepts a zero on either side."""
    # d(u) = 0.175 + 1.4142 * u: +0.30 at u=+0.0884, +0.05 at u=-0.0884
    a = bounds(xmin=(0, 0.0, 10))
    b = bounds(xmin=(0, -1.41421, 10 - 0.175))
    assert orc.same_surface(a, b, O5)


def test_same_surface_crossing_in_time(orc):
    """xmin planes equal at t = 0.5 but with t-slopes 0 and 4: at u = 0 they
    differ by -/+0.18 at t = 0.5 -/+ 0.5/sqrt(128), opposite signs."""
    a = bounds(xmin=(0, 0, 10))
    b = bounds(xmin=(4, 0, 8))
    assert not orc.same_surface(a, b, O5)


# --------------------------------------------------------------- surface scan
def static_boxes(xs, cocs, y=5.0):
    """Screen boxes of static samples at x with CoC c: (x+c, y+c) at u=-1 and
    (x-c, y-c) at u=+1, the same at t=0 and t=1."""
    xs, cocs = np.asarray(xs, np.float32), np.asarray(cocs, np.float32)
    ba = np.stack([xs + cocs, y + cocs], -1)
    bb = np.stack([xs - cocs, y - cocs], -1)
    return ba, bb, ba.copy(), bb.copy()


def test_scan_cocbounds_typo_decides(orc):
    """CoCs 0, 1.6, 0.3, -0.5 at one position.  With the reference's
    cocboundst0[1] = max(coct0, cocboundst0[0]) the upper bound after the
    third sample is 0.3, so the fourth passes the early-out (|c-0|, |c-0.3|
    < 1) and joins; with the running maximum (1.6) it would be tested against
    the second sample, whose boxes differ by 2.1 with opposite signs, and
    start a new surface."""
    ba, bb, ba1, bb1 = static_boxes([5, 5, 5, 5], [0.0, 1.6, 0.3, -0.5])
    assert orc.surface_scan(np.full(4, 0.5), ba, bb, ba1, bb1).tolist() == [0, 0, 0, 0]


def test_scan_time_split(orc):
    ba, bb, ba1, bb1 = static_boxes([5] * 4, [1.0] * 4)
    assert orc.surface_scan([0.5, 0.5, 0.0, 0.0], ba, bb, ba1, bb1).tolist() == [0, 0, 1, 1]


def test_scan_conflict_threshold(orc):
    """Boxes differing by 2.1 (> 2.0) with opposite signs split; by 1.9 not."""
    ba, bb, ba1, bb1 = static_boxes([5, 5], [0.0, 2.1])
    assert orc.surface_scan([0.5, 0.5], ba, bb, ba1, bb1).tolist() == [0, 1]
    ba, bb, ba1, bb1 = static_boxes([5, 5], [0.0, 1.9])
    assert orc.surface_scan([0.5, 0.5], ba, bb, ba1, bb1).tolist() == [0, 0]


def test_scan_crossing_motion(orc):
    """Equal CoCs; one sample moves 5 -> 8, the other 8 -> 5: ordered one way
    at t = 0 and the other at t = 1, so they are different surfaces.  The same
    pair moving in parallel stays one surface."""
    def moving(x0, x1, c=1.0, y=5.0):
        ba = np.array([[x0 + c, y + c]], np.float32)
        bb = np.array([[x0 - c, y - c]], np.float32)
        ba1 = np.array([[x1 + c, y + c]], np.float32)
        bb1 = np.array([[x1 - c, y - c]], np.float32)
        return ba, bb, ba1, bb1
    A, B = moving(5, 8), moving(8, 5)
    arrs = [np.concatenate([a, b]) for a, b in zip(A, B)]
    assert orc.surface_scan([0.5, 0.5], *arrs).tolist() == [0, 1]
    A, B = moving(5, 8), moving(5.5, 8.5)
    arrs = [np.concatenate([a, b]) for a, b in zip(A, B)]
    assert orc.surface_scan([0.5, 0.5], *arrs).tolist() == [0, 0]


# ------------------------------------------------------------ filter control
_oracles = {}


def _built(orc, name):
    if name not in _oracles:
        rec = ps.records(getattr(ps, name))
        o = orc.Oracle(ps.W, ps.H, ps.SPP, ps.C0, ps.C1, rec)
        o.build()
        _oracles[name] = (rec, o)
    return _oracles[name]


def _query(o, px, py):
    return o.query(px, py, 0, out_spp=1, override_uvt=ps.PINHOLE)


def test_fetch2r_front_not_covering_gives_back(orc):
    """Pixel (8,10) beside a front edge at x = 8.3: the front's samples within
    R lie only left of the centre (quadrant mask != 0xF) and the back's are
    within R too, so 2R is fetched; no triangle of front samples can cover the
    centre, so the back surface (the last) is found: blue."""
    rec, o = _built(orc, "edge")
    q = _query(o, 8, 10)
    assert q.fetch2r == 1 and q.num_surfaces_1r == 2 and q.found_surface >= 1 and q.used_nearest == 0
    assert np.allclose(list(q.color)[:3], ps.BLUE, atol=1e-6)
    # well inside the front: one surface, no 2R
    q = _query(o, 6, 10)
    assert q.fetch2r == 0 and q.num_surfaces_1r == 1
    assert np.allclose(list(q.color)[:3], ps.RED, atol=1e-6)


def test_fetch2r_front_covering_gives_front(orc):
    """Pixel (8,10) just inside the front's corner: within R the front misses
    the x>0, y>0 quadrant, so 2R is fetched; at 2R front samples at
    (-0.125,+0.125), (+0.375,-0.125), (-0.125,-0.125) from the centre form a
    triangle with all four half-planes that contains it: red."""
    rec, o = _built(orc, "corner")
    q = _query(o, 8, 10)
    assert q.fetch2r == 1 and q.num_surfaces_1r == 2 and q.found_surface == 0
    assert np.allclose(list(q.color)[:3], ps.RED, atol=1e-6)


def test_no_fetch2r_when_first_surface_has_all_quadrants(orc):
    """Front and back samples interleaved around pixel (8,10): two surfaces
    within R, the front one in all four quadrants, so no 2R fetch and the
    front is the result."""
    rec, o = _built(orc, "mesh")
    q = _query(o, 8, 10)
    assert q.num_surfaces_1r >= 2 and q.fetch2r == 0
    assert np.allclose(list(q.color)[:3], ps.RED, atol=1e-6)


def test_emergency_pass_gives_nearest_sample(orc):
    """No sample within 1.6 px of pixel (20,10): nothing at R or 2R, so the
    first emergency pass doubles the dispersion to 2R and gathers at 4R; its
    tent weights (1 - d/2R) are all <= 0 there, so the colour is the nearest
    sample's."""
    rec, o = _built(orc, "hole")
    q = _query(o, 20, 10)
    want, dmin = ps.nearest_colour(rec, ps.HOLE_CENTRE)
    assert 1.6 <= dmin < 2.0
    assert q.fetch2r == 1 and q.emergency_passes == 1 and q.used_nearest == 1
    assert q.dispersion == np.float32(2 * ps.R)
    assert np.allclose(list(q.color)[:3], want, atol=1e-6)


def test_tiny_front_surface_merges_into_back(orc):
    """Three front samples left of pixel (20,4)'s centre, back samples all
    around: 2R is fetched (the front misses the x>0 half), the 3-sample front
    surface is merged into the back one (fewer than 4 samples), and the
    merged surface, the last, is filtered: the tent mean of every sample
    within R, green and blue together.  Unmerged, the front could not cover
    the centre and the result would be pure blue."""
    rec, o = _built(orc, "tiny")
    q = _query(o, *ps.TINY_PIXEL)
    centre = (ps.TINY_PIXEL[0] + 0.5, ps.TINY_PIXEL[1] + 0.5)
    want = ps.tent_mean(rec, centre)
    assert want[1] > 0.05
    assert q.fetch2r == 1
    assert np.allclose(list(q.color)[:3], want, atol=2e-6), (list(q.color), want)


def test_refgpu_2r_surface_needs_four_points(orc):
    """The GPU-variant 2R stage ends a surface at a leaf unlike an earlier one
    only once the surface holds 4 points: the 3-point front surface of pixel
    (20,4) takes the next (back) leaf in, so green reaches the colour; split
    at 3 points it would be blue only."""
    rec, o = _built(orc, "tiny")
    o5 = np.array([ps.TINY_PIXEL[0] + 0.5, ps.TINY_PIXEL[1] + 0.5, 0, 0, 0.5], np.float32)
    found, col = o.refgpu_failed(o5)
    assert found and col[1] > 0.05, (found, col)


# ------------------------------------------------------------ output pattern
def _pattern_from_primitives(orc, M):
    """generateOutputSamples (Reconstruction.cpp:452-512) composed from the
    oracle's primitives, which test_oracle_primitives pins one by one."""
    L = orc.lib()
    f = np.float32
    rnd = np.zeros(64, np.uint32)
    L.orc_ranrot(242, 64, rnd)
    out = []
    for j in range(64):
        xo = f(f(j) + f(0.5)) / f(64)                      # hammersley(j, 64)
        yo, uo, vo = f(L.orc_lp(j)), f(L.orc_halton(2, j)), f(L.orc_halton(3, j))
        tj = f(f(rnd[j]) * f(1.0 / 4294967296.0))           # Random::getF32
        rows = []
        for i in range(M):
            s2 = np.zeros(2, np.float32)
            L.orc_sobol2d(i + 1, s2)
            x, y = f(L.orc_sobol(3, i + 1)), f(L.orc_lp(i + 1))
            t = f(f(f(i) + tj) / f(M))
            xy = [f(x + xo), f(y + yo)]
            uv = [f(s2[0] + uo), f(s2[1] + vo)]
            for v in (xy, uv):
                for k in range(2):
                    if v[k] >= 1:
                        v[k] = f(v[k] - f(1))
            ub, vb, tb = int(np.floor(f(256) * uv[0])), int(np.floor(f(256) * uv[1])), int(np.floor(f(256) * t))
            key = 0
            for k in range(8):
                key |= ((ub >> k) & 1) << (3 * k + 2) | ((vb >> k) & 1) << (3 * k + 1) | ((tb >> k) & 1) << (3 * k)
            d = np.zeros(2, np.float32)
            L.orc_to_unit_disk(uv[0], uv[1], d)
            if M == 1:
                xy = [f(0.5), f(0.5)]
            rows.append([xy[0], xy[1], d[0], d[1], t, f(key)])
        rows.sort(key=lambda r: r[5])   # keys are distinct: any sort is the reference's order
        out.extend(rows)
    return np.array(out, np.float32)


@pytest.mark.parametrize("M", [1, 4, 16, 32])
def test_output_pattern_composition(orc, M):
    want = _pattern_from_primitives(orc, M)
    got = orc.output_pattern(M)
    assert got.shape == want.shape
    assert len(np.unique(want[:, 5].reshape(64, M), axis=1)) > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
