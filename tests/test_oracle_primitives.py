"""CPU pins of the oracle's primitive restatements (closed forms, textbook
properties, independent computations).  No GPU."""
import math
import random
from fractions import Fraction

import numpy as np
import pytest


def test_sobol_dim3_direction_numbers(orc):
    # First direction numbers of Sobol dimension 4 (Joe-Kuo: s=3, a=1,
    # m=1,3,1): v_c = m_c * 2^(31-c) for c < 3, then the recurrence.
    L = orc.lib()
    expect = [0x80000000, 0xC0000000, 0x20000000, 0x50000000, 0xF8000000, 0x74000000,
              0xA2000000, 0x93000000]
    for c, v in enumerate(expect):
        # sobol(3, 2^c) == v_c * 2^-32 (a single generating-matrix column)
        assert L.orc_sobol(3, 1 << c) == np.float32(v * 2.0 ** -32)


def test_sobol_dim4_direction_numbers(orc):
    # Sobol dimension 5 (index 4; Joe-Kuo: s=3, a=2, m=1,1,1, polynomial
    # x^3+x^2+1), used for the light positions of the shadow path: the
    # recurrence v_c = v_{c-3} ^ (v_{c-3} >> 3) ^ v_{c-1} from the published
    # initial numbers.
    L = orc.lib()
    v = [0x80000000, 0x40000000, 0x20000000]
    for c in range(3, 12):
        v.append(v[c - 3] ^ (v[c - 3] >> 3) ^ v[c - 1])
    assert v[3:8] == [0xB0000000, 0xF8000000, 0xDC000000, 0x7A000000, 0x9D000000]
    for c, x in enumerate(v):
        assert L.orc_sobol(4, 1 << c) == np.float32(x * 2.0 ** -32)


@pytest.mark.parametrize("dim", [0, 3, 4])
def test_sobol_is_a_net(orc, dim):
    # The first 2^m points of a 1-D Sobol dimension are a (0,m,1)-net:
    # exactly one point per interval [k/2^m, (k+1)/2^m).
    L = orc.lib()
    for m in range(1, 11):
        pts = [L.orc_sobol(dim, i) for i in range(1 << m)]
        cells = sorted(int(p * (1 << m)) for p in pts)
        assert cells == list(range(1 << m)), (dim, m)


def test_sobol2d_coordinates(orc):
    # sobol2D (Util.cpp:266-288): the first coordinate runs the all-ones
    # generator matrix (v |= v>>1, the Larcher-Pillichshammer matrix), the
    # second the Pascal matrix of Sobol dimension 1 (v ^= v>>1).  Both are
    # (0,1)-sequences; together they are a (0,2)-sequence in base 2.
    L = orc.lib()
    out = np.zeros(2, np.float32)
    for m in range(1, 9):
        pts = []
        for i in range(1 << m):
            L.orc_sobol2d(i, out)
            pts.append((float(out[0]), float(out[1])))
        for i, (a, _) in enumerate(pts):
            assert a == L.orc_lp(i)
        for j in range(m + 1):
            nx, ny = 1 << j, 1 << (m - j)
            cells = sorted((int(a * nx), int(b * ny)) for a, b in pts)
            assert cells == sorted((x, y) for x in range(nx) for y in range(ny)), (m, j)


def test_lp_is_a_net(orc):
    L = orc.lib()
    for m in range(1, 11):
        cells = sorted(int(L.orc_lp(i) * (1 << m)) for i in range(1 << m))
        assert cells == list(range(1 << m))


@pytest.mark.parametrize("base", [2, 3])
def test_halton_radical_inverse(orc, base):
    L = orc.lib()
    for i in range(200):
        exact, f, k = Fraction(0), Fraction(1, base), i
        while k:
            exact += (k % base) * f
            k //= base
            f /= base
        v = L.orc_halton(base, i)
        assert abs(v - float(exact)) <= 2e-7 * max(1.0, float(exact))
        if base == 2:
            assert v == float(exact)


def test_to_unit_disk_known_points(orc):
    L = orc.lib()
    out = np.zeros(2, np.float32)
    for (sx, sy), (ex, ey) in [((0.5, 0.5), (0, 0)), ((1.0, 0.5), (1, 0)), ((0.5, 1.0), (0, 1)),
                               ((0.0, 0.5), (-1, 0)), ((0.5, 0.0), (0, -1)),
                               ((1.0, 1.0), (math.sqrt(0.5), math.sqrt(0.5)))]:
        L.orc_to_unit_disk(sx, sy, out)
        assert abs(out[0] - ex) < 1e-6 and abs(out[1] - ey) < 1e-6, (sx, sy, out)
    rng = np.random.default_rng(3)
    for sx, sy in rng.random((500, 2)):
        L.orc_to_unit_disk(sx, sy, out)
        # concentric map: radius is the square's Chebyshev radius
        r = max(abs(2 * sx - 1), abs(2 * sy - 1))
        assert abs(math.hypot(out[0], out[1]) - r) < 2e-6


def _jenkins_hash(a, b):
    # Bob Jenkins' lookup2 mix(a, b, c) with c = golden ratio constant, as
    # published (independent transcription).
    M = 0xFFFFFFFF
    c = 0x9E3779B9
    a &= M
    b &= M
    for sa, sb, sc in ((13, 8, 13), (12, 16, 5), (3, 10, 15)):
        a = (a - b - c) & M; a ^= c >> sa
        b = (b - c - a) & M; b ^= (a << sb) & M
        c = (c - a - b) & M; c ^= b >> sc
    return c


def test_hash_bits(orc):
    L = orc.lib()
    for x in range(0, 300, 7):
        for y in range(0, 300, 11):
            assert L.orc_hash_bits(x, y) == _jenkins_hash(x, y)
    # pattern indices are spread over all 64 patterns
    idx = {L.orc_hash_bits(x, y) % 64 for x in range(32) for y in range(32)}
    assert len(idx) == 64


def test_ranrot_properties(orc):
    L = orc.lib()
    out = np.zeros(64, np.uint32)
    L.orc_ranrot(242, 64, out)
    # independent transcription of Agner Fog's RANROT-A (11-word buffer,
    # rotate 13, taps 0 and 7) seeded through xorshift 13/17/5
    s = 242
    buf = []
    for _ in range(11):
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        buf.append(s)
    p1, p2 = 0, 7

    def nxt():
        nonlocal p1, p2
        x = (buf[p1] + buf[p2]) & 0xFFFFFFFF
        x = ((x << 13) | (x >> 19)) & 0xFFFFFFFF
        buf[p1] = x
        p1 = (p1 - 1) % 11
        p2 = (p2 - 1) % 11
        return x

    for _ in range(11):
        nxt()
    assert [nxt() for _ in range(64)] == [int(v) for v in out]


def test_fw_sort_small_is_stable_insertion(orc):
    L = orc.lib()
    rng = random.Random(5)
    for n in range(0, 16):
        keys = np.array([rng.randint(0, 3) for _ in range(n)], np.float32)
        pay = np.arange(n, dtype=np.int32)
        k2, p2 = keys.copy(), pay.copy()
        L.orc_fw_sort(n, k2, p2)
        order = sorted(range(n), key=lambda i: (keys[i], i))  # stable
        assert list(p2) == order


def test_fw_sort_sorts_and_permutes(orc):
    L = orc.lib()
    rng = np.random.default_rng(9)
    for n in (16, 17, 31, 64, 200, 1000):
        for ties in (2, 50, 100000):
            keys = rng.integers(0, ties, n).astype(np.float32)
            pay = np.arange(n, dtype=np.int32)
            k2, p2 = keys.copy(), pay.copy()
            L.orc_fw_sort(n, k2, p2)
            assert np.all(np.diff(k2) >= 0)
            assert sorted(p2.tolist()) == list(range(n))
            assert np.array_equal(keys[p2], k2)


def _py_fw_sort(keys):
    """Independent transcription of FW::sort (Sort.cpp:63-211) for cross-checking tie order."""
    a = list(range(len(keys)))
    k = list(keys)

    def cmp(i, j):
        return -1 if k[i] < k[j] else (1 if k[i] > k[j] else 0)

    def swp(i, j):
        k[i], k[j] = k[j], k[i]
        a[i], a[j] = a[j], a[i]

    def ins(start, size):
        for i in range(1, size):
            j = start + i - 1
            while j >= start and cmp(j, j + 1) > 0:
                swp(j, j + 1)
                j -= 1

    def part(low, high):
        l, c, h = low, (low + high) >> 1, high - 2
        if cmp(l, h) > 0:
            l, h = h, l
        if cmp(l, c) > 0:
            c = l
        swp(h if cmp(c, h) > 0 else c, high - 1)
        i, j = low - 1, high - 1
        while True:
            i += 1
            while cmp(i, high - 1) < 0:
                i += 1
            j -= 1
            while cmp(j, high - 1) > 0:
                j -= 1
            if i >= j:
                break
            swp(i, j)
        swp(i, high - 1)
        return i

    n = len(keys)
    if n < 2:
        return a
    low, stack = 0, [n]
    while stack:
        high = stack.pop()
        if high - low < 16 or len(stack) + 2 > 32:
            ins(low, high - low)
            low = high + 1
            continue
        i = part(low, high)
        if high - i > 2:
            stack.append(high)
        if i - low > 1:
            stack.append(i)
        else:
            low = i + 1
    return a


def test_fw_sort_tie_order_matches_transcription(orc):
    L = orc.lib()
    rng = np.random.default_rng(11)
    for n in (17, 40, 100, 333):
        keys = rng.integers(0, 4, n).astype(np.float32)
        k2, p2 = keys.copy(), np.arange(n, dtype=np.int32)
        L.orc_fw_sort(n, k2, p2)
        assert list(p2) == _py_fw_sort(keys.tolist())


def _exact_orient(ax, ay, bx, by, px, py):
    A, B = Fraction(ay) - Fraction(by), Fraction(bx) - Fraction(ax)
    Cc = -Fraction(ax) * A - Fraction(ay) * B
    return Fraction(px) * A + Fraction(py) * B + Cc


def test_triangle_cover_vs_exact(orc):
    L = orc.lib()
    rng = np.random.default_rng(21)
    n_in = n_out = 0
    for _ in range(3000):
        r = rng.uniform(-2, 2, (3, 2)).astype(np.float32)
        o = rng.uniform(-2, 2, 2).astype(np.float32)
        d = [_exact_orient(*[float(v) for v in r[(i + 1) % 3]], *[float(v) for v in r[(i + 2) % 3]],
                           float(o[0]), float(o[1])) for i in range(3)]
        # skip cases within rounding distance of an edge (tie rule tested below)
        if min(abs(float(x)) for x in d) < 1e-6:
            continue
        inside = all(x > 0 for x in d) or all(x < 0 for x in d)
        got = L.orc_triangle_covers(r[0].copy(), r[1].copy(), r[2].copy(), o.copy())
        assert bool(got) == inside
        n_in += inside
        n_out += not inside
    assert n_in > 100 and n_out > 100


def test_triangle_tie_rule(orc):
    # A point on an edge counts for exactly one of the two triangles sharing
    # it (top-left rule through acceptEqual).
    L = orc.lib()
    a = np.array([0, 0], np.float32)
    b = np.array([2, 0], np.float32)
    c = np.array([1, 1], np.float32)
    d = np.array([1, -1], np.float32)
    o = np.array([1, 0], np.float32)
    t1 = L.orc_triangle_covers(a, b, c, o)
    t2 = L.orc_triangle_covers(b, a, d, o)
    assert t1 + t2 == 1


def test_dispersion_table_through_build(orc):
    # spp rounding (Util.hpp:64-73) and the spp > 256 failure
    rec = np.zeros((2, 2, 3, 16), np.float32)
    rec[..., 0] = np.arange(2)[None, :, None] + 0.5
    rec[..., 1] = np.arange(2)[:, None, None] + 0.5
    rec[..., 3] = 2.0
    rec[..., 10] = 1.0
    o = orc.Oracle(2, 2, 3, 0.0, 1.0, rec)
    assert o.build().spp_pow2 == 4
    rec = np.zeros((1, 1, 300, 16), np.float32)
    rec[..., 3] = 2.0
    o = orc.Oracle(1, 1, 300, 0.0, 1.0, rec)
    with pytest.raises(RuntimeError):
        o.build()
