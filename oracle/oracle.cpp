/*
 * CPU oracle: restatement of the reference CPU reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY (see tlr_oracle.h).  Built with g++ -O2
 * -ffp-contract=off, no fast-math, so every FP32 operation is IEEE
 * round-to-nearest in the order written here, which is the reference's order.
 *
 * Parity pins (what checks each function; tests/ names in brackets):
 *   hash_bits, ranrot, sobol, sobol2d, halton, lp, to_unit_disk
 *       closed forms / textbook sequences        [test_oracle_primitives]
 *   fw_sort replica                     permutation invariants, exhaustive
 *       small-n equality with an independent transcription, tie behaviour
 *                                                [test_oracle_primitives]
 *   tux_planes (computeTUXPlanes)       containment of the exact trajectory
 *       at ta, tb, tmid and u=+-1; static closed form   [test_oracle_bounds]
 *   merge_bounds (mergeTUXPlanes)       merged plane >= both inputs at the
 *       four (t,u) corners; identity on equal planes     [test_oracle_bounds]
 *   reprojection / bucketing            closed form on hand-built records,
 *       round trip to the sample's own (u,v,t)           [test_oracle_build]
 *   leaf build                          partition invariants (each kept
 *       sample in exactly one leaf, one t per leaf, leaf inside one initial
 *       leaf, sorted by (t,w) inside)                    [test_oracle_build]
 *   hierarchy                           leaf sets equal brute force over all
 *       leaves                                           [test_oracle_build]
 *   gather / clustering                 gathered set equals brute force over
 *       leaves + per-sample distance test; surfaces contiguous
 *                                                        [test_oracle_filter]
 *   triangle coverage                   exact rational point-in-triangle on
 *       random small inputs                              [test_oracle_filter]
 *   tent filter / image                 constant-colour scene reconstructs
 *       to that colour; single static plane scene
 *                                                        [test_oracle_filter]
 *   GPU-variant method (RefGpu)         checkTriangles equals a float64
 *       brute force over triples; reprojection equals the closed form; the
 *       nearest search is nearest over every leaf its bounds reach; 1R first
 *       surface and 2R surfaces equal float64 tent means on a single plane;
 *       near plane first with two planes present         [test_oracle_refgpu]
 *   same_surface                        hand-built bounds: parallel within
 *       and beyond 0.1, crossing in (u,v), one-sided difference, crossing in
 *       time                                             [test_oracle_pins]
 *   leaf surface scan                   hand-built boxes: time split, the
 *       2.0 threshold, crossing motion, a CoC sequence the cocbounds[0]
 *       typo decides                                     [test_oracle_pins]
 *   Filterer::reconstruct control       constructed static scenes with
 *       closed-form outcomes: fetch2R + coverage both ways, no fetch2R with
 *       a four-quadrant first surface, emergency pass -> nearest sample,
 *       3-sample surface merged into the next            [test_oracle_pins]
 *   output pattern composition          recomposed from the pinned
 *       primitives, bit for bit                          [test_oracle_pins]
 *   RefGpu 2R surface split             a 3-point front surface takes the
 *       next leaf in                                     [test_oracle_pins]
 *   shadow filter                       constructed light-space planes:
 *       a query on / 0.02 behind / 0.009 behind the plane, a w-gradient
 *       lighting one side, density 4 doubling the radius, the nearest
 *       fallback lit and shadowed, the extended result   [test_oracle_shadow]
 *   FW inverse, lightSpaceDot,          float64 inverse; the geometric
 *   reconstructDofMotionShadows           normal and cosine to a point
 *       light on a tilted plane; a pinhole scene lit from the camera, with
 *       and without an occluder, in closed form          [test_oracle_shadow]
 *   Each of the last rules is broken on purpose in a mutant build
 *   (ORC_MUTANT below) and the pins are checked to fail [test_oracle_mutants].
 * The reference itself cannot be built here (Win32-only headers), so no
 * function is pinned against reference outputs.
 */
#include "tlr_oracle.h"

// Mutants for the pin tests: built with -DORC_MUTANT=k, the oracle breaks one
// rule on purpose (tests/test_oracle_mutants.py checks that the pins catch
// each); the oracle proper is built with ORC_MUTANT 0.
//   1 sameSurface without its time test      2 ternaryEqual needs equal signs
//   3 leaf scan without the cocbounds[0] typo 4 leaf scan ignores t changes
//   5 fetch2R ignores the quadrant mask       6 no tiny-surface merge
//   7 emergency passes keep the dispersion    8 no nearest-sample fallback
//   9 output times without the jitter         10 GPU-variant 2R surfaces
//                                                split at 3 points
//  11 shadow queries keep the dispersion      12 shadow bias added, not
//     (no sqrt(density) scaling)                 subtracted
//  13 shadow nearest fallback skips the light test
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define ORC_M(k) (ORC_MUTANT == (k))

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Scalar helpers with the reference's semantics.
// ---------------------------------------------------------------------------

// rcp(): src/framework/base/Math.hpp:112  (a ? 1/a : 0)
inline float rcpf(float a) { return (a != 0.0f) ? 1.0f / a : 0.0f; }
// min/max: src/framework/base/Defs.hpp:196-197  ((a<b)?a:b, (a>b)?a:b)
inline float mn(float a, float b) { return (a < b) ? a : b; }
inline float mx(float a, float b) { return (a > b) ? a : b; }
// (int)floor(x) as x86 cvttss2si evaluates it: out-of-range and NaN give
// INT_MIN.  Used for every float->int bucket conversion.
inline int f2i(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
    return (int)f;
}
inline int floor_i(float f) { return f2i(std::floor(f)); }

// ---------------------------------------------------------------------------
// FW::sort replica: src/framework/base/Sort.cpp:35-36,63-165,197-211.
// Unstable quicksort, median-of-3 pivot hidden at high-1, insertion sort
// below 16 elements, explicit stack of 32 exclusive ends.
// ---------------------------------------------------------------------------

template <class Cmp, class Swp>
void fw_insertion(int start, int size, Cmp& cmp, Swp& swp) {
    for (int i = 1; i < size; i++) {
        int j = start + i - 1;
        while (j >= start && cmp(j, j + 1) > 0) {
            swp(j, j + 1);
            j--;
        }
    }
}

template <class Cmp>
int fw_median3(int low, int high, Cmp& cmp) {
    int l = low;
    int c = (low + high) >> 1;
    int h = high - 2;
    if (cmp(l, h) > 0) std::swap(l, h);
    if (cmp(l, c) > 0) c = l;
    return (cmp(c, h) > 0) ? h : c;
}

template <class Cmp, class Swp>
int fw_partition(int low, int high, Cmp& cmp, Swp& swp) {
    swp(fw_median3(low, high, cmp), high - 1);
    int i = low - 1;
    int j = high - 1;
    for (;;) {
        do i++; while (cmp(i, high - 1) < 0);
        do j--; while (cmp(j, high - 1) > 0);
        if (i >= j) break;
        swp(i, j);
    }
    swp(i, high - 1);
    return i;
}

template <class Cmp, class Swp>
void fw_sort(int start, int end, Cmp& cmp, Swp& swp) {
    if (end - start < 2) return;
    int low = start;
    int high = end;
    int stack[32];
    int sp = 0;
    stack[sp++] = high;
    while (sp) {
        high = stack[--sp];
        if (high - low < 16 || sp + 2 > 32) {
            fw_insertion(low, high - low, cmp, swp);
            low = high + 1;
            continue;
        }
        int i = fw_partition(low, high, cmp, swp);
        if (high - i > 2) stack[sp++] = high;
        if (i - low > 1)
            stack[sp++] = i;
        else
            low = i + 1;
    }
}

// ---------------------------------------------------------------------------
// Hash, RANROT-A, QMC sequences.
// ---------------------------------------------------------------------------

// hashBits(a,b): src/framework/base/Hash.hpp:169-183 (Jenkins mix, c seeded
// with the golden-ratio constant).
uint32_t hash_bits(uint32_t a, uint32_t b) {
    uint32_t c = 0x9e3779b9u;
    a -= b; a -= c; a ^= (c >> 13);
    b -= c; b -= a; b ^= (a << 8);
    c -= a; c -= b; c ^= (b >> 13);
    a -= b; a -= c; a ^= (c >> 12);
    b -= c; b -= a; b ^= (a << 16);
    c -= a; c -= b; c ^= (b >> 5);
    a -= b; a -= c; a ^= (c >> 3);
    b -= c; b -= a; b ^= (a << 10);
    c -= a; c -= b; c ^= (b >> 15);
    return c;
}

// RANROT-A: src/framework/base/Random.cpp:52-92.
struct Ranrot {
    int p1, p2;
    uint32_t buf[11];
    explicit Ranrot(uint32_t seed) {
        if (seed == 0) seed--;
        for (int i = 0; i < 11; i++) {
            seed ^= seed << 13;
            seed ^= seed >> 17;
            seed ^= seed << 5;
            buf[i] = seed;
        }
        p1 = 0;
        p2 = 7;
        for (int i = 0; i < 11; i++) next();
    }
    uint32_t next() {
        uint32_t x = buf[p1] + buf[p2];
        x = (x << 13) | (x >> 19);
        buf[p1] = x;
        p1--; if (p1 < 0) p1 += 11;
        p2--; if (p2 < 0) p2 += 11;
        return x;
    }
    // Random::getF32: src/framework/base/Random.hpp:57
    float f32() { return (float)next() * (1.0f / 4294967296.0f); }
};

// Sobol direction numbers generated from Joe-Kuo entries (primitive
// polynomial degree s, coefficients a, initial m): index 3 = x^3+x+1 with
// m = 1,3,1 and index 4 = x^3+x^2+1 with m = 1,1,1, the rows the reference
// tabulates in common/Sobol5.inl:97-128 and :129-160 (the output pattern uses
// index 3, the shadow path's light positions indices 3 and 4).  Dimension 0
// is the van der Corput sequence (v_c = 2^(31-c)); indices 1, 2 are unused
// and left zero.
struct SobolTable {
    uint32_t v[5][32];
    static void gen(uint32_t* V, int s, uint32_t a, const uint32_t* m) {
        for (int c = 0; c < s; c++) V[c] = m[c] << (31 - c);
        for (int c = s; c < 32; c++) {
            uint32_t x = V[c - s] ^ (V[c - s] >> s);
            for (int k = 1; k < s; k++)
                if ((a >> (s - 1 - k)) & 1) x ^= V[c - k];
            V[c] = x;
        }
    }
    SobolTable() {
        for (int c = 0; c < 32; c++) v[0][c] = 1u << (31 - c);
        for (int d = 1; d < 5; d++) for (int c = 0; c < 32; c++) v[d][c] = 0;
        const uint32_t m3[3] = {1, 3, 1}, m4[3] = {1, 1, 1};
        gen(v[3], 3, 1, m3);
        gen(v[4], 3, 2, m4);
    }
};
const SobolTable& sobol_table() { static SobolTable t; return t; }

// sobol(dim,i): common/Util.cpp:255-264
float sobol(int dim, int i) {
    const uint32_t* m = sobol_table().v[dim];
    uint32_t r = 0;
    for (unsigned c = 0; i; i >>= 1, ++c)
        if (i & 1) r ^= m[c];
    return (float)r * (1.f / 4294967296.0f);
}

// sobol2D(i): common/Util.cpp:266-288
void sobol2d(int i, float* out) {
    uint32_t r1 = 0, r2 = 0;
    for (uint32_t v1 = 1u << 31, v2 = 3u << 30; i; i >>= 1) {
        if (i & 1) {
            r1 ^= v1;
            r2 ^= v2 << 1;
        }
        v1 |= v1 >> 1;
        v2 ^= v2 >> 1;
    }
    out[0] = (float)r1 * (1.f / 4294967296.0f);
    out[1] = (float)r2 * (1.f / 4294967296.0f);
}

// hammersley(i,n): common/Util.cpp:229-233
float hammersley(int i, int n) { return (i + 0.5f) / n; }

// halton(base,i): common/Util.cpp:235-251
float halton(int base, int i) {
    float h = 0.f;
    float half = rcpf(float(base));
    while (i) {
        int digit = i % base;
        h = h + digit * half;
        i = (i - digit) / base;
        half = half / base;
    }
    return h;
}

// larcherPillichshammer(i,0): common/Util.cpp:290-297
float lp(int i) {
    uint32_t r = 0;
    for (uint32_t v = 1u << 31; i; i >>= 1, v |= v >> 1)
        if (i & 1) r ^= v;
    return float((double)r / (double)0x100000000LL);
}

// ToUnitDisk (Shirley-Chiu): common/Util.cpp:37-74
void to_unit_disk(float sx, float sy, float* out) {
    const float PI = 3.1415926535897932384626433832795f;
    float phi, r;
    float a = 2 * sx - 1;
    float b = 2 * sy - 1;
    if (a > -b) {
        if (a > b) { r = a; phi = (PI / 4) * (b / a); }
        else       { r = b; phi = (PI / 4) * (2 - (a / b)); }
    } else {
        if (a < b) { r = -a; phi = (PI / 4) * (4 + (b / a)); }
        else {
            r = -b;
            if (b != 0) phi = (PI / 4) * (6 - (a / b));
            else phi = 0;
        }
    }
    out[0] = r * (float)cosf(phi);
    out[1] = r * (float)sinf(phi);
}

// ---------------------------------------------------------------------------
// Time/lens bounds: common/TimeBounds.cpp, Reconstruction.hpp:149-167.
// Plane p = (A,B,C): x = A*t + B*u + C.  Index 0 xmin, 1 xmax, 2 ymin, 3 ymax.
// ---------------------------------------------------------------------------

struct TLB {
    float p[4][3];
    float ts[2];
    TLB() { std::memset(this, 0, sizeof(*this)); }
    bool valid() const { return ts[0] < ts[1]; }
    // TimeLensBounds::evaluate: Reconstruction.hpp:162
    float eval(int k, float t, float uv) const { return t * p[k][0] + p[k][1] * uv + p[k][2]; }
};

// computePerspectivePlanes1D: TimeBounds.cpp:228-332
bool perspective_1d(float x0, float w0, float dx, float dw, float xmin, float xmax,
                    float dest[2][2], float span[2]) {
    if (dx == 0.0f && dw == 0.0f) {
        float xow0 = x0 / w0;
        dest[0][0] = 0.0f; dest[0][1] = xow0;
        dest[1][0] = 0.0f; dest[1][1] = xow0;
        span[0] = 0.f; span[1] = 1.f;
        return w0 > 0.0f;
    }
    float tb0 = 0.0f, tb1 = 1.0f;
    if (dw != 0.0f) {
        float tw0 = -w0 / dw;
        if (tw0 >= 0 && tw0 <= 1) {
            if (w0 < 0) tb0 = tw0;
            else tb1 = tw0;
        }
    }
    bool movesright = dx * w0 > dw * x0;
    bool movesforward = dw > 0;

    float tmin = (xmin * w0 - x0) / (dx - xmin * dw);
    if (w0 + tmin * dw >= 0) {
        if (movesright) tb0 = mx(tb0, tmin);
        else tb1 = mn(tb1, tmin);
    }
    float tmax = (xmax * w0 - x0) / (dx - xmax * dw);
    if (w0 + tmax * dw >= 0) {
        if (!movesright) tb0 = mx(tb0, tmax);
        else tb1 = mn(tb1, tmax);
    }
    if (tb0 > tb1) return false;

    float thalf = 0.5f * (tb0 + tb1);
    float px = (x0 + thalf * dx) / (w0 + thalf * dw);
    float dxdt = dx / (w0 + thalf * dw) - (dw * (x0 + thalf * dx)) / ((w0 + thalf * dw) * (w0 + thalf * dw));
    float xa = (x0 + tb0 * dx) / (w0 + tb0 * dw);
    float xb = (x0 + tb1 * dx) / (w0 + tb1 * dw);
    float pl1[2] = {dxdt, (px - dxdt * thalf)};
    float pl2[2] = {-(xb - xa) / (tb0 - tb1), xa + ((xb - xa) * tb0) / (tb0 - tb1)};
    if (movesright == movesforward) {
        dest[1][0] = pl1[0]; dest[1][1] = pl1[1];
        dest[0][0] = pl2[0]; dest[0][1] = pl2[1];
    } else {
        dest[0][0] = pl1[0]; dest[0][1] = pl1[1];
        dest[1][0] = pl2[0]; dest[1][1] = pl2[1];
    }
    span[0] = tb0; span[1] = tb1;
    return true;
}

// computeTUXPlanes: TimeBounds.cpp:73-147 (with computePerspectivePlanes
// :44-62 inlined).  Leaves out->p untouched (zero) when invalid.
bool tux_planes(const float pos[3], const float mot[3], float bbx0, float bby0, float bbx1,
                float bby1, const float cc[2], TLB& out) {
    out.ts[0] = 0.0f; out.ts[1] = 0.0f;
    float pl[4][2];
    float spx[2], spy[2];
    bool bx = perspective_1d(pos[0], pos[2], mot[0], mot[2], bbx0, bbx1, pl + 0, spx);
    bool by = perspective_1d(pos[1], pos[2], mot[1], mot[2], bby0, bby1, pl + 2, spy);
    if (!bx || !by) return false;
    out.ts[0] = mx(spx[0], spy[0]);
    out.ts[1] = mn(spx[1], spy[1]);
    if (!(out.ts[0] < out.ts[1])) return false;

    float ft[3];
    ft[0] = out.ts[0];
    ft[1] = out.ts[1];
    ft[2] = 0.5f * (out.ts[0] + out.ts[1]);
    float xyw[3][3];
    float ep[3][2];
    float coc[3];
    for (int i = 0; i < 3; i++) {
        xyw[i][0] = pos[0] + ft[i] * mot[0];
        xyw[i][1] = pos[1] + ft[i] * mot[1];
        xyw[i][2] = pos[2] + ft[i] * mot[2];
        float c = rcpf(xyw[i][2]);
        ep[i][0] = xyw[i][0] * c;
        ep[i][1] = xyw[i][1] * c;
    }
    for (int i = 0; i < 3; i++) coc[i] = cc[0] / xyw[i][2] + cc[1];
    float avg = (1.f / 3.f) * (coc[0] + coc[1] + coc[2]);

    out.p[0][0] = pl[0][0]; out.p[0][1] = avg; out.p[0][2] = FLT_MAX;
    out.p[1][0] = pl[1][0]; out.p[1][1] = avg; out.p[1][2] = -FLT_MAX;
    out.p[2][0] = pl[2][0]; out.p[2][1] = avg; out.p[2][2] = FLT_MAX;
    out.p[3][0] = pl[3][0]; out.p[3][1] = avg; out.p[3][2] = -FLT_MAX;
    for (int i = 0; i < 3; i++) {
        float x11 = 1.0f * (ep[i][0] - coc[i]) - pl[0][0] * ft[i] - avg * (-1.0f);
        float x21 = 1.0f * (ep[i][0] + coc[i]) - pl[0][0] * ft[i] - avg * (1.0f);
        float x12 = 1.0f * (ep[i][0] - coc[i]) - pl[1][0] * ft[i] - avg * (-1.0f);
        float x22 = 1.0f * (ep[i][0] + coc[i]) - pl[1][0] * ft[i] - avg * (1.0f);
        float y11 = 1.0f * (ep[i][1] - coc[i]) - pl[2][0] * ft[i] - avg * (-1.0f);
        float y21 = 1.0f * (ep[i][1] + coc[i]) - pl[2][0] * ft[i] - avg * (1.0f);
        float y12 = 1.0f * (ep[i][1] - coc[i]) - pl[3][0] * ft[i] - avg * (-1.0f);
        float y22 = 1.0f * (ep[i][1] + coc[i]) - pl[3][0] * ft[i] - avg * (1.0f);
        out.p[0][2] = mn(mn(out.p[0][2], x11), x21);
        out.p[1][2] = mx(mx(out.p[1][2], x12), x22);
        out.p[2][2] = mn(mn(out.p[2][2], y11), y21);
        out.p[3][2] = mx(mx(out.p[3][2], y12), y22);
    }
    return true;
}

// mergeTUXPlanes(plane1, plane2): TimeBounds.cpp:154-208
void merge_plane(const float a[3], const float b[3], float out[3]) {
    float x1[4], x2[4], m[4];
    x1[0] = -a[1] + a[2];
    x1[1] = a[0] - a[1] + a[2];
    x1[2] = a[0] + a[1] + a[2];
    x1[3] = a[1] + a[2];
    x2[0] = -b[1] + b[2];
    x2[1] = b[0] - b[1] + b[2];
    x2[2] = b[0] + b[1] + b[2];
    x2[3] = b[1] + b[2];
    for (int k = 0; k < 4; k++) m[k] = mx(x1[k], x2[k]);
    float c0 = m[1] + m[2];
    float c1 = m[2] + m[3] - m[0] - m[1];
    float c2 = m[0] + m[1] + m[2] + m[3];
    float A = c0 - 0.5f * c2;
    float B = 0.25f * c1;
    float k0 = -B, k1 = A - B, k2 = A + B, k3 = B;
    float d = m[0] - k0;
    d = mx(d, m[1] - k1);
    d = mx(d, m[2] - k2);
    d = mx(d, m[3] - k3);
    out[0] = A; out[1] = B; out[2] = d;
}

// mergeTUXPlanes(planes1, planes2, ...): TimeBounds.cpp:212-224
void merge_planes(const TLB& a, const TLB& b, TLB& out) {
    float na[3], nb[3], r[3];
    for (int k = 0; k < 4; k++) {
        if (k == 0 || k == 2) {
            for (int j = 0; j < 3; j++) { na[j] = -a.p[k][j]; nb[j] = -b.p[k][j]; }
            merge_plane(na, nb, r);
            for (int j = 0; j < 3; j++) out.p[k][j] = -r[j];
        } else {
            merge_plane(a.p[k], b.p[k], r);
            for (int j = 0; j < 3; j++) out.p[k][j] = r[j];
        }
    }
    out.ts[0] = mn(a.ts[0], b.ts[0]);
    out.ts[1] = mx(a.ts[1], b.ts[1]);
}

// TimeLensBounds(tb0, tb1): Reconstruction.hpp:153-160
TLB merge_tlb(const TLB& a, const TLB& b) {
    TLB r;
    if (a.valid() && b.valid()) merge_planes(a, b, r);
    else if (a.valid()) r = a;
    else if (b.valid()) r = b;
    else { r = TLB(); }
    return r;
}

// ---------------------------------------------------------------------------
// sameSurface: Reconstruction.hpp:428-525
// ---------------------------------------------------------------------------

inline int tcmp(float a, float b, float eps) {
    if (std::fabs(a - b) < eps) return 0;
    else if (a < b) return -1;
    else return 1;
}
inline bool teq2(int a, int b) { return ORC_M(2) ? a == b : (a == 0 || b == 0 || a == b); }
inline bool teq4(int a, int b, int c, int d) {
    int lo = std::min(std::min(std::min(a, b), c), d);
    int hi = std::max(std::max(std::max(a, b), c), d);
    return hi - lo < 2;
}

bool same_surface(const TLB& n0, const TLB& n1, const float uvtmin[3], const float uvtmax[3]) {
    const float T = 1 * 0.1f;
    float tmid = 0.5f * (uvtmin[2] + uvtmax[2]);
    float Xmin00 = n0.eval(0, tmid, uvtmin[0]), Xmin10 = n0.eval(0, tmid, uvtmax[0]);
    float Xmax00 = n0.eval(1, tmid, uvtmin[0]), Xmax10 = n0.eval(1, tmid, uvtmax[0]);
    float Ymin00 = n0.eval(2, tmid, uvtmin[1]), Ymin10 = n0.eval(2, tmid, uvtmax[1]);
    float Ymax00 = n0.eval(3, tmid, uvtmin[1]), Ymax10 = n0.eval(3, tmid, uvtmax[1]);
    float Xmin01 = n1.eval(0, tmid, uvtmin[0]), Xmin11 = n1.eval(0, tmid, uvtmax[0]);
    float Xmax01 = n1.eval(1, tmid, uvtmin[0]), Xmax11 = n1.eval(1, tmid, uvtmax[0]);
    float Ymin01 = n1.eval(2, tmid, uvtmin[1]), Ymin11 = n1.eval(2, tmid, uvtmax[1]);
    float Ymax01 = n1.eval(3, tmid, uvtmin[1]), Ymax11 = n1.eval(3, tmid, uvtmax[1]);
    bool coc = teq2(tcmp(Xmin00, Xmin01, T), tcmp(Xmin10, Xmin11, T)) &&
               teq2(tcmp(Ymin00, Ymin01, T), tcmp(Ymin10, Ymin11, T)) &&
               teq2(tcmp(Xmax00, Xmax01, T), tcmp(Xmax10, Xmax11, T)) &&
               teq2(tcmp(Ymax00, Ymax01, T), tcmp(Ymax10, Ymax11, T));
    if (!coc) return false;
    if (ORC_M(1)) return true;
    float xa0 = n0.eval(0, uvtmin[2], 0.0f), xb0 = n0.eval(0, uvtmax[2], 0.0f);
    float ya0 = n0.eval(2, uvtmin[2], 0.0f), yb0 = n0.eval(2, uvtmax[2], 0.0f);
    float Xa0 = n0.eval(1, uvtmin[2], 0.0f), Xb0 = n0.eval(1, uvtmax[2], 0.0f);
    float Ya0 = n0.eval(3, uvtmin[2], 0.0f), Yb0 = n0.eval(3, uvtmax[2], 0.0f);
    float xa1 = n1.eval(0, uvtmin[2], 0.0f), xb1 = n1.eval(0, uvtmax[2], 0.0f);
    float ya1 = n1.eval(2, uvtmin[2], 0.0f), yb1 = n1.eval(2, uvtmax[2], 0.0f);
    float Xa1 = n1.eval(1, uvtmin[2], 0.0f), Xb1 = n1.eval(1, uvtmax[2], 0.0f);
    float Ya1 = n1.eval(3, uvtmin[2], 0.0f), Yb1 = n1.eval(3, uvtmax[2], 0.0f);
    return teq2(tcmp(xa0, xa1, T), tcmp(xb0, xb1, T)) &&
           teq2(tcmp(Xa0, Xa1, T), tcmp(Xb0, Xb1, T)) &&
           teq2(tcmp(ya0, ya1, T), tcmp(yb0, yb1, T)) &&
           teq2(tcmp(Ya0, Ya1, T), tcmp(Yb0, Yb1, T));
}

// ---------------------------------------------------------------------------
// EdgeFunction (FP64): common/EdgeFunction.cpp:39-45, EdgeFunction.hpp:45-47
// ---------------------------------------------------------------------------

struct Edge {
    double A, B, C;
    bool acc;
    Edge(double x0, double y0, double x1, double y1) {
        A = y0 - y1;
        B = x1 - x0;
        C = -x0 * A - y0 * B;
        acc = (A == 0.) ? (B > 0.) : (A > 0.);
    }
    float eval(double x, double y) const { return (float)(x * A + y * B + C); }
    bool test(float d) const { return d > 0.f ? true : (d == 0.f && acc); }
};

bool triangle_covers(float r0x, float r0y, float r1x, float r1y, float r2x, float r2y, float ox,
                     float oy) {
    Edge e0(r1x, r1y, r2x, r2y);
    Edge e1(r2x, r2y, r0x, r0y);
    Edge e2(r0x, r0y, r1x, r1y);
    bool p0 = e0.test(e0.eval(ox, oy));
    bool p1 = e1.test(e1.eval(ox, oy));
    bool p2 = e2.test(e2.eval(ox, oy));
    return (p0 && p1 && p2) || (!p0 && !p1 && !p2);
}

inline float len2(float dx, float dy) { return std::sqrt(dx * dx + dy * dy); }

// ---------------------------------------------------------------------------
// Data of one context.
// ---------------------------------------------------------------------------

struct RS {            // reprojected sample (TreeGather::Sample subset)
    float x, y, w, t, mvx, mvy, mvw, r, g, b;
    float wgx, wgy;    // dw/dx, dw/dy of the input sample (Reconstruction.hpp:80)
    float density;     // LIGHTSPACE_DENSITY channel, 1 without it (Reconstruction.cpp:547)
    int src;
};

struct INode { int x0, y0, x1, y1, c0, c1, n; };

struct Leaf {
    int s0, s1;
    TLB b;
};

struct Inner {
    int c0, c1;        // child ids; leaf ids tagged with 0x80000000
    TLB b;
};

const int LEAF_TAG = (int)0x80000000;

struct Recon;

}  // namespace

struct orc_ctx {
    int W = 0, H = 0, spp = 0;
    float C0 = 0, C1 = 0;
    int N = 0;
    std::vector<float> rec;
    std::string err;

    // build
    bool built = false;
    int spp_pow2 = 1;
    float cc[2] = {0, 0};
    std::vector<float> density;       // per input sample (empty: 1)
    bool coc_override = false;        // overrideRefocusDistance's coefficients
    float coc_ov[2] = {0, 0};
    std::vector<RS> repro;            // kept, input order
    float bbmin[2], bbmax[2], scale = 0;
    int bbmin_i[2], gw = 0, gh = 0, max_hits = 0, num_bucketed = 0;
    std::vector<int> cell_of;         // per kept sample
    std::vector<int> cell_start;      // gw*gh+1
    std::vector<int> cell_items;      // kept indices in (cell, input) order
    std::vector<INode> inodes;
    std::vector<int> ileaf_nodes;     // initial leaves in DFS order
    std::vector<int> ileaf_offset;
    std::vector<RS> samples;          // layout order
    std::vector<Leaf> leaves;         // layout order
    // hierarchy over valid leaves
    std::vector<int> leaf_order;
    std::vector<Inner> inner;
    int root = -1;                    // internal id, leaf id | tag, or -1
};

namespace {

// Sample::reprojectToUVTCenter: Reconstruction.hpp:93-121
bool reproject_center(const float* r, const float cc0[2], RS& s) {
    float x = r[0], y = r[1], w = r[3], u = r[4], v = r[5], t = r[6];
    float mvx = r[11], mvy = r[12], mvw = r[13];
    float coc = cc0[0] / w + cc0[1];
    float px = x - coc * u;
    float py = y - coc * v;
    float Px = px * w, Py = py * w, Pw = 1.0f * w;
    float ax = Px - t * mvx, ay = Py - t * mvy, aw = Pw - t * mvw;
    float nt = 0.5f;
    float X = ax + nt * mvx, Y = ay + nt * mvy, Wt = aw + nt * mvw;
    if (Wt <= 0.f) {
        nt = 0.0f;
        X = ax + nt * mvx; Y = ay + nt * mvy; Wt = aw + nt * mvw;
    }
    if (Wt <= 0.f) {
        nt = 1.0f;
        X = ax + nt * mvx; Y = ay + nt * mvy; Wt = aw + nt * mvw;
    }
    if (Wt <= 0.f) return false;
    float c = rcpf(Wt);
    s.w = Wt;
    s.x = X * c;
    s.y = Y * c;
    s.t = nt;
    s.mvx = mvx; s.mvy = mvy; s.mvw = mvw;
    s.r = r[7]; s.g = r[8]; s.b = r[9];
    s.wgx = r[14]; s.wgy = r[15];
    s.density = 1.0f;
    return true;
}

int dispersion_index_ok(int spp) {
    switch (spp) {
        case 256: case 128: case 64: case 32: case 16: case 8: case 4: case 2: case 1: return 1;
        default: return 0;
    }
}

// Dispersion table: ReconstructionOur.cpp:63-76
float dispersion_for(int spp) {
    switch (spp) {
        case 256: return 0.14f;
        case 128: return 0.18f;
        case 64: return 0.27f;
        case 32: return 0.37f;
        case 16: return 0.50f;
        case 8: return 0.80f;
        case 4: return 1.10f;
        case 2: return 1.34f;
        default: return 2.00f;
    }
}

// roundUpToNearestPowerOfTwo: common/Util.hpp:64-73
uint32_t round_pow2(uint32_t v) {
    v--;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}

// buildInitialRecursive: ReconstructionTreeBuilder.cpp:34-96, with counts
// from a summed-area table.
struct SAT {
    int gw, gh;
    std::vector<long long> s;  // (gw+1)*(gh+1)
    long long rect(int x0, int y0, int x1, int y1) const {
        return s[(size_t)y1 * (gw + 1) + x1] - s[(size_t)y0 * (gw + 1) + x1] -
               s[(size_t)y1 * (gw + 1) + x0] + s[(size_t)y0 * (gw + 1) + x0];
    }
};

int build_initial(orc_ctx* c, const SAT& sat, int x0, int x1, int y0, int y1) {
    const int dx = x1 - x0, dy = y1 - y0;
    int idx = (int)c->inodes.size();
    INode nd;
    nd.x0 = x0; nd.x1 = x1; nd.y0 = y0; nd.y1 = y1; nd.c0 = -1; nd.c1 = -1; nd.n = 0;
    c->inodes.push_back(nd);
    bool leaf = std::max(dx, dy) == 1;
    if (!leaf && dx * dy <= 64) leaf = sat.rect(x0, y0, x1, y1) <= 48;
    int c0 = -1, c1 = -1, n = 0;
    if (leaf) {
        n = (int)sat.rect(x0, y0, x1, y1);
    } else {
        if (dx > dy) {
            int xm = (x1 + x0) / 2;
            c0 = build_initial(c, sat, x0, xm, y0, y1);
            c1 = build_initial(c, sat, xm, x1, y0, y1);
        } else {
            int ym = (y1 + y0) / 2;
            c0 = build_initial(c, sat, x0, x1, y0, ym);
            c1 = build_initial(c, sat, x0, x1, ym, y1);
        }
        n = c->inodes[c0].n + c->inodes[c1].n;
    }
    c->inodes[idx].c0 = c0;
    c->inodes[idx].c1 = c1;
    c->inodes[idx].n = n;
    return idx;
}

void collect_initial_leaves(orc_ctx* c, int node, int& offset) {
    const INode& nd = c->inodes[node];
    if (nd.c0 == -1) {
        c->ileaf_nodes.push_back(node);
        c->ileaf_offset.push_back(offset);
        offset += nd.n;
        return;
    }
    collect_initial_leaves(c, nd.c0, offset);
    collect_initial_leaves(c, nd.c1, offset);
}

// Leaf branch of buildRecursive: ReconstructionTreeBuilder.cpp:127-383.
// Writes samples[offset .. offset+n) and appends leaves to `out`.
struct LeafScratch {
    std::vector<RS> cand;
    std::vector<TLB> bounds;
    std::vector<float> ba, bb, ba1, bb1;   // 2 floats each
    std::vector<int> surf;
};

// Surface scan of one initial leaf's candidates (sorted by (t, w)):
// ReconstructionTreeBuilder.cpp:222-340, keeping the cocbounds[0] typo at
// :334,336.  tof(k): candidate k's time; ba/bb/ba1/bb1: its screen boxes
// (2 floats each) at t=0 and t=1; surf[k] receives its surface id.
template <class TOf>
void surface_scan(int n, TOf tof, const float* ba, const float* bb, const float* ba1, const float* bb1, int* surf) {
    if (n <= 0) return;
    int sb = 0, cur = 0;
    surf[0] = 0;
    bool reset = true;
    float cb0[2] = {0, 0}, cb1[2] = {0, 0}, mvmin[2] = {0, 0}, mvmax[2] = {0, 0};
    float st = 0.0f;
    for (int k = 1; k < n; ++k) {
        if (reset) {
            cb0[0] = cb0[1] = 0.5f * (ba[2 * sb] - bb[2 * sb]);
            cb1[0] = cb1[1] = 0.5f * (ba1[2 * sb] - bb1[2 * sb]);
            mvmin[0] = 0.5f * (ba1[2 * sb] + bb1[2 * sb]) - 0.5f * (ba[2 * sb] + bb[2 * sb]);
            mvmin[1] = 0.5f * (ba1[2 * sb + 1] + bb1[2 * sb + 1]) - 0.5f * (ba[2 * sb + 1] + bb[2 * sb + 1]);
            mvmax[0] = mvmin[0]; mvmax[1] = mvmin[1];
            st = tof(sb);
            reset = false;
        }
        bool conflict = false;
        bool early = false;
        const float T = 1 * 2.0f;
        if (!ORC_M(4) && st != tof(k)) {
            sb = k;
            surf[k] = ++cur;
            reset = true;
            continue;
        }
        float c0 = 0.5f * (ba[2 * k] - bb[2 * k]);
        float c1 = 0.5f * (ba1[2 * k] - bb1[2 * k]);
        float mv[2];
        mv[0] = 0.5f * (ba1[2 * k] + bb1[2 * k]) - 0.5f * (ba[2 * k] + bb[2 * k]);
        mv[1] = 0.5f * (ba1[2 * k + 1] + bb1[2 * k + 1]) - 0.5f * (ba[2 * k + 1] + bb[2 * k + 1]);
        float dmv_a = mx(std::fabs(mvmin[0] - mv[0]), std::fabs(mvmin[1] - mv[1]));
        float dmv_b = mx(std::fabs(mvmax[0] - mv[0]), std::fabs(mvmax[1] - mv[1]));
        if (mx(std::fabs(c0 - cb0[0]), std::fabs(c0 - cb0[1])) < 0.5f * T &&
            mx(std::fabs(c1 - cb1[0]), std::fabs(c1 - cb1[1])) < 0.5f * T &&
            mx(dmv_a, dmv_b) < T)
            early = true;
        if (!early) {
            for (int j = sb; j < k; ++j) {
                bool bX = teq4(tcmp(ba[2 * j], ba[2 * k], T), tcmp(bb[2 * j], bb[2 * k], T),
                               tcmp(ba1[2 * j], ba1[2 * k], T), tcmp(bb1[2 * j], bb1[2 * k], T));
                bool bY = teq4(tcmp(ba[2 * j + 1], ba[2 * k + 1], T), tcmp(bb[2 * j + 1], bb[2 * k + 1], T),
                               tcmp(ba1[2 * j + 1], ba1[2 * k + 1], T),
                               tcmp(bb1[2 * j + 1], bb1[2 * k + 1], T));
                if (!(bX && bY)) {
                    sb = k;
                    surf[k] = ++cur;
                    conflict = true;
                    reset = true;
                    break;
                }
            }
        }
        if (!conflict) {
            surf[k] = cur;
            cb0[0] = mn(c0, cb0[0]);
            cb0[1] = ORC_M(3) ? mx(c0, cb0[1]) : mx(c0, cb0[0]);
            cb1[0] = mn(c1, cb1[0]);
            cb1[1] = ORC_M(3) ? mx(c1, cb1[1]) : mx(c1, cb1[0]);
            mvmin[0] = mn(mvmin[0], mv[0]); mvmin[1] = mn(mvmin[1], mv[1]);
            mvmax[0] = mx(mvmax[0], mv[0]); mvmax[1] = mx(mvmax[1], mv[1]);
        }
    }

}

void build_leaf_cell(orc_ctx* c, int inode, int offset, LeafScratch& sc, std::vector<Leaf>& out) {
    const INode& in = c->inodes[inode];
    std::vector<RS>& cand = sc.cand;
    cand.clear();
    for (int y = in.y0; y < in.y1; y++)
        for (int x = in.x0; x < in.x1; x++) {
            int cell = y * c->gw + x;
            for (int k = c->cell_start[cell]; k < c->cell_start[cell + 1]; k++)
                cand.push_back(c->repro[c->cell_items[k]]);
        }
    const int n = (int)cand.size();
    if (n == 0) return;

    // sampleCompareFuncInc: ReconstructionTreeBuilder.cpp:99-117
    auto cmp = [&](int a, int b) -> int {
        const RS& A = cand[a];
        const RS& B = cand[b];
        if (A.t < B.t) return -1;
        else if (A.t > B.t) return 1;
        else {
            if (A.w < B.w) return -1;
            else if (A.w > B.w) return 1;
            else return 0;
        }
    };
    auto swp = [&](int a, int b) { std::swap(cand[a], cand[b]); };
    fw_sort(0, n, cmp, swp);

    sc.bounds.assign(n, TLB());
    sc.ba.assign(2 * n, 0.f); sc.bb.assign(2 * n, 0.f);
    sc.ba1.assign(2 * n, 0.f); sc.bb1.assign(2 * n, 0.f);
    sc.surf.assign(n, 0);
    const float W = (float)c->W, H = (float)c->H;
    for (int i = 0; i < n; i++) {
        const RS& s = cand[i];
        float pos[3] = {s.x * s.w - s.t * s.mvx, s.y * s.w - s.t * s.mvy, s.w - s.t * s.mvw};
        float mot[3] = {s.mvx, s.mvy, s.mvw};
        tux_planes(pos, mot, 0.0f, 0.0f, W, H, c->cc, sc.bounds[i]);
        if (sc.bounds[i].valid()) {
            float ta = sc.bounds[i].ts[0];
            float tb = sc.bounds[i].ts[1];
            float ax = s.x * s.w + (ta - s.t) * s.mvx;
            float ay = s.y * s.w + (ta - s.t) * s.mvy;
            float aw = s.w + (ta - s.t) * s.mvw;
            float bx = s.x * s.w + (tb - s.t) * s.mvx;
            float by = s.y * s.w + (tb - s.t) * s.mvy;
            float bw = s.w + (tb - s.t) * s.mvw;
            float ca = rcpf(aw), cb = rcpf(bw);
            float xta = ax * ca, yta = ay * ca;
            float xtb = bx * cb, ytb = by * cb;
            float oodt = 1.0f / (tb - ta);
            float dxdt = (xtb - xta) * oodt;
            float dydt = (ytb - yta) * oodt;
            float x0 = xta - ta * dxdt, y0 = yta - ta * dydt;
            float x1 = xtb + (1.0f - tb) * dxdt, y1 = ytb + (1.0f - tb) * dydt;
            float cocta = c->cc[0] / aw + c->cc[1];
            float coctb = c->cc[0] / bw + c->cc[1];
            sc.ba[2 * i] = x0 + cocta;  sc.ba[2 * i + 1] = y0 + cocta;
            sc.bb[2 * i] = x0 - cocta;  sc.bb[2 * i + 1] = y0 - cocta;
            sc.ba1[2 * i] = x1 + coctb; sc.ba1[2 * i + 1] = y1 + coctb;
            sc.bb1[2 * i] = x1 - coctb; sc.bb1[2 * i + 1] = y1 - coctb;
        }
    }

    surface_scan(n, [&](int k) { return cand[k].t; }, sc.ba.data(), sc.bb.data(), sc.ba1.data(), sc.bb1.data(),
                 sc.surf.data());

    // Leaf emission: ReconstructionTreeBuilder.cpp:345-378
    int curIdx = offset;
    for (int i = 0; i < n;) {
        int j;
        for (j = i + 1; j < n && sc.surf[i] == sc.surf[j]; ++j) {}
        Leaf lf;
        lf.s0 = curIdx;
        for (int k = i; k < j; ++k) {
            c->samples[curIdx++] = cand[k];
            lf.b = merge_tlb(sc.bounds[k], lf.b);
        }
        lf.s1 = curIdx;
        out.push_back(lf);
        i = j;
    }
}

// Hierarchy over the valid leaves: a balanced binary tree of recursive
// median splits (this build's own rule; it replaces emitNodes,
// ReconstructionTreeBuilder.cpp:409-451, and which leaves a query finds does
// not depend on it).  Each node splits its leaves at the median of the
// feature of widest extent among a leaf's centre in x at the (t, u) corners
// (0,-1), (1,-1), (1,1), (0,1) and its centre in y at the same (t, v)
// corners.  The order inside a node is a stable sort on that feature
// quantised to kd_qbits(level) bits; internal node g is the one whose split
// falls between sorted positions g and g+1.
constexpr int kKdFeat = 8;
struct KdFeat {
    float f[kKdFeat];
};
// lower/upper x and y bounds of a leaf at the four corners
void kd_corners(const TLB& b, float xl[4], float xh[4], float yl[4], float yh[4]) {
    const float ct[4] = {0.0f, 1.0f, 1.0f, 0.0f};
    const float cu[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
    for (int c = 0; c < 4; c++) {
        xl[c] = b.eval(0, ct[c], cu[c]);
        xh[c] = b.eval(1, ct[c], cu[c]);
        yl[c] = b.eval(2, ct[c], cu[c]);
        yh[c] = b.eval(3, ct[c], cu[c]);
    }
}
KdFeat kd_feature(const TLB& b) {
    float xl[4], xh[4], yl[4], yh[4];
    kd_corners(b, xl, xh, yl, yh);
    KdFeat k;
    for (int c = 0; c < 4; c++) {
        k.f[c] = 0.5f * (xl[c] + xh[c]);
        k.f[4 + c] = 0.5f * (yl[c] + yh[c]);
    }
    return k;
}
inline int kd_qbits(int level) { return std::min(16, 32 - level); }
inline uint32_t kd_quant(float v, float lo, float hi, int bits) {
    if (!(hi > lo)) return 0;
    const uint32_t m = (1u << bits) - 1u;
    const float sc = (float)m / (hi - lo);
    const float f = (v - lo) * sc;
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)m) return m;
    return (uint32_t)f;
}

void kd_order(std::vector<int>& ord, const std::vector<KdFeat>& F, int lo, int hi, int level) {
    if (hi - lo < 2) return;
    float mn[kKdFeat], mxv[kKdFeat];
    for (int d = 0; d < kKdFeat; d++) {
        mn[d] = F[ord[lo]].f[d];
        mxv[d] = mn[d];
    }
    for (int k = lo + 1; k < hi; k++)
        for (int d = 0; d < kKdFeat; d++) {
            mn[d] = std::min(mn[d], F[ord[k]].f[d]);
            mxv[d] = std::max(mxv[d], F[ord[k]].f[d]);
        }
    int best = 0;
    float bext = mxv[0] - mn[0];
    for (int d = 1; d < kKdFeat; d++)
        if (mxv[d] - mn[d] > bext) {
            bext = mxv[d] - mn[d];
            best = d;
        }
    const int bits = kd_qbits(level);
    std::vector<uint32_t> q(hi - lo);
    for (int k = lo; k < hi; k++) q[k - lo] = kd_quant(F[ord[k]].f[best], mn[best], mxv[best], bits);
    std::vector<int> idx(hi - lo);
    for (int k = 0; k < hi - lo; k++) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return q[a] < q[b]; });
    std::vector<int> tmp(hi - lo);
    for (int k = 0; k < hi - lo; k++) tmp[k] = ord[lo + idx[k]];
    std::copy(tmp.begin(), tmp.end(), ord.begin() + lo);
    const int mid = (lo + hi) / 2;
    kd_order(ord, F, lo, mid, level + 1);
    kd_order(ord, F, mid, hi, level + 1);
}

// Internal node g <-> range [lo,hi) whose split mid-1 == g.
void kd_link(orc_ctx* c, int lo, int hi, int n) {
    const int mid = (lo + hi) / 2;
    const int g = mid - 1;
    Inner& nd = c->inner[g];
    nd.c0 = (mid - lo == 1) ? (c->leaf_order[lo] | LEAF_TAG) : ((lo + mid) / 2 - 1);
    nd.c1 = (hi - mid == 1) ? (c->leaf_order[mid] | LEAF_TAG) : ((mid + hi) / 2 - 1);
    if (mid - lo > 1) kd_link(c, lo, mid, n);
    if (hi - mid > 1) kd_link(c, mid, hi, n);
    const TLB& a = (nd.c0 & LEAF_TAG) ? c->leaves[nd.c0 & ~LEAF_TAG].b : c->inner[nd.c0].b;
    const TLB& b = (nd.c1 & LEAF_TAG) ? c->leaves[nd.c1 & ~LEAF_TAG].b : c->inner[nd.c1].b;
    nd.b = merge_tlb(a, b);
}

void build_hierarchy(orc_ctx* c) {
    c->leaf_order.clear();
    c->inner.clear();
    c->root = -1;
    std::vector<int> ids;
    for (int i = 0; i < (int)c->leaves.size(); i++)
        if (c->leaves[i].b.valid()) ids.push_back(i);
    const int n = (int)ids.size();
    if (n == 0) return;
    std::vector<KdFeat> F(n);
    for (int k = 0; k < n; k++) F[k] = kd_feature(c->leaves[ids[k]].b);
    std::vector<int> ord(n);
    for (int k = 0; k < n; k++) ord[k] = k;
    kd_order(ord, F, 0, n, 0);
    c->leaf_order.resize(n);
    for (int k = 0; k < n; k++) c->leaf_order[k] = ids[ord[k]];
    if (n == 1) {
        c->root = c->leaf_order[0] | LEAF_TAG;
        return;
    }
    c->inner.assign(n - 1, Inner());
    kd_link(c, 0, n, n);
    c->root = n / 2 - 1;
}

// Node::intersect: Reconstruction.hpp:204-216
inline bool node_hit(const TLB& b, float ox, float oy, float ou, float ov, float ot, float R) {
    float mnx = b.eval(0, ot, ou) - R;
    float mny = b.eval(2, ot, ov) - R;
    float mxx = b.eval(1, ot, ou) + R;
    float mxy = b.eval(3, ot, ov) + R;
    return (ox >= mnx && oy >= mny && ox <= mxx && oy <= mxy);
}

struct OSample {
    float x, y, u, v, t;
    float w = 0.0f;    // shadow queries: the query's own w (ReconstructionOur.cpp:289)
};

long long g_nodes_tested = 0;  // instrumentation: hierarchy nodes tested

void find_leaves(const orc_ctx* c, const OSample& o, float R, std::vector<int>& out) {
    out.clear();
    if (c->root == -1) return;
    std::vector<int> st;
    st.push_back(c->root);
    while (!st.empty()) {
        int id = st.back();
        st.pop_back();
        g_nodes_tested++;
        if (id & LEAF_TAG) {
            int lf = id & ~LEAF_TAG;
            if (node_hit(c->leaves[lf].b, o.x, o.y, o.u, o.v, o.t, R)) out.push_back(lf);
        } else {
            const Inner& nd = c->inner[id];
            if (!nd.b.valid()) continue;
            if (node_hit(nd.b, o.x, o.y, o.u, o.v, o.t, R)) {
                st.push_back(nd.c0);
                st.push_back(nd.c1);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Filter: ReconstructionOur.cpp:50-511, Reconstruction.cpp:642-720
// ---------------------------------------------------------------------------

struct RSamp {          // ReconSample: Reconstruction.hpp:124-132
    float x, y;
    int index;
    int rpos;
    float key;
};

struct Surface {
    std::vector<RSamp> s;
    float minDist;
    int minIndex;
    int quad;
    void clear() { s.clear(); minDist = FLT_MAX; minIndex = -1; quad = 0; }
};

enum { XPOS = 1, YPOS = 2, XNEG = 4, YNEG = 8 };

struct StatsAcc {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double count = 0;
};

struct Filt {
    const orc_ctx* c;
    std::vector<int> leafs;
    std::vector<std::pair<float, int>> keyed;
    std::vector<Surface> surf;
    int nsurf = 0;
    // debug hooks
    std::vector<int>* dbg_leaves = nullptr;
    std::vector<std::pair<int, int>>* dbg_samples = nullptr;

    Surface& add_surface() {
        if ((int)surf.size() <= nsurf) surf.emplace_back();
        Surface& s = surf[nsurf++];
        s.clear();
        return s;
    }

    // collectInputSamples2: ReconstructionOur.cpp:361-511
    int collect(const OSample& o, float R, StatsAcc& st) {
        find_leaves(c, o, R, leafs);
        st.v[0] += (double)leafs.size();
        keyed.clear();
        for (int lf : leafs) {
            const RS& s = c->samples[c->leaves[lf].s0];
            float key = s.w + (o.t - s.t) * s.mvw;
            keyed.push_back({key, lf});
        }
        // Front-to-back by key; ties by leaf id (canonical order).
        std::sort(keyed.begin(), keyed.end(), [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
            if (a.first < b.first) return true;
            if (b.first < a.first) return false;
            return a.second < b.second;
        });
        if (dbg_leaves) {
            dbg_leaves->clear();
            for (auto& k : keyed) dbg_leaves->push_back(k.second);
        }
        if (dbg_samples) dbg_samples->clear();
        nsurf = 0;
        int total = 0;
        int sstart = 0;
        const float RANGE = 1.0f / std::sqrt(128.0f);
        float umin[3] = {o.u - RANGE, o.v - RANGE, o.t - 0.5f * RANGE};
        float umax[3] = {o.u + RANGE, o.v + RANGE, o.t + 0.5f * RANGE};
        const float R2 = R * R;
        for (int li = 0; li < (int)keyed.size(); li++) {
            bool first = true;
            const Leaf& lf = c->leaves[keyed[li].second];
            st.v[1] += (double)(lf.s1 - lf.s0);
            for (int i = lf.s0; i < lf.s1; i++) {
                const RS& s = c->samples[i];
                float x, y, w;
                if (s.mvx != 0.0f || s.mvy != 0.0f || s.mvw != 0.0f) {
                    float Px = s.x * s.w + (o.t - s.t) * s.mvx;
                    float Py = s.y * s.w + (o.t - s.t) * s.mvy;
                    float Pw = 1.0f * s.w + (o.t - s.t) * s.mvw;
                    w = Pw;
                    float cr = rcpf(Pw);
                    float coc = c->cc[0] / w + c->cc[1];
                    x = Px * cr + coc * o.u;
                    y = Py * cr + coc * o.v;
                } else {
                    w = s.w;
                    float coc = c->cc[0] / s.w + c->cc[1];
                    x = s.x + coc * o.u;
                    y = s.y + coc * o.v;
                }
                float dx = x - o.x, dy = y - o.y;
                float d2 = dx * dx + dy * dy;
                if (d2 <= R2) {
                    RSamp r;
                    r.x = x; r.y = y;
                    r.rpos = (dx >= 0 ? XPOS : XNEG) | (dy >= 0 ? YPOS : YNEG);
                    r.key = w;
                    r.index = i;
                    if (first) {
                        first = false;
                        bool same = true;
                        for (int j = sstart; j < li && same; ++j)
                            same = same_surface(lf.b, c->leaves[keyed[j].second].b, umin, umax);
                        if (!same || nsurf == 0) {
                            sstart = li;
                            add_surface();
                        }
                    }
                    Surface& sf = surf[nsurf - 1];
                    sf.s.push_back(r);
                    sf.quad |= 1 << (r.rpos & (XPOS | YPOS));
                    float d = std::sqrt(d2);
                    if (d < sf.minDist) {
                        sf.minDist = d;
                        sf.minIndex = (int)sf.s.size() - 1;
                    }
                    if (dbg_samples) dbg_samples->push_back({i, nsurf - 1});
                    total++;
                }
            }
        }
        st.v[3] += total;
        return total;
    }

    struct Result {
        float color[4];
        int fetch2r, nsurf1, n1, nfinal, emerg, found_surface, nearest;
        float disp;
    };

    // Does a gathered sample (w = key at the output's t, w-gradient wg, at
    // offset (dx, dy) = o.xy - its xy) light the shadow query?
    // ReconstructionOur.cpp:289 with SHADOW_BIAS 0.01 (:46); FW's dot sums
    // from zero (Math.hpp:182).
    static bool lit(const OSample& o, float dx, float dy, float key, float wgx, float wgy) {
        float d = 0.0f;
        d += dx * wgx;
        d += dy * wgy;
        const float lhs = ORC_M(12) ? o.w + 0.01f : o.w - 0.01f;
        return lhs <= key + d;
    }

    // Filterer::Result's xy, w, wg and density (Reconstruction.hpp:324-334):
    // of the nearest gathered sample, at the output's (u,v,t)
    struct Ext {
        float x, y, w, density, wgx, wgy;
    };

    // Filterer::reconstruct: ReconstructionOur.cpp:50-359.  shadow: the query
    // is a light-space point with its own w and density (reconstructShadow).
    Result reconstruct(const OSample& o, const orc_recon_params* p, StatsAcc& st, bool shadow = false,
                       float density = 1.0f, Ext* ext = nullptr) {
        Result res;
        std::memset(&res, 0, sizeof(res));
        if (ext) std::memset(ext, 0, sizeof(*ext));   // Result::clear, xy zero-constructed
        float disp = dispersion_for(c->spp_pow2);
        // non-uniform view sample density in light space: ReconstructionOur.cpp:78-80
        if (shadow && !ORC_M(11)) disp *= std::sqrt(density);
        int n1 = collect(o, disp, st);
        res.nsurf1 = nsurf;
        res.n1 = n1;
        int n2 = 0, nN = 0;
        const bool f2 = nsurf == 0 || ((ORC_M(5) || surf[0].quad != 0xF) && nsurf > 1);
        if (f2) {
            n2 = collect(o, disp * 2, st);
            if (n2 == 0) {
                st.v[7] += 1;
                nN = n2;
                for (int k = 0; k < 3 && !nN; k++) {
                    if (!ORC_M(7)) disp *= 2.f;
                    nN = collect(o, disp * 2, st);
                    res.emerg++;
                }
            }
        }
        st.v[5] += f2 ? 1 : 0;
        st.v[2] += n1;
        res.fetch2r = f2;
        res.nfinal = std::max(std::max(n1, n2), nN);

        int numSurfaces = nsurf;
        if (f2 && !ORC_M(6))
            for (int si = 0; si < nsurf - 1; si++) {
                Surface& src = surf[si];
                if ((int)src.s.size() < 4) {
                    Surface& dst = surf[si + 1];
                    dst.s.insert(dst.s.end(), src.s.begin(), src.s.end());
                    src.s.clear();
                    if (src.minDist < dst.minDist) {
                        dst.minDist = src.minDist;
                        dst.minIndex = src.minIndex;
                    }
                    numSurfaces--;
                }
            }
        st.v[4] += numSurfaces;
        st.v[6] += (numSurfaces != nsurf) ? 1 : 0;

        bool found = false;
        float col[4] = {0, 0, 0, 0};
        float nearestDist = FLT_MAX;
        int nI0 = -1, nI1 = 1;
        res.found_surface = -1;
        for (int si = 0; si < nsurf && !found; si++) {
            const Surface& sf = surf[si];
            const std::vector<RSamp>& in = sf.s;
            const int ns = (int)in.size();
            if (ns == 0) continue;
            if (sf.minDist < nearestDist) {
                nearestDist = sf.minDist;
                nI0 = si; nI1 = sf.minIndex;
            }
            if (!f2 || sf.minDist < 0.f || si == nsurf - 1) found = true;
            if (ns >= 200) found = true;
            for (int m = 0; m < ns - 2 && !found; m++)
                for (int k = m + 1; k < ns - 1 && !found; k++)
                    for (int l = k + 1; l < ns && !found; l++) {
                        const RSamp& r0 = in[m];
                        const RSamp& r1 = in[k];
                        const RSamp& r2 = in[l];
                        int rp = r0.rpos | r1.rpos | r2.rpos;
                        if (rp != (XPOS | XNEG | YPOS | YNEG)) continue;
                        if (len2(r0.x - r1.x, r0.y - r1.y) > 2 * disp ||
                            len2(r0.x - r2.x, r0.y - r2.y) > 2 * disp ||
                            len2(r1.x - r2.x, r1.y - r2.y) > 2 * disp)
                            continue;
                        if (len2(r0.x - o.x, r0.y - o.y) > disp && len2(r1.x - o.x, r1.y - o.y) > disp &&
                            len2(r2.x - o.x, r2.y - o.y) > disp)
                            continue;
                        found = triangle_covers(r0.x, r0.y, r1.x, r1.y, r2.x, r2.y, o.x, o.y);
                    }
            if (found) {
                col[0] = col[1] = col[2] = col[3] = 0;
                for (int k = 0; k < ns; k++) {
                    const RSamp& r = in[k];
                    const RS& s = c->samples[r.index];
                    float dx = o.x - r.x, dy = o.y - r.y;
                    float wgt = 1 - std::sqrt(dx * dx + dy * dy) / disp;
                    if (wgt <= 0) continue;
                    float sc[4] = {s.r, s.g, s.b, 1.0f};
                    // shadows: per-input-sample thresholding, ReconstructionOur.cpp:283-291
                    if (shadow && !lit(o, dx, dy, r.key, s.wgx, s.wgy)) sc[0] = sc[1] = sc[2] = 0.0f;
                    col[0] += wgt * sc[0];
                    col[1] += wgt * sc[1];
                    col[2] += wgt * sc[2];
                    col[3] += wgt * sc[3];
                }
                float oow = rcpf(col[3]);
                for (int k = 0; k < 4; k++) col[k] *= oow;
                nI0 = si; nI1 = sf.minIndex;
                res.found_surface = si;
            }
        }
        res.disp = disp;
        if (nI0 == -1) {
            res.color[0] = res.color[1] = res.color[2] = res.color[3] = 0;
            return res;
        }
        const RSamp& na = surf[nI0].s[nI1];
        const RS& ns = c->samples[na.index];
        // w, wg and density from the nearest sample: ReconstructionOur.cpp:325-331
        if (ext) {
            ext->x = na.x; ext->y = na.y; ext->w = na.key;
            ext->density = ns.density; ext->wgx = ns.wgx; ext->wgy = ns.wgy;
        }
        if (col[3] == 0 && !ORC_M(8)) {
            col[0] = ns.r; col[1] = ns.g; col[2] = ns.b; col[3] = 1.0f;
            // ReconstructionOur.cpp:341-347
            if (shadow && !ORC_M(13) && !lit(o, o.x - na.x, o.y - na.y, na.key, ns.wgx, ns.wgy))
                col[0] = col[1] = col[2] = 0.0f;
            res.nearest = 1;
        }
        if (shadow) {
            for (int k = 0; k < 4; k++) res.color[k] = col[k];
            return res;
        }
        if (p->lens_gaussian) {
            // Gaussian: Reconstruction.hpp:378
            const float e = 2.718281828f;
            float gu = powf(e, -((o.u - 0) * (o.u - 0)) / (2 * (0.5f * 0.5f)));
            float gv = powf(e, -((o.v - 0) * (o.v - 0)) / (2 * (0.5f * 0.5f)));
            float g = gu * gv;
            for (int k = 0; k < 4; k++) col[k] *= g;
        }
        if (p->time_gaussian) {
            const float e = 2.718281828f;
            float g = powf(e, -((o.t - 0.5f) * (o.t - 0.5f)) / (2 * (0.25f * 0.25f)));
            for (int k = 0; k < 4; k++) col[k] *= g;
        }
        for (int k = 0; k < 4; k++) res.color[k] = col[k];
        return res;
    }
};

// generateOutputSamples: Reconstruction.cpp:452-512 with outSpp a parameter.
struct Tmpl { float x, y, u, v, t, key; };

int make_pattern(int out_spp, int override_uvt, const float* uvt, std::vector<Tmpl>& out) {
    const int M = override_uvt ? 1 : out_spp;
    out.clear();
    Ranrot rnd(242);
    for (int j = 0; j < 64; j++) {
        float xo = hammersley(j, 64), yo = lp(j);
        float uo = halton(2, j), vo = halton(3, j);
        float tj = rnd.f32();
        for (int i = 0; i < M; i++) {
            float s2[2];
            sobol2d(i + 1, s2);
            float x = sobol(3, i + 1);
            float y = lp(i + 1);
            float u = s2[0], v = s2[1];
            float t = (i + (ORC_M(9) ? 0.0f : tj)) / M;
            float xy0 = x + xo, xy1 = y + yo, uv0 = u + uo, uv1 = v + vo;
            if (xy0 >= 1.f) xy0 -= 1.f;
            if (xy1 >= 1.f) xy1 -= 1.f;
            if (uv0 >= 1.f) uv0 -= 1.f;
            if (uv1 >= 1.f) uv1 -= 1.f;
            int ub = (int)std::floor(256 * uv0);
            int vb = (int)std::floor(256 * uv1);
            int tb = (int)std::floor(256 * t);
            int key = 0;
            for (int k = 0; k < 8; k++)
                key |= ((ub >> k) & 1) << (3 * k + 2) | ((vb >> k) & 1) << (3 * k + 1) | ((tb >> k) & 1) << (3 * k + 0);
            float d[2];
            to_unit_disk(uv0, uv1, d);
            if (override_uvt) {
                d[0] = uvt[0]; d[1] = uvt[1];
                t = uvt[2];
            }
            if (M == 1) { xy0 = 0.5f; xy1 = 0.5f; }
            Tmpl s;
            s.x = xy0; s.y = xy1; s.u = d[0]; s.v = d[1]; s.t = t; s.key = (float)key;
            out.push_back(s);
        }
        // Sort<Sample>::increasing over [j*M, (j+1)*M): Util.hpp:75-83
        auto cmp = [&](int a, int b) -> int {
            return (out[a].key < out[b].key) ? -1 : (out[a].key > out[b].key) ? 1 : 0;
        };
        auto swp = [&](int a, int b) { std::swap(out[a], out[b]); };
        fw_sort(j * M, (j + 1) * M, cmp, swp);
    }
    return M;
}

int hw_threads(int t) {
    if (t > 0) return t;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

template <class F>
void parallel_for(int n, int threads, F f) {
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int i = 0; i < n; i++) f(0, i);
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> ts;
    for (int w = 0; w < threads; w++)
        ts.emplace_back([&, w]() {
            for (;;) {
                int i = next.fetch_add(1);
                if (i >= n) break;
                f(w, i);
            }
        });
    for (auto& t : ts) t.join();
}

// ---------------------------------------------------------------------------
// The reference's GPU-variant algorithm (ReconstructionCudaKernels.cu:242-868:
// recKernel, failKernel, nearestKernel), per output sample.  Where the
// reference's CPU replica (ReconstructionCuda.cpp:571-872) differs from the
// kernels, the kernels are followed: reprojection through oow = 1/pw with
// cocr = C0*oow + C1 and no static-sample branch (.cu:425-437); the 1R
// "any point inside" test is strict (.cu:378) and the tent sum starts at
// that point (.cu:411); failKernel's points are per surface (.cu:722) and its
// triangle test is the sign-partition form (.cu:586-632); the nearest search
// shrinks its radius as it walks the hierarchy (.cu:545-549).  Capacity
// limits (MAX_LEAVES_*, MAX_POINTS, .cu:39-43) and the CPU fallback for
// queries beyond them (ReconstructionCuda.cpp:412-434) do not exist here.
// Arithmetic is IEEE (correctly rounded division and square root); the
// reference builds these kernels with -use_fast_math (.cu:35).
// Leaves are ordered by (key, leaf id) (the reference's selection sort breaks
// key ties by traversal order); the nearest walk visits child 0 first, so of
// two points at one distance the later in that walk wins, as in .cu:545.
// ---------------------------------------------------------------------------
struct RefGpu {
    const orc_ctx* c = nullptr;
    float disp = 0.f;
    std::vector<std::pair<float, int>> keyed;
    std::vector<int> leafs;
    std::vector<uint8_t> empty;
    std::vector<float> xy;    // current surface: dx, dy per point
    std::vector<float> good;  // triangle test partition
    std::vector<int> st;

    // .cu:419-437
    void point(const RS& p, const OSample& o, float& dx, float& dy, float& d2) const {
        const float p0x = p.x * p.w, p0y = p.y * p.w, p0w = p.w;
        const float px = p0x + (o.t - p.t) * p.mvx;
        const float py = p0y + (o.t - p.t) * p.mvy;
        const float pw = p0w + (o.t - p.t) * p.mvw;
        const float oow = 1.f / pw;
        const float cocr = c->cc[0] * oow + c->cc[1];
        const float x = px * oow + cocr * o.u;
        const float y = py * oow + cocr * o.v;
        dx = x - o.x;
        dy = y - o.y;
        d2 = dx * dx + dy * dy;
    }

    // .cu:262-330: leaves within radius, front to back by depth at o.t
    void sorted_leaves(const OSample& o, float R) {
        find_leaves(c, o, R, leafs);
        keyed.clear();
        for (int lf : leafs) {
            const RS& s = c->samples[c->leaves[lf].s0];
            keyed.push_back({s.w + (o.t - s.t) * s.mvw, lf});
        }
        std::sort(keyed.begin(), keyed.end(), [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
            if (a.first < b.first) return true;
            if (b.first < a.first) return false;
            return a.second < b.second;
        });
    }

    // angular extent of the points around the output sample (.cu:443-445)
    static void wedge(float* box, float dx, float dy) {
        const float k = dx / dy;
        if (dy >= 0.f) {
            box[0] = fminf(box[0], k);
            box[1] = fmaxf(box[1], k);
        } else {
            box[2] = fminf(box[2], k);
            box[3] = fmaxf(box[3], k);
        }
    }

    // recKernel (.cu:242-474).  Returns 0: colour found, 1: go to 2R,
    // 2: go to the nearest search.
    int first_surface(const OSample& o, float* col) {
        const float radius = disp;
        sorted_leaves(o, radius);
        if (keyed.empty()) return 1;
        const float RANGE = 1.0f / std::sqrt(128.0f);
        const float umin[3] = {o.u - RANGE, o.v - RANGE, o.t - 0.5f * RANGE};
        const float umax[3] = {o.u + RANGE, o.v + RANGE, o.t + 0.5f * RANGE};
        float wbox[4] = {FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
        float color[4] = {0, 0, 0, 0};
        bool multi = false;
        empty.assign(keyed.size(), 0);
        for (size_t li = 0; li < keyed.size(); li++) {
            const Leaf& L = c->leaves[keyed[li].second];
            int i0 = L.s0;
            bool found = false;
            for (; i0 < L.s1; i0++) {
                float dx, dy, d2;
                point(c->samples[i0], o, dx, dy, d2);
                if (d2 < radius * radius) {
                    found = true;
                    break;
                }
            }
            if (!found) {
                empty[li] = 1;
            } else {
                for (size_t j = 0; j < li; j++) {
                    if (empty[j]) continue;
                    if (!same_surface(L.b, c->leaves[keyed[j].second].b, umin, umax)) {
                        multi = true;
                        break;
                    }
                }
            }
            if (multi) break;
            for (; i0 < L.s1; i0++) {
                float dx, dy, d2;
                const RS& p = c->samples[i0];
                point(p, o, dx, dy, d2);
                if (d2 <= radius * radius) {
                    wedge(wbox, dx, dy);
                    const float w = fmaxf(0.f, 1.f - std::sqrt(d2) / radius);
                    color[0] += w * p.r;
                    color[1] += w * p.g;
                    color[2] += w * p.b;
                    color[3] += w * 1.0f;
                }
            }
        }
        if (color[3] > 0.f && wbox[1] > wbox[2] && wbox[0] < wbox[3]) {
            const float s = 1.f / color[3];
            for (int k = 0; k < 4; k++) col[k] = color[k] * s;
            return 0;
        }
        return multi ? 1 : 2;
    }

    // checkTriangles (.cu:586-632) over the current surface's points
    bool triangles(int n) {
        if (n < 3) return false;
        const float dcomp = 4.f * disp * disp;
        good.assign(2 * (size_t)n, 0.f);
        for (int i = 0; i < n - 2; i++) {
            const float p0x = xy[2 * i], p0y = xy[2 * i + 1];
            int pg = 0, ng = n;  // positive side fills from the front, negative from the back
            for (int j = i + 1; j < n; j++) {
                const float p1x = xy[2 * j], p1y = xy[2 * j + 1];
                const float ex = p0x - p1x, ey = p0y - p1y;
                if (ex * ex + ey * ey < dcomp) {
                    const float c0 = p0x * p1y - p1x * p0y;
                    const int idx = c0 >= 0.f ? pg++ : --ng;
                    good[2 * idx] = p1x;
                    good[2 * idx + 1] = p1y;
                }
            }
            if (ng != n)
                for (int j = 0; j < pg; j++) {
                    const float p1x = good[2 * j], p1y = good[2 * j + 1];
                    for (int k = ng; k < n; k++) {
                        const float p2x = good[2 * k], p2y = good[2 * k + 1];
                        const float ex = p1x - p2x, ey = p1y - p2y;
                        if (p1x * p2y >= p2x * p1y && ex * ex + ey * ey < dcomp) return true;
                    }
                }
        }
        return false;
    }

    // failKernel (.cu:634-868).  Returns true when a colour was found.
    bool failed(const OSample& o, float* col) {
        const float radius = disp * 2.f;
        sorted_leaves(o, radius);
        if (keyed.empty()) return false;
        const float RANGE = 1.0f / std::sqrt(128.0f);
        const float umin[3] = {o.u - RANGE, o.v - RANGE, o.t - 0.5f * RANGE};
        const float umax[3] = {o.u + RANGE, o.v + RANGE, o.t + 0.5f * RANGE};
        const int nl = (int)keyed.size();
        float color[4] = {0, 0, 0, 0};
        int lidx = 0;
        do {
            const int sfirst = lidx;
            int np = 0;
            xy.clear();
            for (int k = 0; k < 4; k++) color[k] = 0.f;
            float abox[4] = {FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
            float wbox[4] = {FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
            do {
                const Leaf& L = c->leaves[keyed[lidx].second];
                bool end = false;
                if (np >= (ORC_M(10) ? 3 : 4))
                    for (int j = sfirst; j < lidx; j++)
                        if (!same_surface(L.b, c->leaves[keyed[j].second].b, umin, umax)) {
                            end = true;
                            break;
                        }
                if (end) break;
                for (int i = L.s0; i < L.s1; i++) {
                    float dx, dy, d2;
                    const RS& p = c->samples[i];
                    point(p, o, dx, dy, d2);
                    if (d2 <= radius * radius) {
                        xy.push_back(dx);
                        xy.push_back(dy);
                        np++;
                        const float w = 1.f - std::sqrt(d2) / disp;
                        if (w <= 0.f) {
                            wedge(abox, dx, dy);
                        } else {
                            wedge(wbox, dx, dy);
                            color[0] += w * p.r;
                            color[1] += w * p.g;
                            color[2] += w * p.b;
                            color[3] += w * 1.0f;
                        }
                    }
                }
                lidx++;
            } while (lidx < nl);
            if (lidx < nl) {
                abox[0] = fminf(abox[0], wbox[0]);
                abox[1] = fmaxf(abox[1], wbox[1]);
                abox[2] = fminf(abox[2], wbox[2]);
                abox[3] = fmaxf(abox[3], wbox[3]);
                if (abox[1] > abox[2] && abox[0] < abox[3])
                    if ((wbox[1] > wbox[2] && wbox[0] < wbox[3]) || triangles(np)) break;
            }
        } while (lidx < nl);
        if (color[3] > 0.f) {
            const float s = 1.f / color[3];
            for (int k = 0; k < 4; k++) col[k] = color[k] * s;
            return true;
        }
        return false;
    }

    // nearestKernel (.cu:476-579): radius 1, 2, 4, 8, 16 x dispersion; within
    // a pass the radius shrinks to each nearer point found.  Zero colour when
    // nothing lies within 16 x dispersion.
    int nearest(const OSample& o, float* col) {
        for (int k = 0; k < 4; k++) col[k] = 0.f;
        if (c->root == -1) return -1;
        for (float rmul = 1.f; rmul <= 16.f; rmul *= 2.f) {
            float radius = rmul * disp;
            int best = -1;
            st.clear();
            st.push_back(c->root);
            while (!st.empty()) {
                const int id = st.back();
                st.pop_back();
                if (id & LEAF_TAG) {
                    const Leaf& L = c->leaves[id & ~LEAF_TAG];
                    if (!node_hit(L.b, o.x, o.y, o.u, o.v, o.t, radius)) continue;
                    for (int i = L.s0; i < L.s1; i++) {
                        float dx, dy, d2;
                        point(c->samples[i], o, dx, dy, d2);
                        if (d2 <= radius * radius) {
                            radius = std::sqrt(d2);
                            best = i;
                        }
                    }
                } else {
                    const Inner& nd = c->inner[id];
                    if (!nd.b.valid()) continue;
                    if (node_hit(nd.b, o.x, o.y, o.u, o.v, o.t, radius)) {
                        st.push_back(nd.c1);
                        st.push_back(nd.c0);
                    }
                }
            }
            if (best >= 0) {
                const RS& p = c->samples[best];
                col[0] = p.r;
                col[1] = p.g;
                col[2] = p.b;
                col[3] = 1.0f;
                return best;
            }
        }
        return -1;
    }

    // one output sample: 1R, then 2R or the nearest search.  *stage: the
    // pass that produced the colour (0: 1R, 1: 2R, 2: nearest).
    void query(const OSample& o, float* col, int* stage) {
        const int r = first_surface(o, col);
        if (r == 0) {
            *stage = 0;
            return;
        }
        if (r == 1 && failed(o, col)) {
            *stage = 1;
            return;
        }
        *stage = 2;
        nearest(o, col);
    }
};

// ---------------------------------------------------------------------------
// FW matrix arithmetic (src/framework/base/Math.hpp), row-major m[4*i+j] =
// get(i, j), every sum started at zero as FW does.
// ---------------------------------------------------------------------------
struct M4 {
    float m[16];
    float& operator()(int i, int j) { return m[4 * i + j]; }
    float operator()(int i, int j) const { return m[4 * i + j]; }
    static M4 identity() {
        M4 r;
        for (int k = 0; k < 16; k++) r.m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
        return r;
    }
};

// Matrix * Matrix: Math.hpp:1024-1037
M4 mmul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float rr = 0.0f;
            for (int k = 0; k < 4; k++) rr += a(i, k) * b(k, j);
            r(i, j) = rr;
        }
    return r;
}

// Matrix * Vector: Math.hpp:989-1000
void mvec(const M4& a, const float v[4], float out[4]) {
    float r[4];
    for (int i = 0; i < 4; i++) {
        float rr = 0.0f;
        for (int j = 0; j < 4; j++) rr += a(i, j) * v[j];
        r[i] = rr;
    }
    for (int i = 0; i < 4; i++) out[i] = r[i];
}

// 3x3 determinant: Math.hpp:928-933
float det3(const float v[3][3]) {
    return v[0][0] * v[1][1] * v[2][2] - v[0][0] * v[1][2] * v[2][1] + v[1][0] * v[2][1] * v[0][2] -
           v[1][0] * v[2][2] * v[0][1] + v[2][0] * v[0][1] * v[1][2] - v[2][0] * v[0][2] * v[1][1];
}

// inverted(): Math.hpp:955-976 (cofactors; d accumulates every i, j, so it
// is 4 det, undone by the final * L)
M4 minv(const M4& a) {
    M4 r;
    float d = 0.0f, si = 1.0f;
    for (int i = 0; i < 4; i++) {
        float sj = si;
        for (int j = 0; j < 4; j++) {
            float sub[3][3];
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) sub[k][l] = a((k < j) ? k : k + 1, (l < i) ? l : l + 1);
            const float dd = det3(sub) * sj;
            r(i, j) = dd;
            d += dd * a(j, i);
            sj = -sj;
        }
        si = -si;
    }
    const float rd = rcpf(d);
    for (int k = 0; k < 16; k++) r.m[k] = r.m[k] * rd * 4.0f;
    return r;
}

M4 mtrans(const M4& a) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r(i, j) = a(j, i);
    return r;
}

M4 mscale(float x, float y, float z) {
    M4 r = M4::identity();
    r(0, 0) = x; r(1, 1) = y; r(2, 2) = z;
    return r;
}

M4 mtranslate(float x, float y, float z) {
    M4 r = M4::identity();
    r(0, 3) = x; r(1, 3) = y; r(2, 3) = z;
    return r;
}

M4 m4_in(const float* p) {
    M4 r;
    for (int k = 0; k < 16; k++) r.m[k] = p[k];
    return r;
}

// normalized(): v * (1 * rcp(length())), length = sqrt of the squares summed
// from zero (Math.hpp:139-141)
void normalized3(const float v[3], float out[3]) {
    float l2 = 0.0f;
    for (int k = 0; k < 3; k++) l2 += v[k] * v[k];
    const float f = 1.0f * rcpf(std::sqrt(l2));
    for (int k = 0; k < 3; k++) out[k] = v[k] * f;
}

// viewSpaceNormalFromWGradients: common/Util.cpp:172-198
void view_normal(const float p[3], float dwdu, float dwdv, int vw, int vh, const M4& inv_proj, float n[3]) {
    const float x = p[0], y = p[1], w = p[2];
    const float u = x / w, v = y / w;
    dwdu *= (float)vw / 2.0f;
    dwdv *= (float)vh / 2.0f;
    float b1[4] = {w + u * dwdu, v * dwdu, dwdu, 0.0f};
    float b2[4] = {u * dwdv, w + v * dwdv, dwdv, 0.0f};
    M4 ip = inv_proj;
    ip(2, 0) = 0; ip(2, 1) = 0; ip(2, 2) = -1; ip(2, 3) = 0;
    ip(3, 0) = 0; ip(3, 1) = 0; ip(3, 2) = 0; ip(3, 3) = 1;
    mvec(ip, b1, b1);
    mvec(ip, b2, b2);
    const float c[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
    normalized3(c, n);
}

// lightSpaceDot: common/Util.cpp:202-224, with the two matrix products it
// forms first (invws * cameraProjectedZfromW, cameraToShadow2 *
// invCameraProjection) taken from the setup
float light_space_dot(const float xyw[4], float wgx, float wgy, const orc_shadow_params* sp, float n[3]) {
    float xyzw[4];
    mvec(m4_in(sp->zfromw_ndc), xyw, xyzw);
    const float p3[3] = {xyzw[0], xyzw[1], xyzw[3]};
    view_normal(p3, wgx, wgy, sp->width, sp->height, m4_in(sp->inv_projection), n);
    float n4[4] = {n[0], n[1], n[2], 0.0f}, nl4[4];
    mvec(m4_in(sp->shadow2_inv_t), n4, nl4);
    float pl[4];
    mvec(m4_in(sp->shadow2_inv_proj), xyzw, pl);
    const float c = rcpf(pl[3]);
    const float plight[3] = {pl[0] * c, pl[1] * c, pl[2] * c};
    const float nl[3] = {nl4[0], nl4[1], nl4[2]};
    float a[3], b[3];
    normalized3(plight, a);
    normalized3(nl, b);
    float d = 0.0f;
    for (int k = 0; k < 3; k++) d += a[k] * b[k];
    return -d;
}

}  // namespace

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------

extern "C" {

orc_ctx* orc_create(int W, int H, int spp, float C0, float C1, const float* records) {
    orc_ctx* c = new orc_ctx;
    c->W = W; c->H = H; c->spp = spp; c->C0 = C0; c->C1 = C1;
    c->N = W * H * spp;
    c->rec.assign(records, records + (size_t)c->N * 16);
    return c;
}

// An irregular buffer: n samples in any pixel order (the header's spp 0).
orc_ctx* orc_create_irregular(int W, int H, float C0, float C1, const float* records, int n) {
    orc_ctx* c = new orc_ctx;
    c->W = W; c->H = H; c->spp = 0; c->C0 = C0; c->C1 = C1;
    c->N = n;
    c->rec.assign(records, records + (size_t)n * 16);
    return c;
}

void orc_destroy(orc_ctx* c) { delete c; }
const char* orc_error(const orc_ctx* c) { return c->err.c_str(); }

int orc_build(orc_ctx* c, float ap_adj, float foc_adj, int threads) {
    threads = hw_threads(threads);
    // TreeGather::init preamble: Reconstruction.cpp:63-87; an irregular
    // buffer (spp 0) averages its per-pixel counts, integer division (:65-71)
    int spp = c->spp;
    if (spp == 0) spp = c->N / (c->W * c->H);
    c->spp_pow2 = std::max(1, (int)round_pow2((uint32_t)spp));
    if (!dispersion_index_ok(c->spp_pow2)) {
        c->err = "unsupported spp";
        return 4;  // TLR_E_UNSUPPORTED_SPP
    }
    {
        float cx = c->C0, cy = c->C1;
        float a = -cx;
        float f = -cx * rcpf(cy);
        a *= ap_adj;
        f *= foc_adj;
        cx = -a;
        cy = a * rcpf(f);
        c->cc[0] = cx; c->cc[1] = cy;
        // overrideRefocusDistance: Reconstruction.cpp:88-95
        if (c->coc_override) {
            c->cc[0] = c->coc_ov[0];
            c->cc[1] = c->coc_ov[1];
        }
    }
    const float cc0[2] = {c->C0, c->C1};

    // reprojectToUVTCenter: Reconstruction.cpp:514-640
    c->repro.clear();
    c->repro.reserve(c->N);
    int bad = 0;
    for (int k = 0; k < c->N; k++) {
        RS s;
        if (reproject_center(&c->rec[(size_t)k * 16], cc0, s)) {
            // Non-finite reprojected positions are rejected as invalid input
            // (the reference's bbox/bucketing is undefined for them).
            if (!std::isfinite(s.x) || !std::isfinite(s.y)) {
                bad++;
                continue;
            }
            s.src = k;
            if (!c->density.empty()) s.density = c->density[k];
            c->repro.push_back(s);
        }
    }
    if (bad) {
        c->err = "samples reproject to non-finite positions";
        return 1;  // TLR_E_INVALID
    }
    const int W = c->W, H = c->H;
    float bmin[2] = {FLT_MAX, FLT_MAX}, bmax[2] = {-FLT_MAX, -FLT_MAX};
    std::vector<int> hits((size_t)W * H, 0);
    for (const RS& s : c->repro) {
        bmin[0] = mn(bmin[0], s.x); bmin[1] = mn(bmin[1], s.y);
        bmax[0] = mx(bmax[0], s.x); bmax[1] = mx(bmax[1], s.y);
        int sx = floor_i(s.x), sy = floor_i(s.y);
        if (sx >= 0 && sy >= 0 && sx < W && sy < H) hits[(size_t)sy * W + sx]++;
    }
    int mh = 0;
    for (int h : hits) mh = std::max(mh, h);
    c->max_hits = mh;
    c->bbmin[0] = bmin[0]; c->bbmin[1] = bmin[1];
    c->bbmax[0] = bmax[0]; c->bbmax[1] = bmax[1];
    if (mh == 0) {
        c->err = "All samples were discarded in reprojection to (u,v,t)=0";
        return 3;  // TLR_E_ALL_DISCARDED
    }
    float scale = std::sqrt(1.f * mh / 48);
    float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1];
    float ext = mx(ex, ey);
    if (scale * ext > 4096) scale = 4096 / ext;
    c->scale = scale;
    float smin[2] = {bmin[0] * scale, bmin[1] * scale};
    float smax[2] = {bmax[0] * scale, bmax[1] * scale};
    c->bbmin_i[0] = floor_i(smin[0]);
    c->bbmin_i[1] = floor_i(smin[1]);
    int bmax_i[2] = {floor_i(smax[0]) + 1, floor_i(smax[1]) + 1};
    c->gw = bmax_i[0] - c->bbmin_i[0];
    c->gh = bmax_i[1] - c->bbmin_i[1];
    const int gw = c->gw, gh = c->gh;
    const int K = (int)c->repro.size();
    c->cell_of.assign(K, -1);
    std::vector<int> cnt((size_t)gw * gh, 0);
    int acc = 0;
    for (int k = 0; k < K; k++) {
        const RS& s = c->repro[k];
        int sx = floor_i(scale * s.x) - c->bbmin_i[0];
        int sy = floor_i(scale * s.y) - c->bbmin_i[1];
        if (sx < 0 || sy < 0 || sx >= gw || sy >= gh) continue;
        c->cell_of[k] = sy * gw + sx;
        cnt[(size_t)sy * gw + sx]++;
        acc++;
    }
    c->num_bucketed = acc;
    if (acc == 0) {
        c->err = "All samples were discarded in reprojection to (u,v,t)=0";
        return 3;
    }
    c->cell_start.assign((size_t)gw * gh + 1, 0);
    for (size_t i = 0; i < (size_t)gw * gh; i++) c->cell_start[i + 1] = c->cell_start[i] + cnt[i];
    c->cell_items.assign(acc, 0);
    {
        std::vector<int> fill(c->cell_start.begin(), c->cell_start.end() - 1);
        for (int k = 0; k < K; k++)
            if (c->cell_of[k] >= 0) c->cell_items[fill[c->cell_of[k]]++] = k;
    }

    // Initial hierarchy: ReconstructionTreeBuilder.cpp:34-96
    SAT sat;
    sat.gw = gw; sat.gh = gh;
    sat.s.assign((size_t)(gw + 1) * (gh + 1), 0);
    for (int y = 0; y < gh; y++) {
        long long row = 0;
        for (int x = 0; x < gw; x++) {
            row += cnt[(size_t)y * gw + x];
            sat.s[(size_t)(y + 1) * (gw + 1) + x + 1] = sat.s[(size_t)y * (gw + 1) + x + 1] + row;
        }
    }
    c->inodes.clear();
    c->ileaf_nodes.clear();
    c->ileaf_offset.clear();
    build_initial(c, sat, 0, gw, 0, gh);
    int off = 0;
    collect_initial_leaves(c, 0, off);

    // Leaves: ReconstructionTreeBuilder.cpp:127-383, one task per initial leaf.
    c->samples.assign(acc, RS());
    const int NL = (int)c->ileaf_nodes.size();
    std::vector<LeafScratch> scratch(threads);
    std::vector<std::vector<Leaf>> outs(threads);
    parallel_for(NL, threads, [&](int w, int i) {
        build_leaf_cell(c, c->ileaf_nodes[i], c->ileaf_offset[i], scratch[w], outs[w]);
    });
    c->leaves.clear();
    for (auto& o : outs) c->leaves.insert(c->leaves.end(), o.begin(), o.end());
    std::sort(c->leaves.begin(), c->leaves.end(), [](const Leaf& a, const Leaf& b) { return a.s0 < b.s0; });

    build_hierarchy(c);
    c->built = true;
    return 0;
}

void orc_get_build_info(const orc_ctx* c, orc_build_info* o) {
    std::memset(o, 0, sizeof(*o));
    o->num_input = c->N;
    o->num_kept = (int)c->repro.size();
    o->num_bucketed = c->num_bucketed;
    o->spp_pow2 = c->spp_pow2;
    o->coc_adj[0] = c->cc[0]; o->coc_adj[1] = c->cc[1];
    o->bbmin[0] = c->bbmin[0]; o->bbmin[1] = c->bbmin[1];
    o->bbmax[0] = c->bbmax[0]; o->bbmax[1] = c->bbmax[1];
    o->scale = c->scale;
    o->bbmin_int[0] = c->bbmin_i[0]; o->bbmin_int[1] = c->bbmin_i[1];
    o->grid_w = c->gw; o->grid_h = c->gh;
    o->max_hits = c->max_hits;
    o->num_initial_nodes = (int)c->inodes.size();
    o->num_initial_leaves = (int)c->ileaf_nodes.size();
    o->num_leaves = (int)c->leaves.size();
    o->num_internal = (int)c->inner.size();
    o->root = c->root;
}

void orc_get_reprojected(const orc_ctx* c, float* o, int32_t* src) {
    for (size_t k = 0; k < c->repro.size(); k++) {
        const RS& s = c->repro[k];
        float v[10] = {s.x, s.y, s.w, s.t, s.mvx, s.mvy, s.mvw, s.r, s.g, s.b};
        std::memcpy(o + 10 * k, v, sizeof(v));
        src[k] = s.src;
    }
}

void orc_get_cells(const orc_ctx* c, int32_t* cell) {
    for (size_t k = 0; k < c->cell_of.size(); k++) cell[k] = c->cell_of[k];
}

void orc_get_initial_leaves(const orc_ctx* c, int32_t* o) {
    for (size_t i = 0; i < c->ileaf_nodes.size(); i++) {
        const INode& n = c->inodes[c->ileaf_nodes[i]];
        int32_t v[6] = {n.x0, n.y0, n.x1, n.y1, c->ileaf_offset[i], n.n};
        std::memcpy(o + 6 * i, v, sizeof(v));
    }
}

void orc_get_layout(const orc_ctx* c, int32_t* src) {
    for (size_t k = 0; k < c->samples.size(); k++) src[k] = c->samples[k].src;
}

static void tlb_out(const TLB& b, float* o) {
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 3; j++) o[3 * k + j] = b.p[k][j];
    o[12] = b.ts[0];
    o[13] = b.ts[1];
}
static TLB tlb_in(const float* o) {
    TLB b;
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 3; j++) b.p[k][j] = o[3 * k + j];
    b.ts[0] = o[12];
    b.ts[1] = o[13];
    return b;
}

void orc_get_leaves(const orc_ctx* c, int32_t* s01, float* b14) {
    for (size_t i = 0; i < c->leaves.size(); i++) {
        s01[2 * i] = c->leaves[i].s0;
        s01[2 * i + 1] = c->leaves[i].s1;
        tlb_out(c->leaves[i].b, b14 + 14 * i);
    }
}

void orc_get_hierarchy(const orc_ctx* c, int32_t* order, int32_t* ch, float* b14) {
    for (size_t i = 0; i < c->leaf_order.size(); i++) order[i] = c->leaf_order[i];
    for (size_t i = 0; i < c->inner.size(); i++) {
        ch[2 * i] = c->inner[i].c0;
        ch[2 * i + 1] = c->inner[i].c1;
        tlb_out(c->inner[i].b, b14 + 14 * i);
    }
}

int orc_output_pattern(int out_spp, int override_uvt, const float* uvt, float* out6) {
    std::vector<Tmpl> t;
    int M = make_pattern(out_spp, override_uvt, uvt, t);
    if (out6)
        for (size_t k = 0; k < t.size(); k++) {
            float v[6] = {t[k].x, t[k].y, t[k].u, t[k].v, t[k].t, t[k].key};
            std::memcpy(out6 + 6 * k, v, sizeof(v));
        }
    return M;
}

int orc_reconstruct(orc_ctx* c, const orc_recon_params* p, float* image, float* qcol, orc_stats* stats) {
    return orc_reconstruct_debug(c, p, image, qcol, stats, nullptr);
}

void orc_set_coc_override(orc_ctx* c, const float* cc) {
    c->coc_override = cc != nullptr;
    if (cc) {
        c->coc_ov[0] = cc[0];
        c->coc_ov[1] = cc[1];
    }
}

int orc_reconstruct_debug(orc_ctx* c, const orc_recon_params* p, float* image, float* qcol, orc_stats* stats,
                          float* debug4) {
    if (!c->built) return 1;
    std::vector<Tmpl> pat;
    const int M = make_pattern(p->out_spp, p->override_uvt, p->uvt, pat);
    const int W = c->W, H = c->H;
    const int stride = std::max(1, p->row_stride);
    const int threads = hw_threads(p->threads);
    const int rows = (H + stride - 1) / stride;
    std::vector<StatsAcc> st(threads);
    std::vector<Filt> fl(threads);
    for (auto& f : fl) f.c = c;
    // FilterTask::process: Reconstruction.cpp:642-720, one task per scanline.
    parallel_for(rows, threads, [&](int w, int ri) {
        const int y = ri * stride;
        Filt& F = fl[w];
        for (int x = 0; x < W; x++) {
            const int pat_i = (int)(hash_bits((uint32_t)x, (uint32_t)y) % 64u);
            float pc[4] = {0, 0, 0, 0};
            for (int i = 0; i < M; i++) {
                st[w].count += 1;
                const Tmpl& tp = pat[(size_t)pat_i * M + i];
                OSample o;
                o.x = tp.x + (float)x;
                o.y = tp.y + (float)y;
                o.u = tp.u; o.v = tp.v; o.t = tp.t;
                Filt::Result r = F.reconstruct(o, p, st[w]);
                for (int k = 0; k < 4; k++) pc[k] += r.color[k];
                // debugColor: ReconstructionOur.cpp:312-314, the last output
                // sample's stays (Reconstruction.cpp:276)
                if (debug4 && i == M - 1) {
                    float* d = debug4 + 4 * ((size_t)y * W + x);
                    d[0] = (float)F.nsurf;
                    d[1] = (float)r.n1;
                    d[2] = d[3] = 0.0f;
                }
                if (qcol) {
                    size_t q = ((size_t)y * W + x) * M + i;
                    for (int k = 0; k < 4; k++) qcol[4 * q + k] = r.color[k];
                }
            }
            if (pc[3] == 0) { pc[0] = 1; pc[1] = 0; pc[2] = 0; pc[3] = 1; }
            float rw = rcpf(pc[3]);
            if (image)
                for (int k = 0; k < 4; k++) image[4 * ((size_t)y * W + x) + k] = pc[k] * rw;
        }
    });
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (auto& s : st) {
            stats->leaf_nodes += s.v[0];
            stats->samples_in_leaf_nodes += s.v[1];
            stats->samples_within_1r += s.v[2];
            stats->samples_fetched += s.v[3];
            stats->surfaces += s.v[4];
            stats->num_2r += s.v[5];
            stats->surfaces_merged += s.v[6];
            stats->at_least_one += s.v[7];
            stats->count += s.count;
        }
    }
    return 0;
}

// CudaReconstruction::reconstructGPU (ReconstructionCuda.cpp:327-465): every
// output sample through recKernel / failKernel / nearestKernel, per-pixel sum
// normalised by alpha, red where the sum has no weight (.cpp:446-456).  The
// per-pixel sum is in output-sample order (the reference adds atomically).
int orc_reconstruct_refgpu(orc_ctx* c, const orc_recon_params* p, float* image, float* qcol, int32_t* stage) {
    if (!c->built) return 1;
    std::vector<Tmpl> pat;
    const int M = make_pattern(p->out_spp, p->override_uvt, p->uvt, pat);
    const int W = c->W, H = c->H;
    const int stride = std::max(1, p->row_stride);
    const int threads = hw_threads(p->threads);
    const int rows = (H + stride - 1) / stride;
    std::vector<RefGpu> rg(threads);
    for (auto& r : rg) {
        r.c = c;
        r.disp = dispersion_for(c->spp_pow2);
    }
    parallel_for(rows, threads, [&](int w, int ri) {
        const int y = ri * stride;
        for (int x = 0; x < W; x++) {
            const int pat_i = (int)(hash_bits((uint32_t)x, (uint32_t)y) % 64u);
            float pc[4] = {0, 0, 0, 0};
            for (int i = 0; i < M; i++) {
                const Tmpl& tp = pat[(size_t)pat_i * M + i];
                OSample o;
                o.x = tp.x + (float)x;
                o.y = tp.y + (float)y;
                o.u = tp.u; o.v = tp.v; o.t = tp.t;
                float col[4];
                int stg = 0;
                rg[w].query(o, col, &stg);
                for (int k = 0; k < 4; k++) pc[k] += col[k];
                const size_t q = ((size_t)y * W + x) * M + i;
                if (qcol)
                    for (int k = 0; k < 4; k++) qcol[4 * q + k] = col[k];
                if (stage) stage[q] = stg;
            }
            if (pc[3] == 0) { pc[0] = 1; pc[1] = 0; pc[2] = 0; pc[3] = 1; }
            const float rw = rcpf(pc[3]);
            if (image)
                for (int k = 0; k < 4; k++) image[4 * ((size_t)y * W + x) + k] = pc[k] * rw;
        }
    });
    return 0;
}

// Pieces of the GPU-variant method for the oracle's own checks.
int orc_refgpu_query(orc_ctx* c, const float* o5, float* col4, int32_t* stage) {
    if (!c->built) return 1;
    RefGpu r;
    r.c = c;
    r.disp = dispersion_for(c->spp_pow2);
    OSample o{o5[0], o5[1], o5[2], o5[3], o5[4]};
    int st = 0;
    r.query(o, col4, &st);
    if (stage) *stage = st;
    return 0;
}

// nearestKernel alone: tree-layout index of the sample found, -1 if none.
int orc_refgpu_nearest(orc_ctx* c, const float* o5, float* col4) {
    if (!c->built) return -2;
    RefGpu r;
    r.c = c;
    r.disp = dispersion_for(c->spp_pow2);
    OSample o{o5[0], o5[1], o5[2], o5[3], o5[4]};
    return r.nearest(o, col4);
}

// failKernel alone (2R surfaces): 1 when it produced a colour.
int orc_refgpu_failed(orc_ctx* c, const float* o5, float* col4) {
    if (!c->built) return -2;
    RefGpu r;
    r.c = c;
    r.disp = dispersion_for(c->spp_pow2);
    OSample o{o5[0], o5[1], o5[2], o5[3], o5[4]};
    for (int k = 0; k < 4; k++) col4[k] = 0.f;
    return r.failed(o, col4) ? 1 : 0;
}

// checkTriangles on n points (x, y pairs relative to the output sample).
int orc_refgpu_triangles(const float* xy, int n, float disp) {
    RefGpu r;
    r.disp = disp;
    r.xy.assign(xy, xy + 2 * (size_t)n);
    return r.triangles(n) ? 1 : 0;
}

// The GPU-variant reprojection of every tree sample (layout order) to o:
// dx, dy, dist^2 per sample.
int orc_refgpu_reproject(orc_ctx* c, const float* o5, float* out3) {
    if (!c->built) return 1;
    RefGpu r;
    r.c = c;
    OSample o{o5[0], o5[1], o5[2], o5[3], o5[4]};
    for (size_t i = 0; i < c->samples.size(); i++) r.point(c->samples[i], o, out3[3 * i], out3[3 * i + 1], out3[3 * i + 2]);
    return 0;
}

int orc_gather(orc_ctx* c, const float* o5, float R, int32_t* leaves, int max_leaves, int* num_leaves,
               int32_t* samples, int32_t* surface_of, int max_samples) {
    Filt F;
    F.c = c;
    std::vector<int> dl;
    std::vector<std::pair<int, int>> ds;
    F.dbg_leaves = &dl;
    F.dbg_samples = &ds;
    StatsAcc st;
    OSample o{o5[0], o5[1], o5[2], o5[3], o5[4]};
    F.collect(o, R, st);
    if (num_leaves) *num_leaves = (int)dl.size();
    if (leaves)
        for (int i = 0; i < (int)dl.size() && i < max_leaves; i++) leaves[i] = dl[i];
    if (samples)
        for (int i = 0; i < (int)ds.size() && i < max_samples; i++) {
            samples[i] = ds[i].first;
            if (surface_of) surface_of[i] = ds[i].second;
        }
    return (int)ds.size();
}

int orc_query(orc_ctx* c, const orc_recon_params* p, int x, int y, int i, orc_query_info* out) {
    std::vector<Tmpl> pat;
    const int M = make_pattern(p->out_spp, p->override_uvt, p->uvt, pat);
    const int pat_i = (int)(hash_bits((uint32_t)x, (uint32_t)y) % 64u);
    const Tmpl& tp = pat[(size_t)pat_i * M + i];
    OSample o;
    o.x = tp.x + (float)x;
    o.y = tp.y + (float)y;
    o.u = tp.u; o.v = tp.v; o.t = tp.t;
    Filt F;
    F.c = c;
    StatsAcc st;
    Filt::Result r = F.reconstruct(o, p, st);
    out->fetch2r = r.fetch2r;
    out->num_surfaces_1r = r.nsurf1;
    out->num_samples_1r = r.n1;
    out->num_samples_final = r.nfinal;
    out->emergency_passes = r.emerg;
    out->found_surface = r.found_surface;
    out->used_nearest = r.nearest;
    out->dispersion = r.disp;
    for (int k = 0; k < 4; k++) out->color[k] = r.color[k];
    return 0;
}

long long orc_nodes_tested(int reset) {
    long long v = g_nodes_tested;
    if (reset) g_nodes_tested = 0;
    return v;
}

int orc_leaves_bruteforce(orc_ctx* c, const float* o5, float R, int32_t* leaves, int max_leaves) {
    int n = 0;
    for (int i = 0; i < (int)c->leaves.size(); i++) {
        const Leaf& lf = c->leaves[i];
        if (!lf.b.valid()) continue;
        if (node_hit(lf.b, o5[0], o5[1], o5[2], o5[3], o5[4], R)) {
            if (leaves && n < max_leaves) leaves[n] = i;
            n++;
        }
    }
    return n;
}

uint32_t orc_hash_bits(uint32_t a, uint32_t b) { return hash_bits(a, b); }
void orc_ranrot(uint32_t seed, int n, uint32_t* out) {
    Ranrot r(seed);
    for (int i = 0; i < n; i++) out[i] = r.next();
}
float orc_sobol(int dim, int i) { return sobol(dim, i); }
void orc_sobol2d(int i, float* o) { sobol2d(i, o); }
float orc_halton(int base, int i) { return halton(base, i); }
float orc_lp(int i) { return lp(i); }
void orc_to_unit_disk(float x, float y, float* o) { to_unit_disk(x, y, o); }

int orc_tux_planes(const float* pos, const float* mot, float W, float H, const float* cc, float* o14) {
    TLB b;
    bool v = tux_planes(pos, mot, 0.0f, 0.0f, W, H, cc, b);
    tlb_out(b, o14);
    return v ? 1 : 0;
}

void orc_merge_bounds(const float* a, const float* b, float* o) {
    TLB r = merge_tlb(tlb_in(a), tlb_in(b));
    tlb_out(r, o);
}

void orc_fw_sort(int n, float* keys, int32_t* payload) {
    auto cmp = [&](int a, int b) -> int { return (keys[a] < keys[b]) ? -1 : (keys[a] > keys[b]) ? 1 : 0; };
    auto swp = [&](int a, int b) {
        std::swap(keys[a], keys[b]);
        std::swap(payload[a], payload[b]);
    };
    fw_sort(0, n, cmp, swp);
}

int orc_triangle_covers(const float* r0, const float* r1, const float* r2, const float* o) {
    return triangle_covers(r0[0], r0[1], r1[0], r1[1], r2[0], r2[1], o[0], o[1]) ? 1 : 0;
}

int orc_surface_scan(int n, const float* t, const float* ba, const float* bb, const float* ba1, const float* bb1,
                     int32_t* surf) {
    surface_scan(n, [&](int k) { return t[k]; }, ba, bb, ba1, bb1, surf);
    return n;
}

int orc_same_surface(const float* n0, const float* n1, const float* o5) {
    const float RANGE = 1.0f / std::sqrt(128.0f);
    float umin[3] = {o5[2] - RANGE, o5[3] - RANGE, o5[4] - 0.5f * RANGE};
    float umax[3] = {o5[2] + RANGE, o5[3] + RANGE, o5[4] + 0.5f * RANGE};
    return same_surface(tlb_in(n0), tlb_in(n1), umin, umax) ? 1 : 0;
}

void orc_set_density(orc_ctx* c, const float* d) {
    if (d)
        c->density.assign(d, d + c->N);
    else
        c->density.clear();
    c->built = false;
}

// TreeGather::reconstructShadows / FilterTask::process with a query buffer
// (Reconstruction.cpp:302-347, :648-719): every query is an output sample of
// its own; shadow selects Filterer::reconstruct's reconstructShadow.
int orc_reconstruct_queries(orc_ctx* c, const orc_recon_params* p, const float* q8, int n, int shadow, float* rgba,
                            float* ext8, orc_stats* stats) {
    if (!c->built) return 1;
    const int threads = hw_threads(p->threads);
    std::vector<StatsAcc> st(threads);
    std::vector<Filt> fl(threads);
    for (auto& f : fl) f.c = c;
    const int chunk = 256;
    parallel_for((n + chunk - 1) / chunk, threads, [&](int w, int ci) {
        for (int qi = ci * chunk; qi < std::min(n, (ci + 1) * chunk); qi++) {
            const float* q = q8 + 8 * (size_t)qi;
            OSample o;
            o.x = q[0]; o.y = q[1]; o.u = q[2]; o.v = q[3]; o.t = q[4]; o.w = q[5];
            Filt::Ext e;
            st[w].count += 1;
            Filt::Result r = fl[w].reconstruct(o, p, st[w], shadow != 0, q[6], &e);
            if (rgba)
                for (int k = 0; k < 4; k++) rgba[4 * (size_t)qi + k] = r.color[k];
            if (ext8) {
                float* x = ext8 + 8 * (size_t)qi;
                x[0] = e.x; x[1] = e.y; x[2] = e.w; x[3] = e.density; x[4] = e.wgx; x[5] = e.wgy;
                x[6] = x[7] = 0.0f;
            }
        }
    });
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (auto& a : st) {
            stats->leaf_nodes += a.v[0];
            stats->samples_in_leaf_nodes += a.v[1];
            stats->samples_within_1r += a.v[2];
            stats->samples_fetched += a.v[3];
            stats->surfaces += a.v[4];
            stats->num_2r += a.v[5];
            stats->surfaces_merged += a.v[6];
            stats->at_least_one += a.v[7];
            stats->count += a.count;
        }
    }
    return 0;
}

// The matrices of TreeGather::reconstructDofMotionShadows
// (Reconstruction.cpp:365-388) from the two cameras.
void orc_shadow_setup(const float* cam_proj, const float* cam_to_world, const float* light_proj,
                      const float* light_world_to_cam, int w, int h, const float* cam_coc, const float* light_coc,
                      float cam_far, orc_shadow_params* out) {
    const M4 P = m4_in(cam_proj), C2W = m4_in(cam_to_world), PL = m4_in(light_proj), W2L = m4_in(light_world_to_cam);
    const M4 invP = minv(P);
    const M4 c2s = mmul(mmul(mmul(PL, W2L), C2W), invP);
    const M4 c2s2 = mmul(W2L, C2W);
    const M4 c2s2_it = mtrans(minv(c2s2));
    const M4 c2w_it = mtrans(minv(C2W));
    const M4 ws = mmul(mscale(0.5f * w, 0.5f * h, 0.5f), mtranslate(1.0f, 1.0f, 1.0f));
    const M4 invws = mmul(mtranslate(-1.0f, -1.0f, -1.0f), mscale(2.f / w, 2.f / h, 2.f));
    const M4 wsP = mmul(ws, P);
    M4 zw = M4::identity();
    zw(2, 2) = wsP(2, 3);
    zw(2, 3) = -wsP(2, 2);
    const M4 c2l = mmul(mmul(mmul(ws, c2s), invws), zw);
    const M4 zndc = mmul(invws, zw);
    const M4 s2ip = mmul(c2s2, invP);
    out->width = w;
    out->height = h;
    out->camera_coc[0] = cam_coc[0]; out->camera_coc[1] = cam_coc[1];
    out->light_coc[0] = light_coc[0]; out->light_coc[1] = light_coc[1];
    out->camera_far = cam_far;
    std::memcpy(out->c2l, c2l.m, 64);
    std::memcpy(out->zfromw_ndc, zndc.m, 64);
    std::memcpy(out->inv_projection, invP.m, 64);
    std::memcpy(out->shadow2_inv_proj, s2ip.m, 64);
    std::memcpy(out->shadow2_inv_t, c2s2_it.m, 64);
    std::memcpy(out->world_inv_t, c2w_it.m, 64);
}

void orc_mat4_inverse(const float* m, float* out) {
    const M4 r = minv(m4_in(m));
    std::memcpy(out, r.m, 64);
}

float orc_light_space_dot(const float* xyw, const float* wg, const orc_shadow_params* sp, float* n3) {
    return light_space_dot(xyw, wg[0], wg[1], sp, n3);
}

// FilterTask::process2 (Reconstruction.cpp:722-833): the camera's output
// samples, each one's shadow "ray" into the light's buffer, the shadow query
// there, the Lambert term (light-space dot with the normal from the
// w-gradients) and the ambient term; per pixel, red where a sample finds no
// surface.  uvoffset takes the generator's first and second draws as x, y.
int orc_reconstruct_dof_motion_shadows(orc_ctx* cam, orc_ctx* light, const orc_recon_params* p,
                                       const orc_shadow_params* sp, float* image) {
    if (!cam->built || !light->built) return 1;
    if (cam->W != light->W || cam->H != light->H || cam->W != sp->width || cam->H != sp->height) return 2;
    std::vector<Tmpl> pat;
    const int M = make_pattern(p->out_spp, p->override_uvt, p->uvt, pat);
    const int W = cam->W, H = cam->H;
    const int lw = light->W, lh = light->H;
    const int threads = hw_threads(p->threads);
    std::vector<StatsAcc> st(threads);
    std::vector<Filt> fc(threads), fli(threads);
    for (auto& f : fc) f.c = cam;
    for (auto& f : fli) f.c = light;
    const M4 c2l = m4_in(sp->c2l), w_it = m4_in(sp->world_inv_t);
    const float LIGHTING_SCALE = 1.5f, AMBIENT_SCALE = 0.5f;   // Reconstruction.cpp:34-35
    parallel_for(H, threads, [&](int w, int y) {
        for (int x = 0; x < W; x++) {
            const uint32_t h = hash_bits((uint32_t)x, (uint32_t)y);
            const int pat_i = (int)(h % 64u);
            Ranrot rnd(h);
            const float uo0 = rnd.f32();
            const float uo1 = rnd.f32();
            const float camFarW = sp->camera_far - 1e-2f;
            float pc[4] = {0, 0, 0, 0};
            for (int i = 0; i < M; i++) {
                st[w].count += 1;
                const Tmpl& tp = pat[(size_t)pat_i * M + i];
                OSample cs;
                cs.x = tp.x + (float)x;
                cs.y = tp.y + (float)y;
                cs.u = tp.u; cs.v = tp.v; cs.t = tp.t;
                Filt::Ext e;
                Filt::Result r = fc[w].reconstruct(cs, p, st[w], false, 1.0f, &e);
                float cx = e.x, cy = e.y;
                const float cw_ = e.w;
                const float cc[4] = {r.color[0], r.color[1], r.color[2], r.color[3]};
                if (cc[0] == 0 && cc[1] == 0 && cc[2] == 0 && cc[3] == 0) {
                    pc[0] = 1; pc[1] = 0; pc[2] = 0; pc[3] = 1;
                    break;
                }
                float col[4] = {cc[0], cc[1], cc[2], cc[3]};
                if (cw_ <= camFarW) {
                    // reprojectToUVCenter: Reconstruction.hpp:87-91
                    const float coc = sp->camera_coc[0] / cw_ + sp->camera_coc[1];
                    cx = cx - coc * cs.u;
                    cy = cy - coc * cs.v;
                    const float v4[4] = {cx * cw_, cy * cw_, 1.0f, cw_};
                    float lp[4];
                    mvec(c2l, v4, lp);
                    const float lc = rcpf(lp[3]);
                    const float lx = lp[0] * lc, ly = lp[1] * lc, lpw = lp[3];
                    if (lx >= 0 && ly >= 0 && lx <= (float)lw && ly <= (float)lh && lpw > 0) {
                        OSample ls;
                        ls.x = lx; ls.y = ly; ls.w = lpw; ls.t = cs.t;
                        float uu = sobol(3, i + 1) + uo0, vv = sobol(4, i + 1) + uo1;
                        if (uu >= 1.f) uu -= 1.f;
                        if (vv >= 1.f) vv -= 1.f;
                        ls.u = 2.f * uu - 1.f;
                        ls.v = 2.f * vv - 1.f;
                        const float lcoc = sp->light_coc[0] / ls.w + sp->light_coc[1];
                        ls.x += ls.u * lcoc;
                        ls.y += ls.v * lcoc;
                        Filt::Result sr = fli[w].reconstruct(ls, p, st[w], true, e.density, nullptr);
                        for (int k = 0; k < 4; k++) col[k] = sr.color[k];
                        float n[3];
                        const float ldotn = light_space_dot(v4, e.wgx, e.wgy, sp, n);
                        float n4[4] = {n[0], n[1], n[2], 0.0f}, wn[4];
                        mvec(w_it, n4, wn);
                        const float ambient = wn[1] * 0.5f + 0.5f;
                        // max(ldotn, 0) (fmaxf) with +0 for -0 and NaN
                        const float cosine = ldotn > 0.0f ? ldotn : 0.0f;
                        const float sc[4] = {LIGHTING_SCALE * cosine, LIGHTING_SCALE * cosine,
                                             LIGHTING_SCALE * cosine, LIGHTING_SCALE * 1.0f};
                        for (int k = 0; k < 4; k++) col[k] *= sc[k];
                        const float a = ambient * AMBIENT_SCALE;
                        for (int k = 0; k < 4; k++) col[k] += a * cc[k];
                        col[3] = 1;
                    }
                }
                for (int k = 0; k < 4; k++) pc[k] += col[k];
            }
            if (pc[3] == 0) { pc[0] = 1; pc[1] = 0; pc[2] = 0; pc[3] = 1; }
            const float rw = rcpf(pc[3]);
            for (int k = 0; k < 4; k++) image[4 * ((size_t)y * W + x) + k] = pc[k] * rw;
        }
    });
    return 0;
}

}  // extern "C"
