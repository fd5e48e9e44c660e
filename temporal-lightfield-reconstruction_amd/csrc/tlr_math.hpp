// Device arithmetic of the reconstruction path.
//
// Every function here reproduces the reference's FP32 operation order exactly
// (the build uses -ffp-contract=off and IEEE division/sqrt), so device results
// are bit-identical to an IEEE restatement of the reference CPU code.  Each
// function cites the reference file:line it follows.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TLR_DEV __device__ __forceinline__
#define TLR_HD __host__ __device__ __forceinline__

namespace tlr {

constexpr float kFltMax = 3.402823466e+38f;

// rcp(): src/framework/base/Math.hpp:112 (a ? 1/a : 0)
TLR_HD float rcp(float a) { return (a != 0.0f) ? 1.0f / a : 0.0f; }
// FW::min / FW::max: src/framework/base/Defs.hpp:196-197
TLR_HD float fmn(float a, float b) { return (a < b) ? a : b; }
TLR_HD float fmx(float a, float b) { return (a > b) ? a : b; }

// float -> int the way x86 cvttss2si evaluates (int)x: values outside the
// int range and NaN become INT_MIN.  Used for every bucket index so device and
// host agree on degenerate inputs.
TLR_HD int f2i(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return (int)0x80000000;
    return (int)f;
}
TLR_HD int floor_i(float f) { return f2i(floorf(f)); }

// ---------------------------------------------------------------------------
// Hyperplane bounds (TimeLensBounds, Reconstruction.hpp:149-167).
// p[3k+0..2] = (A,B,C) of plane k (0 xmin, 1 xmax, 2 ymin, 3 ymax):
// x = A*t + B*u + C.
// ---------------------------------------------------------------------------
// 64-byte records so the 12 plane coefficients load as three 16-byte words.
struct __align__(16) Bounds {
    float p[12];
    float ts0, ts1;
    float pad[2];
};

TLR_HD bool bounds_valid(const Bounds& b) { return b.ts0 < b.ts1; }
// TimeLensBounds::evaluate: Reconstruction.hpp:162
TLR_HD float plane_eval(const float* p, float t, float uv) { return t * p[0] + p[1] * uv + p[2]; }

TLR_HD void bounds_zero(Bounds& b) {
#pragma unroll
    for (int k = 0; k < 12; k++) b.p[k] = 0.0f;
    b.ts0 = 0.0f;
    b.ts1 = 0.0f;
}

// computePerspectivePlanes1D: common/TimeBounds.cpp:228-332
TLR_HD bool persp_1d(float x0, float w0, float dx, float dw, float xmin, float xmax, float* lo,
                     float* hi, float* span) {
    if (dx == 0.0f && dw == 0.0f) {
        float xow0 = x0 / w0;
        lo[0] = 0.0f; lo[1] = xow0;
        hi[0] = 0.0f; hi[1] = xow0;
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
    const bool right = dx * w0 > dw * x0;
    const bool fwd = dw > 0;
    float tmin = (xmin * w0 - x0) / (dx - xmin * dw);
    if (w0 + tmin * dw >= 0) {
        if (right) tb0 = fmx(tb0, tmin);
        else tb1 = fmn(tb1, tmin);
    }
    float tmax = (xmax * w0 - x0) / (dx - xmax * dw);
    if (w0 + tmax * dw >= 0) {
        if (!right) tb0 = fmx(tb0, tmax);
        else tb1 = fmn(tb1, tmax);
    }
    if (tb0 > tb1) return false;
    float th = 0.5f * (tb0 + tb1);
    float px = (x0 + th * dx) / (w0 + th * dw);
    float dxdt = dx / (w0 + th * dw) - (dw * (x0 + th * dx)) / ((w0 + th * dw) * (w0 + th * dw));
    float xa = (x0 + tb0 * dx) / (w0 + tb0 * dw);
    float xb = (x0 + tb1 * dx) / (w0 + tb1 * dw);
    float t1a = dxdt, t1b = (px - dxdt * th);                       // tangent at mid span
    float t2a = -(xb - xa) / (tb0 - tb1);                           // chord
    float t2b = xa + ((xb - xa) * tb0) / (tb0 - tb1);
    if (right == fwd) {
        hi[0] = t1a; hi[1] = t1b;
        lo[0] = t2a; lo[1] = t2b;
    } else {
        lo[0] = t1a; lo[1] = t1b;
        hi[0] = t2a; hi[1] = t2b;
    }
    span[0] = tb0; span[1] = tb1;
    return true;
}

// computeTUXPlanes (+ computePerspectivePlanes): common/TimeBounds.cpp:44-147.
// Screen rectangle is [0,W]x[0,H].  Planes stay zero when invalid.
TLR_HD bool tux_bounds(float px, float py, float pw, float mx_, float my_, float mw, float W, float H,
                       float cc0, float cc1, Bounds& b) {
    bounds_zero(b);
    float lx[2], hx[2], ly[2], hy[2], sx[2], sy[2];
    bool bx = persp_1d(px, pw, mx_, mw, 0.0f, W, lx, hx, sx);
    bool by = persp_1d(py, pw, my_, mw, 0.0f, H, ly, hy, sy);
    if (!bx || !by) return false;
    b.ts0 = fmx(sx[0], sy[0]);
    b.ts1 = fmn(sx[1], sy[1]);
    if (!(b.ts0 < b.ts1)) return false;
    float ft[3] = {b.ts0, b.ts1, 0.5f * (b.ts0 + b.ts1)};
    float ex[3], ey[3], coc[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float X = px + ft[i] * mx_;
        float Y = py + ft[i] * my_;
        float Wt = pw + ft[i] * mw;
        float c = rcp(Wt);
        ex[i] = X * c;
        ey[i] = Y * c;
        coc[i] = cc0 / Wt + cc1;
    }
    float avg = (1.f / 3.f) * (coc[0] + coc[1] + coc[2]);
    float z0 = kFltMax, z1 = -kFltMax, z2 = kFltMax, z3 = -kFltMax;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float x11 = 1.0f * (ex[i] - coc[i]) - lx[0] * ft[i] - avg * (-1.0f);
        float x21 = 1.0f * (ex[i] + coc[i]) - lx[0] * ft[i] - avg * (1.0f);
        float x12 = 1.0f * (ex[i] - coc[i]) - hx[0] * ft[i] - avg * (-1.0f);
        float x22 = 1.0f * (ex[i] + coc[i]) - hx[0] * ft[i] - avg * (1.0f);
        float y11 = 1.0f * (ey[i] - coc[i]) - ly[0] * ft[i] - avg * (-1.0f);
        float y21 = 1.0f * (ey[i] + coc[i]) - ly[0] * ft[i] - avg * (1.0f);
        float y12 = 1.0f * (ey[i] - coc[i]) - hy[0] * ft[i] - avg * (-1.0f);
        float y22 = 1.0f * (ey[i] + coc[i]) - hy[0] * ft[i] - avg * (1.0f);
        z0 = fmn(fmn(z0, x11), x21);
        z1 = fmx(fmx(z1, x12), x22);
        z2 = fmn(fmn(z2, y11), y21);
        z3 = fmx(fmx(z3, y12), y22);
    }
    b.p[0] = lx[0]; b.p[1] = avg; b.p[2] = z0;
    b.p[3] = hx[0]; b.p[4] = avg; b.p[5] = z1;
    b.p[6] = ly[0]; b.p[7] = avg; b.p[8] = z2;
    b.p[9] = hy[0]; b.p[10] = avg; b.p[11] = z3;
    return true;
}

// mergeTUXPlanes(plane1, plane2): common/TimeBounds.cpp:154-208
TLR_HD void merge_upper(float a0, float a1, float a2, float b0, float b1, float b2, float* o) {
    float x10 = -a1 + a2, x11 = a0 - a1 + a2, x12 = a0 + a1 + a2, x13 = a1 + a2;
    float x20 = -b1 + b2, x21 = b0 - b1 + b2, x22 = b0 + b1 + b2, x23 = b1 + b2;
    float m0 = fmx(x10, x20), m1 = fmx(x11, x21), m2 = fmx(x12, x22), m3 = fmx(x13, x23);
    float c0 = m1 + m2;
    float c1 = m2 + m3 - m0 - m1;
    float c2 = m0 + m1 + m2 + m3;
    float A = c0 - 0.5f * c2;
    float B = 0.25f * c1;
    float d = m0 - (-B);
    d = fmx(d, m1 - (A - B));
    d = fmx(d, m2 - (A + B));
    d = fmx(d, m3 - B);
    o[0] = A; o[1] = B; o[2] = d;
}

// TimeLensBounds(tb0, tb1) with mergeTUXPlanes(planes...):
// Reconstruction.hpp:153-160, common/TimeBounds.cpp:212-224
TLR_HD void bounds_merge(const Bounds& a, const Bounds& b, Bounds& o) {
    const bool va = bounds_valid(a), vb = bounds_valid(b);
    if (va && vb) {
        float r[3];
        merge_upper(-a.p[0], -a.p[1], -a.p[2], -b.p[0], -b.p[1], -b.p[2], r);
        o.p[0] = -r[0]; o.p[1] = -r[1]; o.p[2] = -r[2];
        merge_upper(a.p[3], a.p[4], a.p[5], b.p[3], b.p[4], b.p[5], r);
        o.p[3] = r[0]; o.p[4] = r[1]; o.p[5] = r[2];
        merge_upper(-a.p[6], -a.p[7], -a.p[8], -b.p[6], -b.p[7], -b.p[8], r);
        o.p[6] = -r[0]; o.p[7] = -r[1]; o.p[8] = -r[2];
        merge_upper(a.p[9], a.p[10], a.p[11], b.p[9], b.p[10], b.p[11], r);
        o.p[9] = r[0]; o.p[10] = r[1]; o.p[11] = r[2];
        o.ts0 = fmn(a.ts0, b.ts0);
        o.ts1 = fmx(a.ts1, b.ts1);
    } else if (va) {
        o = a;
    } else if (vb) {
        o = b;
    } else {
        bounds_zero(o);
    }
}

// ---------------------------------------------------------------------------
// sameSurface heuristic: Reconstruction.hpp:428-525
// ---------------------------------------------------------------------------
TLR_HD int tern_cmp(float a, float b, float eps) {
    if (fabsf(a - b) < eps) return 0;
    return (a < b) ? -1 : 1;
}
TLR_HD bool tern_eq2(int a, int b) { return a == 0 || b == 0 || a == b; }
TLR_HD bool tern_eq4(int a, int b, int c, int d) {
    int lo = a < b ? a : b;
    lo = lo < c ? lo : c;
    lo = lo < d ? lo : d;
    int hi = a > b ? a : b;
    hi = hi > c ? hi : c;
    hi = hi > d ? hi : d;
    return hi - lo < 2;
}

// n0, n1: 12 plane coefficients.  umin/umax: (u, v, t) box around the output.
TLR_HD bool same_surface(const float* n0, const float* n1, float u0, float v0, float t0, float u1, float v1,
                         float t1) {
    const float T = 1 * 0.1f;
    const float tm = 0.5f * (t0 + t1);
    bool coc = tern_eq2(tern_cmp(plane_eval(n0 + 0, tm, u0), plane_eval(n1 + 0, tm, u0), T),
                        tern_cmp(plane_eval(n0 + 0, tm, u1), plane_eval(n1 + 0, tm, u1), T)) &&
               tern_eq2(tern_cmp(plane_eval(n0 + 6, tm, v0), plane_eval(n1 + 6, tm, v0), T),
                        tern_cmp(plane_eval(n0 + 6, tm, v1), plane_eval(n1 + 6, tm, v1), T)) &&
               tern_eq2(tern_cmp(plane_eval(n0 + 3, tm, u0), plane_eval(n1 + 3, tm, u0), T),
                        tern_cmp(plane_eval(n0 + 3, tm, u1), plane_eval(n1 + 3, tm, u1), T)) &&
               tern_eq2(tern_cmp(plane_eval(n0 + 9, tm, v0), plane_eval(n1 + 9, tm, v0), T),
                        tern_cmp(plane_eval(n0 + 9, tm, v1), plane_eval(n1 + 9, tm, v1), T));
    if (!coc) return false;
    return tern_eq2(tern_cmp(plane_eval(n0 + 0, t0, 0.0f), plane_eval(n1 + 0, t0, 0.0f), T),
                    tern_cmp(plane_eval(n0 + 0, t1, 0.0f), plane_eval(n1 + 0, t1, 0.0f), T)) &&
           tern_eq2(tern_cmp(plane_eval(n0 + 3, t0, 0.0f), plane_eval(n1 + 3, t0, 0.0f), T),
                    tern_cmp(plane_eval(n0 + 3, t1, 0.0f), plane_eval(n1 + 3, t1, 0.0f), T)) &&
           tern_eq2(tern_cmp(plane_eval(n0 + 6, t0, 0.0f), plane_eval(n1 + 6, t0, 0.0f), T),
                    tern_cmp(plane_eval(n0 + 6, t1, 0.0f), plane_eval(n1 + 6, t1, 0.0f), T)) &&
           tern_eq2(tern_cmp(plane_eval(n0 + 9, t0, 0.0f), plane_eval(n1 + 9, t0, 0.0f), T),
                    tern_cmp(plane_eval(n0 + 9, t1, 0.0f), plane_eval(n1 + 9, t1, 0.0f), T));
}

// sameSurface split into a per-leaf part and a pair test.  surf_sig
// evaluates the 16 plane values same_surface compares (each plane at the two
// u (or v) ends at the mid time, then at u = 0 and the two ends of the time
// range); same_surface_sig(a, b) == same_surface(na, nb) for the signatures
// of the same planes and box.
struct SurfSig {
    float v[16];
};
TLR_HD void surf_sig(const float* n, float u0, float v0, float t0, float u1, float v1, float t1, SurfSig& s) {
    const float tm = 0.5f * (t0 + t1);
    s.v[0] = plane_eval(n + 0, tm, u0);
    s.v[1] = plane_eval(n + 0, tm, u1);
    s.v[2] = plane_eval(n + 6, tm, v0);
    s.v[3] = plane_eval(n + 6, tm, v1);
    s.v[4] = plane_eval(n + 3, tm, u0);
    s.v[5] = plane_eval(n + 3, tm, u1);
    s.v[6] = plane_eval(n + 9, tm, v0);
    s.v[7] = plane_eval(n + 9, tm, v1);
    s.v[8] = plane_eval(n + 0, t0, 0.0f);
    s.v[9] = plane_eval(n + 0, t1, 0.0f);
    s.v[10] = plane_eval(n + 3, t0, 0.0f);
    s.v[11] = plane_eval(n + 3, t1, 0.0f);
    s.v[12] = plane_eval(n + 6, t0, 0.0f);
    s.v[13] = plane_eval(n + 6, t1, 0.0f);
    s.v[14] = plane_eval(n + 9, t0, 0.0f);
    s.v[15] = plane_eval(n + 9, t1, 0.0f);
}
TLR_HD bool same_surface_sig(const SurfSig& a, const SurfSig& b) {
    const float T = 1 * 0.1f;
    bool r = true;
#pragma unroll
    for (int k = 0; k < 16; k += 2) r = r && tern_eq2(tern_cmp(a.v[k], b.v[k], T), tern_cmp(a.v[k + 1], b.v[k + 1], T));
    return r;
}

// same_surface evaluated without branches over both leaves' signatures (the
// same 8 pair tests, all required): the 24 coefficients load together, where
// testing plane by plane lets the compiler load each plane only after the one
// before it passed.  Faster where registers are to spare (path 2's 2R stage);
// the general kernel keeps same_surface, whose registers are not.
TLR_HD bool same_surface_flat(const float* n0, const float* n1, float u0, float v0, float t0, float u1, float v1,
                              float t1) {
    float a[12], b[12];
#pragma unroll
    for (int q = 0; q < 12; q++) {
        a[q] = n0[q];
        b[q] = n1[q];
    }
    SurfSig sa, sb;
    surf_sig(a, u0, v0, t0, u1, v1, t1, sa);
    surf_sig(b, u0, v0, t0, u1, v1, t1, sb);
    const float T = 1 * 0.1f;
    bool r = true;
#pragma unroll
    for (int k = 0; k < 16; k += 2) r = r & tern_eq2(tern_cmp(sa.v[k], sb.v[k], T), tern_cmp(sa.v[k + 1], sb.v[k + 1], T));
    return r;
}

// ---------------------------------------------------------------------------
// Point-in-triangle with FP64 edge functions: common/EdgeFunction.cpp:39-45,
// EdgeFunction.hpp:45-47; test at ReconstructionOur.cpp:249-258.
// ---------------------------------------------------------------------------
TLR_HD bool edge_side(double x0, double y0, double x1, double y1, double px, double py) {
    double A = y0 - y1;
    double B = x1 - x0;
    double C = -x0 * A - y0 * B;
    bool acc = (A == 0.) ? (B > 0.) : (A > 0.);
    float d = (float)(px * A + py * B + C);
    return d > 0.f ? true : (d == 0.f && acc);
}

TLR_HD bool tri_covers(float r0x, float r0y, float r1x, float r1y, float r2x, float r2y, float ox, float oy) {
    bool p0 = edge_side(r1x, r1y, r2x, r2y, ox, oy);
    bool p1 = edge_side(r2x, r2y, r0x, r0y, ox, oy);
    bool p2 = edge_side(r0x, r0y, r1x, r1y, ox, oy);
    return (p0 && p1 && p2) || (!p0 && !p1 && !p2);
}

// Dispersion (filter radius) by power-of-two spp: ReconstructionOur.cpp:63-76
TLR_HD float dispersion_for_spp(int spp) {
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

// hashBits(x, y): src/framework/base/Hash.hpp:169-183
TLR_HD uint32_t hash_bits(uint32_t a, uint32_t b) {
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

}  // namespace tlr
