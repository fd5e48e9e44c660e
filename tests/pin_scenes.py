"""Hand-built sample buffers whose reconstruction the reference's rules fix in
closed form (ReconstructionOur.cpp:50-359): static samples on a regular 4x4
grid per pixel (16 spp, so the filter radius R = dispersion = 0.5), lens
position (0,0), time 0.5, with a tiny deterministic jitter so that no two
samples lie at exactly the same distance from a pixel centre.  Queries are
pinhole output samples (override (u,v,t) = (0,0,0.5): one output sample at
each pixel centre).

Depths: FRONT w = 1 (CoC 3.81 px with C0=-15.24, C1=19.05) and BACK w = 20
(CoC 18.29 px): their bounds cross inside every query's (u,v) box, so leaves
of the two are different surfaces (Reconstruction.hpp:457-525).
"""
import numpy as np

W, H, SPP = 32, 24, 16
C0, C1 = -15.24, 19.05
R = 0.5
FRONT, BACK = 1.0, 20.0
RED = (1.0, 0.0, 0.0)
GREEN = (0.0, 1.0, 0.0)
BLUE = (0.0, 0.0, 1.0)
PINHOLE = (0.0, 0.0, 0.5)


def positions():
    """(H, W, 16, 2) sample positions: 4x4 grid per pixel plus a jitter of at
    most 0.002 px from a fixed integer hash."""
    ys, xs, k = np.meshgrid(np.arange(H), np.arange(W), np.arange(SPP), indexing="ij")
    a, b = k % 4, k // 4
    h = (xs * 73856093) ^ (ys * 19349663) ^ (k * 83492791)
    jx = ((h % 1000) / 1000.0 - 0.5) * 0.004
    jy = (((h // 1000) % 1000) / 1000.0 - 0.5) * 0.004
    x = xs + (a + 0.5) / 4 + jx
    y = ys + (b + 0.5) / 4 + jy
    return np.stack([x, y], -1).astype(np.float32)


def records(fn):
    """Records from fn(x, y, pixel_x, pixel_y, k) -> (w, (r, g, b)) or None
    (discarded: behind the camera, no motion)."""
    pos = positions()
    rec = np.zeros((H, W, SPP, 16), np.float32)
    for py in range(H):
        for px in range(W):
            for k in range(SPP):
                x, y = float(pos[py, px, k, 0]), float(pos[py, px, k, 1])
                r = fn(x, y, px, py, k)
                out = rec[py, px, k]
                out[0], out[1] = x, y
                out[6] = 0.5
                out[10] = 1.0
                if r is None:
                    out[3] = -1.0
                else:
                    out[3] = r[0]
                    out[7:10] = r[1]
    return rec


def edge(x, y, px, py, k):
    """Front (red) left of x = 8.3, back (blue) right of it."""
    return (FRONT, RED) if x < 8.3 else (BACK, BLUE)


def corner(x, y, px, py, k):
    """Front (red) everywhere except the quadrant x > 8.6, y > 10.6 (back,
    blue): the pixel (8,10) centre sits just inside the front's corner."""
    return (BACK, BLUE) if (x > 8.6 and y > 10.6) else (FRONT, RED)


def mesh(x, y, px, py, k):
    """Front (red) with every other sample column behind it (back, blue)
    around pixel (8,10): both surfaces within R of the centre, the front in
    all four quadrants."""
    if 6 <= px < 11 and 8 <= py < 13 and (k % 4) in (1, 3):
        return (BACK, BLUE)
    return (FRONT, RED)


HOLE_CENTRE = (20.5, 10.5)
HOLE_RADIUS = 1.6


def hole_colour(x, y):
    return (x / W, y / H, 0.25)


def hole(x, y, px, py, k):
    """One plane (w = 5) with a colour per position and no sample within 1.6
    px of the pixel (20,10) centre: nothing within R or 2R there."""
    if np.hypot(x - HOLE_CENTRE[0], y - HOLE_CENTRE[1]) < HOLE_RADIUS:
        return None
    return (5.0, hole_colour(x, y))


TINY_PIXEL = (20, 4)
TINY_SLOTS = (4 * 1 + 0, 4 * 2 + 0, 4 * 1 + 1)   # grid (a, b) = (0,1), (0,2), (1,1)


def tiny(x, y, px, py, k):
    """Back plane (blue) with three front samples (green) just left of the
    pixel (20,4) centre: a front surface of 3 samples, all in the x < 0
    half-plane of the output sample."""
    if (px, py) == TINY_PIXEL and k in TINY_SLOTS:
        return (FRONT, GREEN)
    return (BACK, BLUE)


def tent_mean(rec, o, radius=R):
    """Tent-weighted mean colour (float64) of every kept sample within
    `radius` of o (weights 1 - d / radius)."""
    r = rec.reshape(-1, 16).astype(np.float64)
    keep = r[:, 3] > 0
    d = np.hypot(r[:, 0] - o[0], r[:, 1] - o[1])
    sel = keep & (d < radius)
    w = 1.0 - d[sel] / radius
    return (w[:, None] * r[sel, 7:10]).sum(0) / w.sum()


def nearest_colour(rec, o):
    r = rec.reshape(-1, 16).astype(np.float64)
    keep = r[:, 3] > 0
    d = np.hypot(r[:, 0] - o[0], r[:, 1] - o[1])
    d[~keep] = np.inf
    return r[np.argmin(d), 7:10], float(np.min(d))
