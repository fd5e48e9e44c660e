/*
 * Synthetic sample-buffer generator (see include/tlr_scene.h).
 *
 * Scene: a static background quad behind everything plus `num_quads` large
 * and `num_occluders` small screen-aligned quads.  Quad q is the set of
 * centre-lens screen points (X,Y) in [x0,x1)x[y0,y1) at T=0 and depth w0; all
 * its points share the homogeneous motion V, so at time t a point sits at
 *   P(t) = (X*w0 + t*Vx, Y*w0 + t*Vy, w0 + t*Vw)
 * and is seen through lens position (u,v) at
 *   xy = P(t).xy / P(t).w + (C0/P(t).w + C1) * (u,v).
 * A sample (x,y,u,v,t) hits q if inverting that map lands inside the rect;
 * the hit with the smallest w(t) wins.  Albedo is a two-colour checker in
 * the quad's (X,Y).
 */
#include "../../../include/tlr_scene.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(mix64(seed)) {}
    uint64_t next() { s = mix64(s); return s; }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * u01(); }
};

struct Quad {
    double x0, y0, x1, y1;   // T=0 centre-lens screen rect
    double w0;               // depth at T=0
    double vx, vy, vw;       // homogeneous motion
    float c0[3], c1[3];      // checker colours
};

struct Scene {
    tlr_scene_desc d;
    std::vector<Quad> quads;
    int tw, th, tile;
    std::vector<std::vector<int>> bins;
};

double draw_depth(Rng& r, const tlr_scene_desc& d) {
    if (d.w_log_uniform) return std::exp(r.uni(std::log((double)d.w_min), std::log((double)d.w_max)));
    return r.uni(d.w_min, d.w_max);
}

void make_quad(Rng& r, const tlr_scene_desc& d, double side_lo, double side_hi, Quad& q) {
    const double W = d.width, H = d.height;
    double sx = r.uni(side_lo, side_hi), sy = r.uni(side_lo, side_hi);
    double cx = r.uni(-0.1 * W, 1.1 * W), cy = r.uni(-0.1 * H, 1.1 * H);
    q.x0 = cx - 0.5 * sx; q.x1 = cx + 0.5 * sx;
    q.y0 = cy - 0.5 * sy; q.y1 = cy + 0.5 * sy;
    q.w0 = draw_depth(r, d);
    bool stat = r.u01() < d.static_frac;
    if (stat || (d.mv_xy_px == 0.0f && d.mv_w_frac == 0.0f)) {
        q.vx = q.vy = q.vw = 0.0;
    } else {
        q.vx = r.uni(-d.mv_xy_px, d.mv_xy_px) * q.w0;
        q.vy = r.uni(-d.mv_xy_px, d.mv_xy_px) * q.w0;
        q.vw = r.uni(-d.mv_w_frac, d.mv_w_frac) * q.w0;
    }
    for (int k = 0; k < 3; k++) {
        q.c0[k] = (float)r.u01();
        q.c1[k] = (float)r.u01();
    }
}

// Conservative screen rect of everything a quad can cover for t in [0,1] and
// any lens position: the centre rect at both shutter ends, dilated by the
// largest |CoC| over the depths it passes through.
void quad_extent(const Quad& q, const tlr_scene_desc& d, double out[4]) {
    double lo_x = 1e300, hi_x = -1e300, lo_y = 1e300, hi_y = -1e300, cmax = 0;
    for (int e = 0; e <= 8; e++) {
        double t = e / 8.0;
        double w = q.w0 + t * q.vw;
        double ax = (q.x0 * q.w0 + t * q.vx) / w, bx = (q.x1 * q.w0 + t * q.vx) / w;
        double ay = (q.y0 * q.w0 + t * q.vy) / w, by = (q.y1 * q.w0 + t * q.vy) / w;
        lo_x = std::min(lo_x, std::min(ax, bx)); hi_x = std::max(hi_x, std::max(ax, bx));
        lo_y = std::min(lo_y, std::min(ay, by)); hi_y = std::max(hi_y, std::max(ay, by));
        cmax = std::max(cmax, std::fabs(d.coc_c0 / w + d.coc_c1));
    }
    double m = cmax * 1.5 + 2.0;  // lens radius is 1; margin for the disk edge
    out[0] = lo_x - m; out[1] = lo_y - m; out[2] = hi_x + m; out[3] = hi_y + m;
}

void build_scene(const tlr_scene_desc& d, Scene& s) {
    s.d = d;
    Rng r(d.seed * 0x2545F4914F6CDD1Dull + 17);
    const double W = d.width, H = d.height;
    const double mside = std::min(W, H);
    // Background: static, behind every other surface, far larger than the
    // screen so that every sample hits something.
    Quad bg;
    bg.w0 = d.w_max * 1.25;
    double margin = 4.0 * (std::fabs(d.coc_c0 / bg.w0 + d.coc_c1) + d.mv_xy_px + 16.0);
    bg.x0 = -margin; bg.x1 = W + margin; bg.y0 = -margin; bg.y1 = H + margin;
    bg.vx = bg.vy = bg.vw = 0.0;
    for (int k = 0; k < 3; k++) { bg.c0[k] = 0.15f + 0.1f * k; bg.c1[k] = 0.85f - 0.2f * k; }
    s.quads.push_back(bg);
    for (int i = 0; i < d.num_quads; i++) {
        Quad q;
        make_quad(r, d, d.quad_min_frac * mside, d.quad_max_frac * mside, q);
        s.quads.push_back(q);
    }
    for (int i = 0; i < d.num_occluders; i++) {
        Quad q;
        make_quad(r, d, d.occluder_min_px, d.occluder_max_px, q);
        s.quads.push_back(q);
    }
    s.tile = 32;
    s.tw = (d.width + s.tile - 1) / s.tile;
    s.th = (d.height + s.tile - 1) / s.tile;
    s.bins.assign((size_t)s.tw * s.th, std::vector<int>());
    for (int qi = 0; qi < (int)s.quads.size(); qi++) {
        double e[4];
        quad_extent(s.quads[qi], d, e);
        int tx0 = std::max(0, (int)std::floor(e[0] / s.tile));
        int ty0 = std::max(0, (int)std::floor(e[1] / s.tile));
        int tx1 = std::min(s.tw - 1, (int)std::floor(e[2] / s.tile));
        int ty1 = std::min(s.th - 1, (int)std::floor(e[3] / s.tile));
        for (int ty = ty0; ty <= ty1; ty++)
            for (int tx = tx0; tx <= tx1; tx++) s.bins[(size_t)ty * s.tw + tx].push_back(qi);
    }
}

inline bool hit(const Quad& q, const tlr_scene_desc& d, double x, double y, double u, double v, double t,
                double& wt, double& X, double& Y) {
    wt = q.w0 + t * q.vw;
    if (wt <= 0) return false;
    double coc = d.coc_c0 / wt + d.coc_c1;
    double xc = x - coc * u, yc = y - coc * v;
    X = (xc * wt - t * q.vx) / q.w0;
    Y = (yc * wt - t * q.vy) / q.w0;
    return X >= q.x0 && X < q.x1 && Y >= q.y0 && Y < q.y1;
}

void render_pixel(const Scene& s, int px, int py, float* out) {
    const tlr_scene_desc& d = s.d;
    const int spp = d.spp;
    Rng r(d.seed ^ mix64(((uint64_t)py << 32) | (uint64_t)px));
    // Stratification: t in spp strata; (u,v) in a g x g grid (g = ceil(sqrt
    // spp)) with strata shuffled against t; (x,y) jittered uniformly.
    int g = (int)std::ceil(std::sqrt((double)spp));
    std::vector<int> perm(g * g);
    for (int k = 0; k < g * g; k++) perm[k] = k;
    for (int k = g * g - 1; k > 0; k--) {
        int j = (int)(r.next() % (uint64_t)(k + 1));
        std::swap(perm[k], perm[j]);
    }
    const std::vector<int>& bin = s.bins[(size_t)(py / s.tile) * s.tw + (px / s.tile)];
    for (int i = 0; i < spp; i++) {
        double x = px + r.u01();
        double y = py + r.u01();
        int cell = perm[i];
        double u = -1.0 + 2.0 * ((cell % g) + r.u01()) / g;
        double v = -1.0 + 2.0 * ((cell / g) + r.u01()) / g;
        double t = (i + r.u01()) / spp;
        if (t >= 1.0) t = std::nextafter(1.0, 0.0);
        int best = -1;
        double bw = 1e300, bX = 0, bY = 0;
        for (int qi : bin) {
            double wt, X, Y;
            if (hit(s.quads[qi], d, x, y, u, v, t, wt, X, Y) && wt < bw) {
                bw = wt; best = qi; bX = X; bY = Y;
            }
        }
        float* rec = out + 16 * (size_t)i;
        float col[3] = {0.5f, 0.5f, 0.5f};
        float mv[3] = {0, 0, 0};
        if (best >= 0) {
            const Quad& q = s.quads[best];
            double per = d.checker_period_px > 0 ? d.checker_period_px : 8.0;
            long cx = (long)std::floor(bX / (0.5 * per)), cy = (long)std::floor(bY / (0.5 * per));
            const float* c = ((cx + cy) & 1) ? q.c1 : q.c0;
            col[0] = c[0]; col[1] = c[1]; col[2] = c[2];
            mv[0] = (float)q.vx; mv[1] = (float)q.vy; mv[2] = (float)q.vw;
        } else {
            bw = d.w_max * 1.25;
        }
        rec[0] = (float)x; rec[1] = (float)y; rec[2] = 0.0f; rec[3] = (float)bw;
        rec[4] = (float)u; rec[5] = (float)v; rec[6] = (float)t;
        rec[7] = col[0]; rec[8] = col[1]; rec[9] = col[2]; rec[10] = 1.0f;
        rec[11] = mv[0]; rec[12] = mv[1]; rec[13] = mv[2];
        rec[14] = 0.0f; rec[15] = 0.0f;
    }
}

}  // namespace

extern "C" {

int tlr_scene_preset(int preset, int width, int height, int spp, uint64_t seed, tlr_scene_desc* o) {
    std::memset(o, 0, sizeof(*o));
    o->width = width; o->height = height; o->spp = spp;
    o->seed = seed;
    o->checker_period_px = 8.0f;
    o->quad_min_frac = 0.05f;
    o->quad_max_frac = 0.35f;
    o->occluder_min_px = 2.0f;
    o->occluder_max_px = 8.0f;
    switch (preset) {
        case 1:  // defocus only, three quads
            o->coc_c0 = -15.24f; o->coc_c1 = 19.05f;
            o->num_quads = 3; o->quad_min_frac = 0.25f; o->quad_max_frac = 0.6f;
            o->w_min = 2.0f; o->w_max = 12.0f; o->w_log_uniform = 0;
            break;
        case 2: case 4: case 5:  // defocus + perspective motion
            o->coc_c0 = -15.24f; o->coc_c1 = 19.05f;
            o->num_quads = 50;
            o->w_min = 1.0f; o->w_max = 20.0f; o->w_log_uniform = 1;
            o->mv_xy_px = 30.0f; o->mv_w_frac = 0.2f;
            break;
        case 3:  // strong defocus, occluders, mixed motion
            o->coc_c0 = -40.0f; o->coc_c1 = 50.0f;
            o->num_quads = 400; o->num_occluders = 5000;
            o->w_min = 0.5f; o->w_max = 30.0f; o->w_log_uniform = 1;
            o->mv_xy_px = 60.0f; o->mv_w_frac = 0.0f; o->static_frac = 0.3f;
            break;
        default:
            return -1;
    }
    return 0;
}

int tlr_scene_generate(const tlr_scene_desc* d, float* records) {
    if (!d || !records || d->width <= 0 || d->height <= 0 || d->spp <= 0) return -1;
    Scene s;
    build_scene(*d, s);
    int threads = d->threads > 0 ? d->threads : (int)std::max(1u, std::thread::hardware_concurrency());
    const int H = d->height, W = d->width;
    std::atomic<int> next(0);
    std::vector<std::thread> ts;
    for (int k = 0; k < threads; k++)
        ts.emplace_back([&]() {
            for (;;) {
                int y = next.fetch_add(1);
                if (y >= H) break;
                for (int x = 0; x < W; x++)
                    render_pixel(s, x, y, records + 16 * ((size_t)y * W + x) * (size_t)d->spp);
            }
        });
    for (auto& t : ts) t.join();
    return 0;
}

}  // extern "C"
