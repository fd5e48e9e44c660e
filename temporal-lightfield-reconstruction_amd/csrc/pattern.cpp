// Output sampling pattern (host): 64 patterns x M output samples of
// (x, y, u, v, t), as TreeGather::generateOutputSamples builds them
// (Reconstruction.cpp:452-512) with the output spp a parameter instead of the
// hard-wired 128.  Tiny (64*M entries) and computed once per reconstruction
// setting, so it stays on the host and is uploaded.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#include "pattern.hpp"

namespace tlr {
namespace {

// RANROT-A generator: src/framework/base/Random.cpp:52-92
class RanrotA {
  public:
    explicit RanrotA(uint32_t seed) {
        if (seed == 0) seed--;
        for (uint32_t& b : buf_) {
            seed ^= seed << 13;
            seed ^= seed >> 17;
            seed ^= seed << 5;
            b = seed;
        }
        p1_ = 0;
        p2_ = 7;
        for (int i = 0; i < 11; i++) next();
    }
    uint32_t next() {
        uint32_t x = buf_[p1_] + buf_[p2_];
        x = (x << 13) | (x >> 19);
        buf_[p1_] = x;
        p1_ = (p1_ == 0) ? 10 : p1_ - 1;
        p2_ = (p2_ == 0) ? 10 : p2_ - 1;
        return x;
    }
    // Random::getF32: src/framework/base/Random.hpp:57
    float unit() { return (float)next() * (1.0f / 4294967296.0f); }

  private:
    uint32_t buf_[11];
    int p1_, p2_;
};

// Direction numbers of Sobol dimension 3 (primitive polynomial x^3+x+1,
// m = 1,3,1), the table row common/Sobol5.inl:97-128 holds.
struct SobolDim3 {
    uint32_t v[32];
    SobolDim3() {
        v[0] = 1u << 31;
        v[1] = 3u << 30;
        v[2] = 1u << 29;
        for (int c = 3; c < 32; c++) v[c] = v[c - 2] ^ v[c - 3] ^ (v[c - 3] >> 3);
    }
};

float sobol_dim3(int i) {
    static const SobolDim3 tab;
    uint32_t r = 0;
    for (int c = 0; i; i >>= 1, c++)
        if (i & 1) r ^= tab.v[c];
    return (float)r * (1.f / 4294967296.0f);
}

// Direction numbers of Sobol dimension 4 (primitive polynomial x^3+x^2+1,
// m = 1,1,1), the table row common/Sobol5.inl:129-160 holds.
struct SobolDim4 {
    uint32_t v[32];
    SobolDim4() {
        v[0] = 1u << 31;
        v[1] = 1u << 30;
        v[2] = 1u << 29;
        for (int c = 3; c < 32; c++) v[c] = v[c - 1] ^ v[c - 3] ^ (v[c - 3] >> 3);
    }
};

float sobol_dim4(int i) {
    static const SobolDim4 tab;
    uint32_t r = 0;
    for (int c = 0; i; i >>= 1, c++)
        if (i & 1) r ^= tab.v[c];
    return (float)r * (1.f / 4294967296.0f);
}

// sobol2D: common/Util.cpp:266-288
void sobol_2d(int i, float& a, float& b) {
    uint32_t r1 = 0, r2 = 0;
    uint32_t v1 = 1u << 31, v2 = 3u << 30;
    for (; i; i >>= 1) {
        if (i & 1) {
            r1 ^= v1;
            r2 ^= v2 << 1;
        }
        v1 |= v1 >> 1;
        v2 ^= v2 >> 1;
    }
    a = (float)r1 * (1.f / 4294967296.0f);
    b = (float)r2 * (1.f / 4294967296.0f);
}

// larcherPillichshammer(i, 0): common/Util.cpp:290-297
float larcher_pillichshammer(int i) {
    uint32_t r = 0;
    for (uint32_t v = 1u << 31; i; i >>= 1, v |= v >> 1)
        if (i & 1) r ^= v;
    return float((double)r / 4294967296.0);
}

// halton(base, i): common/Util.cpp:235-251
float radical_inverse(int base, int i) {
    float h = 0.f;
    float half = 1.0f / float(base);
    while (i) {
        const int digit = i % base;
        h = h + digit * half;
        i = (i - digit) / base;
        half = half / base;
    }
    return h;
}

// Shirley-Chiu concentric map: common/Util.cpp:37-74
void square_to_disk(float sx, float sy, float& u, float& v) {
    const float PI = 3.1415926535897932384626433832795f;
    const float a = 2 * sx - 1;
    const float b = 2 * sy - 1;
    float r, phi;
    if (a > -b) {
        if (a > b) {
            r = a;
            phi = (PI / 4) * (b / a);
        } else {
            r = b;
            phi = (PI / 4) * (2 - (a / b));
        }
    } else if (a < b) {
        r = -a;
        phi = (PI / 4) * (4 + (b / a));
    } else {
        r = -b;
        phi = (b != 0) ? (PI / 4) * (6 - (a / b)) : 0;
    }
    u = r * cosf(phi);
    v = r * sinf(phi);
}

struct Tmpl {
    float x, y, u, v, t, key;
};

// FW::sort (src/framework/base/Sort.cpp:63-211) over [lo, hi) by key:
// median-of-3 quicksort with an explicit stack and insertion sort below 16.
// Its tie order is part of the pattern, so the steps are the same.
class KeySorter {
  public:
    explicit KeySorter(std::vector<Tmpl>& a) : a_(a) {}
    void sort(int start, int end) {
        if (end - start < 2) return;
        int low = start, high = end;
        int stack[32];
        int sp = 0;
        stack[sp++] = high;
        while (sp) {
            high = stack[--sp];
            if (high - low < 16 || sp + 2 > 32) {
                insertion(low, high - low);
                low = high + 1;
                continue;
            }
            const int i = partition(low, high);
            if (high - i > 2) stack[sp++] = high;
            if (i - low > 1)
                stack[sp++] = i;
            else
                low = i + 1;
        }
    }

  private:
    int cmp(int i, int j) const {
        return (a_[i].key < a_[j].key) ? -1 : (a_[i].key > a_[j].key) ? 1 : 0;
    }
    void swp(int i, int j) { std::swap(a_[i], a_[j]); }
    void insertion(int start, int size) {
        for (int i = 1; i < size; i++)
            for (int j = start + i - 1; j >= start && cmp(j, j + 1) > 0; j--) swp(j, j + 1);
    }
    int partition(int low, int high) {
        int l = low, c = (low + high) >> 1, h = high - 2;
        if (cmp(l, h) > 0) std::swap(l, h);
        if (cmp(l, c) > 0) c = l;
        swp((cmp(c, h) > 0) ? h : c, high - 1);
        int i = low - 1, j = high - 1;
        for (;;) {
            do i++; while (cmp(i, high - 1) < 0);
            do j--; while (cmp(j, high - 1) > 0);
            if (i >= j) break;
            swp(i, j);
        }
        swp(i, high - 1);
        return i;
    }
    std::vector<Tmpl>& a_;
};

}  // namespace

int make_output_pattern(int out_spp, bool override_uvt, const float* uvt, std::vector<float>& out5) {
    const int M = override_uvt ? 1 : out_spp;
    std::vector<Tmpl> all;
    all.reserve((size_t)64 * M);
    RanrotA rng(242);
    for (int j = 0; j < 64; j++) {
        const float xo = (j + 0.5f) / 64, yo = larcher_pillichshammer(j);
        const float uo = radical_inverse(2, j), vo = radical_inverse(3, j);
        const float tj = rng.unit();
        for (int i = 0; i < M; i++) {
            float u, v;
            sobol_2d(i + 1, u, v);
            float x = sobol_dim3(i + 1) + xo;
            float y = larcher_pillichshammer(i + 1) + yo;
            u += uo;
            v += vo;
            float t = (i + tj) / M;
            if (x >= 1.f) x -= 1.f;
            if (y >= 1.f) y -= 1.f;
            if (u >= 1.f) u -= 1.f;
            if (v >= 1.f) v -= 1.f;
            const int ub = (int)std::floor(256 * u);
            const int vb = (int)std::floor(256 * v);
            const int tb = (int)std::floor(256 * t);
            int key = 0;
            for (int k = 0; k < 8; k++)
                key |= ((ub >> k) & 1) << (3 * k + 2) | ((vb >> k) & 1) << (3 * k + 1) | ((tb >> k) & 1) << (3 * k);
            Tmpl s;
            square_to_disk(u, v, s.u, s.v);
            s.t = t;
            if (override_uvt) {
                s.u = uvt[0];
                s.v = uvt[1];
                s.t = uvt[2];
            }
            s.x = (M == 1) ? 0.5f : x;
            s.y = (M == 1) ? 0.5f : y;
            s.key = (float)key;
            all.push_back(s);
        }
        KeySorter(all).sort(j * M, (j + 1) * M);
    }
    out5.resize((size_t)5 * all.size());
    for (size_t k = 0; k < all.size(); k++) {
        out5[5 * k + 0] = all[k].x;
        out5[5 * k + 1] = all[k].y;
        out5[5 * k + 2] = all[k].u;
        out5[5 * k + 3] = all[k].v;
        out5[5 * k + 4] = all[k].t;
    }
    return M;
}

void light_sample_positions(int M, std::vector<float>& out2) {
    out2.resize(2 * (size_t)std::max(M, 1));
    for (int i = 0; i < M; i++) {
        out2[2 * i] = sobol_dim3(i + 1);
        out2[2 * i + 1] = sobol_dim4(i + 1);
    }
}

}  // namespace tlr
