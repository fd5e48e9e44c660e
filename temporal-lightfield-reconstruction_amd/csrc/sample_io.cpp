// v1.3 sample-buffer files: the reader of UVTSampleBuffer(const char*)
// (SampleBuffer.cpp:230-387) and the writer of UVTSampleBuffer::serialize
// (SampleBuffer.cpp:404-564), with their quirks: alpha is forced to 1 on
// read; the writer stores b in the alpha slot; binary payloads need the
// separate "<file>.header".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tlr.h"

namespace {

bool take_literal(const char*& p, const char* lit) {
    const char* t = p;
    while (*lit && *t == *lit) {
        t++;
        lit++;
    }
    if (*lit) return false;
    p = t;
    return true;
}

float float_of_bits(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}

// FW::parseFloat (src/framework/base/String.cpp:428-487) step for step, in
// float arithmetic with every operation rounded (no contraction): an optional
// '+' or '-', integer digits accumulated as v = v * 10 + d, fraction digits
// added as v += scale * d with scale *= 0.1f before each, the literals #INF,
// #SNAN, #QNAN, #IND after the digits, then an optional exponent (e or E and
// an integer) applied as v *= powf(10, e).  No whitespace is skipped.  The
// result is not the correctly rounded value of the decimal string in general
// (strtof's).  Without a digit the reference leaves the value unset and the
// position unchanged; here the value is 0.
const char* parse_float(const char* p, float& out) {
    const char* t = p;
    bool neg = false;
    if (*t == '+') {
        t++;
    } else if (*t == '-') {
        neg = true;
        t++;
    }
    float v = 0.0f;
    int digits = 0;
    while (*t >= '0' && *t <= '9') {
        v = v * 10.0f + (float)(*t++ - '0');
        digits++;
    }
    if (*t == '.') {
        t++;
        float scale = 1.0f;
        while (*t >= '0' && *t <= '9') {
            scale *= 0.1f;
            v += scale * (float)(*t++ - '0');
            digits++;
        }
    }
    if (!digits) {
        out = 0.0f;
        return p;
    }
    p = t;
    if (*p == '#') {
        if (take_literal(p, "#INF")) {
            out = float_of_bits(neg ? 0xFF800000u : 0x7F800000u);
            return p;
        }
        if (take_literal(p, "#SNAN")) {
            out = float_of_bits(neg ? 0xFF800001u : 0x7F800001u);
            return p;
        }
        if (take_literal(p, "#QNAN")) {
            out = float_of_bits(neg ? 0xFFC00001u : 0x7FC00001u);
            return p;
        }
        if (take_literal(p, "#IND")) {
            out = float_of_bits(neg ? 0xFFC00000u : 0x7FC00000u);
            return p;
        }
    }
    if (*t == 'e' || *t == 'E') {
        // FW::parseInt (String.cpp:372-385): optional sign, at least one digit
        const char* q = t + 1;
        bool eneg = false;
        if (*q == '+') {
            q++;
        } else if (*q == '-') {
            eneg = true;
            q++;
        }
        if (*q >= '0' && *q <= '9') {
            int32_t e = 0;
            while (*q >= '0' && *q <= '9') e = (int32_t)((uint32_t)e * 10u + (uint32_t)(*q++ - '0'));
            if (eneg) e = -e;
            p = q;
            if (e) v *= powf(10.0f, (float)e);
        }
    }
    out = neg ? -v : v;
    return p;
}

}  // namespace

extern "C" {

tlr_status tlr_load_sample_file(const char* path, tlr_header* h, float** records, int64_t* n) {
    if (!path || !h || !records || !n) return TLR_E_INVALID;
    *records = nullptr;
    *n = 0;
    FILE* fp = fopen(path, "rb");
    if (!fp) return TLR_E_IO;
    FILE* fph = fopen((std::string(path) + ".header").c_str(), "rt");
    const bool separate = fph != nullptr;
    if (!separate) fph = fp;
    float version = 0;
    int W = 0, H = 0, spp = 0;
    if (fscanf(fph, "Version %f\n", &version) != 1 || fscanf(fph, "Width %d\n", &W) != 1 ||
        fscanf(fph, "Height %d\n", &H) != 1 || fscanf(fph, "Samples per pixel %d\n", &spp) != 1 ||
        version != 1.3f || W <= 0 || H <= 0 || spp <= 0) {
        fclose(fp);
        if (separate) fclose(fph);
        return TLR_E_IO;
    }
    char motion[1024] = {0};
    if (fscanf(fph, "Motion model: %1023s\n", motion) != 1) { /* optional, ignored */ }
    float c0 = 0, c1 = 0;
    if (fscanf(fph, "CoC coefficients (coc radius = C0/w+C1): %f,%f\n", &c0, &c1) != 2) {
        fclose(fp);
        if (separate) fclose(fph);
        return TLR_E_IO;
    }
    bool binary = false;
    char enc[1024] = {0};
    if (fscanf(fph, "Encoding = %1023s\n", enc) == 1) binary = std::string(enc) == "binary";
    if (fscanf(fph, "\n") < 0) { /* end of header */ }
    char desc[1024] = {0};
    if (fscanf(fph, "%1023s\n", desc) != 1) { /* descriptor line is informative */ }

    const int64_t N = (int64_t)W * H * spp;
    float* rec = (float*)malloc((size_t)N * 16 * sizeof(float));
    if (!rec) {
        fclose(fp);
        if (separate) fclose(fph);
        return TLR_E_IO;
    }
    tlr_status st = TLR_OK;
    if (!binary) {
        std::vector<char> line(4096);
        for (int64_t k = 0; k < N; k++) {
            if (!fgets(line.data(), (int)line.size(), fp)) {
                st = TLR_E_IO;
                break;
            }
            const char* p = line.data();
            float* r = rec + 16 * k;
            for (int v = 0; v < 16; v++) {
                if (v > 0) {
                    while (*p != ',' && *p != '\n' && *p != 0) p++;
                    if (*p) p++;
                }
                p = parse_float(p, r[v]);
            }
            r[10] = 1.0f;  // alpha forced to 1 (SampleBuffer.cpp:343)
        }
    } else {
        if (fread(rec, sizeof(float) * 16, (size_t)N, fp) != (size_t)N) st = TLR_E_IO;
        for (int64_t k = 0; k < N && st == TLR_OK; k++) rec[16 * k + 10] = 1.0f;  // :372
    }
    fclose(fp);
    if (separate) fclose(fph);
    if (st != TLR_OK) {
        free(rec);
        return st;
    }
    h->width = W;
    h->height = H;
    h->spp = spp;
    h->coc_c0 = c0;
    h->coc_c1 = c1;
    *records = rec;
    *n = N;
    return TLR_OK;
}

tlr_status tlr_save_sample_file(const char* path, const tlr_header* h, const float* rec, int binary,
                                int separate_header) {
    if (!path || !h || !rec) return TLR_E_INVALID;
    if (binary && !separate_header) return TLR_E_INVALID;  // SampleBuffer.cpp:406-407
    FILE* fp = fopen(path, binary ? "wb" : "wt");
    if (!fp) return TLR_E_IO;
    FILE* fph = fp;
    if (separate_header) {
        fph = fopen((std::string(path) + ".header").c_str(), "wt");
        if (!fph) {
            fclose(fp);
            return TLR_E_IO;
        }
    }
    fprintf(fph, "Version 1.3\n");
    fprintf(fph, "Width %d\n", h->width);
    fprintf(fph, "Height %d\n", h->height);
    fprintf(fph, "Samples per pixel %d\n", h->spp);
    fprintf(fph, "CoC coefficients (coc radius = C0/w+C1): %f,%f\n", h->coc_c0, h->coc_c1);
    fprintf(fph, "Encoding = %s\n", binary ? "binary" : "text");
    fprintf(fph, "x,y,z/w,w,u,v,t,r,g,b,a,mv_x,mv_y,mv_w,dwdx,dwdy\n");
    const int64_t N = (int64_t)h->width * h->height * h->spp;
    for (int64_t k = 0; k < N; k++) {
        float r[16];
        memcpy(r, rec + 16 * k, sizeof(r));
        r[10] = r[9];  // the reference writes b into the alpha slot (SampleBuffer.cpp:498,542)
        if (binary) {
            fwrite(r, sizeof(r), 1, fp);
        } else {
            for (int v = 0; v < 16; v++) fprintf(fp, "%f,", r[v]);
            fprintf(fp, "\n");
        }
    }
    if (separate_header) fclose(fph);
    fclose(fp);
    return TLR_OK;
}

void tlr_free(void* p) { free(p); }

}  // extern "C"
